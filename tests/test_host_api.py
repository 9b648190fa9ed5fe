"""The C++ host layer (Tortoise:: types over the C ABI) builds, and its CPU
side (mesh numbering, DG <-> nodal, integrals) holds against closed forms."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "tortoise_amd", "host")


def test_host_selftest_builds_and_passes(tmp_path):
    exe = str(tmp_path / "host_selftest")
    subprocess.run(["g++", "-O1", "-std=c++20", "-Wall", "-Werror", "-I", HOST, "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "host_selftest.cpp"),
                    "-L", os.path.join(ROOT, "tortoise_amd"), "-ltortoise_gpu",
                    "-Wl,-rpath," + os.path.join(ROOT, "tortoise_amd")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout


def test_example_builds():
    subprocess.run(["make", "-C", HOST, "-B"], check=True, capture_output=True)
    assert os.path.exists(os.path.join(HOST, "examples", "thermalisation_square"))
