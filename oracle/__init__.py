"""ctypes front-end of the CPU FP64 oracle (oracle/tortoise_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker the HIP path is compared
against.  The product package never imports it.

Reference files restated (paths relative to the TORTOISE tree):
  mesh numbering       src/Geometry/Structured/DirectSpace/Mesh.cpp:55-87,222-267,384-433
  error norm           src/Physics/Material/Band.cpp:87-101, MaterialStatus.cpp:189-196
  DP5(4) driver        include/Generics/Algorithms/DormandPrince54.hpp:479-665
  amplitude convention src/Physics/Scattering/ScatteringChannel.cpp:399-415
  tuple operator       include/Physics/Scattering/IntegratorCore.hpp:260-357,
                       src/Physics/Scattering/ScatteringChannel.cpp:285-538,618-688
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")


class EGrid(C.Structure):
    _fields_ = [("e0", C.c_double), ("de", C.c_double), ("nbins", C.c_int32)]


class DP5Stats(C.Structure):
    _fields_ = [("attempted", C.c_int64), ("accepted", C.c_int64), ("rejected", C.c_int64),
                ("evals", C.c_int64), ("failed", C.c_int32), ("n_dense", C.c_int32)]


RHS_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double))
NORM_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))


def build(force: bool = False) -> str:
    """Compiles liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "all"], cwd=_HERE, check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_num_nodes.restype = C.c_int64
        L.or_num_nodes.argtypes = [C.c_int, _i32p]
        L.or_num_elements.restype = C.c_int64
        L.or_num_elements.argtypes = [C.c_int, _i32p]
        L.or_element_volume.restype = C.c_double
        L.or_element_volume.argtypes = [C.c_int, _i32p, _f64p]
        L.or_mesh_nodes.argtypes = [C.c_int, _i32p, _f64p, _f64p, _f64p]
        L.or_mesh_elements.argtypes = [C.c_int, _i32p, _i32p]
        L.or_node_volumes.argtypes = [C.c_int, _i32p, _f64p, _f64p]
        L.or_egrid_make.argtypes = [_f64p, C.c_int64, C.c_double, C.POINTER(EGrid)]
        L.or_projection.restype = C.c_int32
        L.or_projection.argtypes = [C.c_int, _i32p, _f64p, _f64p, C.POINTER(EGrid), _i32p, _i32p,
                                    C.c_void_p, C.c_int32]
        L.or_simplex_density.restype = C.c_double
        L.or_simplex_density.argtypes = [C.c_int, _f64p, C.c_int, C.c_double, C.c_double]
        L.or_gauss_table.restype = C.c_int32
        L.or_gauss_table.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.POINTER(C.c_int32), C.c_void_p, C.c_int32]
        L.or_kernel_full.argtypes = [C.c_int64, _i32p, _i32p, _f64p, C.c_int32, _f64p, C.c_int32,
                                     C.c_int32, C.c_double, _f64p]
        L.or_kernel_row.argtypes = [C.c_int64, C.c_int64, _i32p, _i32p, _f64p, C.c_int32, _f64p,
                                    C.c_int32, C.c_int32, C.c_double, _f64p]
        L.or_collision_apply.argtypes = [C.c_int64, _f64p, _f64p, C.c_double, _f64p, _f64p, C.c_int]
        L.or_hat_create.restype = C.c_void_p
        L.or_hat_create.argtypes = [C.c_int, _i32p]
        L.or_hat_free.argtypes = [C.c_void_p]
        L.or_hat_apply.argtypes = [C.c_void_p, _f64p, _f64p]
        L.or_hat_apply_t.argtypes = [C.c_void_p, _f64p, _f64p]
        L.or_hat_stencil.restype = C.c_int64
        L.or_hat_stencil.argtypes = [C.c_int, _i32p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_collision_apply_occ.argtypes = [C.c_int64, _f64p, _f64p, C.c_double, _f64p, C.c_void_p, _f64p,
                                             C.c_int]
        L.or_collision_general_occ.argtypes = [C.c_int64, _f64p, _f64p, C.c_double, C.c_double, _f64p,
                                               C.c_void_p, _f64p]
        L.or_collision_vjp_occ.argtypes = [C.c_int64, _f64p, _f64p, C.c_double, _f64p, C.c_void_p, _f64p,
                                           _f64p]
        L.or_collision_vjp_ex.argtypes = [C.c_int64, _f64p, _f64p, C.c_double, _f64p, C.c_void_p, C.c_int,
                                          _f64p, _f64p]
        L.or_conservation.argtypes = [C.c_int64, _f64p, _f64p, _f64p, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
        L.or_relative_error.restype = C.c_double
        L.or_relative_error.argtypes = [C.c_int, C.c_int64, _i32p, _f64p, _f64p, C.c_double,
                                        C.c_double, C.c_double, C.c_double]
        L.or_dp5_integrate.argtypes = [C.c_int64, _f64p, C.c_double, C.c_double, C.c_double,
                                       RHS_FN, NORM_FN, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_int32, C.c_void_p, C.c_void_p, C.c_int64,
                                       C.POINTER(DP5Stats)]
        L.or_dp5_collision.argtypes = [C.c_int, _i32p, C.c_int64, _f64p, _f64p, C.c_double, _f64p,
                                       C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_int, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                       C.c_void_p, C.c_int64, C.POINTER(DP5Stats)]
        L.or_dp5_collision_ex.argtypes = [C.c_int, _i32p, C.c_int64, _f64p, _f64p, C.c_void_p,
                                          C.c_double, C.c_void_p, C.c_int32, C.c_double, C.c_double,
                                          C.c_double, C.c_double, _f64p, C.c_double, C.c_double,
                                          C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p,
                                          C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64,
                                          C.POINTER(DP5Stats), C.POINTER(C.c_double)]
        L.or_dp5_collision_replay_ex.argtypes = [C.c_int, _i32p, C.c_int64, _f64p, _f64p, C.c_void_p,
                                                 C.c_double, C.c_void_p, C.c_int32, C.c_double,
                                                 C.c_double, _f64p, C.c_double, C.c_double,
                                                 C.c_double, C.c_int, _f64p, _i8p, C.c_int64,
                                                 C.c_void_p, C.c_void_p, C.c_int32,
                                                 C.POINTER(C.c_double)]
        L.or_dp5_mc_eph_replay.argtypes = [C.c_int, _i32p, _f64p, _f64p, C.c_void_p,
                                           C.c_double, C.c_double, C.c_int32, _f64p, C.c_int32,
                                           _i32p, C.c_int64, C.c_int32, C.c_double, C.c_void_p,
                                           C.c_int32, C.c_double, C.c_double, _f64p, C.c_double,
                                           C.c_double, C.c_double, _f64p, C.c_int64, C.c_void_p,
                                           C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
        _i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.or_collision_apply_band.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                              _i32p, C.c_double, _f64p, _f64p, _f64p, C.c_int]
        L.or_collision_apply_band_occ.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                                  _i32p, C.c_double, _f64p, C.c_void_p, C.c_int, _f64p,
                                                  _f64p, C.c_int]
        L.or_dp5_replay_ex.argtypes = [C.c_int64, _f64p, C.c_double, C.c_double, C.c_double, RHS_FN,
                                       C.c_void_p, C.c_void_p, C.c_void_p, _f64p, _i8p, C.c_int64,
                                       C.c_void_p, C.c_void_p, C.c_int32]
        L.or_sched_value.restype = C.c_double
        L.or_sched_value.argtypes = [_f64p, C.c_int32, C.c_double, C.c_double, C.c_double]
        L.or_step_vetoer.restype = C.c_double
        L.or_step_vetoer.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.or_energy_dg.restype = C.c_double
        L.or_energy_dg.argtypes = [C.c_int, _i32p, _f64p, _f64p, _f64p]
        L.or_dp5_mc_eph_ex.argtypes = [C.c_int, _i32p, _f64p, _f64p, C.c_void_p, C.c_double,
                                       C.c_double, C.c_int32, _f64p, C.c_int32, _i32p, C.c_int64,
                                       C.c_int32, C.c_double, C.c_void_p, C.c_int32, C.c_double,
                                       C.c_double, C.c_double, C.c_double, _f64p, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_double, C.c_void_p,
                                       C.c_void_p, C.c_int32, C.c_int32, C.POINTER(DP5Stats),
                                       C.POINTER(C.c_double)]
        L.or_dp5_collision_replay.argtypes = [C.c_int, _i32p, C.c_int64, _f64p, _f64p, C.c_double,
                                              _f64p, C.c_double, C.c_double, C.c_double, C.c_int,
                                              _f64p, _i8p, C.c_int64, C.c_void_p, C.c_void_p,
                                              C.c_int32]
        L.or_kernel_domega_full.argtypes = [C.c_int64, _i32p, _i32p, _f64p, C.c_int32, C.c_double,
                                            C.c_double, C.c_double, C.c_double, C.c_double, _f64p]
        L.or_collision_general.argtypes = [C.c_int64, _f64p, _f64p, C.c_double, C.c_double, _f64p,
                                           _f64p]
        L.or_collision_vjp.argtypes = [C.c_int64, _f64p, _f64p, C.c_double, _f64p, _f64p, _f64p]
        L.or_dp5_adjoint.argtypes = [C.c_int64, _f64p, _f64p, _f64p, C.c_double, C.c_double,
                                     C.c_double, _f64p, C.c_double, C.c_double, C.c_double, _f64p,
                                     C.c_int64, C.c_int32, _f64p, _f64p, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        L.or_dp5_adjoint_occ.argtypes = [C.c_int64, _f64p, _f64p, _f64p, C.c_double, C.c_double,
                                         C.c_double, C.c_void_p, C.c_int, _f64p, C.c_double, C.c_double,
                                         C.c_double, _f64p, C.c_int64, C.c_int32, _f64p, _f64p,
                                         C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_mc_rows.restype = C.c_int32
        L.or_mc_rows.argtypes = [C.c_int, C.c_int]
        L.or_mc_points.argtypes = [C.c_int, C.c_int, C.c_int32, C.c_uint64, _f64p]
        L.or_mc_invert.restype = C.c_double
        L.or_mc_invert.argtypes = [C.c_int, _f64p, _f64p]
        L.or_mc_integrate.restype = C.c_int
        L.or_mc_integrate.argtypes = [C.c_int, C.c_int, _f64p, _f64p, _f64p, _f64p, _f64p,
                                      C.c_int32, _f64p]
        L.or_mc_tuples_eph.restype = C.c_int64
        L.or_mc_tuples_eph.argtypes = [C.c_int, _i32p, _f64p, _f64p, C.c_void_p, C.c_double,
                                       C.c_int32, _f64p, C.c_int32, C.c_double, C.c_double,
                                       C.c_void_p, C.c_int64]
        L.or_mc_inverse_mass.argtypes = [C.c_int, C.c_double, _f64p]
        L.or_mc_apply_eph.argtypes = [C.c_int, _i32p, _f64p, _f64p, C.c_void_p, C.c_double,
                                      C.c_double, C.c_int32, _f64p, C.c_int32, _i32p, C.c_int64,
                                      _f64p, _f64p, _f64p, C.c_void_p, C.c_int32]
        L.or_mc_nodal_to_dg.argtypes = [C.c_int, _i32p, _f64p, _f64p]
        L.or_mc_tuples_ee.restype = C.c_int64
        L.or_mc_tuples_ee.argtypes = [C.c_int, _i32p, _f64p, _f64p, C.c_int32, _f64p, C.c_int32,
                                      C.c_double, C.c_double, C.c_int32, C.c_void_p, C.c_int64]
        L.or_mc_tuple_ee.restype = C.c_int
        L.or_mc_tuple_ee.argtypes = [C.c_int, _i32p, _f64p, _f64p, C.c_int32, _f64p, C.c_int32,
                                     C.c_double, _i32p, _f64p, _f64p, _f64p, _f64p]
        L.or_dp5_mc_ee.argtypes = [C.c_int, _i32p, _f64p, _f64p, C.c_int32, _f64p, C.c_int32, C.c_double,
                                   C.c_double, C.c_double, _i32p, C.c_int64, _f64p, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_int32,
                                   C.c_void_p]
        L.or_mc_apply_ee.argtypes = [C.c_int, _i32p, _f64p, _f64p, C.c_int32, _f64p, C.c_int32,
                                     C.c_double, C.c_double, C.c_double, _i32p, C.c_int64, _f64p,
                                     C.c_int32, _f64p]
        L.or_mc_rates_eph.argtypes = [C.c_int, _i32p, _f64p, _f64p, C.c_void_p, C.c_double,
                                      C.c_double, C.c_int32, _f64p, C.c_int32, _i32p, C.c_int64,
                                      _f64p, _f64p, _f64p, C.c_int32]
        L.or_relative_error_dg.restype = C.c_double
        L.or_relative_error_dg.argtypes = [C.c_int, C.c_int64, _f64p, _f64p, C.c_double, C.c_double,
                                           C.c_double, C.c_double]
        L.or_dp5_mc_eph.argtypes = [C.c_int, _i32p, _f64p, _f64p, C.c_void_p, C.c_double,
                                    C.c_double, C.c_int32, _f64p, C.c_int32, _i32p, C.c_int64,
                                    C.c_int32, C.c_double, _f64p, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_int32,
                                    C.c_int32, C.POINTER(DP5Stats)]
        L.or_ee_grid.argtypes = [C.c_int, _i32p, _f64p, C.c_double, C.POINTER(C.c_int32)]
        L.or_ee_apply.argtypes = [C.c_int, _i32p, _f64p, _f64p, _f64p, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_double, _f64p]
        L.or_ee_apply_direct.argtypes = [C.c_int, _i32p, _f64p, _f64p, _f64p, C.c_double,
                                         C.c_double, C.c_double, C.c_double, _f64p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ------------------------------------------------------------------ mesh
def _nsec(nsec):
    a = np.zeros(3, np.int32)
    a[: len(nsec)] = nsec
    return a


def _vec3(v):
    a = np.zeros(3, np.float64)
    a[: len(v)] = v
    return a


def num_nodes(dim, nsec):
    return int(lib().or_num_nodes(dim, _nsec(nsec)))


def num_elements(dim, nsec):
    return int(lib().or_num_elements(dim, _nsec(nsec)))


def element_volume(dim, nsec, extent):
    return float(lib().or_element_volume(dim, _nsec(nsec), _vec3(extent)))


def mesh_nodes(dim, nsec, origin, extent):
    n = num_nodes(dim, nsec)
    out = np.zeros(n * dim)
    lib().or_mesh_nodes(dim, _nsec(nsec), _vec3(origin), _vec3(extent), out)
    return out.reshape(n, dim)


def mesh_elements(dim, nsec):
    ne = num_elements(dim, nsec)
    out = np.zeros(ne * (dim + 1), np.int32)
    lib().or_mesh_elements(dim, _nsec(nsec), out)
    return out.reshape(ne, dim + 1)


def node_volumes(dim, nsec, extent):
    out = np.zeros(num_nodes(dim, nsec))
    lib().or_node_volumes(dim, _nsec(nsec), _vec3(extent), out)
    return out


# ------------------------------------------------------ energy projection
def egrid(eps, de):
    g = EGrid()
    e = np.ascontiguousarray(eps, np.float64)
    lib().or_egrid_make(e, e.size, de, C.byref(g))
    return g


def projection(dim, nsec, extent, eps, grid):
    """Returns (off, len, R) with R of shape (n, maxw)."""
    n = num_nodes(dim, nsec)
    eps = np.ascontiguousarray(eps, np.float64)
    off = np.zeros(n, np.int32)
    ln = np.zeros(n, np.int32)
    maxw = lib().or_projection(dim, _nsec(nsec), _vec3(extent), eps, C.byref(grid), off, ln, None, 0)
    R = np.zeros((n, maxw))
    lib().or_projection(dim, _nsec(nsec), _vec3(extent), eps, C.byref(grid), off, ln, _ptr(R), maxw)
    return off, ln, R


def simplex_density(dim, evert, a, vol, x):
    return float(lib().or_simplex_density(dim, np.ascontiguousarray(evert, np.float64), a, vol, x))


def gauss_table(de, shift, sigma, cutoff):
    mlo = C.c_int32()
    cnt = lib().or_gauss_table(de, shift, sigma, cutoff, C.byref(mlo), None, 0)
    gm = np.zeros(cnt)
    lib().or_gauss_table(de, shift, sigma, cutoff, C.byref(mlo), _ptr(gm), cnt)
    return int(mlo.value), gm


class Kernel:
    """Emission kernel K of the nodal W-form on a structured mesh."""

    def __init__(self, dim, nsec, origin, extent, eps, g, omega0, sigma, de=None, cutoff=8.0):
        self.dim, self.nsec = dim, list(nsec)
        self.origin, self.extent = list(origin), list(extent)
        self.eps = np.ascontiguousarray(eps, np.float64)
        self.g, self.omega0, self.sigma = g, omega0, sigma
        self.de = sigma / 8.0 if de is None else de
        self.cutoff = cutoff
        self.n = num_nodes(dim, nsec)
        self.V = node_volumes(dim, nsec, extent)
        self.grid = egrid(self.eps, self.de)
        self.off, self.len, self.R = projection(dim, nsec, extent, self.eps, self.grid)
        self.maxw = self.R.shape[1]
        self.m_lo, self.gm = gauss_table(self.de, omega0, sigma, cutoff)
        self.amp = g * g

    def domega_full(self):
        """dK/domega0 on the same windows (the sampled G' table)."""
        Kd = np.zeros((self.n, self.n))
        lib().or_kernel_domega_full(self.n, self.off, self.len, self.R, self.maxw, self.de,
                                    self.omega0, self.sigma, self.cutoff, self.amp, Kd)
        return Kd

    def row(self, i):
        out = np.zeros(self.n)
        lib().or_kernel_row(self.n, i, self.off, self.len, self.R, self.maxw, self.gm,
                            self.m_lo, self.gm.size, self.amp, out)
        return out

    def full(self):
        K = np.zeros((self.n, self.n))
        lib().or_kernel_full(self.n, self.off, self.len, self.R, self.maxw, self.gm,
                             self.m_lo, self.gm.size, self.amp, K)
        return K


# ------------------------------------------------- hat-averaged occupation
OCC_CONSERVE = 1   # bits of the DP5 wrappers' mode argument (OR_CONSERVE, OR_OCC_HAT)
OCC_HAT = 2


class HatStencil:
    """fbar = S f, fbar_i = int phi_i f / int phi_i of the P1 interpolant on the
    structured mesh (consistent mass over lumped mass)."""

    def __init__(self, dim, nsec):
        self.dim, self.nsec = dim, list(nsec)
        self.n = num_nodes(dim, nsec)
        self.h = lib().or_hat_create(dim, _nsec(nsec))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_hat_free(self.h)
            self.h = None

    def apply(self, f):
        out = np.zeros(self.n)
        lib().or_hat_apply(self.h, np.ascontiguousarray(f, np.float64), out)
        return out

    def apply_t(self, z):
        out = np.zeros(self.n)
        lib().or_hat_apply_t(self.h, np.ascontiguousarray(z, np.float64), out)
        return out

    def csr(self):
        nnz = lib().or_hat_stencil(self.dim, _nsec(self.nsec), None, None, None)
        ptr = np.zeros(self.n + 1, np.int64)
        col = np.zeros(nnz, np.int32)
        w = np.zeros(nnz)
        lib().or_hat_stencil(self.dim, _nsec(self.nsec), _ptr(ptr), _ptr(col), _ptr(w))
        return ptr, col, w


def _occ_mode(conserve, occupation):
    if occupation not in ("nodal", "hat"):
        raise ValueError(f"occupation must be 'nodal' or 'hat', not {occupation!r}")
    return (OCC_CONSERVE if conserve else 0) | (OCC_HAT if occupation == "hat" else 0)


# ---------------------------------------------------------------- collision
def collision_apply(K, V, bath_n, f, conserve=False, hat=None):
    """The W-form; hat (a HatStencil): the test function's occupation factors
    are the hat averages of f instead of the nodal values."""
    K = np.ascontiguousarray(K, np.float64)
    f = np.ascontiguousarray(f, np.float64)
    out = np.zeros_like(f)
    lib().or_collision_apply_occ(f.size, K, np.ascontiguousarray(V, np.float64), bath_n, f,
                                 None if hat is None else hat.h, out, int(bool(conserve)))
    return out


class KernelBand:
    """K's band in energy order (the CPU W-form that streams the same bytes as
    the GPU's plain evaluation).  Built from a dense K (node order) in row
    blocks: perm sorts the nodes by energy (stable), row p keeps its columns
    from the first to the last nonzero."""

    def __init__(self, K_rows, eps, V, block=1024):
        n = eps.size
        self.n = n
        self.perm = np.ascontiguousarray(np.argsort(eps, kind="stable"), np.int32)
        lo = np.zeros(n, np.int32)
        ln = np.zeros(n, np.int64)
        chunks = []
        rk = np.zeros(n)
        ck = np.zeros(n)
        for p0 in range(0, n, block):
            rows = K_rows(self.perm[p0:p0 + block])          # node-order rows
            Ks = rows[:, self.perm]
            rk[p0:p0 + len(rows)] = Ks.sum(1)
            ck += Ks.sum(0)
            for q in range(len(Ks)):
                nz = np.flatnonzero(Ks[q])
                if nz.size:
                    lo[p0 + q] = nz[0]
                    ln[p0 + q] = nz[-1] + 1 - nz[0]
                    chunks.append(Ks[q, nz[0]:nz[-1] + 1].copy())
        self.lo = lo
        self.rowptr = np.concatenate([[0], np.cumsum(ln)]).astype(np.int64)
        self.vals = np.ascontiguousarray(np.concatenate(chunks) if chunks else np.zeros(1))
        self.rk, self.ck = rk, ck
        self.Vs = np.ascontiguousarray(V[self.perm], np.float64)
        self.entries = int(self.rowptr[-1])

    def apply(self, bath_n, f, nthreads=1, scratch=None, hat=None, conserve=False):
        f = np.ascontiguousarray(f, np.float64)
        out = np.empty_like(f)
        if scratch is None:
            scratch = np.empty((nthreads + 3) * self.n)
        lib().or_collision_apply_band_occ(self.n, self.rowptr, self.lo, self.vals, self.rk, self.ck, self.Vs,
                                          self.perm, bath_n, f, None if hat is None else hat.h,
                                          int(bool(conserve)), out, scratch, nthreads)
        return out


def conservation(V, eps, f):
    num, en = C.c_double(), C.c_double()
    lib().or_conservation(f.size, np.ascontiguousarray(V, np.float64),
                          np.ascontiguousarray(eps, np.float64),
                          np.ascontiguousarray(f, np.float64), C.byref(num), C.byref(en))
    return num.value, en.value


def relative_error(dim, conn, err, pop, stat=-1.0, rtol=1e-3, atol=1e-3, penalty=0.0):
    conn = np.ascontiguousarray(conn, np.int32)
    return float(lib().or_relative_error(dim, conn.shape[0], conn.reshape(-1),
                                         np.ascontiguousarray(err, np.float64),
                                         np.ascontiguousarray(pop, np.float64),
                                         stat, rtol, atol, penalty))


# ------------------------------------------------------------------- DP5
def dp5_integrate(rhs, norm, y0, t0, t_final, dense_dt, max_dense=100000, log_cap=1 << 20):
    """Generic DP5(4) with Python callbacks: rhs(t, y) -> dydt, norm(err, y5) -> float."""
    y = np.array(y0, np.float64, copy=True)
    n = y.size

    def _rhs(_u, t, yp, dp):
        ya = np.ctypeslib.as_array(yp, shape=(n,))
        np.ctypeslib.as_array(dp, shape=(n,))[:] = rhs(t, ya.copy())

    def _norm(_u, ep, yp):
        return float(norm(np.ctypeslib.as_array(ep, shape=(n,)).copy(),
                          np.ctypeslib.as_array(yp, shape=(n,)).copy()))

    cr, cn = RHS_FN(_rhs), NORM_FN(_norm)
    dense = np.zeros((max_dense, n))
    dt = np.zeros(max_dense)
    hl = np.zeros(log_cap)
    al = np.zeros(log_cap, np.int8)
    st = DP5Stats()
    lib().or_dp5_integrate(n, y, t0, t_final, dense_dt, cr, cn, None, _ptr(dense), _ptr(dt),
                           max_dense, _ptr(hl), _ptr(al), log_cap, C.byref(st))
    nd = min(st.n_dense, max_dense)
    na = min(st.attempted, log_cap)
    return dict(y=y, dense=dense[:nd], dense_t=dt[:nd], h=hl[:na], accepted=al[:na].astype(bool),
                stats=st)


class BathSchedule:
    """Samples of a uniform constrained bath n(t) at t_start + k dt (linear
    between samples, held outside) and the band's error tolerances for the
    step vetoer (Band::constrainedPopulation, MaterialTimeStatus::timeStepVetoer)."""

    def __init__(self, samples, t_start, dt, atol=1e-3, rtol=1e-3):
        self.samples = np.ascontiguousarray(samples, np.float64)
        self.t_start, self.dt, self.atol, self.rtol = float(t_start), float(dt), float(atol), float(rtol)

    def _struct(self):
        return _BathSchedC(self.samples.ctypes.data_as(C.POINTER(C.c_double)), self.samples.size,
                           self.t_start, self.dt, self.atol, self.rtol)

    def value(self, t):
        return lib().or_sched_value(self.samples, self.samples.size, self.t_start, self.dt, t)


class _BathSchedC(C.Structure):
    _fields_ = [("s", C.POINTER(C.c_double)), ("ns", C.c_int32), ("t_start", C.c_double),
                ("dt", C.c_double), ("atol", C.c_double), ("rtol", C.c_double)]


def step_vetoer(sched, t, h, dense_dt):
    """MaterialTimeStatus::timeStepVetoer on timesToTest(t, h); sched None: no constrained band."""
    if sched is None:
        return lib().or_step_vetoer(None, t, h, dense_dt)
    cs = sched._struct()
    return lib().or_step_vetoer(C.cast(C.pointer(cs), C.c_void_p), t, h, dense_dt)


def dp5_replay(rhs, y0, t0, t_final, dense_dt, steps, max_dense=100000):
    """Fixed-step DP5 with a Python rhs(t, y) -> dydt on the accepted steps
    `steps` (the stage arithmetic and dense output of dp5_integrate)."""
    y = np.array(y0, np.float64, copy=True)
    n = y.size

    def _rhs(_u, t, yp, dp):
        np.ctypeslib.as_array(dp, shape=(n,))[:] = rhs(t, np.ctypeslib.as_array(yp, shape=(n,)).copy())

    cr = RHS_FN(_rhs)
    h = np.ascontiguousarray(steps, np.float64)
    acc = np.ones(h.size, np.int8)
    dense = np.zeros((max_dense, n))
    dt = np.zeros(max_dense)
    lib().or_dp5_replay_ex(n, y, t0, t_final, dense_dt, cr, None, None, None, h, acc, h.size,
                           _ptr(dense), _ptr(dt), max_dense)
    nd = int(np.count_nonzero(dt))
    return dict(y=y, dense=dense[:nd], dense_t=dt[:nd])


def dp5_collision(dim, nsec, K, V, bath_n, f0, t0, t_final, dense_dt, rtol, atol,
                  conserve=False, max_dense=100000, log_cap=1 << 20, sched=None, eps=None,
                  occupation="nodal"):
    """Adaptive DP5 on the W-form.  sched (BathSchedule): time-dependent bath
    instead of bath_n.  eps (nodal energies): also return the bath energy
    ledger ("ledger": the energy the bath received)."""
    f = np.array(f0, np.float64, copy=True)
    n = f.size
    dense = np.zeros((max_dense, n)) if max_dense else None
    dt = np.zeros(max(max_dense, 1))
    hl = np.zeros(log_cap)
    al = np.zeros(log_cap, np.int8)
    st = DP5Stats()
    led = C.c_double(0.0)
    e = None if eps is None else np.ascontiguousarray(eps, np.float64)
    sp = None if sched is None else sched.samples
    lib().or_dp5_collision_ex(dim, _nsec(nsec), n, np.ascontiguousarray(K, np.float64),
                              np.ascontiguousarray(V, np.float64), _ptr(e), bath_n, _ptr(sp),
                              0 if sched is None else sp.size,
                              0.0 if sched is None else sched.t_start,
                              1.0 if sched is None else sched.dt,
                              0.0 if sched is None else sched.atol,
                              0.0 if sched is None else sched.rtol, f, t0, t_final, dense_dt,
                              rtol, atol, _occ_mode(conserve, occupation), _ptr(dense), _ptr(dt), max_dense,
                              _ptr(hl), _ptr(al), log_cap, C.byref(st),
                              C.byref(led) if e is not None else None)
    nd = min(st.n_dense, max_dense)
    na = min(st.attempted, log_cap)
    return dict(y=f, dense=None if dense is None else dense[:nd], dense_t=dt[:nd], h=hl[:na],
                accepted=al[:na].astype(bool), stats=st, ledger=led.value)


def dp5_collision_replay(dim, nsec, K, V, bath_n, f0, t0, t_final, dense_dt, h, accepted,
                         conserve=False, max_dense=100000, sched=None, eps=None, occupation="nodal"):
    """Fixed-step DP5 on the recorded attempts (h, accepted); sched: a
    BathSchedule instead of bath_n; with eps the bath energy ledger of the
    replayed steps is returned too ("ledger")."""
    f = np.array(f0, np.float64, copy=True)
    n = f.size
    dense = np.zeros((max_dense, n))
    dt = np.zeros(max_dense)
    h = np.ascontiguousarray(h, np.float64)
    acc = np.ascontiguousarray(np.asarray(accepted).astype(np.int8))
    led = C.c_double(0.0)
    e = None if eps is None else np.ascontiguousarray(eps, np.float64)
    sp = None if sched is None else sched.samples
    lib().or_dp5_collision_replay_ex(dim, _nsec(nsec), n, np.ascontiguousarray(K, np.float64),
                                     np.ascontiguousarray(V, np.float64), _ptr(e), bath_n, _ptr(sp),
                                     0 if sched is None else sp.size,
                                     0.0 if sched is None else sched.t_start,
                                     1.0 if sched is None else sched.dt, f, t0, t_final,
                                     dense_dt, _occ_mode(conserve, occupation), h, acc, h.size, _ptr(dense),
                                     _ptr(dt), max_dense, C.byref(led) if e is not None else None)
    nd = int(np.count_nonzero(dt))
    return dict(y=f, dense=dense[:nd], dense_t=dt[:nd], ledger=led.value)


def collision_vjp(K, V, bath_n, f, mu, hat=None, conserve=False):
    out = np.zeros_like(np.asarray(f, np.float64))
    lib().or_collision_vjp_ex(out.size, np.ascontiguousarray(K, np.float64),
                              np.ascontiguousarray(V, np.float64), bath_n,
                              np.ascontiguousarray(f, np.float64), None if hat is None else hat.h,
                              int(bool(conserve)), np.ascontiguousarray(mu, np.float64), out)
    return out


def dp5_adjoint(K, Kd, V, bath_n, dbath, g, f0, t0, t_final, dense_dt, steps, dL_df, ckpt=32, hat=None,
                conserve=False):
    """Discrete adjoint of the fixed-step DP5 with dense output at t_final.
    Returns (dL/df0, dL/dg, dL/domega0)."""
    n = len(f0)
    out = np.zeros(n)
    dg, dw = C.c_double(), C.c_double()
    steps = np.ascontiguousarray(steps, np.float64)
    lib().or_dp5_adjoint_occ(n, np.ascontiguousarray(K, np.float64), np.ascontiguousarray(Kd, np.float64),
                             np.ascontiguousarray(V, np.float64), bath_n, dbath, g,
                             None if hat is None else hat.h, int(bool(conserve)),
                             np.ascontiguousarray(f0, np.float64), t0, t_final, dense_dt, steps,
                             steps.size, ckpt, np.ascontiguousarray(dL_df, np.float64), out,
                             C.byref(dg), C.byref(dw))
    return out, dg.value, dw.value


# ------------------------------------------------------------ e-e channel
def ee_apply(dim, nsec, extent, eps, f, g, kappa2, sigma, dw, cutoff=8.0):
    """Electron-electron operator with the energy delta projected on the
    grid m*dw (ee_oracle.c); one value per mesh node."""
    nsec = np.ascontiguousarray(nsec, np.int32)
    ext = np.ascontiguousarray(list(extent) + [1.0] * (3 - len(extent)), np.float64)
    f = np.ascontiguousarray(f, np.float64)
    out = np.zeros_like(f)
    lib().or_ee_apply(dim, nsec, ext, np.ascontiguousarray(eps, np.float64), f, g, kappa2, sigma,
                      dw, cutoff, out)
    return out


def ee_apply_direct(dim, nsec, extent, eps, f, g, kappa2, sigma, cutoff=8.0):
    """The same operator with the unprojected Gaussian delta (O(P^3))."""
    nsec = np.ascontiguousarray(nsec, np.int32)
    ext = np.ascontiguousarray(list(extent) + [1.0] * (3 - len(extent)), np.float64)
    f = np.ascontiguousarray(f, np.float64)
    out = np.zeros_like(f)
    lib().or_ee_apply_direct(dim, nsec, ext, np.ascontiguousarray(eps, np.float64), f, g, kappa2,
                             sigma, cutoff, out)
    return out


def ee_grid(dim, nsec, eps, dw):
    mh = C.c_int32()
    lib().or_ee_grid(dim, np.ascontiguousarray(nsec, np.int32),
                     np.ascontiguousarray(eps, np.float64), dw, C.byref(mh))
    return mh.value


# ------------------------------------------------ reference tuple operator
def _pad3(x, fill):
    return np.ascontiguousarray(list(x) + [fill] * (3 - len(x)), np.float64)


def mc_points(dim, nlegs, nmc, seed):
    """MC point matrix [rows][nmc] of a channel (mc_oracle.c)."""
    rows = lib().or_mc_rows(dim, nlegs)
    pts = np.zeros((rows, nmc))
    lib().or_mc_points(dim, nlegs, nmc, seed, pts)
    return pts


def mc_invert(A):
    A = np.ascontiguousarray(A, np.float64)
    inv = np.zeros_like(A)
    det = lib().or_mc_invert(A.shape[0], A, inv)
    return det, inv


def mc_integrate(dim, nlegs, D, N, fa, fb, pts):
    """scatteringIntegrationTypeB with the standard output forms."""
    out = np.zeros(nlegs * (dim + 1))
    pts = np.ascontiguousarray(pts, np.float64)
    rc = lib().or_mc_integrate(dim, nlegs, np.ascontiguousarray(D, np.float64),
                               np.ascontiguousarray(N, np.float64),
                               np.ascontiguousarray(fa, np.float64),
                               np.ascontiguousarray(fb, np.float64), pts, pts.shape[1], out)
    if rc != 0:
        raise ValueError("singular D_C")
    return out


def _eps_ph_ptr(eps_ph):
    if eps_ph is None:
        return None, None
    a = np.ascontiguousarray(eps_ph, np.float64)
    return a, a.ctypes.data


def zone_geometry(dim, origin, extent, sides=None, zone=None):
    """Geometry block of the tuple-operator functions (21 doubles): the
    origin, the mesh's side vectors (side c: node index i_c along it at
    origin + sum_c (i_c / nsec_c) side_c), the zone's side vectors (the
    umklapp basis).  Default: the axis-aligned box origin + extent, which is
    also the zone; with sides and no zone the mesh covers the zone."""
    g = np.zeros(21)
    g[:dim] = np.asarray(origin, np.float64)[:dim]
    S = np.diag(np.asarray(extent, np.float64)[:dim]) if sides is None else np.asarray(sides, np.float64)
    Z = S if zone is None else np.asarray(zone, np.float64)
    for c in range(dim):
        g[3 + 3 * c: 3 + 3 * c + dim] = S[c][:dim]
        g[12 + 3 * c: 12 + 3 * c + dim] = Z[c][:dim]
    return g


def mc_tuples_eph(dim, nsec, origin, extent, eps, omega0, nu, pts, min_det=1e-9,
                  threshold=1e-9, eps_ph=None, sides=None, zone=None):
    """Tuple list int32 [T][4] = (e0, e1, e2, umklapp id) of the e-ph channel
    (phonon band: nodal energies eps_ph, or the constant omega0)."""
    nsec = np.ascontiguousarray(nsec, np.int32)
    geo = zone_geometry(dim, origin, extent, sides, zone)
    eps = np.ascontiguousarray(eps, np.float64)
    pts = np.ascontiguousarray(pts, np.float64)
    keep, ph = _eps_ph_ptr(eps_ph)
    L = lib()
    n = L.or_mc_tuples_eph(dim, nsec, geo, eps, ph, omega0, nu, pts, pts.shape[1], min_det,
                           threshold, None, 0)
    idx = np.zeros((max(n, 1), 4), np.int32)
    m = L.or_mc_tuples_eph(dim, nsec, geo, eps, ph, omega0, nu, pts, pts.shape[1], min_det,
                           threshold, idx.ctypes.data, n)
    assert m == n
    return idx[:n]


def mc_tuples_ee(dim, nsec, origin, extent, eps, nu, pts, min_det=1e-9, threshold=1e-9,
                 brute=False, sides=None, zone=None):
    """Tuple list int32 [T][5] = (e0, e1, e2, e3, umklapp id) of the 4-leg
    channel {el in, el in, el out, el out} (mc_oracle.c); brute: every leg-3
    element tested, without the box and energy pruning."""
    nsec = np.ascontiguousarray(nsec, np.int32)
    geo = zone_geometry(dim, origin, extent, sides, zone)
    eps = np.ascontiguousarray(eps, np.float64)
    pts = np.ascontiguousarray(pts, np.float64)
    L = lib()
    n = L.or_mc_tuples_ee(dim, nsec, geo, eps, nu, pts, pts.shape[1], min_det, threshold,
                          int(brute), None, 0)
    idx = np.zeros((max(n, 1), 5), np.int32)
    m = L.or_mc_tuples_ee(dim, nsec, geo, eps, nu, pts, pts.shape[1], min_det, threshold,
                          int(brute), idx.ctypes.data, n)
    assert m == n
    return idx[:n]


def mc_tuple_ee(dim, nsec, origin, extent, eps, nu, pts, tup, c, min_det=1e-9, sides=None,
                zone=None):
    """One 4-leg tuple's integral: (res [4][dim+1], D, N, ok) for the legs'
    DG coefficient blocks c [4][dim+1] (no amplitude, no leg signs)."""
    nsec = np.ascontiguousarray(nsec, np.int32)
    geo = zone_geometry(dim, origin, extent, sides, zone)
    pts = np.ascontiguousarray(pts, np.float64)
    res = np.zeros((4, dim + 1))
    D = np.zeros((dim + 1, 4 * dim))
    N = np.zeros(dim + 1)
    rc = lib().or_mc_tuple_ee(dim, nsec, geo, np.ascontiguousarray(eps, np.float64), nu, pts,
                              pts.shape[1], min_det, np.ascontiguousarray(tup, np.int32),
                              np.ascontiguousarray(c, np.float64), res, D, N)
    return res, D, N, rc == 0


def mc_apply_ee(dim, nsec, origin, extent, eps, nu, pts, idx, f, amplitude=1.0, kappa=0.0,
                min_det=1e-9, sides=None, zone=None):
    """DG derivative [batch][nel][dim+1] of the 4-leg channel over a tuple
    list, after the inverse mass."""
    nsec = np.ascontiguousarray(nsec, np.int32)
    geo = zone_geometry(dim, origin, extent, sides, zone)
    pts = np.ascontiguousarray(pts, np.float64)
    f = np.ascontiguousarray(f, np.float64)
    batch = f.shape[0] if f.ndim == 3 else 1
    out = np.zeros_like(f)
    idx = np.ascontiguousarray(idx, np.int32).reshape(-1, 5)
    lib().or_mc_apply_ee(dim, nsec, geo, np.ascontiguousarray(eps, np.float64), nu, pts,
                         pts.shape[1], min_det, amplitude, kappa, idx, idx.shape[0], f, batch, out)
    return out


def dp5_mc_ee(dim, nsec, origin, extent, eps, nu, pts, idx, f0, t0, t_final, dense_dt, rtol, atol,
              amplitude=1.0, kappa=0.0, min_det=1e-9, max_dense=0, sides=None, zone=None):
    """DP5(4) of the electron band's DG state [nel][dim+1] with the 4-leg channel."""
    nsec = np.ascontiguousarray(nsec, np.int32)
    geo = zone_geometry(dim, origin, extent, sides, zone)
    pts = np.ascontiguousarray(pts, np.float64)
    idx = np.ascontiguousarray(idx, np.int32).reshape(-1, 5)
    f = np.ascontiguousarray(f0, np.float64).copy()
    dense = np.zeros((max(max_dense, 1),) + f.shape)
    dense_t = np.zeros(max(max_dense, 1))
    st = DP5Stats()
    lib().or_dp5_mc_ee(dim, nsec, geo, np.ascontiguousarray(eps, np.float64), nu, pts, pts.shape[1], min_det,
                       amplitude, kappa, idx, idx.shape[0], f, t0, t_final, dense_dt, rtol, atol,
                       dense.ctypes.data if max_dense else None, dense_t.ctypes.data if max_dense else None,
                       max_dense, C.byref(st))
    return {"f": f, "dense": dense[:st.n_dense] if max_dense else None,
            "dense_t": dense_t[:st.n_dense] if max_dense else None, "stats": st}


def mc_inverse_mass(dim, vol):
    m = np.zeros((dim + 1, dim + 1))
    lib().or_mc_inverse_mass(dim, vol, m)
    return m


def mc_apply_eph(dim, nsec, origin, extent, eps, omega0, amplitude, nu, pts, idx, f_el, f_ph,
                 nthreads=1, phonon_out=False, eps_ph=None, sides=None, zone=None):
    """Electron (and optionally phonon) DG derivative of the e-ph channel."""
    nsec = np.ascontiguousarray(nsec, np.int32)
    geo = zone_geometry(dim, origin, extent, sides, zone)
    pts = np.ascontiguousarray(pts, np.float64)
    f_el = np.ascontiguousarray(f_el, np.float64)
    f_ph = np.ascontiguousarray(f_ph, np.float64)
    out_el = np.zeros_like(f_el)
    out_ph = np.zeros_like(f_ph) if phonon_out else None
    idx = np.ascontiguousarray(idx, np.int32)
    keep, ph = _eps_ph_ptr(eps_ph)
    lib().or_mc_apply_eph(dim, nsec, geo, np.ascontiguousarray(eps, np.float64), ph, omega0,
                          amplitude, nu, pts, pts.shape[1], idx, idx.shape[0], f_el, f_ph, out_el,
                          None if out_ph is None else out_ph.ctypes.data, nthreads)
    return (out_el, out_ph) if phonon_out else out_el


def mc_nodal_to_dg(dim, nsec, f):
    nsec = np.ascontiguousarray(nsec, np.int32)
    nel = int(np.prod(nsec)) * (2 if dim == 2 else 6 if dim == 3 else 1)
    dg = np.zeros((nel, dim + 1))
    lib().or_mc_nodal_to_dg(dim, nsec, np.ascontiguousarray(f, np.float64), dg)
    return dg


def relative_error_dg(dim, err, pop, stat=-1.0, rtol=1e-8, atol=1e-12, penalty=0.0):
    """Band::relativeError on DG coefficients [nel][dim+1]."""
    err = np.ascontiguousarray(err, np.float64)
    pop = np.ascontiguousarray(pop, np.float64)
    return lib().or_relative_error_dg(dim, err.shape[0], err, pop, stat, rtol, atol, penalty)


def energy_dg(dim, nsec, extent, eps, f, sides=None):
    """int eps f dk of DG coefficients [nel][dim+1] (exact per element)."""
    geo = zone_geometry(dim, [0.0] * dim, extent, sides, None)
    return lib().or_energy_dg(dim, np.ascontiguousarray(nsec, np.int32), geo,
                              np.ascontiguousarray(eps, np.float64), np.ascontiguousarray(f, np.float64))


def dp5_mc_eph(dim, nsec, origin, extent, eps, omega0, amplitude, nu, pts, idx, bath_n, f0, t0,
               t_final, dense_dt, rtol, atol, max_dense=0, nthreads=8, dynamic=False, eps_ph=None,
               sched=None, sides=None, zone=None):
    """DP5(4) of the electron band's DG state with the tuple operator; with
    dynamic=True the state is [electron DG | phonon DG] and bath_n unused.
    sched (BathSchedule): time-dependent bath.  The result holds the bath
    energy ledger ("ledger")."""
    nsec = np.ascontiguousarray(nsec, np.int32)
    geo = zone_geometry(dim, origin, extent, sides, zone)
    pts = np.ascontiguousarray(pts, np.float64)
    idx = np.ascontiguousarray(idx, np.int32)
    f = np.ascontiguousarray(f0, np.float64).copy()
    dense = np.zeros((max(max_dense, 1),) + f.shape)
    dense_t = np.zeros(max(max_dense, 1))
    st = DP5Stats()
    keep, ph = _eps_ph_ptr(eps_ph)
    led = C.c_double(0.0)
    sp = None if sched is None else sched.samples
    lib().or_dp5_mc_eph_ex(dim, nsec, geo, np.ascontiguousarray(eps, np.float64), ph, omega0,
                           amplitude, nu, pts, pts.shape[1], idx, idx.shape[0], 1 if dynamic else 0,
                           bath_n, _ptr(sp), 0 if sched is None else sp.size,
                           0.0 if sched is None else sched.t_start,
                           1.0 if sched is None else sched.dt,
                           0.0 if sched is None else sched.atol,
                           0.0 if sched is None else sched.rtol, f, t0, t_final,
                           dense_dt, rtol, atol, dense.ctypes.data if max_dense else None,
                           dense_t.ctypes.data if max_dense else None, max_dense, nthreads,
                           C.byref(st), C.byref(led))
    return {"f": f, "dense": dense[:st.n_dense] if max_dense else None,
            "dense_t": dense_t[:st.n_dense] if max_dense else None, "stats": st,
            "ledger": led.value}


def dp5_mc_eph_replay(dim, nsec, origin, extent, eps, omega0, amplitude, nu, pts, idx, bath_n, f0,
                      t0, t_final, dense_dt, steps, max_dense=0, nthreads=8, dynamic=False,
                      eps_ph=None, sched=None, sides=None, zone=None):
    """Fixed-step DP5 of the tuple operator on the accepted steps `steps`;
    returns the final state, dense outputs and the bath energy ledger."""
    nsec = np.ascontiguousarray(nsec, np.int32)
    geo = zone_geometry(dim, origin, extent, sides, zone)
    pts = np.ascontiguousarray(pts, np.float64)
    idx = np.ascontiguousarray(idx, np.int32)
    steps = np.ascontiguousarray(steps, np.float64)
    f = np.ascontiguousarray(f0, np.float64).copy()
    dense = np.zeros((max(max_dense, 1),) + f.shape)
    dense_t = np.zeros(max(max_dense, 1))
    keep, ph = _eps_ph_ptr(eps_ph)
    led = C.c_double(0.0)
    sp = None if sched is None else sched.samples
    lib().or_dp5_mc_eph_replay(dim, nsec, geo, np.ascontiguousarray(eps, np.float64), ph, omega0,
                               amplitude, nu, pts, pts.shape[1], idx, idx.shape[0], 1 if dynamic else 0,
                               bath_n, _ptr(sp), 0 if sched is None else sp.size,
                               0.0 if sched is None else sched.t_start,
                               1.0 if sched is None else sched.dt, f, t0, t_final, dense_dt, steps,
                               steps.size, dense.ctypes.data if max_dense else None,
                               dense_t.ctypes.data if max_dense else None, max_dense, nthreads,
                               C.byref(led))
    nd = int(np.count_nonzero(dense_t)) if max_dense else 0
    return {"f": f, "dense": dense[:nd] if max_dense else None, "ledger": led.value}


def mc_rates_eph(dim, nsec, origin, extent, eps, omega0, amplitude, nu, pts, idx, f_el, f_ph,
                 nthreads=1, eps_ph=None, sides=None, zone=None):
    """ScatteringChannel::scatteringRates of the electron band (DG, inverse-massed)."""
    nsec = np.ascontiguousarray(nsec, np.int32)
    geo = zone_geometry(dim, origin, extent, sides, zone)
    pts = np.ascontiguousarray(pts, np.float64)
    f_el = np.ascontiguousarray(f_el, np.float64)
    f_ph = np.ascontiguousarray(f_ph, np.float64)
    out = np.zeros_like(f_el)
    idx = np.ascontiguousarray(idx, np.int32)
    keep, ph = _eps_ph_ptr(eps_ph)
    lib().or_mc_rates_eph(dim, nsec, geo, np.ascontiguousarray(eps, np.float64), ph, omega0,
                          amplitude, nu, pts, pts.shape[1], idx, idx.shape[0], f_el, f_ph, out,
                          nthreads)
    return out
