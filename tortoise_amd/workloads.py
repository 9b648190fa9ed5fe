"""Named synthetic workloads of the scattering-integral path.

Each workload fixes a structured momentum mesh over the square (cube) zone
[-pi, pi]^d, a parabolic band eps = |k|^2/2 sampled at the lattice nodes, one
dispersionless optical phonon treated as a bath, and an initial distribution
(Fermi-Dirac plus a Gaussian hot-electron bump in energy).  The names say what
the workload is; the numbers are the ones the measurements and parity tests
use.

Spec decisions recorded here (they are not fixed by the reference):
  * the hot-electron bump width is a standard deviation in energy;
  * the phonon bath occupation is Bose-Einstein at the electron temperature,
    n = 1/(exp(omega0/T) - 1) (0.08943 at omega0 = 0.05, T = 0.02);
  * dense_dt = t_final / 100, which also sets the first trial step and the
    step-underflow limit dense_dt/1e4 (DormandPrince54.hpp:530-532);
  * the energy grid of the overlap quadrature is sigma/8 and the broadened
    delta is cut at 8 sigma.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass(frozen=True)
class Workload:
    name: str
    dim: int
    nsec: tuple
    g: float = 0.1
    omega0: float = 0.05
    sigma: float = 0.01
    mu: float = 0.5
    temperature: float = 0.02
    pump_center: float = 0.9
    pump_width: float = 0.03
    pump_amp: float = 0.2
    t_final: float = 100.0
    rtol: float = 1e-8
    atol: float = 1e-12
    batch: int = 1
    # electron-electron channel (off when ee_g == 0): amplitude
    # ee_g^2 / (|q|^2 + ee_kappa2)^2, Gaussian energy delta of width ee_sigma
    ee_g: float = 0.0
    ee_kappa2: float = 0.09
    ee_sigma: float = 0.02

    @property
    def origin(self):
        return tuple([-math.pi] * self.dim)

    @property
    def extent(self):
        return tuple([2.0 * math.pi] * self.dim)

    @property
    def n_nodes(self):
        return int(np.prod([s + 1 for s in self.nsec]))

    @property
    def dense_dt(self):
        return self.t_final / 100.0

    @property
    def bath_n(self):
        return bose_einstein(self.omega0, self.temperature)

    def node_coords(self):
        axes = [self.origin[d] + (self.extent[d] / self.nsec[d]) * np.arange(self.nsec[d] + 1)
                for d in range(self.dim)]
        grids = np.meshgrid(*axes, indexing="ij")
        # node id = ix + (nx+1)*(iy + (ny+1)*iz): x runs fastest
        return np.stack([g.transpose(*reversed(range(self.dim))).reshape(-1) for g in grids], axis=1)

    def eps(self):
        k = self.node_coords()
        return 0.5 * (k * k).sum(axis=1)

    def initial_f(self, eps=None, pump_center=None, pump_amp=None, temperature=None):
        e = self.eps() if eps is None else eps
        return initial_distribution(
            e, self.mu, self.temperature if temperature is None else temperature,
            self.pump_center if pump_center is None else pump_center, self.pump_width,
            self.pump_amp if pump_amp is None else pump_amp)


def bose_einstein(omega0, temperature):
    return 1.0 / math.expm1(omega0 / temperature)


def initial_distribution(eps, mu, temperature, pump_center, pump_width, pump_amp):
    x = np.clip((eps - mu) / temperature, -700.0, 700.0)
    fd = 1.0 / (np.exp(x) + 1.0)
    return fd + pump_amp * np.exp(-0.5 * ((eps - pump_center) / pump_width) ** 2)


def ensemble_members(batch, seed=0):
    """Independent initial conditions of the batched workload: pump centre
    U[0.2, 1.0], pump amplitude lognormal(mu=-2, sigma=0.5), temperature
    U[0.01, 0.05], drawn from numpy's PCG64 with the given seed."""
    rng = np.random.default_rng(seed)
    centre = rng.uniform(0.2, 1.0, batch)
    amp = rng.lognormal(-2.0, 0.5, batch)
    temp = rng.uniform(0.01, 0.05, batch)
    return centre, amp, temp


# --- the workloads -----------------------------------------------------------
# 32x32 sections, 2048 triangles, 1089 nodes; FP64; t = 100, rtol 1e-8
SQUARE32 = Workload("square32_fp64", 2, (32, 32), t_final=100.0, rtol=1e-8, atol=1e-12)
# 128x128 sections, 32768 triangles, 16641 nodes; FP64; t = 200, rtol 1e-8
SQUARE128 = Workload("square128_fp64", 2, (128, 128), t_final=200.0, rtol=1e-8, atol=1e-12)
# same mesh, 256 independent members, fp16 hi+lo split / FP32 MFMA; rtol 1e-6
SQUARE128_ENSEMBLE = Workload("square128_ensemble256_mixed", 2, (128, 128), t_final=200.0,
                              rtol=1e-6, atol=1e-12, batch=256)
# 64x64 sections, 8192 triangles, 4225 nodes; gradients at t = 50
SQUARE64 = Workload("square64_adjoint", 2, (64, 64), t_final=50.0, rtol=1e-8, atol=1e-12)
# 20^3 sections x 6 = 48000 tetrahedra, 9261 nodes (8000 periodic points);
# e-ph plus e-e with V(q) = 1/(|q|^2 + 0.09), sigma_ee = 0.02; 16 members;
# t = 50, rtol 1e-6
CUBE20 = Workload("cube20_ensemble16", 3, (20, 20, 20), t_final=50.0, rtol=1e-6, atol=1e-12,
                  batch=16, ee_g=1.0)

ALL = {w.name: w for w in (SQUARE32, SQUARE128, SQUARE128_ENSEMBLE, SQUARE64, CUBE20)}
