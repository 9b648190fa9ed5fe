"""Python front-end of libtortoise_gpu.so (the C ABI in include/tortoise_gpu.h).

The library is hand-written HIP for gfx950.  This module only binds it with
ctypes; there is no fallback: importing works without the library, but every
call raises if libtortoise_gpu.so is missing or no GPU is present.

    ctx = Context.create(workload)        # tg_ctx_create
    ctx.assemble()                        # tg_kernel_assemble
    I = ctx.apply(f)                      # tg_collision_apply
    f_end, dense, stats = ctx.dp5(f0, 0.0, t_final, dense_dt, rtol, atol)

Device pointers (torch tensors) are accepted by the *_device methods, which
pass TG_PTR_DEVICE.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import workloads  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtortoise_gpu.so")

TG_OK = 0
TG_ERR_ARG = -1
TG_ERR_HIP = -2
TG_ERR_NONFINITE = -3
TG_ERR_STEP_UNDERFLOW = -4
TG_ERR_STATE = -5
TG_ERR_NOMEM = -6

TG_PTR_DEVICE = 0x1
TG_PREC_MIXED = 0x2
TG_CONSERVE = 0x4
TG_OCC_NODAL = 0
TG_OCC_HAT = 1
OCCUPATIONS = {"nodal": TG_OCC_NODAL, "hat": TG_OCC_HAT}
TG_QUIET = 0x8
TG_TIME_KERNEL = 0x10
TG_BATH_SCHEDULE = 0x20

# every symbol include/tortoise_gpu.h declares
EXPORTS = [
    "tg_num_nodes", "tg_ctx_create", "tg_ctx_destroy", "tg_last_error", "tg_ctx_set_stream",
    "tg_ctx_stream", "tg_mesh_info", "tg_node_volumes", "tg_kernel_assemble",
    "tg_kernel_copy_rows", "tg_collision_apply", "tg_dp5_integrate", "tg_conservation",
    "tg_relative_error", "tg_kernel_stats_get", "tg_last_kernel_ms", "tg_dp5_replay",
    "tg_dp5_adjoint",
    "tg_tuple_create", "tg_tuple_destroy", "tg_tuple_last_error", "tg_tuple_point_rows",
    "tg_tuple_set_points", "tg_tuple_count", "tg_tuple_copy_indices", "tg_tuple_set_indices",
    "tg_tuple_save_indices", "tg_tuple_load_indices", "tg_tuple_set_amplitude",
    "tg_tuple_propagate", "tg_tuple_last_kernel_ms", "tg_tuple_stream", "tg_tuple_dp5",
    "tg_tuple_rates", "tg_set_bath_schedule", "tg_tuple_set_bath_schedule",
    "tg_collision_apply_noalloc", "tg_tuple4_create",
]


class TgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tortoise_gpu status {code}: {msg}")
        self.code = code


class ProblemDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("nsec", C.c_int32 * 3), ("origin", C.c_double * 3),
        ("extent", C.c_double * 3), ("eps", C.POINTER(C.c_double)), ("g", C.c_double),
        ("omega0", C.c_double), ("sigma", C.c_double), ("bath_n", C.c_double),
        ("de", C.c_double), ("cutoff", C.c_double), ("max_batch", C.c_int32),
        ("occupation", C.c_int32), ("ee_g", C.c_double), ("ee_kappa2", C.c_double),
        ("ee_sigma", C.c_double), ("ee_dw", C.c_double),
    ]


class KernelStats(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64), ("plain_band", C.c_int64),
                ("mixed_band", C.c_int64),
                ("mixed_splits", C.c_int32), ("mixed_width", C.c_int32),
                ("mixed_scale", C.c_int32), ("window", C.c_int32), ("kmax", C.c_double),
                ("mixed_tile", C.c_int32), ("mixed_fold", C.c_int32),
                ("mixed_waves", C.c_int32), ("mixed_depth", C.c_int32),
                ("mixed_tuned_ms", C.c_float), ("mixed_persist", C.c_int32),
                ("mixed_skip", C.c_int32), ("mixed_last_band", C.c_int64), ("mixed_cuts", C.c_int32),
                ("mixed_nnz", C.c_int64), ("mixed_last_nnz", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepStats(C.Structure):
    _fields_ = [("attempted", C.c_int64), ("accepted", C.c_int64), ("rejected", C.c_int64),
                ("evals", C.c_int64), ("failed", C.c_int32), ("n_dense", C.c_int32),
                ("bath_energy", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BathScheduleDesc(C.Structure):
    _fields_ = [("samples", C.POINTER(C.c_double)), ("n_samples", C.c_int32),
                ("batch", C.c_int32), ("t_start", C.c_double), ("dt", C.c_double),
                ("atol", C.c_double), ("rtol", C.c_double)]


def _schedule_desc(samples, t_start, dt, atol, rtol):
    """samples [ns] or [batch, ns] -> (descriptor, array kept alive)."""
    a = np.ascontiguousarray(samples, np.float64)
    if a.ndim == 1:
        a = a[None]
    d = BathScheduleDesc()
    d.samples = a.ctypes.data_as(C.POINTER(C.c_double))
    d.n_samples = a.shape[1]
    d.batch = a.shape[0]
    d.t_start, d.dt, d.atol, d.rtol = t_start, dt, atol, rtol
    return d, a


class TupleDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("nsec", C.c_int32 * 3), ("origin", C.c_double * 3),
                ("extent", C.c_double * 3), ("eps", C.POINTER(C.c_double)), ("omega0", C.c_double),
                ("amplitude", C.c_double), ("n_umklapp", C.c_int32), ("num_mc", C.c_int32),
                ("min_det", C.c_double), ("threshold", C.c_double), ("max_batch", C.c_int32),
                ("dynamic_phonons", C.c_int32), ("eps_ph", C.POINTER(C.c_double)),
                ("sides", (C.c_double * 3) * 3), ("zone", (C.c_double * 3) * 3)]


_lib = None


def library_path():
    return LIB_PATH


def load():
    """Loads libtortoise_gpu.so; raises if it has not been built.

    In a process that also uses torch on the GPU, import torch first: torch's
    libraries ask for `libamdhip64.so` by a name the ROCm runtime this library
    links (`libamdhip64.so.7`) does not answer to, so torch loaded second
    brings a second HIP runtime into the process, which then sees no device.
    Loaded first, torch's runtime is the one both use."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (hipcc --offload-arch=gfx950)")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.tg_num_nodes.restype = i64
    L.tg_num_nodes.argtypes = [i32, C.POINTER(i32)]
    L.tg_ctx_create.restype = i32
    L.tg_ctx_create.argtypes = [C.POINTER(ProblemDesc), C.POINTER(vp)]
    L.tg_ctx_destroy.restype = None
    L.tg_ctx_destroy.argtypes = [vp]
    L.tg_last_error.restype = C.c_char_p
    L.tg_last_error.argtypes = [vp]
    L.tg_ctx_set_stream.restype = i32
    L.tg_ctx_set_stream.argtypes = [vp, vp]
    L.tg_ctx_stream.restype = vp
    L.tg_ctx_stream.argtypes = [vp]
    L.tg_mesh_info.restype = i32
    L.tg_mesh_info.argtypes = [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(f64)]
    L.tg_node_volumes.restype = i32
    L.tg_node_volumes.argtypes = [vp, vp]
    L.tg_kernel_assemble.restype = i32
    L.tg_kernel_assemble.argtypes = [vp, i32]
    L.tg_kernel_stats_get.restype = i32
    L.tg_kernel_stats_get.argtypes = [vp, C.POINTER(KernelStats)]
    L.tg_last_kernel_ms.restype = i32
    L.tg_last_kernel_ms.argtypes = [vp, C.POINTER(f64)]
    L.tg_kernel_copy_rows.restype = i32
    L.tg_kernel_copy_rows.argtypes = [vp, i64, i64, vp]
    L.tg_collision_apply.restype = i32
    L.tg_collision_apply.argtypes = [vp, vp, vp, i64, i32, vp, i32]
    L.tg_collision_apply_noalloc.restype = i32
    L.tg_collision_apply_noalloc.argtypes = [vp, vp, vp, i64, i32, vp, i32]
    L.tg_dp5_integrate.restype = i32
    L.tg_dp5_integrate.argtypes = [vp, vp, i64, i32, f64, f64, f64, f64, f64, vp, i32, vp, i32,
                                   vp, vp, i32]
    L.tg_dp5_replay.restype = i32
    L.tg_dp5_replay.argtypes = [vp, vp, i64, i32, f64, f64, f64, vp, vp, i32, vp, i32, vp, i32]
    L.tg_dp5_adjoint.restype = i32
    L.tg_dp5_adjoint.argtypes = [vp, vp, i64, f64, f64, f64, vp, i32, i32, f64, f64, vp, vp, vp, i32]
    L.tg_conservation.restype = i32
    L.tg_conservation.argtypes = [vp, vp, i64, i32, vp, vp, i32]
    L.tg_relative_error.restype = i32
    L.tg_relative_error.argtypes = [vp, vp, vp, i64, i32, f64, f64, vp, i32]
    L.tg_tuple_create.restype = i32
    L.tg_tuple_create.argtypes = [C.POINTER(TupleDesc), C.POINTER(vp)]
    L.tg_tuple4_create.restype = i32
    L.tg_tuple4_create.argtypes = [C.POINTER(TupleDesc), f64, C.POINTER(vp)]
    L.tg_tuple_destroy.restype = None
    L.tg_tuple_destroy.argtypes = [vp]
    L.tg_tuple_last_error.restype = C.c_char_p
    L.tg_tuple_last_error.argtypes = [vp]
    L.tg_tuple_point_rows.restype = i32
    L.tg_tuple_point_rows.argtypes = [i32, i32]
    for name, args in (("tg_tuple_set_points", [vp, vp, i32]),
                       ("tg_tuple_count", [vp, C.POINTER(i64)]),
                       ("tg_tuple_copy_indices", [vp, vp, i64]),
                       ("tg_tuple_set_indices", [vp, vp, i64]),
                       ("tg_tuple_save_indices", [vp, C.c_char_p]),
                       ("tg_tuple_load_indices", [vp, C.c_char_p]),
                       ("tg_tuple_set_amplitude", [vp, f64]),
                       ("tg_tuple_propagate", [vp, vp, vp, vp, i64, i32, i32]),
                       ("tg_tuple_rates", [vp, vp, vp, vp, i64, i32, i32]),
                       ("tg_tuple_last_kernel_ms", [vp, C.POINTER(f64)])):
        fn = getattr(L, name)
        fn.restype = i32
        fn.argtypes = args
    L.tg_tuple_stream.restype = vp
    L.tg_tuple_stream.argtypes = [vp]
    L.tg_set_bath_schedule.restype = i32
    L.tg_set_bath_schedule.argtypes = [vp, C.POINTER(BathScheduleDesc)]
    L.tg_tuple_set_bath_schedule.restype = i32
    L.tg_tuple_set_bath_schedule.argtypes = [vp, C.POINTER(BathScheduleDesc)]
    L.tg_tuple_dp5.restype = i32
    L.tg_tuple_dp5.argtypes = [vp, vp, i64, i32, f64, f64, f64, f64, f64, vp, i32, vp, i32,
                               C.POINTER(StepStats), vp, i32]
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One tg_ctx: mesh + e-ph channel + assembled kernel on the current device."""

    def __init__(self, handle, n, workload=None):
        self._h = handle
        self.n = n
        self.workload = workload

    @classmethod
    def create(cls, workload, max_batch=None, eps=None, g=None, omega0=None, sigma=None,
               bath_n=None, de=0.0, cutoff=0.0, ee_g=None, ee_dw=0.0, occupation="nodal"):
        """occupation: the test node's occupation factors, "nodal" (f_i, the
        lumped W-form) or "hat" (the hat average of the P1 f; pair it with
        conserve=True for exact number conservation)."""
        if occupation not in OCCUPATIONS:
            raise ValueError(f"occupation must be one of {sorted(OCCUPATIONS)}, not {occupation!r}")
        L = load()
        w = workload
        e = np.ascontiguousarray(w.eps() if eps is None else eps, np.float64)
        d = ProblemDesc()
        d.dim = w.dim
        for q in range(3):
            d.nsec[q] = w.nsec[q] if q < w.dim else 1
            d.origin[q] = w.origin[q] if q < w.dim else 0.0
            d.extent[q] = w.extent[q] if q < w.dim else 1.0
        d.eps = e.ctypes.data_as(C.POINTER(C.c_double))
        d.g = w.g if g is None else g
        d.omega0 = w.omega0 if omega0 is None else omega0
        d.sigma = w.sigma if sigma is None else sigma
        d.bath_n = w.bath_n if bath_n is None else bath_n
        d.de = de
        d.cutoff = cutoff
        d.max_batch = max_batch or w.batch
        d.ee_g = w.ee_g if ee_g is None else ee_g
        d.ee_kappa2 = w.ee_kappa2
        d.ee_sigma = w.ee_sigma
        d.ee_dw = ee_dw
        d.occupation = OCCUPATIONS[occupation]
        h = C.c_void_p()
        rc = L.tg_ctx_create(C.byref(d), C.byref(h))
        if rc != TG_OK:
            msg = L.tg_last_error(h).decode() if h.value else "tg_ctx_create failed"
            if h.value:
                L.tg_ctx_destroy(h)
            raise TgError(rc, msg)
        ctx = cls(h, e.size, w)
        ctx.eps = e
        ctx.occupation = occupation
        return ctx

    def _check(self, rc):
        if rc != TG_OK:
            raise TgError(rc, load().tg_last_error(self._h).decode())

    def close(self):
        if self._h is not None and self._h.value:
            load().tg_ctx_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- setup
    def set_stream(self, stream_ptr):
        self._check(load().tg_ctx_set_stream(self._h, C.c_void_p(stream_ptr)))

    def stream(self):
        return load().tg_ctx_stream(self._h)

    def mesh_info(self):
        n, ne, v = C.c_int64(), C.c_int64(), C.c_double()
        self._check(load().tg_mesh_info(self._h, C.byref(n), C.byref(ne), C.byref(v)))
        return n.value, ne.value, v.value

    def node_volumes(self):
        out = np.zeros(self.n)
        self._check(load().tg_node_volumes(self._h, _p(out)))
        return out

    def assemble(self, mixed=False):
        self._check(load().tg_kernel_assemble(self._h, TG_PREC_MIXED if mixed else 0))

    def kernel_stats(self):
        st = KernelStats()
        self._check(load().tg_kernel_stats_get(self._h, C.byref(st)))
        return st.as_dict()

    def kernel_rows(self, row0, nrows):
        out = np.zeros((nrows, self.n))
        self._check(load().tg_kernel_copy_rows(self._h, row0, nrows, _p(out)))
        return out

    # ---- hot path, host arrays
    def apply(self, f, bath_n=None, mixed=False, conserve=False):
        f = np.ascontiguousarray(f, np.float64)
        batch = 1 if f.ndim == 1 else f.shape[1]
        fc = np.ascontiguousarray(f.T if f.ndim == 2 else f)  # member-major
        out = np.zeros_like(fc)
        bn = None if bath_n is None else np.ascontiguousarray(np.broadcast_to(bath_n, (batch,)),
                                                              np.float64)
        flags = (TG_PREC_MIXED if mixed else 0) | (TG_CONSERVE if conserve else 0)
        self._check(load().tg_collision_apply(self._h, _p(fc), _p(out), self.n, batch,
                                              None if bn is None else _p(bn), flags))
        return out if f.ndim == 1 else out.T

    def conservation(self, f):
        f = np.ascontiguousarray(f, np.float64)
        batch = 1 if f.ndim == 1 else f.shape[1]
        fc = np.ascontiguousarray(f.T if f.ndim == 2 else f)
        num = np.zeros(batch)
        en = np.zeros(batch)
        self._check(load().tg_conservation(self._h, _p(fc), self.n, batch, _p(num), _p(en), 0))
        return (num[0], en[0]) if f.ndim == 1 else (num, en)

    def relative_error(self, err, pop, rtol, atol):
        err = np.ascontiguousarray(err, np.float64)
        pop = np.ascontiguousarray(pop, np.float64)
        batch = 1 if err.ndim == 1 else err.shape[1]
        e = np.ascontiguousarray(err.T if err.ndim == 2 else err)
        p = np.ascontiguousarray(pop.T if pop.ndim == 2 else pop)
        out = np.zeros(batch)
        self._check(load().tg_relative_error(self._h, _p(e), _p(p), self.n, batch, rtol, atol,
                                             _p(out), 0))
        return out[0] if err.ndim == 1 else out

    def set_bath_schedule(self, samples=None, t_start=0.0, dt=1.0, atol=1e-3, rtol=1e-3):
        """Time-dependent bath n_b(t): samples [ns] or [batch, ns] at
        t_start + k dt (tg_set_bath_schedule); samples None clears it."""
        if samples is None:
            self._check(load().tg_set_bath_schedule(self._h, None))
            return
        d, keep = _schedule_desc(samples, t_start, dt, atol, rtol)
        self._check(load().tg_set_bath_schedule(self._h, C.byref(d)))
        del keep

    def dp5(self, f0, t0, t1, dense_dt, rtol, atol, bath_n=None, mixed=False, conserve=False,
            max_dense=0, quiet=True, raise_on_fail=True, h_cap=0, schedule=False):
        """Returns (f_last, dense[(batch,) max_dense, n], stats) and, with
        h_cap > 0, the accepted step sizes per member as a 4th element.
        schedule=True: the bath follows set_bath_schedule (TG_BATH_SCHEDULE).
        stats[b]["bath_energy"] is the bath's energy ledger."""
        f0 = np.ascontiguousarray(f0, np.float64)
        batch = 1 if f0.ndim == 1 else f0.shape[1]
        f = np.ascontiguousarray(f0.T if f0.ndim == 2 else f0).copy()
        dense = np.zeros((batch, max_dense, self.n)) if max_dense else None
        st = (StepStats * batch)()
        bn = None if bath_n is None else np.ascontiguousarray(np.broadcast_to(bath_n, (batch,)),
                                                              np.float64)
        hl = np.zeros((batch, h_cap)) if h_cap else None
        flags = ((TG_PREC_MIXED if mixed else 0) | (TG_CONSERVE if conserve else 0)
                 | (TG_QUIET if quiet else 0) | (TG_BATH_SCHEDULE if schedule else 0))
        rc = load().tg_dp5_integrate(self._h, _p(f), self.n, batch, t0, t1, dense_dt, rtol, atol,
                                     None if bn is None else _p(bn), flags,
                                     None if dense is None else _p(dense), max_dense, st,
                                     None if hl is None else _p(hl), h_cap)
        if rc != TG_OK and raise_on_fail:
            self._check(rc)
        stats = [s.as_dict() for s in st]
        steps = None
        if hl is not None:
            steps = [hl[b, : min(stats[b]["accepted"], h_cap)].copy() for b in range(batch)]
        if f0.ndim == 1:
            out = (f, (None if dense is None else dense[0]), stats[0])
            return out + ((steps[0],) if steps is not None else ())
        out = (f.T, dense, stats)
        return out + ((steps,) if steps is not None else ())

    def dp5_device(self, f_ptr, batch, t0, t1, dense_dt, rtol, atol, bath_n=None, mixed=False,
                   conserve=False, h_cap=0, schedule=False, raise_on_fail=True):
        """DP5 of device-resident states: f_ptr (device address, batch x n,
        member-major) is integrated in place (TG_PTR_DEVICE).  Returns the
        per-member stats and, with h_cap > 0, the accepted step sizes."""
        st = (StepStats * batch)()
        bn = None if bath_n is None else np.ascontiguousarray(np.broadcast_to(bath_n, (batch,)),
                                                              np.float64)
        hl = np.zeros((batch, h_cap)) if h_cap else None
        flags = (TG_PTR_DEVICE | TG_QUIET | (TG_PREC_MIXED if mixed else 0)
                 | (TG_CONSERVE if conserve else 0) | (TG_BATH_SCHEDULE if schedule else 0))
        rc = load().tg_dp5_integrate(self._h, C.c_void_p(f_ptr), self.n, batch, t0, t1, dense_dt, rtol,
                                     atol, None if bn is None else _p(bn), flags, None, 0, st,
                                     None if hl is None else _p(hl), h_cap)
        if rc != TG_OK and raise_on_fail:
            self._check(rc)
        stats = [s.as_dict() for s in st]
        if hl is None:
            return stats
        return stats, [hl[b, : min(stats[b]["accepted"], h_cap)].copy() for b in range(batch)]

    def dp5_replay(self, f0, t0, t1, dense_dt, steps, bath_n=None, mixed=False, conserve=False,
                   max_dense=0, schedule=False):
        """Fixed-step DP5 on per-member accepted step lists; returns (f_last, dense).
        schedule=True: the bath follows set_bath_schedule."""
        f0 = np.ascontiguousarray(f0, np.float64)
        batch = 1 if f0.ndim == 1 else f0.shape[1]
        if f0.ndim == 1:
            steps = [steps]
        f = np.ascontiguousarray(f0.T if f0.ndim == 2 else f0).copy()
        cap = max(1, max(len(s) for s in steps))
        H = np.zeros((batch, cap))
        ns = np.zeros(batch, np.int32)
        for b, s in enumerate(steps):
            H[b, : len(s)] = s
            ns[b] = len(s)
        dense = np.zeros((batch, max_dense, self.n)) if max_dense else None
        bn = None if bath_n is None else np.ascontiguousarray(np.broadcast_to(bath_n, (batch,)),
                                                              np.float64)
        flags = ((TG_PREC_MIXED if mixed else 0) | (TG_CONSERVE if conserve else 0)
                 | (TG_BATH_SCHEDULE if schedule else 0))
        self._check(load().tg_dp5_replay(self._h, _p(f), self.n, batch, t0, t1, dense_dt, _p(H),
                                         _p(ns), cap, None if bn is None else _p(bn), flags,
                                         None if dense is None else _p(dense), max_dense))
        if f0.ndim == 1:
            return f, (None if dense is None else dense[0])
        return f.T, dense

    def dp5_adjoint(self, f0, t0, t1, dense_dt, steps, dL_df, bath_n, dbath_domega0, ckpt=32,
                    conserve=False):
        """Gradients of L(f(t1)) through the fixed-step DP5: (dL/df0, dL/dg, dL/domega0).
        conserve: the forward map removes the number defect in every stage."""
        f0 = np.ascontiguousarray(f0, np.float64)
        steps = np.ascontiguousarray(steps, np.float64)
        g = np.ascontiguousarray(dL_df, np.float64)
        out = np.zeros_like(f0)
        par = np.zeros(2)
        self._check(load().tg_dp5_adjoint(self._h, _p(f0), self.n, t0, t1, dense_dt, _p(steps),
                                          steps.size, ckpt, bath_n, dbath_domega0, _p(g), _p(out),
                                          _p(par), TG_CONSERVE if conserve else 0))
        return out, par[0], par[1]

    # ---- hot path, device pointers (e.g. torch tensors' data_ptr())
    def last_kernel_ms(self):
        v = C.c_double()
        self._check(load().tg_last_kernel_ms(self._h, C.byref(v)))
        return v.value

    def apply_device(self, f_ptr, I_ptr, batch=1, bath_ptr=None, mixed=False, conserve=False,
                     bath_host=None, time_kernel=False):
        flags = (TG_PTR_DEVICE | (TG_PREC_MIXED if mixed else 0) | (TG_CONSERVE if conserve else 0)
                 | (TG_TIME_KERNEL if time_kernel else 0))
        bn = None
        if bath_host is not None:
            bn = np.ascontiguousarray(np.broadcast_to(bath_host, (batch,)), np.float64)
        self._check(load().tg_collision_apply(self._h, C.c_void_p(f_ptr), C.c_void_p(I_ptr),
                                              self.n, batch, None if bn is None else _p(bn),
                                              flags))


def collision_apply_noalloc(ctx, f_ptr, I_ptr, batch=1, bath_host=None, conserve=False,
                            flags=None):
    """tg_collision_apply_noalloc on device pointers; returns the status code
    (no exception, so callers can check refusals)."""
    L = load()
    bn = None
    if bath_host is not None:
        bn = np.ascontiguousarray(np.broadcast_to(bath_host, (batch,)), np.float64)
    fl = (TG_PTR_DEVICE | (TG_CONSERVE if conserve else 0)) if flags is None else flags
    return L.tg_collision_apply_noalloc(ctx._h, C.c_void_p(f_ptr), C.c_void_p(I_ptr), ctx.n, batch,
                                        None if bn is None else _p(bn), fl)


class TupleOperator:
    """The reference's own e-ph operator (tg_tuple_*): element tuples of the
    channel {el in, pn in, el out}, the MC delta-shell integral per tuple and
    the propagator on DG linear-form coefficients [batch][nel][dim+1]
    (ScatteringChannel.cpp:421-538,618-688; IntegratorCore.hpp:260-357)."""

    def __init__(self, handle, dim, nel, max_batch):
        self._h = handle
        self.dim = dim
        self.nel = nel
        self.max_batch = max_batch
        self.dynamic = False
        self.nlegs = 3

    @classmethod
    def create(cls, workload, n_umklapp=1, num_mc=50, min_det=1e-9, threshold=1e-9,
               max_batch=1, amplitude=None, eps=None, eps_ph=None, dynamic=False, origin=None,
               sides=None, zone=None, legs=3, kappa=0.0):
        """sides (dim x dim, row c = side c) with origin: a general
        parallelepiped mesh (the reference's Mesh over a Region); zone (row c
        = zone side c): the umklapp basis, by default the mesh's sides.
        Without sides: the workload's axis-aligned box.  legs=4: the channel
        {el in, el in, el out, el out} of the electron band
        (ScatteringChannel<NDim,4>, tg_tuple4_create) with functAmplitude 1
        (kappa <= 0) or 1/(|q|^2 + kappa^2)^2; tuples are [n][5] and
        propagate takes no bath."""
        if legs not in (3, 4):
            raise ValueError("legs must be 3 or 4")
        L = load()
        w = workload
        e = np.ascontiguousarray(w.eps() if eps is None else eps, np.float64)
        d = TupleDesc()
        d.dim = w.dim
        for q in range(3):
            d.nsec[q] = w.nsec[q] if q < w.dim else 1
            d.origin[q] = w.origin[q] if q < w.dim else 0.0
            d.extent[q] = w.extent[q] if q < w.dim else 1.0
        if sides is not None:
            S = np.asarray(sides, np.float64)
            Z = S if zone is None else np.asarray(zone, np.float64)
            for c in range(w.dim):
                d.origin[c] = float(origin[c]) if origin is not None else w.origin[c]
                for q in range(w.dim):
                    d.sides[c][q] = S[c][q]
                    d.zone[c][q] = Z[c][q]
        d.eps = e.ctypes.data_as(C.POINTER(C.c_double))
        d.omega0 = w.omega0
        d.amplitude = w.g ** 2 if amplitude is None else amplitude
        d.n_umklapp = n_umklapp
        d.num_mc = num_mc
        d.min_det = min_det
        d.threshold = threshold
        d.max_batch = max_batch
        d.dynamic_phonons = 1 if dynamic else 0
        ph = None
        if eps_ph is not None:
            ph = np.ascontiguousarray(eps_ph, np.float64)
            d.eps_ph = ph.ctypes.data_as(C.POINTER(C.c_double))
        h = C.c_void_p()
        if legs == 4:
            rc = L.tg_tuple4_create(C.byref(d), float(kappa), C.byref(h))
        else:
            rc = L.tg_tuple_create(C.byref(d), C.byref(h))
        if rc != TG_OK:
            msg = L.tg_tuple_last_error(h).decode() if h.value else "tg_tuple_create failed"
            if h.value:
                L.tg_tuple_destroy(h)
            raise TgError(rc, msg)
        nel = int(np.prod(w.nsec[:w.dim])) * (2 if w.dim == 2 else 6)
        op = cls(h, w.dim, nel, max_batch)
        op.rows = L.tg_tuple_point_rows(w.dim, legs)
        op.nlegs = legs
        op.num_mc = num_mc
        op.dynamic = bool(dynamic)
        return op

    def _check(self, rc):
        if rc != TG_OK:
            raise TgError(rc, load().tg_tuple_last_error(self._h).decode())

    def close(self):
        if self._h is not None and self._h.value:
            load().tg_tuple_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_points(self, pts):
        pts = np.ascontiguousarray(pts, np.float64)
        if pts.shape != (self.rows, self.num_mc):
            raise ValueError(f"points must be {self.rows} x {self.num_mc}")
        self._check(load().tg_tuple_set_points(self._h, _p(pts), 0))

    def count(self):
        n = C.c_int64()
        self._check(load().tg_tuple_count(self._h, C.byref(n)))
        return n.value

    def indices(self):
        n = self.count()
        idx = np.zeros((max(n, 1), self.nlegs + 1), np.int32)
        self._check(load().tg_tuple_copy_indices(self._h, _p(idx), max(n, 0)))
        return idx[:max(n, 0)]

    def set_indices(self, idx):
        idx = np.ascontiguousarray(idx, np.int32).reshape(-1, self.nlegs + 1)
        self._check(load().tg_tuple_set_indices(self._h, _p(idx), idx.shape[0]))

    def save_indices(self, path):
        self._check(load().tg_tuple_save_indices(self._h, os.fsencode(path)))

    def load_indices(self, path):
        self._check(load().tg_tuple_load_indices(self._h, os.fsencode(path)))

    def set_amplitude(self, amplitude):
        self._check(load().tg_tuple_set_amplitude(self._h, amplitude))

    def propagate(self, f, bath_n=None, time_kernel=False):
        """f: [nel, dim+1] or [batch, nel, dim+1] (with dynamical phonons
        [2, nel, dim+1] or [batch, 2, nel, dim+1]: electrons, then phonons);
        bath_n scalar or [batch] (unused with dynamical phonons)."""
        f = np.ascontiguousarray(f, np.float64)
        single = f.ndim == (3 if self.dynamic else 2)
        fb = f[None] if single else f
        batch = fb.shape[0]
        bn = None
        if bath_n is not None:
            bn = np.ascontiguousarray(np.broadcast_to(bath_n, (batch,)), np.float64)
        out = np.zeros_like(fb)
        self._check(load().tg_tuple_propagate(self._h, _p(fb), None if bn is None else _p(bn), _p(out),
                                              self.nel, batch, TG_TIME_KERNEL if time_kernel else 0))
        return out[0] if single else out

    def rates(self, f, bath_n=None):
        """ScatteringChannel::scatteringRates of the electron band: [(batch,) nel, dim+1]."""
        f = np.ascontiguousarray(f, np.float64)
        single = f.ndim == (3 if self.dynamic else 2)
        fb = f[None] if single else f
        batch = fb.shape[0]
        bn = None
        if bath_n is not None:
            bn = np.ascontiguousarray(np.broadcast_to(bath_n, (batch,)), np.float64)
        out = np.zeros((batch, self.nel, self.dim + 1))
        self._check(load().tg_tuple_rates(self._h, _p(fb), None if bn is None else _p(bn), _p(out),
                                          self.nel, batch, 0))
        return out[0] if single else out

    def propagate_device(self, f, bath_n, out, batch, time_kernel=False):
        """Device pointers (ints, e.g. torch data_ptr()) on the operator's device."""
        flags = TG_PTR_DEVICE | (TG_TIME_KERNEL if time_kernel else 0)
        self._check(load().tg_tuple_propagate(self._h, C.c_void_p(f), C.c_void_p(bath_n),
                                              C.c_void_p(out), self.nel, batch, flags))

    def last_kernel_ms(self):
        ms = C.c_double()
        self._check(load().tg_tuple_last_kernel_ms(self._h, C.byref(ms)))
        return ms.value

    def stream(self):
        return load().tg_tuple_stream(self._h)

    def set_bath_schedule(self, samples=None, t_start=0.0, dt=1.0, atol=1e-3, rtol=1e-3):
        """Time-dependent bath (tg_tuple_set_bath_schedule); samples None clears it."""
        if samples is None:
            self._check(load().tg_tuple_set_bath_schedule(self._h, None))
            return
        d, keep = _schedule_desc(samples, t_start, dt, atol, rtol)
        self._check(load().tg_tuple_set_bath_schedule(self._h, C.byref(d)))
        del keep

    def dp5(self, f0, t0, t1, dense_dt, rtol, atol, bath_n=None, max_dense=0, quiet=True,
            schedule=False, h_cap=0):
        """DP5(4) of DG states f0 [nel, dim+1] or [batch, nel, dim+1] (with
        dynamical phonons [(batch,) 2, nel, dim+1]); returns (f_last,
        dense [(batch,) max_dense, ...], stats) and, with h_cap > 0, the
        accepted step sizes per member.  schedule=True: the bath follows
        set_bath_schedule."""
        f = np.ascontiguousarray(f0, np.float64).copy()
        single = f.ndim == (3 if self.dynamic else 2)
        fb = f[None] if single else f
        batch = fb.shape[0]
        bn = None
        if bath_n is not None:
            bn = np.ascontiguousarray(np.broadcast_to(bath_n, (batch,)), np.float64)
        dense = np.zeros((batch, max_dense) + fb.shape[1:]) if max_dense else None
        st = (StepStats * batch)()
        flags = (TG_QUIET if quiet else 0) | (TG_BATH_SCHEDULE if schedule else 0)
        hl = np.zeros((batch, h_cap)) if h_cap else None
        rc = load().tg_tuple_dp5(self._h, _p(fb), self.nel, batch, t0, t1, dense_dt, rtol, atol,
                                 None if bn is None else _p(bn), flags,
                                 None if dense is None else _p(dense), max_dense, st,
                                 None if hl is None else _p(hl), h_cap)
        self._check(rc)
        stats = [{k: getattr(s, k) for k, _ in StepStats._fields_} for s in st]
        steps = None
        if hl is not None:
            steps = [hl[b, : min(stats[b]["accepted"], h_cap)].copy() for b in range(batch)]
        if single:
            out = (fb[0], (None if dense is None else dense[0][:stats[0]["n_dense"]]), stats[0])
            return out + ((steps[0],) if steps is not None else ())
        out = (fb, dense, stats)
        return out + ((steps,) if steps is not None else ())
