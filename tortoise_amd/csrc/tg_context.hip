// tg_context.hip -- context lifetime, mesh upload, errors, staging buffers,
// and the FP64 reductions of the boundary (conservation sums, error norm).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "tg_internal.hpp"
#include "tg_mesh.hpp"

namespace tg {

tg_status set_error(tg_ctx* c, tg_status code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

void time_begin(tg_ctx* c) {
    if (!c->time_kernel) return;
    if (!c->ev_k0) {
        hipEventCreate(&c->ev_k0);
        hipEventCreate(&c->ev_k1);
    }
    hipEventRecord(c->ev_k0, c->stream);
}

void time_end(tg_ctx* c) {
    if (!c->time_kernel || !c->ev_k1) return;
    hipEventRecord(c->ev_k1, c->stream);
    c->timed = true;
}

tg_status ensure_staging(tg_ctx* c, int32_t batch) {
    if (batch < 1 || batch > c->max_batch)
        return set_error(c, TG_ERR_ARG, "batch %d outside [1, max_batch=%d]", batch, c->max_batch);
    TG_TRY(dev_alloc(c, &c->d_f, (size_t)c->n * c->max_batch));
    TG_TRY(dev_alloc(c, &c->d_I, (size_t)c->n * c->max_batch));
    return TG_OK;
}

// ---------------------------------------------------------------- kernels
// Per-member FP64 sums: block b of member m writes partial sums; a second
// one-block pass sums the partials in index order (deterministic).
constexpr int kRedThreads = 256;

__device__ inline double block_sum(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) r += sh[q];
    __syncthreads();
    return r;
}

__global__ void conservation_partial_kernel(const double* __restrict__ f,
                                            const double* __restrict__ V,
                                            const double* __restrict__ eps, int64_t n,
                                            double* __restrict__ part, int nblk) {
    __shared__ double sh[kRedThreads / 64];
    const int m = blockIdx.y;
    const double* fm = f + (int64_t)m * n;
    double s = 0.0, e = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double w = V[i] * fm[i];
        s += w;
        e += w * eps[i];
    }
    double bs = block_sum(s, sh);
    double be = block_sum(e, sh);
    if (threadIdx.x == 0) {
        part[((int64_t)m * nblk + blockIdx.x) * 2 + 0] = bs;
        part[((int64_t)m * nblk + blockIdx.x) * 2 + 1] = be;
    }
}

__global__ void sum_partials_kernel(const double* __restrict__ part, int nblk, int ncomp,
                                    double* __restrict__ out) {
    // one thread per (member, component); partials summed in block order
    const int m = blockIdx.x;
    const int q = threadIdx.x;
    if (q >= ncomp) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[((int64_t)m * nblk + b) * ncomp + q];
    out[m * ncomp + q] = s;
}

// Band::relativeError (Band.cpp:87-101) for a fermion band, nodal values:
// r = err/scale with scale = rtol*min(|y|,|1-y|) + atol; per element the exact
// P1 integral vol/((d+1)(d+2)) (sum r_a^2 + (sum r_a)^2); the element volume
// cancels against the Nel*vol normalisation.
__global__ void relerr_partial_kernel(const double* __restrict__ err,
                                      const double* __restrict__ pop,
                                      const int32_t* __restrict__ conn, int dim, int64_t nel,
                                      int64_t n, double rtol, double atol,
                                      double* __restrict__ part, int nblk) {
    __shared__ double sh[kRedThreads / 64];
    const int m = blockIdx.y;
    const double* em = err + (int64_t)m * n;
    const double* pm = pop + (int64_t)m * n;
    double s = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nel;
         e += (int64_t)gridDim.x * blockDim.x) {
        double sr = 0.0, sr2 = 0.0;
        for (int a = 0; a <= dim; ++a) {
            int32_t i = conn[e * (dim + 1) + a];
            double y = pm[i];
            double ay = fabs(y), a1 = fabs(1.0 - y);
            double scale = rtol * (ay < a1 ? ay : a1) + atol;
            double r = em[i] / scale;
            sr += r;
            sr2 += r * r;
        }
        s += sr2 + sr * sr;
    }
    double bs = block_sum(s, sh);
    if (threadIdx.x == 0) part[(int64_t)m * nblk + blockIdx.x] = bs;
}

static int red_blocks(int64_t work) {
    int64_t b = (work + kRedThreads - 1) / kRedThreads;
    if (b > 256) b = 256;
    if (b < 1) b = 1;
    return (int)b;
}

static tg_status ensure_red(tg_ctx* c, size_t count) {
    if ((size_t)c->red_cap >= count) return TG_OK;
    dev_free(&c->d_red);
    TG_TRY(dev_alloc(c, &c->d_red, count));
    c->red_cap = (int32_t)count;
    return TG_OK;
}

tg_status conservation_device(tg_ctx* c, const double* d_f, int32_t batch, double* number,
                              double* energy) {
    int nblk = red_blocks(c->n);
    TG_TRY(ensure_red(c, (size_t)batch * nblk * 2 + (size_t)batch * 2));
    double* part = c->d_red;
    double* out = c->d_red + (size_t)batch * nblk * 2;
    hipLaunchKernelGGL(conservation_partial_kernel, dim3(nblk, batch), dim3(kRedThreads), 0,
                       c->stream, d_f, c->d_V, c->d_eps, c->n, part, nblk);
    TG_CHECK_LAUNCH(c);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(batch), dim3(64), 0, c->stream, part, nblk, 2,
                       out);
    TG_CHECK_LAUNCH(c);
    std::vector<double> h((size_t)batch * 2);
    TG_HIP(c, hipMemcpyAsync(h.data(), out, sizeof(double) * batch * 2, hipMemcpyDeviceToHost,
                             c->stream));
    TG_HIP(c, hipStreamSynchronize(c->stream));
    for (int b = 0; b < batch; ++b) {
        if (number) number[b] = h[2 * b];
        if (energy) energy[b] = h[2 * b + 1];
    }
    return TG_OK;
}

tg_status relative_error_device(tg_ctx* c, const double* d_err, const double* d_pop,
                                int32_t batch, double rtol, double atol, double* h_out) {
    int nblk = red_blocks(c->nel);
    TG_TRY(ensure_red(c, (size_t)batch * nblk + (size_t)batch));
    double* part = c->d_red;
    double* out = c->d_red + (size_t)batch * nblk;
    hipLaunchKernelGGL(relerr_partial_kernel, dim3(nblk, batch), dim3(kRedThreads), 0, c->stream,
                       d_err, d_pop, c->d_conn, c->dim, c->nel, c->n, rtol, atol, part, nblk);
    TG_CHECK_LAUNCH(c);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(batch), dim3(64), 0, c->stream, part, nblk, 1,
                       out);
    TG_CHECK_LAUNCH(c);
    std::vector<double> h(batch);
    TG_HIP(c, hipMemcpyAsync(h.data(), out, sizeof(double) * batch, hipMemcpyDeviceToHost,
                             c->stream));
    TG_HIP(c, hipStreamSynchronize(c->stream));
    const double denom = (double)((c->dim + 1) * (c->dim + 2)) * (double)c->nel;
    for (int b = 0; b < batch; ++b) h_out[b] = std::sqrt(h[b] / denom);
    return TG_OK;
}

int relerr_blocks(tg_ctx* c) { return red_blocks(c->nel); }

tg_status relerr_sums_device(tg_ctx* c, const double* d_err, const double* d_pop, int32_t batch,
                             double rtol, double atol, double* part, double* out) {
    const int nblk = red_blocks(c->nel);
    hipLaunchKernelGGL(relerr_partial_kernel, dim3(nblk, batch), dim3(kRedThreads), 0, c->stream,
                       d_err, d_pop, c->d_conn, c->dim, c->nel, c->n, rtol, atol, part, nblk);
    TG_CHECK_LAUNCH(c);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(batch), dim3(64), 0, c->stream, part, nblk, 1, out);
    TG_CHECK_LAUNCH(c);
    return TG_OK;
}

}  // namespace tg

using namespace tg;

// =================================================================== C ABI
extern "C" int64_t tg_num_nodes(int32_t dim, const int32_t* nsec) {
    if (!nsec || dim < 1 || dim > 3) return -1;
    int64_t n = 1;
    for (int d = 0; d < dim; ++d) {
        if (nsec[d] < 1) return -1;
        n *= (int64_t)(nsec[d] + 1);
    }
    return n;
}

extern "C" const char* tg_last_error(const tg_ctx* ctx) {
    return ctx ? ctx->err.c_str() : "null context";
}

extern "C" tg_status tg_ctx_create(const tg_problem_desc* d, tg_ctx** out) {
    if (!out) return TG_ERR_ARG;
    *out = nullptr;
    if (!d || (d->dim != 2 && d->dim != 3) || !d->eps) return TG_ERR_ARG;
    for (int q = 0; q < d->dim; ++q)
        if (d->nsec[q] < 1 || !(d->extent[q] > 0.0)) return TG_ERR_ARG;
    if (!(d->sigma > 0.0) || d->max_batch < 1) return TG_ERR_ARG;

    tg_ctx* c = new tg_ctx();
    c->dim = d->dim;
    for (int q = 0; q < 3; ++q) {
        c->nsec[q] = q < d->dim ? d->nsec[q] : 1;
        c->origin[q] = d->origin[q];
        c->extent[q] = q < d->dim ? d->extent[q] : 1.0;
    }
    c->n = tg_num_nodes(c->dim, c->nsec);
    c->nel = factorial(c->dim);
    for (int q = 0; q < c->dim; ++q) c->nel *= c->nsec[q];
    double v = 1.0;
    for (int q = 0; q < c->dim; ++q) v *= c->extent[q] / (double)c->nsec[q];
    c->vol = std::fabs(v / (double)factorial(c->dim));
    c->g = d->g;
    c->omega0 = d->omega0;
    c->sigma = d->sigma;
    c->bath_n = d->bath_n;
    c->de = d->de > 0.0 ? d->de : d->sigma / 8.0;
    c->cutoff = d->cutoff > 0.0 ? d->cutoff : 8.0;
    c->max_batch = d->max_batch;
    if (d->occupation != TG_OCC_NODAL && d->occupation != TG_OCC_HAT) {
        c->err = "occupation must be TG_OCC_NODAL or TG_OCC_HAT";
        *out = c;
        return TG_ERR_ARG;
    }
    c->occ = d->occupation;

    auto fail = [&](tg_status s) {
        *out = c;  // keep the context so the caller can read the error
        return s;
    };
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) return fail(set_error(c, TG_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e)));
    {
        const tg_status sn = noalloc_init(c);   // the device's scratch address, once per device
        if (sn != TG_OK) return fail(sn);
    }
    e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(set_error(c, TG_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)));
    c->stream = c->own_stream;

    // host mesh data
    MeshDims md{c->dim, {c->nsec[0], c->nsec[1], c->nsec[2]}};
    std::vector<int32_t> conn((size_t)c->nel * (c->dim + 1));
    std::vector<double> V(c->n, 0.0);
    for (int64_t el = 0; el < c->nel; ++el) {
        int64_t nodes[4];
        element_nodes(md, el, nodes);
        for (int a = 0; a <= c->dim; ++a) {
            conn[el * (c->dim + 1) + a] = (int32_t)nodes[a];
            V[nodes[a]] += c->vol / (double)(c->dim + 1);
        }
    }
    for (int64_t i = 0; i < c->n; ++i) {
        if (!std::isfinite(d->eps[i]))
            return fail(set_error(c, TG_ERR_NONFINITE, "eps[%lld] is not finite", (long long)i));
    }
    tg_status s;
    if ((s = dev_alloc(c, &c->d_eps, c->n)) != TG_OK) return fail(s);
    if ((s = dev_alloc(c, &c->d_V, c->n)) != TG_OK) return fail(s);
    if ((s = dev_alloc(c, &c->d_conn, conn.size())) != TG_OK) return fail(s);
    e = hipMemcpy(c->d_eps, d->eps, sizeof(double) * c->n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_V, V.data(), sizeof(double) * c->n, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(c->d_conn, conn.data(), sizeof(int32_t) * conn.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(set_error(c, TG_ERR_HIP, "upload: %s", hipGetErrorString(e)));
    if (c->occ == TG_OCC_HAT) {
        std::vector<uint32_t> hs;
        if (!hat_build(md, c->n, c->nel, conn, hs))
            return fail(set_error(c, TG_ERR_ARG, "mesh too large for the hat stencil"));
        if ((s = dev_alloc(c, &c->d_hat, hs.size())) != TG_OK) return fail(s);
        if ((s = dev_alloc(c, &c->d_xb, c->n)) != TG_OK) return fail(s);
        e = hipMemcpy(c->d_hat, hs.data(), sizeof(uint32_t) * hs.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(set_error(c, TG_ERR_HIP, "upload: %s", hipGetErrorString(e)));
    }

    // energy order of the nodes (stable: equal energies keep the node order)
    c->h_perm.resize(c->n);
    for (int64_t i = 0; i < c->n; ++i) c->h_perm[i] = (int32_t)i;
    std::stable_sort(c->h_perm.begin(), c->h_perm.end(),
                     [&](int32_t a, int32_t b) { return d->eps[a] < d->eps[b]; });
    std::vector<double> Vs(c->n), Es(c->n);
    for (int64_t p = 0; p < c->n; ++p) {
        Vs[p] = V[c->h_perm[p]];
        Es[p] = d->eps[c->h_perm[p]];
    }
    if ((s = dev_alloc(c, &c->d_perm, c->n)) != TG_OK) return fail(s);
    if ((s = dev_alloc(c, &c->d_V_s, c->n)) != TG_OK) return fail(s);
    if ((s = dev_alloc(c, &c->d_eps_s, c->n)) != TG_OK) return fail(s);
    if ((s = dev_alloc(c, &c->d_fs, c->n)) != TG_OK) return fail(s);
    if ((s = dev_alloc(c, &c->d_bath, (size_t)c->max_batch)) != TG_OK) return fail(s);
    e = hipMemcpy(c->d_perm, c->h_perm.data(), sizeof(int32_t) * c->n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_V_s, Vs.data(), sizeof(double) * c->n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_eps_s, Es.data(), sizeof(double) * c->n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(set_error(c, TG_ERR_HIP, "upload: %s", hipGetErrorString(e)));

    // energy grid (same rule as the CPU restatement: e0 = min - de)
    double lo = d->eps[0], hi = d->eps[0];
    for (int64_t i = 1; i < c->n; ++i) {
        lo = std::fmin(lo, d->eps[i]);
        hi = std::fmax(hi, d->eps[i]);
    }
    c->e0 = lo - c->de;
    c->nbins = (int32_t)std::floor((hi - c->e0) / c->de) + 3;
    if (d->ee_g != 0.0) {
        if (!(d->ee_sigma > 0.0) || !(d->ee_kappa2 > 0.0) || !std::isfinite(d->ee_g))
            return fail(set_error(c, TG_ERR_ARG, "e-e channel needs ee_sigma > 0 and ee_kappa2 > 0"));
        c->ee_g = d->ee_g;
        c->ee_kappa2 = d->ee_kappa2;
        c->ee_sigma = d->ee_sigma;
        c->ee_dw = d->ee_dw;
        if ((s = ee_prepare(c, d->eps)) != TG_OK) return fail(s);
    }
    *out = c;
    return TG_OK;
}

extern "C" void tg_ctx_destroy(tg_ctx* c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    mixed_destroy(c);
    stack_destroy(c);
    dp5_destroy(c);
    adjoint_destroy(c);
    ee_destroy(c);
    dev_free(&c->d_rkd);
    dev_free(&c->d_ckd);
    dev_free(&c->d_eps);
    dev_free(&c->d_V);
    dev_free(&c->d_conn);
    dev_free(&c->d_hat);
    dev_free(&c->d_xb);
    dev_free(&c->d_off);
    dev_free(&c->d_len);
    dev_free(&c->d_R);
    dev_free(&c->d_Q);
    dev_free(&c->d_gm);
    dev_free(&c->d_rk);
    dev_free(&c->d_ck);
    dev_free(&c->d_perm);
    dev_free(&c->d_V_s);
    dev_free(&c->d_eps_s);
    dev_free(&c->d_bath);
    dev_free(&c->d_bsched);
    dev_free(&c->d_rk_s);
    dev_free(&c->d_ck_s);
    dev_free(&c->d_fs);
    dev_free(&c->d_f);
    dev_free(&c->d_I);
    dev_free(&c->d_red);
    if (c->h_pinned) hipHostFree(c->h_pinned);
    if (c->ev_k0) hipEventDestroy(c->ev_k0);
    if (c->ev_k1) hipEventDestroy(c->ev_k1);
    if (c->own_stream) hipStreamDestroy(c->own_stream);
    delete c;
}

extern "C" tg_status tg_ctx_set_stream(tg_ctx* c, void* s) {
    if (!c) return TG_ERR_ARG;
    c->stream = s ? (hipStream_t)s : c->own_stream;
    return TG_OK;
}

extern "C" void* tg_ctx_stream(const tg_ctx* c) { return c ? (void*)c->stream : nullptr; }

extern "C" tg_status tg_mesh_info(const tg_ctx* c, int64_t* n_nodes, int64_t* n_elements,
                                  double* elem_volume) {
    if (!c) return TG_ERR_ARG;
    if (n_nodes) *n_nodes = c->n;
    if (n_elements) *n_elements = c->nel;
    if (elem_volume) *elem_volume = c->vol;
    return TG_OK;
}

extern "C" tg_status tg_node_volumes(tg_ctx* c, double* V) {
    if (!c || !V) return TG_ERR_ARG;
    TG_HIP(c, hipMemcpyAsync(V, c->d_V, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
    TG_HIP(c, hipStreamSynchronize(c->stream));
    return TG_OK;
}

extern "C" tg_status tg_conservation(tg_ctx* c, const double* f, int64_t n, int32_t batch,
                                     double* number, double* energy, int32_t flags) {
    if (!c || !f) return TG_ERR_ARG;
    if (n != c->n) return set_error(c, TG_ERR_ARG, "n=%lld, mesh has %lld nodes", (long long)n, (long long)c->n);
    if (batch < 1) return set_error(c, TG_ERR_ARG, "batch < 1");
    const double* d_f = f;
    if (!(flags & TG_PTR_DEVICE)) {
        TG_TRY(ensure_staging(c, batch));
        TG_HIP(c, hipMemcpyAsync(c->d_f, f, sizeof(double) * n * batch, hipMemcpyHostToDevice, c->stream));
        d_f = c->d_f;
    }
    return conservation_device(c, d_f, batch, number, energy);
}

extern "C" tg_status tg_relative_error(tg_ctx* c, const double* err, const double* pop, int64_t n,
                                       int32_t batch, double rtol, double atol, double* out,
                                       int32_t flags) {
    if (!c || !err || !pop || !out) return TG_ERR_ARG;
    if (n != c->n) return set_error(c, TG_ERR_ARG, "n mismatch");
    if (batch < 1) return set_error(c, TG_ERR_ARG, "batch < 1");
    const double* d_e = err;
    const double* d_p = pop;
    if (!(flags & TG_PTR_DEVICE)) {
        TG_TRY(ensure_staging(c, batch));
        TG_HIP(c, hipMemcpyAsync(c->d_f, err, sizeof(double) * n * batch, hipMemcpyHostToDevice, c->stream));
        TG_HIP(c, hipMemcpyAsync(c->d_I, pop, sizeof(double) * n * batch, hipMemcpyHostToDevice, c->stream));
        d_e = c->d_f;
        d_p = c->d_I;
    }
    return relative_error_device(c, d_e, d_p, batch, rtol, atol, out);
}

extern "C" tg_status tg_last_kernel_ms(tg_ctx* c, double* ms) {
    if (!c || !ms) return TG_ERR_ARG;
    if (!c->timed) return set_error(c, TG_ERR_STATE, "no apply with TG_TIME_KERNEL yet");
    TG_HIP(c, hipEventSynchronize(c->ev_k1));
    float v = 0.f;
    TG_HIP(c, hipEventElapsedTime(&v, c->ev_k0, c->ev_k1));
    *ms = v;
    return TG_OK;
}

extern "C" tg_status tg_set_bath_schedule(tg_ctx* c, const tg_bath_schedule* s) {
    if (!c) return TG_ERR_ARG;
    if (!s) {
        c->h_bsched.clear();
        c->bs_ns = c->bs_batch = 0;
        return TG_OK;
    }
    if (!s->samples || s->n_samples < 2 || s->batch < 1 || s->batch > c->max_batch)
        return set_error(c, TG_ERR_ARG, "bath schedule needs samples, n_samples >= 2 and batch in [1, max_batch]");
    if (!(s->dt > 0.0) || !std::isfinite(s->t_start) || !std::isfinite(s->dt) || !(s->atol >= 0.0) ||
        !(s->rtol >= 0.0))
        return set_error(c, TG_ERR_ARG, "bath schedule needs dt > 0, finite t_start and atol, rtol >= 0");
    if (!(s->atol > 0.0) && !(s->rtol > 0.0))
        return set_error(c, TG_ERR_ARG, "bath schedule: atol and rtol both 0 make the vetoer divide by 0");
    const size_t cnt = (size_t)s->batch * s->n_samples;
    for (size_t i = 0; i < cnt; ++i)
        if (!std::isfinite(s->samples[i]))
            return set_error(c, TG_ERR_NONFINITE, "bath schedule sample %zu is not finite", i);
    dev_free(&c->d_bsched);
    TG_TRY(dev_alloc(c, &c->d_bsched, cnt));
    TG_HIP(c, hipMemcpy(c->d_bsched, s->samples, sizeof(double) * cnt, hipMemcpyHostToDevice));
    c->h_bsched.assign(s->samples, s->samples + cnt);
    c->bs_ns = s->n_samples;
    c->bs_batch = s->batch;
    c->bs_t0 = s->t_start;
    c->bs_dt = s->dt;
    c->bs_atol = s->atol;
    c->bs_rtol = s->rtol;
    return TG_OK;
}
