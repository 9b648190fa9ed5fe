// tg_stack.hip -- plain FP64 collision operator as one streaming GEMV over
// the stacked band of [K; K^T].
//
// The products the operator needs are u = K f and v = K^T f (f in energy
// order).  Both are row products of the stacked matrix S = [K; K^T]
// (2n x n, energy order; the K^T half starts at the 128-aligned row kt), so
// every output row is one dot product and no partial sums cross rows:
//
//   - S is stored once, in FP64, band-major: each 128-row panel keeps only the
//     32-column iterations its rows touch (exact FP64 nonzero extents of K's
//     rows and columns), each iteration as [column 0..31][row 0..127], so one
//     column of a panel is 1 KB of contiguous memory;
//   - a workgroup takes an item of at most kStackChunk iterations of one
//     panel; lane l of a wave holds rows 2l and 2l+1 and streams columns
//     (16 bytes per lane, kStackUnroll loads in flight), the four waves take
//     the item's columns round robin, and their row sums are combined in LDS
//     in wave order;
//   - the finish kernel adds a position's item partials of its K panel (u)
//     and of its K^T panel (v) in item order, applies the Pauli-blocked
//     assembly with the member's bath and scatters to the caller's order.
// Every sum runs in a fixed order, so evaluations are bit-reproducible.
//
// It streams 8 bytes per stored band entry of S (about 1.6 times K's own
// band) and writes no per-column partials; the dual kernel (tg_plain.hip)
// reads K's band once but reduces column partials across workgroups, and
// stays in use for products with other matrices (dK/domega0).
#include <algorithm>
#include <vector>

#include "tg_internal.hpp"

namespace tg {

constexpr int kStackRows = 128;     // rows per panel
constexpr int kStackIt = 32;        // columns per iteration
constexpr int kStackChunk = 8;      // iterations per work item
constexpr int kStackThreads = 256;  // four waves
constexpr int kStackUnroll = 8;     // columns in flight per lane

struct StackData {
    int64_t kt = 0;                 // first row of the K^T half
    int32_t npanel = 0;
    int32_t nitems = 0;
    int64_t elems = 0;              // stored entries of S
    std::vector<int2> h_kr;         // per panel: [first, end) iteration
    double* Sb = nullptr;           // band-major S
    int64_t* off = nullptr;         // per panel: first entry
    int2* kr = nullptr;
    int4* items = nullptr;          // (panel, first iteration, end iteration, 0)
    int2* pitems = nullptr;         // per panel: [first item, end item)
    double* part = nullptr;         // nitems x 128
};

static StackData* sd_of(tg_ctx* c) { return (StackData*)c->stack; }

// first / last nonzero column of every row of K (-1 when empty)
__global__ void stack_row_extent_kernel(const double* __restrict__ K, int64_t n, int64_t ld,
                                        int32_t* __restrict__ first, int32_t* __restrict__ last) {
    __shared__ int32_t sf, sl;
    const int64_t i = blockIdx.x;
    if (threadIdx.x == 0) {
        sf = INT32_MAX;
        sl = -1;
    }
    __syncthreads();
    int32_t f = INT32_MAX, l = -1;
    for (int64_t j = threadIdx.x; j < n; j += blockDim.x)
        if (K[i * ld + j] != 0.0) {
            f = min(f, (int32_t)j);
            l = max(l, (int32_t)j);
        }
    atomicMin(&sf, f);
    atomicMax(&sl, l);
    __syncthreads();
    if (threadIdx.x == 0) {
        first[i] = sl < 0 ? -1 : sf;
        last[i] = sl;
    }
}

// fs[p] = f[perm[p]]
__global__ void stack_gather_kernel(const double* __restrict__ f, const int32_t* __restrict__ perm,
                                    int64_t n, double* __restrict__ fs) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p < n) fs[p] = f[perm[p]];
}

// first / last nonzero row of every column of K (-1 when empty)
__global__ void col_nz_extent_kernel(const double* __restrict__ K, int64_t n, int64_t ld,
                                     int32_t* __restrict__ first, int32_t* __restrict__ last) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    int32_t f = -1, l = -1;
    for (int64_t i = 0; i < n; ++i)
        if (K[i * ld + j] != 0.0) {
            if (f < 0) f = (int32_t)i;
            l = (int32_t)i;
        }
    first[j] = f;
    last[j] = l;
}

// block (panel, local iteration): entries [column 0..31][row 0..127] of S
__global__ void stack_pack_kernel(const double* __restrict__ K, int64_t n, int64_t ld, int64_t kt,
                                  const int2* __restrict__ kr, const int64_t* __restrict__ off,
                                  double* __restrict__ Sb) {
    const int panel = blockIdx.x;
    const int itl = blockIdx.y;
    const int2 r = kr[panel];
    if (itl >= r.y - r.x) return;
    double* dst = Sb + off[panel] + (int64_t)itl * kStackIt * kStackRows;
    for (int t = threadIdx.x; t < kStackIt * kStackRows; t += blockDim.x) {
        const int col = t / kStackRows, row = t % kStackRows;
        const int64_t R = (int64_t)panel * kStackRows + row;
        const int64_t C = (int64_t)(r.x + itl) * kStackIt + col;
        double x = 0.0;
        if (C < n) {
            if (R < n) x = K[R * ld + C];
            else if (R >= kt && R - kt < n) x = K[C * ld + (R - kt)];
        }
        dst[t] = x;
    }
}

// item partials: part[item][row] = sum over the item's columns of S[row][c] fs[c]
__global__ void __launch_bounds__(kStackThreads)
stack_gemv_kernel(const double* __restrict__ Sb, const int64_t* __restrict__ off,
                  const int2* __restrict__ kr, const int4* __restrict__ items, int64_t n,
                  const double* __restrict__ fs, double* __restrict__ part) {
    __shared__ double red[4][kStackRows];
    typedef double dbl2 __attribute__((ext_vector_type(2)));
    const int4 it = items[blockIdx.x];
    const int panel = it.x;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int2 r = kr[panel];
    const double* base = Sb + off[panel] + (int64_t)(it.y - r.x) * kStackIt * kStackRows;
    const int ncol = (it.z - it.y) * kStackIt;
    const int64_t col0 = (int64_t)it.y * kStackIt;
    double a0 = 0.0, a1 = 0.0;
    // columns wave, wave + 4, ...: kStackUnroll loads issued before use
    for (int c0 = wave; c0 < ncol; c0 += 4 * kStackUnroll) {
        dbl2 k[kStackUnroll];
        double fv[kStackUnroll];
#pragma unroll
        for (int q = 0; q < kStackUnroll; ++q) {
            const int c = min(c0 + 4 * q, ncol - 1);
            k[q] = *reinterpret_cast<const dbl2*>(base + (int64_t)c * kStackRows + 2 * lane);
            const int64_t C = col0 + c;
            fv[q] = C < n ? fs[C] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kStackUnroll; ++q) {
            if (c0 + 4 * q < ncol) {
                a0 += k[q].x * fv[q];
                a1 += k[q].y * fv[q];
            }
        }
    }
    red[wave][2 * lane] = a0;
    red[wave][2 * lane + 1] = a1;
    __syncthreads();
    if (threadIdx.x < kStackRows) {
        const int t = threadIdx.x;
        part[(int64_t)blockIdx.x * kStackRows + t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
    }
}

// per sorted position: u, v from the item partials (item order), Pauli
// assembly, scatter to the caller's order; block sums of V I and V |f(1-f)|
// for the number-defect removal
constexpr int kStackFinishT = 256;
__global__ void __launch_bounds__(kStackFinishT)
stack_finish_kernel(const double* __restrict__ part, const int2* __restrict__ pitems, int64_t n,
                    int64_t kt, const int32_t* __restrict__ perm, const double* __restrict__ fs,
                    const double* __restrict__ rk, const double* __restrict__ ck,
                    const double* __restrict__ V, double bath_n, double* __restrict__ out,
                    double* __restrict__ dpart) {
    const int64_t p = blockIdx.x * (int64_t)kStackFinishT + threadIdx.x;
    double dI = 0.0, dw = 0.0;
    if (p < n) {
        double u = 0.0, v = 0.0;
        const int2 iu = pitems[p / kStackRows];
        for (int i = iu.x; i < iu.y; ++i) u += part[(int64_t)i * kStackRows + p % kStackRows];
        const int64_t q = kt + p;
        const int2 iv = pitems[q / kStackRows];
        for (int i = iv.x; i < iv.y; ++i) v += part[(int64_t)i * kStackRows + q % kStackRows];
        const double np1 = bath_n + 1.0;
        const double fi = fs[p];
        const double wf = np1 * u + bath_n * v;
        const double wt1 = np1 * ck[p] + bath_n * rk[p];
        const double wtf = np1 * v + bath_n * u;
        const double Ii = ((1.0 - fi) * wf - fi * (wt1 - wtf)) / V[p];
        out[perm[p]] = Ii;
        dI = V[p] * Ii;
        dw = V[p] * fabs(fi * (1.0 - fi));
    }
    if (dpart) {
        __shared__ double red[2][kStackFinishT / 64];
        for (int o = 32; o > 0; o >>= 1) {
            dI += __shfl_down(dI, o, 64);
            dw += __shfl_down(dw, o, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            red[0][threadIdx.x >> 6] = dI;
            red[1][threadIdx.x >> 6] = dw;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = 0.0, b = 0.0;
            for (int w = 0; w < kStackFinishT / 64; ++w) {
                a += red[0][w];
                b += red[1][w];
            }
            dpart[2 * blockIdx.x] = a;
            dpart[2 * blockIdx.x + 1] = b;
        }
    }
}

// builds the stacked band from the assembled K (energy order): extents,
// band-major copy, work items
tg_status stack_prepare(tg_ctx* c) {
    if (c->stack) return TG_OK;
    StackData* s = new StackData();
    c->stack = s;
    const int64_t n = c->n;
    s->kt = (n + kStackRows - 1) / kStackRows * kStackRows;
    s->npanel = (int32_t)(2 * s->kt / kStackRows);
    hipStream_t st = c->stream;
    int32_t *d_rf = nullptr, *d_rl = nullptr, *d_cf = nullptr, *d_cl = nullptr;
    TG_TRY(dev_alloc(c, &d_rf, (size_t)n));
    TG_TRY(dev_alloc(c, &d_rl, (size_t)n));
    TG_TRY(dev_alloc(c, &d_cf, (size_t)n));
    TG_TRY(dev_alloc(c, &d_cl, (size_t)n));
    hipLaunchKernelGGL(stack_row_extent_kernel, dim3(n), dim3(256), 0, st, c->d_K, n, c->ld, d_rf, d_rl);
    hipLaunchKernelGGL(col_nz_extent_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, c->d_K, n, c->ld,
                       d_cf, d_cl);
    TG_CHECK_LAUNCH(c);
    std::vector<int32_t> rf(n), rl(n), cf(n), cl(n);
    TG_HIP(c, hipMemcpyAsync(rf.data(), d_rf, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipMemcpyAsync(rl.data(), d_rl, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipMemcpyAsync(cf.data(), d_cf, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipMemcpyAsync(cl.data(), d_cl, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipStreamSynchronize(st));
    dev_free(&d_rf);
    dev_free(&d_rl);
    dev_free(&d_cf);
    dev_free(&d_cl);
    // panel extents in iterations; a K^T row kt + j is column j of K
    s->h_kr.assign(s->npanel, int2{0, 0});
    std::vector<int64_t> off(s->npanel, 0);
    std::vector<int4> items;
    std::vector<int2> pitems(s->npanel, int2{0, 0});
    int64_t elems = 0;
    int32_t maxit = 0;
    for (int pn = 0; pn < s->npanel; ++pn) {
        int32_t lo = INT32_MAX, hi = -1;
        for (int r = 0; r < kStackRows; ++r) {
            const int64_t R = (int64_t)pn * kStackRows + r;
            int32_t f = -1, l = -1;
            if (R < n) {
                f = rf[R];
                l = rl[R];
            } else if (R >= s->kt && R - s->kt < n) {
                f = cf[R - s->kt];
                l = cl[R - s->kt];
            }
            if (f >= 0) {
                lo = std::min(lo, f);
                hi = std::max(hi, l);
            }
        }
        int2 k{0, 0};
        if (hi >= 0) k = int2{lo / kStackIt, hi / kStackIt + 1};
        s->h_kr[pn] = k;
        off[pn] = elems;
        elems += (int64_t)(k.y - k.x) * kStackIt * kStackRows;
        maxit = std::max(maxit, k.y - k.x);
        pitems[pn].x = (int32_t)items.size();
        for (int a = k.x; a < k.y; a += kStackChunk)
            items.push_back(int4{pn, a, std::min(a + kStackChunk, k.y), 0});
        pitems[pn].y = (int32_t)items.size();
    }
    if (items.empty()) items.push_back(int4{0, 0, 0, 0});
    s->elems = elems;
    s->nitems = (int32_t)items.size();
    c->band_elems = elems;
    TG_TRY(dev_alloc(c, &s->Sb, (size_t)std::max<int64_t>(elems, 1)));
    TG_TRY(dev_alloc(c, &s->off, (size_t)s->npanel));
    TG_TRY(dev_alloc(c, &s->kr, (size_t)s->npanel));
    TG_TRY(dev_alloc(c, &s->items, items.size()));
    TG_TRY(dev_alloc(c, &s->pitems, (size_t)s->npanel));
    TG_TRY(dev_alloc(c, &s->part, (size_t)s->nitems * kStackRows));
    TG_HIP(c, hipMemcpyAsync(s->off, off.data(), sizeof(int64_t) * s->npanel, hipMemcpyHostToDevice, st));
    TG_HIP(c, hipMemcpyAsync(s->kr, s->h_kr.data(), sizeof(int2) * s->npanel, hipMemcpyHostToDevice, st));
    TG_HIP(c, hipMemcpyAsync(s->items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice, st));
    TG_HIP(c, hipMemcpyAsync(s->pitems, pitems.data(), sizeof(int2) * s->npanel, hipMemcpyHostToDevice, st));
    if (maxit > 0) {
        hipLaunchKernelGGL(stack_pack_kernel, dim3(s->npanel, maxit), dim3(256), 0, st, c->d_K, n, c->ld,
                           s->kt, s->kr, s->off, s->Sb);
        TG_CHECK_LAUNCH(c);
    }
    TG_HIP(c, hipStreamSynchronize(st));
    return TG_OK;
}

void stack_destroy(tg_ctx* c) {
    StackData* s = sd_of(c);
    if (!s) return;
    dev_free(&s->Sb);
    dev_free(&s->off);
    dev_free(&s->kr);
    dev_free(&s->items);
    dev_free(&s->pitems);
    dev_free(&s->part);
    delete s;
    c->stack = nullptr;
}

// one plain FP64 evaluation (caller order in and out); kernel launches only
tg_status stack_apply_device(tg_ctx* c, const double* d_f, double* d_I, double bath_n, int conserve) {
    StackData* s = sd_of(c);
    if (!s) return set_error(c, TG_ERR_STATE, "stacked band not prepared");
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(stack_gather_kernel, dim3(ceil_div(c->n, 256)), dim3(256), 0, st, d_f, c->d_perm,
                       c->n, c->d_fs);
    TG_CHECK_LAUNCH(c);
    time_begin(c);
    hipLaunchKernelGGL(stack_gemv_kernel, dim3(s->nitems), dim3(kStackThreads), 0, st, s->Sb, s->off, s->kr,
                       s->items, c->n, c->d_fs, s->part);
    time_end(c);
    TG_CHECK_LAUNCH(c);
    const unsigned nb = ceil_div(c->n, kStackFinishT);
    double* dpart = nullptr;
    if (conserve) {
        if ((size_t)c->red_cap < 2 * (size_t)nb) {
            dev_free(&c->d_red);
            TG_TRY(dev_alloc(c, &c->d_red, 2 * (size_t)nb));
            c->red_cap = 2 * (int32_t)nb;
        }
        dpart = c->d_red;
    }
    hipLaunchKernelGGL(stack_finish_kernel, dim3(nb), dim3(kStackFinishT), 0, st, s->part, s->pitems, c->n,
                       s->kt, c->d_perm, c->d_fs, c->d_rk_s, c->d_ck_s, c->d_V_s, bath_n, d_I, dpart);
    TG_CHECK_LAUNCH(c);
    if (conserve) TG_TRY(conserve_fix_device(c, dpart, (int)nb, d_f, d_I));
    return TG_OK;
}

}  // namespace tg
