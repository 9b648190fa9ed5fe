// tg_stack.hip -- plain FP64 collision operator as one streaming pass over
// the band of K that yields both products.
//
// The operator needs u = K f and v = K^T f (f in energy order).  K's band is
// read once:
//
//   - K is stored once, in FP64, band-major: each 128-row panel keeps only the
//     32-column iterations its rows touch (exact FP64 nonzero extents of K's
//     rows), each iteration as [column 0..31][row 0..127], so one column of a
//     panel is 1 KB of contiguous memory;
//   - a workgroup takes an item of at most kStackChunk iterations of one
//     panel; lane l of a wave holds rows 2l and 2l+1 and streams columns
//     (16 bytes per lane, kStackUnroll loads in flight), the four waves take
//     the item's columns round robin;
//   - u: each lane accumulates its two rows over its columns, and the waves'
//     row sums are combined in LDS in wave order (one partial per item and
//     row);
//   - v: the column's 128 products K[r][c] f[r] are summed across the wave at
//     once (an 8-column butterfly: three exchange steps halve the columns a
//     lane holds, three more finish one column's sum), one partial per item
//     and column; a column is covered by the items of the few panels whose
//     band reaches it;
//   - the finish kernel adds a position's row partials (u) and column partials
//     (v, from a per-iteration list of the items covering it) in item order,
//     applies the Pauli-blocked assembly with the member's bath and scatters
//     to the caller's order.
// Every sum runs in a fixed order, so evaluations are bit-reproducible.
//
// It streams 8 bytes per stored band entry of K and writes 8 bytes per item
// row and per item column of partials (about 1% of the band).
#include <algorithm>
#include <mutex>
#include <vector>

#include "tg_internal.hpp"

namespace tg {

constexpr int kStackRows = 128;     // rows per panel
constexpr int kStackIt = 32;        // columns per iteration
constexpr int kStackChunk = 8;      // iterations per work item
constexpr int kStackThreads = 256;  // four waves
constexpr int kStackUnroll = 8;     // columns in flight per lane
static_assert(kStackUnroll == 8, "the column butterfly sums 8 columns at a time");

struct StackData {
    int64_t kt = 0;                 // n rounded up to whole panels
    int32_t npanel = 0;
    int32_t nitems = 0;
    int32_t maxit = 0;              // most iterations of a panel
    int64_t elems = 0;              // stored entries of S
    std::vector<int2> h_kr;         // per panel: [first, end) iteration
    double* Sb = nullptr;           // band-major S
    double* Sd = nullptr;           // dK/domega0 in the same layout (adjoint; null until needed)
    int64_t* off = nullptr;         // per panel: first entry
    int2* kr = nullptr;
    int4* items = nullptr;          // (panel, first iteration, end iteration, 0)
    int2* pitems = nullptr;         // per panel: [first item, end item)
    double* part = nullptr;         // nitems x 128: row partials
    double* vpart = nullptr;        // nitems x kStackChunk*32: column partials
    int32_t* cit_off = nullptr;     // per 32-column iteration: [cit_off[I], cit_off[I+1]) ...
    int32_t* cit = nullptr;         // ... offsets in vpart of the iteration's first column in the
                                    //     items covering it, in item order
};

static StackData* sd_of(tg_ctx* c) { return (StackData*)c->stack; }

// fs[p] = f[perm[p]]
// fs = f in energy order; with a hat stencil hs also the occupation factors
// xb = S f in the caller's order (thread p computes node p)
__global__ void stack_gather_kernel(const double* __restrict__ f, const int32_t* __restrict__ perm,
                                    int64_t n, double* __restrict__ fs, int dim,
                                    const uint32_t* __restrict__ hs, double* __restrict__ xb) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    fs[p] = f[perm[p]];
    if (hs) xb[p] = hat_value_dim(dim, hs, p, [&](int64_t j) { return f[j]; });
}

// block (panel, local iteration): entries [column 0..31][row 0..127] of the
// energy-ordered matrix whose entries k gives (K, or dK/domega0)
__global__ void stack_pack_kernel(KSrc k, const int32_t* __restrict__ perm, int64_t n,
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
        dst[t] = (C < n && R < n) ? k_entry(k, perm[R], perm[C]) : 0.0;
    }
}

// item partials: part[item][row] = sum over the item's columns of K[row][c] fs[c]
// and vpart[item][c] = sum over the panel's rows of K[row][c] fs[row]
__global__ void __launch_bounds__(kStackThreads)
stack_gemv_kernel(const double* __restrict__ Sb, const int64_t* __restrict__ off,
                  const int2* __restrict__ kr, const int4* __restrict__ items, int64_t n,
                  const double* __restrict__ fs, double* __restrict__ part,
                  double* __restrict__ vpart) {
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
    const int64_t row0 = (int64_t)panel * kStackRows + 2 * lane;
    const double fr0 = row0 < n ? fs[row0] : 0.0, fr1 = row0 + 1 < n ? fs[row0 + 1] : 0.0;
    double* vp = vpart + (int64_t)blockIdx.x * (kStackChunk * kStackIt);
    const int b5 = (lane >> 5) & 1, b4 = (lane >> 4) & 1, b3 = (lane >> 3) & 1;
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
        double t[kStackUnroll];
#pragma unroll
        for (int q = 0; q < kStackUnroll; ++q) {
            if (c0 + 4 * q < ncol) {
                a0 += k[q].x * fv[q];
                a1 += k[q].y * fv[q];
            }
            t[q] = k[q].x * fr0 + k[q].y * fr1;
        }
        // column sums over the wave: after the exchanges across lanes 32, 16
        // and 8 apart, lane l holds column 4 b5 + 2 b4 + b3 summed over its
        // group of 8 lanes; three more finish the sum over all 64
        double t4[4], t2[2];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            t4[j] = (b5 ? t[j + 4] : t[j]) + __shfl_xor(b5 ? t[j] : t[j + 4], 32, 64);
#pragma unroll
        for (int j = 0; j < 2; ++j)
            t2[j] = (b4 ? t4[j + 2] : t4[j]) + __shfl_xor(b4 ? t4[j] : t4[j + 2], 16, 64);
        double t1 = (b3 ? t2[1] : t2[0]) + __shfl_xor(b3 ? t2[0] : t2[1], 8, 64);
        t1 += __shfl_xor(t1, 4, 64);
        t1 += __shfl_xor(t1, 2, 64);
        t1 += __shfl_xor(t1, 1, 64);
        const int q = lane >> 3;
        if ((lane & 7) == 0 && c0 + 4 * q < ncol) vp[c0 + 4 * q] = t1;
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
constexpr int kStackFinishT = 64;

// u = (K fs)[p], v = (K^T fs)[p] from the item partials, in item order, four
// loads in flight at a time
__device__ inline void stack_sums(const double* __restrict__ part, const double* __restrict__ vpart,
                                  const int2* __restrict__ pitems, const int32_t* __restrict__ cit_off,
                                  const int32_t* __restrict__ cit, int64_t p, double& u, double& v) {
    u = 0.0;
    v = 0.0;
    const int2 iu = pitems[p / kStackRows];
    for (int i = iu.x; i < iu.y; i += 4) {
        double x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            x[j] = i + j < iu.y ? part[(int64_t)(i + j) * kStackRows + p % kStackRows] : 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (i + j < iu.y) u += x[j];
    }
    const int I = (int)(p / kStackIt);
    const int k1 = cit_off[I + 1];
    for (int k = cit_off[I]; k < k1; k += 4) {
        int o[4];
        double x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) o[j] = k + j < k1 ? cit[k + j] : -1;
#pragma unroll
        for (int j = 0; j < 4; ++j) x[j] = o[j] >= 0 ? vpart[(int64_t)o[j] + p % kStackIt] : 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (k + j < k1) v += x[j];
    }
}

// the products alone, in the caller's order: u[perm[p]], v[perm[p]]
__global__ void __launch_bounds__(256)
stack_products_kernel(const double* __restrict__ part, const double* __restrict__ vpart,
                      const int2* __restrict__ pitems, const int32_t* __restrict__ cit_off,
                      const int32_t* __restrict__ cit, int64_t n, const int32_t* __restrict__ perm,
                      double* __restrict__ u_out, double* __restrict__ v_out) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    double u, v;
    stack_sums(part, vpart, pitems, cit_off, cit, p, u, v);
    const int32_t node = perm[p];
    u_out[node] = u;
    v_out[node] = v;
}

// One position per thread: the operator with the member's bath (bath_p[0]
// when given, else bath_n) in the caller's order; the occupation factors of
// the row's node are xb (caller order; hat form) when given, else the nodal
// f; with dpart, per-block sums of V I, V w, V eps I and V eps w for the
// number-defect removal with weights w and the energy flux
// (conserve_fix_kernel): w = |f(1-f)|, or with xb the hat form's
// hat_defect_weight, which replaces the node's entry of xb.
__global__ void __launch_bounds__(kStackFinishT)
stack_finish_kernel(const double* __restrict__ part, const double* __restrict__ vpart,
                    const int2* __restrict__ pitems, const int32_t* __restrict__ cit_off,
                    const int32_t* __restrict__ cit, int64_t n,
                    const int32_t* __restrict__ perm, const double* __restrict__ fs,
                    double* __restrict__ xb,
                    const double* __restrict__ rk, const double* __restrict__ ck,
                    const double* __restrict__ V, const double* __restrict__ eps_s, double bath_n,
                    const double* __restrict__ bath_p, double* __restrict__ out,
                    double* __restrict__ dpart) {
    const int64_t p = blockIdx.x * (int64_t)kStackFinishT + threadIdx.x;
    const double nb = bath_p ? bath_p[0] : bath_n;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (p < n) {
        double u, v;
        stack_sums(part, vpart, pitems, cit_off, cit, p, u, v);
        const double np1 = nb + 1.0;
        const double fi = xb ? xb[perm[p]] : fs[p];
        const double wf = np1 * u + nb * v;
        const double wt1 = np1 * ck[p] + nb * rk[p];
        const double wtf = np1 * v + nb * u;
        const double Ii = ((1.0 - fi) * wf - fi * (wt1 - wtf)) / V[p];
        out[perm[p]] = Ii;
        acc[0] = V[p] * Ii;
        if (xb) {
            // the hat form's number-defect weight, kept for the removal in
            // the slot of this node's factor (read above, not read again)
            const double w = hat_defect_weight(fi, np1, nb, rk[p], ck[p], V[p]);
            xb[perm[p]] = w;
            acc[1] = V[p] * w;
        } else {
            acc[1] = V[p] * fabs(fi * (1.0 - fi));
        }
        acc[2] = eps_s[p] * acc[0];
        acc[3] = eps_s[p] * acc[1];
    }
    if (dpart) {
        __shared__ double red[4][kStackFinishT / 64];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_down(acc[q], o, 64);
            if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = acc[q];
        }
        __syncthreads();
        if (threadIdx.x < 4) {
            double a = 0.0;
            for (int w = 0; w < kStackFinishT / 64; ++w) a += red[threadIdx.x][w];
            dpart[4 * blockIdx.x + threadIdx.x] = a;
        }
    }
}

// builds K's band (energy order) from the exact row extents h_rf / h_rl:
// band-major entries (computed from the projection data), work items and
// the items covering each column iteration
tg_status stack_prepare(tg_ctx* c) {
    if (c->stack) return TG_OK;
    StackData* s = new StackData();
    c->stack = s;
    const int64_t n = c->n;
    s->kt = (n + kStackRows - 1) / kStackRows * kStackRows;
    s->npanel = (int32_t)(s->kt / kStackRows);
    hipStream_t st = c->stream;
    if ((int64_t)c->h_rf.size() != n || (int64_t)c->h_rl.size() != n)
        return set_error(c, TG_ERR_STATE, "row extents of K not computed");
    const std::vector<int32_t>& rf = c->h_rf;
    const std::vector<int32_t>& rl = c->h_rl;
    // panel extents in iterations
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
            if (R < n && rf[R] >= 0) {
                lo = std::min(lo, rf[R]);
                hi = std::max(hi, rl[R]);
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
    // items covering each column iteration, in item order
    const int64_t nit = (n + kStackIt - 1) / kStackIt;
    std::vector<int32_t> cit_off(nit + 1, 0), cit;
    for (const int4& q : items)
        for (int I = q.y; I < q.z; ++I) ++cit_off[I + 1];
    for (int64_t I = 0; I < nit; ++I) cit_off[I + 1] += cit_off[I];
    cit.assign(std::max<int32_t>(cit_off[nit], 1), 0);
    {
        std::vector<int32_t> fill(cit_off.begin(), cit_off.end() - 1);
        for (int32_t i = 0; i < (int32_t)items.size(); ++i)
            for (int I = items[i].y; I < items[i].z; ++I)
                cit[fill[I]++] = i * (kStackChunk * kStackIt) + (I - items[i].y) * kStackIt;
    }
    s->elems = elems;
    s->nitems = (int32_t)items.size();
    c->band_elems = elems;
    TG_TRY(dev_alloc(c, &s->Sb, (size_t)std::max<int64_t>(elems, 1)));
    TG_TRY(dev_alloc(c, &s->off, (size_t)s->npanel));
    TG_TRY(dev_alloc(c, &s->kr, (size_t)s->npanel));
    TG_TRY(dev_alloc(c, &s->items, items.size()));
    TG_TRY(dev_alloc(c, &s->pitems, (size_t)s->npanel));
    TG_TRY(dev_alloc(c, &s->part, (size_t)s->nitems * kStackRows));
    TG_TRY(dev_alloc(c, &s->vpart, (size_t)s->nitems * kStackChunk * kStackIt));
    TG_TRY(dev_alloc(c, &s->cit_off, (size_t)nit + 1));
    TG_TRY(dev_alloc(c, &s->cit, cit.size()));
    TG_HIP(c, hipMemcpyAsync(s->off, off.data(), sizeof(int64_t) * s->npanel, hipMemcpyHostToDevice, st));
    TG_HIP(c, hipMemcpyAsync(s->kr, s->h_kr.data(), sizeof(int2) * s->npanel, hipMemcpyHostToDevice, st));
    TG_HIP(c, hipMemcpyAsync(s->items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice, st));
    TG_HIP(c, hipMemcpyAsync(s->pitems, pitems.data(), sizeof(int2) * s->npanel, hipMemcpyHostToDevice, st));
    TG_HIP(c, hipMemcpyAsync(s->cit_off, cit_off.data(), sizeof(int32_t) * (nit + 1), hipMemcpyHostToDevice, st));
    TG_HIP(c, hipMemcpyAsync(s->cit, cit.data(), sizeof(int32_t) * cit.size(), hipMemcpyHostToDevice, st));
    s->maxit = maxit;
    if (maxit > 0) {
        hipLaunchKernelGGL(stack_pack_kernel, dim3(s->npanel, maxit), dim3(256), 0, st, ksrc_of(c, c->d_Q),
                           c->d_perm, n, s->kr, s->off, s->Sb);
        TG_CHECK_LAUNCH(c);
    }
    TG_HIP(c, hipStreamSynchronize(st));
    return TG_OK;
}

// dK/domega0 in the band layout of K (entries from the derivative table's
// convolved windows Qd)
tg_status stack_prepare_domega(tg_ctx* c, const double* d_Qd) {
    StackData* s = sd_of(c);
    if (!s) return set_error(c, TG_ERR_STATE, "band of K not prepared");
    if (s->Sd) return TG_OK;
    TG_TRY(dev_alloc(c, &s->Sd, (size_t)std::max<int64_t>(s->elems, 1)));
    if (s->maxit > 0) {
        hipLaunchKernelGGL(stack_pack_kernel, dim3(s->npanel, s->maxit), dim3(256), 0, c->stream, ksrc_of(c, d_Qd),
                           c->d_perm, c->n, s->kr, s->off, s->Sd);
        TG_CHECK_LAUNCH(c);
    }
    TG_HIP(c, hipStreamSynchronize(c->stream));
    return TG_OK;
}

void stack_destroy(tg_ctx* c) {
    StackData* s = sd_of(c);
    if (!s) return;
    dev_free(&s->Sb);
    dev_free(&s->Sd);
    dev_free(&s->off);
    dev_free(&s->kr);
    dev_free(&s->items);
    dev_free(&s->pitems);
    dev_free(&s->part);
    dev_free(&s->vpart);
    dev_free(&s->cit_off);
    dev_free(&s->cit);
    delete s;
    c->stack = nullptr;
}

// u = M f, v = M^T f (caller order in and out) from one pass over the band
// of M = K (or dK/domega0 with domega); kernel launches only
tg_status stack_products_device(tg_ctx* c, const double* d_f, double* d_u, double* d_v, bool domega) {
    StackData* s = sd_of(c);
    if (!s) return set_error(c, TG_ERR_STATE, "band of K not prepared");
    if (domega && !s->Sd) return set_error(c, TG_ERR_STATE, "band of dK/domega0 not prepared");
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(stack_gather_kernel, dim3(ceil_div(c->n, 256)), dim3(256), 0, st, d_f, c->d_perm,
                       c->n, c->d_fs, c->dim, (const uint32_t*)nullptr, (double*)nullptr);
    TG_CHECK_LAUNCH(c);
    hipLaunchKernelGGL(stack_gemv_kernel, dim3(s->nitems), dim3(kStackThreads), 0, st, domega ? s->Sd : s->Sb,
                       s->off, s->kr, s->items, c->n, c->d_fs, s->part, s->vpart);
    TG_CHECK_LAUNCH(c);
    hipLaunchKernelGGL(stack_products_kernel, dim3(ceil_div(c->n, 256)), dim3(256), 0, st, s->part, s->vpart,
                       s->pitems, s->cit_off, s->cit, c->n, c->d_perm, d_u, d_v);
    TG_CHECK_LAUNCH(c);
    return TG_OK;
}

// one plain FP64 evaluation (caller order in and out); kernel launches only.
// The bath is d_bath[0] (device) when given, else bath_n; d_flux (device,
// nullable) receives the e-ph electron energy flux sum_i V_i eps_i I_i.
tg_status stack_apply_device(tg_ctx* c, const double* d_f, double* d_I, double bath_n, int conserve,
                             const double* d_bath, double* d_flux) {
    StackData* s = sd_of(c);
    if (!s) return set_error(c, TG_ERR_STATE, "band of K not prepared");
    hipStream_t st = c->stream;
    const bool hat = c->occ == TG_OCC_HAT;
    hipLaunchKernelGGL(stack_gather_kernel, dim3(ceil_div(c->n, 256)), dim3(256), 0, st, d_f, c->d_perm,
                       c->n, c->d_fs, c->dim, hat ? (const uint32_t*)c->d_hat : nullptr, c->d_xb);
    TG_CHECK_LAUNCH(c);
    time_begin(c);
    hipLaunchKernelGGL(stack_gemv_kernel, dim3(s->nitems), dim3(kStackThreads), 0, st, s->Sb, s->off, s->kr,
                       s->items, c->n, c->d_fs, s->part, s->vpart);
    time_end(c);
    TG_CHECK_LAUNCH(c);
    const unsigned nb = ceil_div(c->n, kStackFinishT);
    double* dpart = nullptr;
    if (conserve || d_flux) {
        if ((size_t)c->red_cap < 4 * (size_t)nb) {
            dev_free(&c->d_red);
            TG_TRY(dev_alloc(c, &c->d_red, 4 * (size_t)nb));
            c->red_cap = 4 * (int32_t)nb;
        }
        dpart = c->d_red;
    }
    hipLaunchKernelGGL(stack_finish_kernel, dim3(nb), dim3(kStackFinishT), 0, st, s->part, s->vpart, s->pitems,
                       s->cit_off, s->cit, c->n, c->d_perm, c->d_fs, hat ? c->d_xb : nullptr,
                       c->d_rk_s, c->d_ck_s, c->d_V_s, c->d_eps_s, bath_n, d_bath, d_I, dpart);
    TG_CHECK_LAUNCH(c);
    if (dpart) TG_TRY(conserve_fix_device(c, dpart, (int)nb, d_f, hat ? c->d_xb : nullptr, d_I, conserve, d_flux));
    return TG_OK;
}

// ------------------------------------------------ allocation-free variant
// tg_collision_apply_noalloc runs the same kernels as the plain evaluation
// with their scratch (the gathered state, the row and column partials, the
// number-defect partials) in one file-scope device array instead of buffers
// the library allocates.  Each GPU holds its own copy of the array for the
// life of the process; the host reaches it only through the device address
// hipGetSymbolAddress returns (recorded per device when a context is created
// on it): the array's name in host code is not a device address.
constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;   // 16 MiB per device
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];
static double* g_scratch[kNoAllocMaxDevices] = {};
static std::mutex g_scratch_mu;

// called by tg_ctx_create with the context's device current
tg_status noalloc_init(tg_ctx* c) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c->device < 0 || c->device >= kNoAllocMaxDevices)
        return set_error(c, TG_ERR_ARG, "device %d beyond the %d the scratch table holds", c->device,
                         kNoAllocMaxDevices);
    if (g_scratch[c->device]) return TG_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(c, TG_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[c->device] = (double*)p;
    return TG_OK;
}

static size_t round256(size_t doubles) { return (doubles + 31) / 32 * 32; }

}  // namespace tg

using namespace tg;

extern "C" tg_status tg_collision_apply_noalloc(tg_ctx* c, const double* f, double* I, int64_t n, int32_t batch,
                                                const double* bath_n, int32_t flags) {
    if (!c || !f || !I) return TG_ERR_ARG;
    if (n != c->n)
        return set_error(c, TG_ERR_ARG, "n=%lld but the mesh has %lld nodes", (long long)n, (long long)c->n);
    if (batch < 1) return set_error(c, TG_ERR_ARG, "batch < 1");
    if (!(flags & TG_PTR_DEVICE))
        return set_error(c, TG_ERR_ARG, "the allocation-free evaluation takes device pointers (TG_PTR_DEVICE)");
    if (flags & (TG_PREC_MIXED | TG_TIME_KERNEL))
        return set_error(c, TG_ERR_ARG, "the allocation-free evaluation is the plain FP64 one (no TG_PREC_MIXED, "
                                        "no TG_TIME_KERNEL)");
    if (c->ee) return set_error(c, TG_ERR_STATE, "the allocation-free evaluation has no e-e channel");
    if (!c->assembled) return set_error(c, TG_ERR_STATE, "tg_kernel_assemble not called");
    StackData* sd = sd_of(c);
    if (!sd) return set_error(c, TG_ERR_STATE, "band of K not prepared");
    int dev = -1;
    TG_HIP(c, hipGetDevice(&dev));
    if (dev != c->device) return set_error(c, TG_ERR_STATE, "current device %d is not the context's %d", dev, c->device);
    double* base = g_scratch[dev];
    if (!base) return set_error(c, TG_ERR_STATE, "no scratch address recorded for device %d", dev);
    // scratch layout: gathered state | occupation factors | row partials |
    // column partials | block sums
    const unsigned nb = ceil_div(n, kStackFinishT);
    const bool hat = c->occ == TG_OCC_HAT;
    const size_t o_fs = 0;
    const size_t o_xb = o_fs + round256((size_t)n);
    const size_t o_part = o_xb + (hat ? round256((size_t)n) : 0);
    const size_t o_vpart = o_part + round256((size_t)sd->nitems * kStackRows);
    const size_t o_dpart = o_vpart + round256((size_t)sd->nitems * kStackChunk * kStackIt);
    const size_t need = o_dpart + round256(4 * (size_t)nb);
    if (need * sizeof(double) > kNoAllocScratchBytes)
        return set_error(c, TG_ERR_ARG, "this mesh needs %zu bytes of scratch; the allocation-free evaluation has %zu",
                         need * sizeof(double), kNoAllocScratchBytes);
    double* fs = base + o_fs;
    double* xb = hat ? base + o_xb : nullptr;
    double* part = base + o_part;
    double* vpart = base + o_vpart;
    double* dpart = base + o_dpart;
    const int conserve = (flags & TG_CONSERVE) ? 1 : 0;
    hipStream_t st = c->stream;
    for (int b = 0; b < batch; ++b) {
        const double* fb = f + (int64_t)b * n;
        double* Ib = I + (int64_t)b * n;
        const double nbath = bath_n ? bath_n[b] : c->bath_n;
        hipLaunchKernelGGL(stack_gather_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, fb, c->d_perm, n, fs,
                           c->dim, hat ? (const uint32_t*)c->d_hat : nullptr, xb);
        TG_CHECK_LAUNCH(c);
        hipLaunchKernelGGL(stack_gemv_kernel, dim3(sd->nitems), dim3(kStackThreads), 0, st, sd->Sb, sd->off, sd->kr,
                           sd->items, n, fs, part, vpart);
        TG_CHECK_LAUNCH(c);
        hipLaunchKernelGGL(stack_finish_kernel, dim3(nb), dim3(kStackFinishT), 0, st, part, vpart, sd->pitems,
                           sd->cit_off, sd->cit, n, c->d_perm, fs, xb, c->d_rk_s, c->d_ck_s,
                           c->d_V_s, c->d_eps_s, nbath, (const double*)nullptr, Ib,
                           conserve ? dpart : (double*)nullptr);
        TG_CHECK_LAUNCH(c);
        if (conserve) TG_TRY(conserve_fix_device(c, dpart, (int)nb, fb, xb, Ib, 1, nullptr));
    }
    return TG_OK;
}

namespace tg {
}  // namespace tg
