// tg_mixed.hip -- batched mixed-precision collision operator on MFMA.
//
// For B members at once the operator needs U = K F and V = K^T F (F = n x B).
// Both come from one GEMM with the stacked matrix A = [K; K^T] (2n x n):
//
//   C = A F,  A = 2^-sa (A_hi + A_lo),  F = 2^-sf (F_hi + F_lo)   (fp16 parts)
//   C ~ 2^-(sa+sf) (A_hi F_hi + A_hi F_lo + A_lo F_hi)
//
// three v_mfma_f32_32x32x16_f16 products per tile, FP32 accumulation inside
// a K-split, FP64 sum of the splits and FP64 Pauli assembly in the epilogue
// (per-member bath occupation, number-defect removal).  The row and column
// sums K 1, K^T 1 used by the epilogue are exact FP64 sums of the *split*
// matrix, so the split error cancels in K^T 1 - K^T f for Pauli-blocked nodes.
//
// Node order: K is stored in energy order (see tg_assemble.hip), which makes
// the emission kernel a band: every 128-row panel of A has a contiguous
// column range outside which it is exactly zero, and the GEMM only streams
// that range.  Inputs and outputs stay in the caller's node order; the
// F-packing kernel gathers through the permutation and the epilogue scatters
// back.
//
// Packed layouts (16-byte units of eight fp16 values along the reduction k):
//   A_{hi,lo}[kg][row]   kg = k/8, row in [0, rows_pad)
//   F_{hi,lo}[kg][col]   col = member in [0, NT)
// A workgroup's LDS tile and every MFMA fragment are contiguous 16-byte
// reads in these layouts (lanes 0-31: k-group 2s, lanes 32-63: k-group 2s+1).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "tg_internal.hpp"

namespace tg {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kMT = 128;        // rows of A per workgroup
constexpr int kKS = 32;         // reduction depth per iteration (two MFMA k-steps)
constexpr int kKG = kKS / 8;    // 16-byte k-groups per iteration
constexpr int kMixThreads = 256;
constexpr int kSF = 12;         // F scale exponent: f in [-8, 8] stays inside fp16

struct MixedData {
    int64_t rows = 0;          // 2n
    int64_t kt = 0;            // first row of the K^T half (n rounded up to kMT)
    int64_t rows_pad = 0;      // 2 kt
    int64_t kpad = 0;          // n rounded up to kKS
    int nt = 256;              // members per GEMM tile (padded batch)
    int sa = 0;                // A scale exponent
    int npanel = 0;
    uint4* A_hi = nullptr;
    uint4* A_lo = nullptr;
    uint4* F_hi = nullptr;
    uint4* F_lo = nullptr;
    uint4* G_hi = nullptr;     // operand g = 1 - f
    uint4* G_lo = nullptr;
    uint8_t* useg = nullptr;   // npanel x nt operand choice
    int2* krange = nullptr;    // per panel [kit_begin, kit_end) in iterations
    std::vector<int2> h_krange;
    const int32_t* perm = nullptr;  // sorted position -> node (owned by the context)
    double* rk_s = nullptr;    // split-matrix K 1 (sorted order)
    double* ck_s = nullptr;    // split-matrix K^T 1 (sorted order)
    const double* V_s = nullptr;    // node volumes (sorted order, owned by the context)
    float* P = nullptr;        // nsplit x nt x rows_pad partial sums
    int nsplit = 1;
    size_t P_cap = 0;
    double* bath = nullptr;    // per-member bath occupation (device)
    double* part = nullptr;    // per-member conservation partials
    int part_cap = 0;
    int64_t work_items = 0;
};

static MixedData* md_of(tg_ctx* c) { return (MixedData*)c->mixed; }

// ------------------------------------------------------------------ packing
__device__ inline void split_fp16(double x, double scale, _Float16& hi, _Float16& lo) {
    const double s = x * scale;
    hi = (_Float16)s;
    lo = (_Float16)(s - (double)hi);
}

// A[r][q] from K stored in energy order: r < n -> K[r][q]; the K^T half
// starts at the panel-aligned row kt: kt <= r < kt + n -> K[q][r-kt]
__global__ void pack_A_kernel(const double* __restrict__ K, int64_t n, int64_t ld, int64_t kt,
                              int64_t rows_pad, int64_t kgroups, double scale,
                              uint4* __restrict__ Ahi, uint4* __restrict__ Alo) {
    const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t kg = blockIdx.y;
    if (row >= rows_pad || kg >= kgroups) return;
    union { half8 v; uint4 u; } h, l;
    for (int q = 0; q < 8; ++q) {
        const int64_t col = kg * 8 + q;
        double x = 0.0;
        if (col < n) {
            if (row < n) x = K[row * ld + col];
            else if (row >= kt && row < kt + n) x = K[col * ld + (row - kt)];
        }
        _Float16 a, b;
        split_fp16(x, scale, a, b);
        h.v[q] = a;
        l.v[q] = b;
    }
    Ahi[kg * rows_pad + row] = h.u;
    Alo[kg * rows_pad + row] = l.u;
}

// first / last nonzero column (sorted order) of every A row; -1 when empty
__global__ void row_extent_kernel(const uint4* __restrict__ Ahi, const uint4* __restrict__ Alo,
                                  int64_t rows_pad, int64_t kgroups, int32_t* __restrict__ first,
                                  int32_t* __restrict__ last) {
    const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (row >= rows_pad) return;
    int32_t f = -1, l = -1;
    for (int64_t kg = 0; kg < kgroups; ++kg) {
        uint4 a = Ahi[kg * rows_pad + row], b = Alo[kg * rows_pad + row];
        if ((a.x | a.y | a.z | a.w | b.x | b.y | b.z | b.w) != 0u) {
            if (f < 0) f = (int32_t)kg;
            l = (int32_t)kg;
        }
    }
    first[row] = f;
    last[row] = l;
}

// exact FP64 row sums of the split matrix: rows [0, n) give K 1, rows
// [kt, kt + n) give K^T 1 (both in sorted order)
__global__ void split_rowsum_kernel(const uint4* __restrict__ Ahi, const uint4* __restrict__ Alo,
                                    int64_t rows_pad, int64_t kgroups, int64_t n, int64_t kt,
                                    double inv_scale, double* __restrict__ rk,
                                    double* __restrict__ ck) {
    const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (row >= rows_pad || (row >= n && row < kt) || row >= kt + n) return;
    double s = 0.0;
    for (int64_t kg = 0; kg < kgroups; ++kg) {
        union { half8 v; uint4 u; } h, l;
        h.u = Ahi[kg * rows_pad + row];
        l.u = Alo[kg * rows_pad + row];
        for (int q = 0; q < 8; ++q) s += (double)h.v[q] + (double)l.v[q];
    }
    s *= inv_scale;
    if (row < n) rk[row] = s;
    else ck[row - kt] = s;
}

// F (caller order, member-major f[b*n + i], FP64) -> packed sorted fp16 hi/lo
// of both operands z = f and z = g = 1 - f (the GEMM streams one of them per
// (panel, member), see choose_operand_kernel)
__global__ void pack_F_kernel(const double* __restrict__ f, int64_t n, int B,
                              const int32_t* __restrict__ perm, int64_t kgroups, int nt,
                              uint4* __restrict__ Fhi, uint4* __restrict__ Flo,
                              uint4* __restrict__ Ghi, uint4* __restrict__ Glo) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;  // member
    const int64_t kg = blockIdx.y;
    if (col >= nt || kg >= kgroups) return;
    union { half8 v; uint4 u; } fh, fl, gh, gl;
    const double scale = (double)(1 << kSF);
    for (int q = 0; q < 8; ++q) {
        const int64_t p = kg * 8 + q;
        double x = 0.0, y = 0.0;
        if (col < B && p < n) {
            x = f[(int64_t)col * n + perm[p]];
            y = 1.0 - x;
        }
        _Float16 a, b;
        split_fp16(x, scale, a, b);
        fh.v[q] = a;
        fl.v[q] = b;
        split_fp16(y, scale, a, b);
        gh.v[q] = a;
        gl.v[q] = b;
    }
    Fhi[kg * nt + col] = fh.u;
    Flo[kg * nt + col] = fl.u;
    Ghi[kg * nt + col] = gh.u;
    Glo[kg * nt + col] = gl.u;
}

// Operand of (panel, member): g = 1 - f when the panel's middle node is in
// the filled part of the member's distribution (f > 1/2), else f.  The
// products the epilogue needs come back as K z and K 1 - K z, so the
// Pauli-blocked out-scattering K^T(1 - f) of filled nodes is a sum of small
// terms instead of the difference of two large ones.
__global__ void choose_operand_kernel(const double* __restrict__ f, int64_t n, int64_t kt, int B,
                                      int npanel, const int32_t* __restrict__ perm,
                                      uint8_t* __restrict__ useg, int nt) {
    const int pn = blockIdx.x;
    const int b = threadIdx.x;
    if (pn >= npanel || b >= nt) return;
    uint8_t v = 0;
    if (b < B) {
        // K rows and K^T rows index the same sorted nodes
        const int64_t row = (int64_t)pn * kMT + kMT / 2;
        int64_t p = row < kt ? row : row - kt;
        if (p > n - 1) p = n - 1;
        v = f[(int64_t)b * n + perm[p]] > 0.5 ? 1 : 0;
    }
    useg[(int64_t)pn * nt + b] = v;
}

// --------------------------------------------------------------------- GEMM
// Workgroup: 128 rows x NT members, four waves in a 2 x 2 grid; a wave owns
// 64 rows x NT/2 members = 2 x (NT/64) tiles of 32 x 32.  One iteration
// stages A (4 k-groups x 128 rows, hi and lo) and F (4 k-groups x NT, hi and
// lo) in LDS, prefetching the next iteration's tile into registers while the
// MFMAs run.
template <int NT>
__global__ void __launch_bounds__(kMixThreads)
mixed_gemm_kernel(const uint4* __restrict__ Ahi, const uint4* __restrict__ Alo,
                  const uint4* __restrict__ Fhi, const uint4* __restrict__ Flo,
                  const uint4* __restrict__ Ghi, const uint4* __restrict__ Glo,
                  const uint8_t* __restrict__ useg, int64_t rows_pad,
                  const int2* __restrict__ krange, int nsplit, float* __restrict__ P) {
    constexpr int TN = NT / 64;                  // 32-wide member tiles per wave
    constexpr int A_UNITS = 2 * kKG * kMT;       // 1024 uint4
    constexpr int F_UNITS = 2 * kKG * NT;
    constexpr int A_PER_T = A_UNITS / kMixThreads;
    constexpr int F_PER_T = F_UNITS / kMixThreads;
    __shared__ uint4 sA[2][kKG][kMT];
    __shared__ uint4 sF[2][kKG][NT];

    const int panel = blockIdx.x;
    const int split = blockIdx.y;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int64_t row0 = (int64_t)panel * kMT;

    const int2 kr = krange[panel];
    const int nit = kr.y - kr.x;
    const int it0 = kr.x + (int)((int64_t)nit * split / nsplit);
    const int it1 = kr.x + (int)((int64_t)nit * (split + 1) / nsplit);

    floatx16 acc[2][TN];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < TN; ++u) acc[t][u] = floatx16{0};

    uint4 ra[A_PER_T], rf[F_PER_T];
    // every F unit this thread stages belongs to member tid % NT
    const bool my_g = useg[(int64_t)panel * NT + (tid % NT)] != 0;
    const uint4* Zhi = my_g ? Ghi : Fhi;
    const uint4* Zlo = my_g ? Glo : Flo;
    auto gload = [&](int it) {
        const int64_t kg0 = (int64_t)it * kKG;
#pragma unroll
        for (int j = 0; j < A_PER_T; ++j) {
            const int q = tid + j * kMixThreads;
            const int which = q / (kKG * kMT);
            const int kg = (q / kMT) % kKG;
            const int r = q % kMT;
            const uint4* src = which ? Alo : Ahi;
            ra[j] = src[(kg0 + kg) * rows_pad + row0 + r];
        }
#pragma unroll
        for (int j = 0; j < F_PER_T; ++j) {
            const int q = tid + j * kMixThreads;
            const int which = q / (kKG * NT);
            const int kg = (q / NT) % kKG;
            const int col = q % NT;
            const uint4* src = which ? Zlo : Zhi;
            rf[j] = src[(kg0 + kg) * NT + col];
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int j = 0; j < A_PER_T; ++j) {
            const int q = tid + j * kMixThreads;
            sA[q / (kKG * kMT)][(q / kMT) % kKG][q % kMT] = ra[j];
        }
#pragma unroll
        for (int j = 0; j < F_PER_T; ++j) {
            const int q = tid + j * kMixThreads;
            sF[q / (kKG * NT)][(q / NT) % kKG][q % NT] = rf[j];
        }
    };

    if (it0 < it1) gload(it0);
    for (int it = it0; it < it1; ++it) {
        lstore();
        __syncthreads();
        if (it + 1 < it1) gload(it + 1);
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int kg = 2 * ks + (lane >> 5);
            union { uint4 u; half8 v; } ah[2], al[2], fh[TN], fl[TN];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int r = wr * 64 + t * 32 + (lane & 31);
                ah[t].u = sA[0][kg][r];
                al[t].u = sA[1][kg][r];
            }
#pragma unroll
            for (int u = 0; u < TN; ++u) {
                const int col = wc * (NT / 2) + u * 32 + (lane & 31);
                fh[u].u = sF[0][kg][col];
                fl[u].u = sF[1][kg][col];
            }
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int u = 0; u < TN; ++u) {
                    acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[t].v, fh[u].v, acc[t][u], 0, 0, 0);
                    acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[t].v, fl[u].v, acc[t][u], 0, 0, 0);
                    acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[t].v, fh[u].v, acc[t][u], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    // C tile -> P[split][member][row]; lane holds member (lane & 31) of each
    // 32-wide tile and rows 8*(r/4) + 4*(lane/32) + (r%4) of each 32-row tile
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < TN; ++u) {
            const int col = wc * (NT / 2) + u * 32 + (lane & 31);
            float* dst = P + ((int64_t)split * NT + col) * rows_pad + row0 + wr * 64 + t * 32;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                floatx4 v = {acc[t][u][4 * g + 0], acc[t][u][4 * g + 1], acc[t][u][4 * g + 2],
                             acc[t][u][4 * g + 3]};
                *reinterpret_cast<floatx4*>(dst + 8 * g + 4 * (lane >> 5)) = v;
            }
        }
}

// ------------------------------------------------------------------ epilogue
// One thread per (sorted node p, member b): FP64 sum of the split partials,
// Pauli-blocked assembly with the member's bath occupation, result written
// to the caller's node order; per-block partials of sum V I and
// sum V |f(1-f)| for the number-defect removal.
constexpr int kFinT = 256;
__global__ void __launch_bounds__(kFinT)
mixed_finish_kernel(const float* __restrict__ P, int nsplit, int nt, int64_t rows_pad,
                    int64_t n, int64_t kt, const double* __restrict__ f, const int32_t* __restrict__ perm,
                    const double* __restrict__ rk, const double* __restrict__ ck,
                    const double* __restrict__ V, const double* __restrict__ bath,
                    const uint8_t* __restrict__ useg, double inv_scale, double* __restrict__ I,
                    double* __restrict__ part, int nblk) {
    const int b = blockIdx.y;
    const int64_t p = blockIdx.x * (int64_t)kFinT + threadIdx.x;
    double dI = 0.0, dw = 0.0;
    if (p < n) {
        double u = 0.0, v = 0.0;
        for (int s = 0; s < nsplit; ++s) {
            const float* Ps = P + ((int64_t)s * nt + b) * rows_pad;
            u += (double)Ps[p];
            v += (double)Ps[kt + p];
        }
        u *= inv_scale;
        v *= inv_scale;
        // u = K z, v = K^T z' with z, z' = f or g = 1 - f per panel
        const bool gu = useg[(p / kMT) * (int64_t)nt + b] != 0;
        const bool gv = useg[((kt + p) / kMT) * (int64_t)nt + b] != 0;
        const double Kf = gu ? rk[p] - u : u, Kg = gu ? u : rk[p] - u;
        const double KTf = gv ? ck[p] - v : v, KTg = gv ? v : ck[p] - v;
        const double nb = bath[b];
        const double np1 = nb + 1.0;
        const int64_t node = perm[p];
        const double fi = f[(int64_t)b * n + node];
        // (1-f)(W f) - f (W^T g),  W = (n+1) K + n K^T
        const double Ii = ((1.0 - fi) * (np1 * Kf + nb * KTf) - fi * (np1 * KTg + nb * Kg)) / V[p];
        I[(int64_t)b * n + node] = Ii;
        dI = V[p] * Ii;
        dw = V[p] * fabs(fi * (1.0 - fi));
    }
    if (part) {
        __shared__ double sh[2][kFinT / 64];
        for (int o = 32; o > 0; o >>= 1) {
            dI += __shfl_down(dI, o, 64);
            dw += __shfl_down(dw, o, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            sh[0][threadIdx.x >> 6] = dI;
            sh[1][threadIdx.x >> 6] = dw;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = 0.0, c = 0.0;
            for (int w = 0; w < kFinT / 64; ++w) {
                a += sh[0][w];
                c += sh[1][w];
            }
            part[((int64_t)b * nblk + blockIdx.x) * 2] = a;
            part[((int64_t)b * nblk + blockIdx.x) * 2 + 1] = c;
        }
    }
}

// per member: s_b = sum V I / sum V |f(1-f)|; I -= s_b |f(1-f)|
__global__ void mixed_conserve_kernel(const double* __restrict__ part, int nblk, int64_t n,
                                      const double* __restrict__ f, double* __restrict__ I) {
    const int b = blockIdx.y;
    __shared__ double s_b;
    if (threadIdx.x == 0) {
        double d = 0.0, w = 0.0;
        for (int k = 0; k < nblk; ++k) {
            d += part[((int64_t)b * nblk + k) * 2];
            w += part[((int64_t)b * nblk + k) * 2 + 1];
        }
        s_b = w > 0.0 ? d / w : 0.0;
    }
    __syncthreads();
    const double s = s_b;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double fi = f[(int64_t)b * n + i];
        I[(int64_t)b * n + i] -= s * fabs(fi * (1.0 - fi));
    }
}

// ------------------------------------------------------------------- host
static int nt_for(int B) {
    if (B <= 64) return 64;
    if (B <= 128) return 128;
    return 256;
}

tg_status mixed_prepare(tg_ctx* c) {
    if (c->mixed) return TG_OK;
    if (!c->assembled) return set_error(c, TG_ERR_STATE, "kernel not assembled");
    if (c->max_batch > 256) return set_error(c, TG_ERR_ARG, "mixed path supports max_batch <= 256");
    MixedData* m = new MixedData();
    c->mixed = m;
    const int64_t n = c->n;
    m->rows = 2 * n;
    m->kt = (n + kMT - 1) / kMT * kMT;
    m->rows_pad = 2 * m->kt;
    m->kpad = (n + kKS - 1) / kKS * kKS;
    m->npanel = (int)(m->rows_pad / kMT);
    m->nt = nt_for(c->max_batch);
    const int64_t kgroups = m->kpad / 8;
    hipStream_t st = c->stream;

    // node order shared with the plain path (energy order, built at context creation)
    m->perm = c->d_perm;

    // scale: largest |K| maps to 2^14 (fp16 max 65504)
    double kmax = 0.0;
    {
        // K 1 bounds every entry of a nonnegative K; use the exact max instead
        std::vector<double> rk(n);
        TG_HIP(c, hipMemcpyAsync(rk.data(), c->d_rk, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        TG_HIP(c, hipStreamSynchronize(st));
        for (double v : rk) kmax = std::max(kmax, std::fabs(v));
    }
    m->sa = kmax > 0.0 ? (int)std::floor(std::log2(16384.0 / kmax)) : 0;
    const double scale = std::ldexp(1.0, m->sa);

    TG_TRY(dev_alloc(c, &m->A_hi, (size_t)kgroups * m->rows_pad));
    TG_TRY(dev_alloc(c, &m->A_lo, (size_t)kgroups * m->rows_pad));
    hipLaunchKernelGGL(pack_A_kernel, dim3(ceil_div(m->rows_pad, 256), kgroups), dim3(256), 0, st,
                       c->d_K, n, c->ld, m->kt, m->rows_pad, kgroups, scale, m->A_hi, m->A_lo);
    TG_CHECK_LAUNCH(c);

    // split-matrix row sums and sorted node volumes
    TG_TRY(dev_alloc(c, &m->rk_s, (size_t)n));
    TG_TRY(dev_alloc(c, &m->ck_s, (size_t)n));
    m->V_s = c->d_V_s;
    hipLaunchKernelGGL(split_rowsum_kernel, dim3(ceil_div(m->rows_pad, 256)), dim3(256), 0, st,
                       m->A_hi, m->A_lo, m->rows_pad, kgroups, n, m->kt, 1.0 / scale, m->rk_s,
                       m->ck_s);
    TG_CHECK_LAUNCH(c);

    // band extent of every panel
    int32_t *d_first = nullptr, *d_last = nullptr;
    TG_TRY(dev_alloc(c, &d_first, (size_t)m->rows_pad));
    TG_TRY(dev_alloc(c, &d_last, (size_t)m->rows_pad));
    hipLaunchKernelGGL(row_extent_kernel, dim3(ceil_div(m->rows_pad, 256)), dim3(256), 0, st,
                       m->A_hi, m->A_lo, m->rows_pad, kgroups, d_first, d_last);
    TG_CHECK_LAUNCH(c);
    std::vector<int32_t> first(m->rows_pad), last(m->rows_pad);
    TG_HIP(c, hipMemcpyAsync(first.data(), d_first, sizeof(int32_t) * m->rows_pad, hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipMemcpyAsync(last.data(), d_last, sizeof(int32_t) * m->rows_pad, hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipStreamSynchronize(st));
    dev_free(&d_first);
    dev_free(&d_last);
    m->h_krange.resize(m->npanel);
    int64_t items = 0;
    for (int pn = 0; pn < m->npanel; ++pn) {
        int32_t lo = INT32_MAX, hi = -1;
        for (int r = 0; r < kMT; ++r) {
            const int64_t row = (int64_t)pn * kMT + r;
            if (first[row] >= 0) {
                lo = std::min(lo, first[row]);
                hi = std::max(hi, last[row]);
            }
        }
        int2 kr;
        if (hi < 0) {
            kr.x = kr.y = 0;
        } else {
            kr.x = lo / kKG;             // k-groups -> iterations
            kr.y = hi / kKG + 1;
        }
        m->h_krange[pn] = kr;
        items += kr.y - kr.x;
    }
    m->work_items = items;
    TG_TRY(dev_alloc(c, &m->krange, (size_t)m->npanel));
    TG_HIP(c, hipMemcpyAsync(m->krange, m->h_krange.data(), sizeof(int2) * m->npanel,
                             hipMemcpyHostToDevice, st));
    // K-splits: ~16 workgroups per CU (fills the 256 CUs and shortens every
    // FP32 accumulation chain; the splits are summed in FP64), each split a
    // few iterations long
    const double avg = (double)items / std::max(1, m->npanel);
    int ns = (int)std::ceil(4096.0 / m->npanel);
    ns = std::max(1, std::min(ns, (int)std::max(1.0, avg / 3.0)));
    m->nsplit = std::min(ns, 32);
    // explicit override (measurements of the split count)
    if (const char* e = std::getenv("TG_MIXED_SPLITS")) {
        const int v = std::atoi(e);
        if (v >= 1 && v <= 64) m->nsplit = v;
    }
    TG_TRY(dev_alloc(c, &m->F_hi, (size_t)kgroups * m->nt));
    TG_TRY(dev_alloc(c, &m->F_lo, (size_t)kgroups * m->nt));
    TG_TRY(dev_alloc(c, &m->G_hi, (size_t)kgroups * m->nt));
    TG_TRY(dev_alloc(c, &m->G_lo, (size_t)kgroups * m->nt));
    TG_TRY(dev_alloc(c, &m->useg, (size_t)m->npanel * m->nt));
    TG_TRY(dev_alloc(c, &m->bath, (size_t)c->max_batch));
    TG_HIP(c, hipStreamSynchronize(st));
    return TG_OK;
}

void mixed_stats(tg_ctx* c, tg_kernel_stats* out) {
    MixedData* m = md_of(c);
    if (!m) return;
    out->mixed_band = m->work_items * (int64_t)kMT * kKS;
    out->mixed_splits = m->nsplit;
    out->mixed_width = m->nt;
    out->mixed_scale = m->sa;
}

void mixed_destroy(tg_ctx* c) {
    MixedData* m = md_of(c);
    if (!m) return;
    dev_free(&m->A_hi);
    dev_free(&m->A_lo);
    dev_free(&m->F_hi);
    dev_free(&m->F_lo);
    dev_free(&m->G_hi);
    dev_free(&m->G_lo);
    dev_free(&m->useg);
    dev_free(&m->krange);
    dev_free(&m->rk_s);
    dev_free(&m->ck_s);
    dev_free(&m->P);
    dev_free(&m->bath);
    dev_free(&m->part);
    delete m;
    c->mixed = nullptr;
}

tg_status mixed_apply_device(tg_ctx* c, const double* d_f, double* d_I, int32_t B,
                             const double* h_bath, int conserve) {
    TG_TRY(mixed_prepare(c));
    MixedData* m = md_of(c);
    if (B < 1 || B > m->nt) return set_error(c, TG_ERR_ARG, "batch %d exceeds the packed width %d", B, m->nt);
    hipStream_t st = c->stream;
    const int64_t n = c->n;
    const int64_t kgroups = m->kpad / 8;
    const size_t need = (size_t)m->nsplit * m->nt * m->rows_pad;
    if (m->P_cap < need) {
        dev_free(&m->P);
        TG_TRY(dev_alloc(c, &m->P, need));
        m->P_cap = need;
    }
    std::vector<double> bath(B);
    for (int b = 0; b < B; ++b) bath[b] = h_bath ? h_bath[b] : c->bath_n;
    TG_HIP(c, hipMemcpyAsync(m->bath, bath.data(), sizeof(double) * B, hipMemcpyHostToDevice, st));

    hipLaunchKernelGGL(pack_F_kernel, dim3(ceil_div(m->nt, 64), kgroups), dim3(64), 0, st, d_f, n,
                       B, m->perm, kgroups, m->nt, m->F_hi, m->F_lo, m->G_hi, m->G_lo);
    TG_CHECK_LAUNCH(c);
    hipLaunchKernelGGL(choose_operand_kernel, dim3(m->npanel), dim3(m->nt), 0, st, d_f, n, m->kt, B,
                       m->npanel, m->perm, m->useg, m->nt);
    TG_CHECK_LAUNCH(c);
    const dim3 grid(m->npanel, m->nsplit);
    time_begin(c);
    switch (m->nt) {
        case 64:
            hipLaunchKernelGGL(mixed_gemm_kernel<64>, grid, dim3(kMixThreads), 0, st, m->A_hi,
                               m->A_lo, m->F_hi, m->F_lo, m->G_hi, m->G_lo, m->useg, m->rows_pad,
                               m->krange, m->nsplit, m->P);
            break;
        case 128:
            hipLaunchKernelGGL(mixed_gemm_kernel<128>, grid, dim3(kMixThreads), 0, st, m->A_hi,
                               m->A_lo, m->F_hi, m->F_lo, m->G_hi, m->G_lo, m->useg, m->rows_pad,
                               m->krange, m->nsplit, m->P);
            break;
        default:
            hipLaunchKernelGGL(mixed_gemm_kernel<256>, grid, dim3(kMixThreads), 0, st, m->A_hi,
                               m->A_lo, m->F_hi, m->F_lo, m->G_hi, m->G_lo, m->useg, m->rows_pad,
                               m->krange, m->nsplit, m->P);
            break;
    }
    time_end(c);
    TG_CHECK_LAUNCH(c);
    const int nblk = (int)ceil_div(n, kFinT);
    double* part = nullptr;
    if (conserve) {
        if (m->part_cap < nblk * B) {
            dev_free(&m->part);
            TG_TRY(dev_alloc(c, &m->part, (size_t)2 * nblk * B));
            m->part_cap = nblk * B;
        }
        part = m->part;
    }
    const double inv_scale = std::ldexp(1.0, -(m->sa + kSF));
    hipLaunchKernelGGL(mixed_finish_kernel, dim3(nblk, B), dim3(kFinT), 0, st, m->P, m->nsplit, m->nt,
                       m->rows_pad, n, m->kt, d_f, m->perm, m->rk_s, m->ck_s, m->V_s, m->bath, m->useg,
                       inv_scale, d_I, part, nblk);
    TG_CHECK_LAUNCH(c);
    if (conserve) {
        hipLaunchKernelGGL(mixed_conserve_kernel, dim3(ceil_div(n, 1024), B), dim3(256), 0, st,
                           part, nblk, n, d_f, d_I);
        TG_CHECK_LAUNCH(c);
    }
    return TG_OK;
}

}  // namespace tg
