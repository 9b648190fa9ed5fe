// tg_ee.hip -- electron-electron (4-leg) collision operator on the periodic
// momentum lattice of the mesh box.
//
//   I_1 = sum_{2,3} w^2 A(k1-k3) delta(e1 + e2 - e3 - e4)
//                   [ f3 f4 (1-f1)(1-f2) - f1 f2 (1-f3)(1-f4) ],   k4 = k1 + k2 - k3
//
// over the P = prod(nsec) periodic points (the mesh's duplicate boundary
// nodes carry their periodic image's value), w = volume per point,
// A(q) = g^2/(|q|^2 + kappa2)^2 (minimum image), the Gaussian energy delta
// projected on the grid m*dw with hat functions (the reference counterpart is
// ScatteringChannel<3,4> with functAmplitude = V(q)^2,
// src/Physics/Scattering/ScatteringChannel.cpp:399-415,618-650).
//
// Grouping by the momentum transfer q = k1 - k3 (then k4 = k2 + q) makes the
// raw O(P^3) sum O(P^2 + P M J): per q, the pair spectra
//   Ha_q(m) = sum_2 w h_m(e(2+q) - e(2)) f(2+q)(1 - f(2))   (gain)
//   Hb_q(m) = sum_2 w h_m(e(2+q) - e(2)) f(2)(1 - f(2+q))   (loss)
// are convolved with the Gaussian on the grid and read back at e1 - e3 for
// every k1.  Substituting 2 -> 2 + q shows Ha_{-q}(m) = Hb_q(-m) (bins
// mirrored about 0; the grid and the Gaussian are symmetric), so only the
// loss spectra are built: the convolved S_q = G * Hb_q serves
//   - the loss term of q at k1:  -A(q) f1 (1 - f(k1-q)) S_q(e1 - e(k1-q)),
//   - the gain term of -q at k1: +A(q) (1 - f1) f(k1+q) S_q(e(k1+q) - e1)
// (A(-q) = A(q)).  Both read-back positions are deposited differences
// e(k2 + q) - e(k2) (k2 = k1 - q and k2 = k1), so the convolution is only
// needed on the bins the spectrum occupies.  A workgroup owns 16 values of q
// and one member:
//   - the spectrum is accumulated in LDS as int64 fixed point from one sweep
//     over k2 (integer adds are exact, so the sum does not depend on their
//     order; lanes of a wave that hit the same bin are summed by a segmented
//     scan before one atomic), with its occupied bin range, then converted to
//     FP64 in place;
//   - the convolution runs in FP64 over the occupied range, 9 grid outputs
//     per thread from a register window over the spectrum (zero guard bins
//     on both sides, so no read is bounds-checked; an odd count keeps the
//     wave's LDS reads free of bank conflicts), and overwrites the spectrum
//     in place once every thread holds its outputs;
//   - each thread keeps the sums of its k1 values in registers across the
//     16 transfers and writes them once;
// one spectrum per workgroup (M doubles of LDS) and at most 128 registers
// let two workgroups share a CU, so one's convolution overlaps another's
// deposit.  A last kernel sums the workgroups' partials per k1 in a fixed
// order (tile by tile) and adds the result to the caller's output on every
// mesh node.
#include <algorithm>
#include <cmath>
#include <vector>

#include "tg_internal.hpp"

namespace tg {

constexpr int kEeT = 512;       // threads of the main kernel
constexpr int kEeAccMax = 32;   // k1 values per thread: P <= kEeT * kEeAccMax
constexpr int kEeQTile = 16;    // momentum transfers per workgroup
// grid outputs per thread in the convolution: odd, so the lanes of a wave,
// kEeConvR doubles apart, read LDS without bank conflicts (with 8 they are
// 64 bytes apart and 16 lanes share each bank pair)
constexpr int kEeConvR = 9;
constexpr int kEeMaxJ = 256;    // largest Gaussian half-width in grid cells
constexpr int kEeMaxM = 8192;    // largest energy grid (the convolution's chunks per thread cover it)

struct EeData {
    int dim = 2;
    int ns[3] = {1, 1, 1};
    int64_t P = 0;
    double w = 0.0, dw = 0.0;
    int32_t Mh = 0, M = 0, J = 0;
    int32_t ntile = 0, cap = 0;
    double* gtab = nullptr;      // G_sigma(j dw), |j| <= J
    double* ep = nullptr;        // energies on the lattice (P)
    int32_t* node_of_p = nullptr;
    int32_t* p_of_node = nullptr;
    int32_t* coords = nullptr;   // packed lattice coordinates (P)
    double* amp = nullptr;       // A(q) (P)
    double* fp = nullptr;        // f on the lattice, cap x P
    int32_t* fexp = nullptr;     // fixed-point exponent per member
    double* part = nullptr;      // cap x ntile x P partial sums
    // vector-Jacobian product (one member): mu summed onto the lattice and
    // the fixed-point exponent of the mu-weighted spectrum
    double* mup = nullptr;       // P
    int32_t* mexp = nullptr;     // 1
};

struct EeLattice {
    int dim;
    int ns0, ns1, ns2;
};

// f on the lattice per member and the fixed-point exponent E of its spectra:
// every bin sum is bounded by P w m (1 + m), m = max|f|, and E puts that
// bound at 2^61
__global__ void __launch_bounds__(kEeT)
ee_prep_kernel(const double* __restrict__ f, int64_t n, int64_t P,
               const int32_t* __restrict__ node_of_p, double w, double* __restrict__ fp,
               int32_t* __restrict__ fexp, const double* __restrict__ gate) {
    __shared__ double red[kEeT / 64];
    const int b = blockIdx.x;
    if (gate && gate[b] == 0.0) return;
    double m = 0.0;
    for (int64_t p = threadIdx.x; p < P; p += kEeT) {
        const double x = f[(int64_t)b * n + node_of_p[p]];
        fp[(int64_t)b * P + p] = x;
        m = fmax(m, fabs(x));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0.0;
        for (int i = 0; i < kEeT / 64; ++i) mm = fmax(mm, red[i]);
        const double bound = (double)P * w * mm * (1.0 + mm);
        int32_t e = 61;
        if (!isfinite(bound)) e = 0;
        else if (bound > 0.0) e = (int32_t)floor(61.0 - log2(bound));
        fexp[b] = max(-1000, min(1000, e));
    }
}

// lattice coordinates are packed as c0 | c1 << 10 | c2 << 20 (at most 1023 per axis)
// index of the point a + sign*q (periodic) from the linear indices p of a
// and q of the transfer: p + sign*q, corrected on each axis that wraps
// (sign +: c + q >= ns; sign -: c < q); no divisions or re-linearisation
__device__ inline int ee_shift(const EeLattice& L, int p, int a, int q, int qc, int sign) {
    const int c0 = a & 1023, c1 = (a >> 10) & 1023, c2 = a >> 20;
    const int q0 = qc & 1023, q1 = (qc >> 10) & 1023, q2 = qc >> 20;
    const int s1 = L.ns0 * L.ns1, s2 = s1 * L.ns2;
    if (sign > 0)
        return p + q - (c0 >= L.ns0 - q0 ? L.ns0 : 0) - (c1 >= L.ns1 - q1 ? s1 : 0) -
               (c2 >= L.ns2 - q2 ? s2 : 0);
    return p - q + (c0 < q0 ? L.ns0 : 0) + (c1 < q1 ? s1 : 0) + (c2 < q2 ? s2 : 0);
}

// Adds the two fixed-point deposits of every lane to its bins m, m + 1 of
// the spectrum.  When lanes of one wave hit the same bin (consecutive lanes
// with equal m: common when the energy difference does not depend on the
// fastest lattice axis) they are summed first by a segmented scan, so each
// run costs one pair of atomics.  Integer sums: the result is exact and does
// not depend on the path taken.
__device__ inline void ee_deposit(int m, long long a0, long long a1, long long* __restrict__ H) {
    const int lane = threadIdx.x & 63;
    const int mprev = __shfl_up(m, 1, 64);
    // no two neighbouring lanes share a bin (the usual case when the transfer
    // has a component along the fastest axis): one pair of atomics per lane
    if (__ballot(lane > 0 && mprev == m && m >= 0) == 0) {
        if (m >= 0) {
            atomicAdd(reinterpret_cast<unsigned long long*>(&H[m]), (unsigned long long)a0);
            atomicAdd(reinterpret_cast<unsigned long long*>(&H[m + 1]), (unsigned long long)a1);
        }
        return;
    }
    int start = (lane == 0 || mprev != m) ? lane : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int s = __shfl_up(start, o, 64);
        if (lane >= o) start = max(start, s);
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long x0 = __shfl_up(a0, o, 64), x1 = __shfl_up(a1, o, 64);
        if (lane - o >= start) {
            a0 += x0;
            a1 += x1;
        }
    }
    const int mnext = __shfl_down(m, 1, 64);
    if ((lane == 63 || mnext != m) && m >= 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&H[m]), (unsigned long long)a0);
        atomicAdd(reinterpret_cast<unsigned long long*>(&H[m + 1]), (unsigned long long)a1);
    }
}

// S = G * H on the bins [lo, hi], in place: every thread first computes its
// outputs (kEeConvR consecutive bins per chunk, the taps walked from +J to
// -J with a window of H in registers; one LDS read of H and one scalar load
// of the tap per tap), then, after a barrier, overwrites H with them.  The
// window is not shifted: the taps run in unrolled blocks of kEeConvR, and at
// step s of a block output r reads slot (r + s) % kEeConvR, the slot the
// step then refills.  H is zero outside [lo, hi], including its guard bins
// H[-J .. -1] and H[M .. M + J + kEeConvR - 1], so no read is bounds-checked.
template <int NT>
__device__ inline void ee_convolve_inplace(double* __restrict__ H, const double* __restrict__ gtab,
                                           int lo, int hi, int J) {
    constexpr int CH = (kEeMaxM + NT * kEeConvR - 1) / (NT * kEeConvR);   // chunks per thread
    double out[CH][kEeConvR];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        const int m0 = lo + (threadIdx.x + c * NT) * kEeConvR;
#pragma unroll
        for (int r = 0; r < kEeConvR; ++r) out[c][r] = 0.0;
        if (m0 > hi) continue;
        double h[kEeConvR];
#pragma unroll
        for (int r = 0; r < kEeConvR; ++r) h[r] = H[m0 + r - J];
        for (int j0 = J; j0 >= -J; j0 -= kEeConvR) {
#pragma unroll
            for (int s = 0; s < kEeConvR; ++s) {
                const int j = j0 - s;
                if (j < -J) break;
                const double gj = gtab[j + J];
#pragma unroll
                for (int r = 0; r < kEeConvR; ++r) out[c][r] = fma(gj, h[(r + s) % kEeConvR], out[c][r]);
                h[s] = H[m0 + kEeConvR - j];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        const int m0 = lo + (threadIdx.x + c * NT) * kEeConvR;
#pragma unroll
        for (int r = 0; r < kEeConvR; ++r)
            if (m0 + r <= hi) H[m0 + r] = out[c][r];
    }
}

// grid (q tile, member).  MODE 0: the operator.  MODES 1 and 2: the two
// passes of the vector-Jacobian product lambda = J(f)^T mu (one member, mu on
// the lattice in mup).  With W(1,2,3,4) the kernel of the sum (symmetric under
// 1<->2, 3<->4 and under 12<->34),
//   lambda_1 = sum_{2,3} W (mu3 + mu4 - mu1 - mu2) [f3 f4 (1-f2) + f2 (1-f3)(1-f4)],
// which splits like the operator into pair spectra over k2 read back at k1:
//   MODE 1 deposits the loss spectrum Hb_q (as MODE 0) and reads back
//     (1 - f3)(mu3 - mu1) S_q(e1 - e3), k3 = k1 - q, and f3 (mu3 - mu1) S_q(e3 - e1),
//     k3 = k1 + q (the gain spectrum of -q, mirrored);
//   MODE 2 deposits Hbmu_q(m) = sum_2 w h_m(e4 - e2) (mu4 - mu2) f2 (1 - f4)
//     (the mirror of -Habmu_{-q}) and reads back (1 - f3) S_q(e1 - e3), k3 = k1 - q,
//     and -f3 S_q(e3 - e1), k3 = k1 + q.
template <int ACC, int MODE>
__global__ void __launch_bounds__(kEeT, 4)
ee_main_kernel(EeLattice L, int P, int Mh, int M, int J, double inv_dw, double w,
               const double* __restrict__ gtab, const double* __restrict__ ep,
               const int32_t* __restrict__ coords,
               const double* __restrict__ amp, const double* __restrict__ fp,
               const int32_t* __restrict__ fexp, int ntile, const double* __restrict__ gate,
               const double* __restrict__ mup, double* __restrict__ part) {
    extern __shared__ double lds[];
    __shared__ int s_lo, s_hi;
    if (gate && gate[blockIdx.y] == 0.0) return;   // uniform over the workgroup
    // spectrum bins 0 .. M-1 with zero guard bins -J .. -1 and M .. M+J+kEeConvR-1
    long long* H = reinterpret_cast<long long*>(lds) + J;   // fixed point
    double* Hd = lds + J;                                    // the same bins as FP64
    const int tid = threadIdx.x;
    const int b = blockIdx.y;
    const int tile = blockIdx.x;
    const int q0 = tile * kEeQTile;
    const int q1 = min(q0 + kEeQTile, P);
    const double* fb = fp + (int64_t)b * P;
    const double scale = ldexp(1.0, fexp[b]);
    const double inv = ldexp(1.0, -fexp[b]);
    for (int m = tid - J; m < M + J + kEeConvR; m += kEeT) H[m] = 0;
    if (tid == 0) {
        s_lo = M;
        s_hi = -1;
    }

    // this thread's points p = tid + i*kEeT serve both as k2 (spectrum) and as
    // k1 (evaluation); only the sums of the k1 values live in registers
    double acc[ACC];
#pragma unroll
    for (int i = 0; i < ACC; ++i) acc[i] = 0.0;
    __syncthreads();

    for (int q = q0; q < q1; ++q) {
        const double aq = w * amp[q];
        const int qc = coords[q];
        // loss spectrum of transfer q: k4 = k2 + q (every lane takes part in
        // every iteration: the deposit scans across the wave)
        int mlo = M, mhi = -1;
#pragma unroll 1
        for (int i = 0; i < ACC; ++i) {
            const int p2 = tid + i * kEeT;
            int m = -1;
            long long a0 = 0, a1 = 0;
            if (p2 < P) {
                const int p4 = ee_shift(L, p2, coords[p2], q, qc, +1);
                const double f2 = fb[p2], f4 = fb[p4];
                const double y = (ep[p4] - ep[p2]) * inv_dw;
                const double fl = floor(y);
                const double t = y - fl;
                m = (int)fl + Mh;
                double F = w * f2 * (1.0 - f4);
                if constexpr (MODE == 2) F *= mup[p4] - mup[p2];
                a0 = llrint((1.0 - t) * F * scale);
                a1 = llrint(t * F * scale);
                mlo = min(mlo, m);
                mhi = max(mhi, m);
            }
            ee_deposit(m, a0, a1, H);
        }
        for (int o = 32; o > 0; o >>= 1) {
            mlo = min(mlo, __shfl_xor(mlo, o, 64));
            mhi = max(mhi, __shfl_xor(mhi, o, 64));
        }
        if ((tid & 63) == 0) {
            atomicMin(&s_lo, mlo);
            atomicMax(&s_hi, mhi);
        }
        __syncthreads();
        // nonzero bins: [lo, hi] (deposits go to m and m + 1)
        const int lo = s_lo, hi = s_hi + 1;
        for (int m = lo + tid; m <= hi; m += kEeT) Hd[m] = (double)H[m] * inv;
        __syncthreads();
        ee_convolve_inplace<kEeT>(Hd, gtab, lo, hi, J);
        __syncthreads();
        // k1 = p1: the loss term of q (k3 = k1 - q) and the gain term of -q
        // (k3 = k1 + q), both read within [lo, hi] (x and z are computed by
        // the same operations as the deposits' y at k2 = k1 - q and k2 = k1)
#pragma unroll 1
        for (int i = 0; i < ACC; ++i) {
            const int p1 = tid + i * kEeT;
            if (p1 < P) {
                const int c1 = coords[p1];
                const int pm = ee_shift(L, p1, c1, q, qc, -1), pp = ee_shift(L, p1, c1, q, qc, +1);
                const double e1 = ep[p1];
                const double x = (e1 - ep[pm]) * inv_dw, z = (ep[pp] - e1) * inv_dw;
                const double fx = floor(x), fz = floor(z);
                const double tx = x - fx, tz = z - fz;
                const int mx = (int)fx + Mh, mz = (int)fz + Mh;
                const double sx = (1.0 - tx) * Hd[mx] + tx * Hd[mx + 1];
                const double sz = (1.0 - tz) * Hd[mz] + tz * Hd[mz + 1];
                const double f1 = fb[p1], fm = fb[pm], fp_ = fb[pp];
                if constexpr (MODE == 0) {
                    acc[i] += aq * ((1.0 - f1) * fp_ * sz - f1 * (1.0 - fm) * sx);
                } else if constexpr (MODE == 1) {
                    const double m1 = mup[p1];
                    acc[i] += aq * ((1.0 - fm) * (mup[pm] - m1) * sx + fp_ * (mup[pp] - m1) * sz);
                } else {
                    acc[i] += aq * ((1.0 - fm) * sx - fp_ * sz);
                }
            }
        }
        __syncthreads();
        for (int m = lo + tid; m <= hi; m += kEeT) H[m] = 0;
        if (tid == 0) {
            s_lo = M;
            s_hi = -1;
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < ACC; ++i) {
        const int p1 = tid + i * kEeT;
        if (p1 < P) part[((int64_t)b * ntile + tile) * P + p1] = acc[i];
    }
}

// out[b][node] += sum over workgroups (fixed order) of the node's periodic point
__global__ void ee_reduce_kernel(const double* __restrict__ part, int ntile, int64_t P,
                                 const int32_t* __restrict__ p_of_node, int64_t n,
                                 const double* __restrict__ gate, double* __restrict__ out) {
    const int b = blockIdx.y;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n || (gate && gate[b] == 0.0)) return;
    const int64_t p = p_of_node[i];
    double s = 0.0;
    for (int t = 0; t < ntile; ++t) s += part[((int64_t)b * ntile + t) * P + p];
    out[(int64_t)b * n + i] += s;
}

// mu (nodes) -> mup (lattice, one member): the fixed-order sum over the
// nodes of each periodic point (a point on a face, edge or corner of the box
// has 2, 4 or 8 nodes: coordinate 0 and nsec on each axis where it is 0);
// and the fixed-point exponent of the mu-weighted spectrum, whose bin sums
// are bounded by P w max|f| (1 + max|f|) 2 max|mup|.  One workgroup.
__global__ void __launch_bounds__(kEeT)
ee_vjp_prep_kernel(EeLattice L, int64_t P, const double* __restrict__ mu, const double* __restrict__ fp,
                   double w, double* __restrict__ mup, int32_t* __restrict__ mexp) {
    __shared__ double red[2][kEeT / 64];
    const int s0 = L.ns0 + 1, s1 = (L.ns0 + 1) * (L.ns1 + 1);
    double mf = 0.0, mm = 0.0;
    for (int64_t p = threadIdx.x; p < P; p += kEeT) {
        const int c0 = (int)(p % L.ns0), c1 = (int)((p / L.ns0) % L.ns1), c2 = (int)(p / ((int64_t)L.ns0 * L.ns1));
        const int n0 = c0 == 0 ? 2 : 1, n1 = (L.dim > 1 && c1 == 0) ? 2 : 1, n2 = (L.dim > 2 && c2 == 0) ? 2 : 1;
        double sum = 0.0;
        for (int a2 = 0; a2 < n2; ++a2)
            for (int a1 = 0; a1 < n1; ++a1)
                for (int a0 = 0; a0 < n0; ++a0)
                    sum += mu[(c0 + a0 * L.ns0) + (int64_t)(c1 + a1 * L.ns1) * s0 + (int64_t)(c2 + a2 * L.ns2) * s1];
        mup[p] = sum;
        mm = fmax(mm, fabs(sum));
        mf = fmax(mf, fabs(fp[p]));
    }
    for (int o = 32; o > 0; o >>= 1) {
        mf = fmax(mf, __shfl_down(mf, o, 64));
        mm = fmax(mm, __shfl_down(mm, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = mf;
        red[1][threadIdx.x >> 6] = mm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < kEeT / 64; ++i) {
            a = fmax(a, red[0][i]);
            b = fmax(b, red[1][i]);
        }
        const double bound = (double)P * w * a * (1.0 + a) * 2.0 * b;
        int32_t e = 61;
        if (!isfinite(bound)) e = 0;
        else if (bound > 0.0) e = (int32_t)floor(61.0 - log2(bound));
        *mexp = max(-1000, min(1000, e));
    }
}

// out[i] += sum over workgroups (fixed order) of the point's partials, on the
// node that carries the point's value (node_of_p): the other nodes of a point
// do not enter the operator
__global__ void ee_vjp_reduce_kernel(const double* __restrict__ part, int ntile, int64_t P,
                                     const int32_t* __restrict__ node_of_p, double* __restrict__ out) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= P) return;
    double s = 0.0;
    for (int t = 0; t < ntile; ++t) s += part[(int64_t)t * P + p];
    out[node_of_p[p]] += s;
}

static EeData* ee_of(tg_ctx* c) { return (EeData*)c->ee; }

tg_status ee_prepare(tg_ctx* c, const double* h_eps) {
    if (c->ee) return TG_OK;
    EeData* e = new EeData();
    c->ee = e;
    e->dim = c->dim;
    e->P = 1;
    e->w = 1.0;
    for (int d = 0; d < 3; ++d) e->ns[d] = d < c->dim ? c->nsec[d] : 1;
    for (int d = 0; d < c->dim; ++d) {
        e->P *= e->ns[d];
        e->w *= c->extent[d] / e->ns[d];
    }
    for (int d = 0; d < c->dim; ++d)
        if (e->ns[d] > 1023) return set_error(c, TG_ERR_ARG, "e-e channel: more than 1023 points per axis");
    if (e->P > (int64_t)kEeT * kEeAccMax)
        return set_error(c, TG_ERR_ARG, "e-e channel: %lld lattice points exceed %d",
                         (long long)e->P, kEeT * kEeAccMax);
    e->dw = c->ee_dw > 0.0 ? c->ee_dw : c->ee_sigma / 4.0;
    double lo = h_eps[0], hi = h_eps[0];
    for (int64_t i = 1; i < c->n; ++i) {
        lo = std::fmin(lo, h_eps[i]);
        hi = std::fmax(hi, h_eps[i]);
    }
    e->Mh = (int32_t)std::ceil((hi - lo) / e->dw) + 2;
    e->M = 2 * e->Mh + 1;
    e->J = (int32_t)std::floor(c->cutoff * c->ee_sigma / e->dw);
    if (e->J > kEeMaxJ)
        return set_error(c, TG_ERR_ARG, "e-e channel: Gaussian spans %d grid cells (max %d)", e->J,
                         kEeMaxJ);
    hipDeviceProp_t prop;
    TG_HIP(c, hipGetDeviceProperties(&prop, c->device));
    const size_t lds = sizeof(double) * ((size_t)e->M + 2 * e->J + kEeConvR);
    if (e->M > kEeMaxM || lds + 64 > (size_t)prop.sharedMemPerBlock)
        return set_error(c, TG_ERR_ARG, "e-e channel: energy grid of %d cells exceeds %d", e->M, kEeMaxM);
    for (const void* k : {(const void*)ee_main_kernel<8, 0>, (const void*)ee_main_kernel<16, 0>,
                          (const void*)ee_main_kernel<32, 0>, (const void*)ee_main_kernel<8, 1>,
                          (const void*)ee_main_kernel<16, 1>, (const void*)ee_main_kernel<32, 1>,
                          (const void*)ee_main_kernel<8, 2>, (const void*)ee_main_kernel<16, 2>,
                          (const void*)ee_main_kernel<32, 2>})
        TG_HIP(c, hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    e->ntile = (int32_t)((e->P + kEeQTile - 1) / kEeQTile);
    e->cap = c->max_batch;

    // lattice tables (host)
    const int64_t P = e->P;
    std::vector<int32_t> node_of_p(P), p_of_node(c->n), coords(P);
    std::vector<double> ep(P), amp(P);
    for (int64_t p = 0; p < P; ++p) {
        int64_t r = p, id = 0, stride = 1;
        for (int d = 0; d < c->dim; ++d) {
            const int64_t cd = r % e->ns[d];
            r /= e->ns[d];
            id += cd * stride;
            stride *= c->nsec[d] + 1;
        }
        node_of_p[p] = (int32_t)id;
        ep[p] = h_eps[id];
        {
            int64_t rr = p;
            int32_t pk = 0;
            for (int d = 0; d < c->dim; ++d) {
                pk |= (int32_t)(rr % e->ns[d]) << (10 * d);
                rr /= e->ns[d];
            }
            coords[p] = pk;
        }
        // A(q) for the lattice vector q = p (minimum image)
        double q2 = 0.0;
        r = p;
        for (int d = 0; d < c->dim; ++d) {
            const int64_t cd = r % e->ns[d];
            r /= e->ns[d];
            const double box = c->extent[d];
            double qd = cd * (box / e->ns[d]);
            if (qd >= 0.5 * box) qd -= box;
            q2 += qd * qd;
        }
        const double v = 1.0 / (q2 + c->ee_kappa2);
        amp[p] = c->ee_g * c->ee_g * v * v;
    }
    for (int64_t i = 0; i < c->n; ++i) {
        int64_t r = i, p = 0, stride = 1;
        for (int d = 0; d < c->dim; ++d) {
            const int64_t cd = r % (c->nsec[d] + 1);
            r /= c->nsec[d] + 1;
            p += (cd % e->ns[d]) * stride;
            stride *= e->ns[d];
        }
        p_of_node[i] = (int32_t)p;
    }
    std::vector<double> gt(2 * e->J + 1);
    const double nrm = 1.0 / (c->ee_sigma * std::sqrt(2.0 * M_PI));
    for (int j = -e->J; j <= e->J; ++j) {
        const double x = j * e->dw;
        gt[j + e->J] = nrm * std::exp(-0.5 * x * x / (c->ee_sigma * c->ee_sigma));
    }
    TG_TRY(dev_alloc(c, &e->gtab, gt.size()));
    TG_TRY(dev_alloc(c, &e->ep, (size_t)P));
    TG_TRY(dev_alloc(c, &e->amp, (size_t)P));
    TG_TRY(dev_alloc(c, &e->node_of_p, (size_t)P));
    TG_TRY(dev_alloc(c, &e->p_of_node, (size_t)c->n));
    TG_TRY(dev_alloc(c, &e->coords, (size_t)P));
    TG_TRY(dev_alloc(c, &e->fp, (size_t)P * e->cap));
    TG_TRY(dev_alloc(c, &e->fexp, (size_t)e->cap));
    TG_TRY(dev_alloc(c, &e->part, (size_t)P * e->ntile * e->cap));
    TG_HIP(c, hipMemcpy(e->gtab, gt.data(), sizeof(double) * gt.size(), hipMemcpyHostToDevice));
    TG_HIP(c, hipMemcpy(e->ep, ep.data(), sizeof(double) * P, hipMemcpyHostToDevice));
    TG_HIP(c, hipMemcpy(e->amp, amp.data(), sizeof(double) * P, hipMemcpyHostToDevice));
    TG_HIP(c, hipMemcpy(e->node_of_p, node_of_p.data(), sizeof(int32_t) * P, hipMemcpyHostToDevice));
    TG_HIP(c, hipMemcpy(e->p_of_node, p_of_node.data(), sizeof(int32_t) * c->n, hipMemcpyHostToDevice));
    TG_HIP(c, hipMemcpy(e->coords, coords.data(), sizeof(int32_t) * P, hipMemcpyHostToDevice));
    return TG_OK;
}

void ee_destroy(tg_ctx* c) {
    EeData* e = ee_of(c);
    if (!e) return;
    dev_free(&e->gtab);
    dev_free(&e->ep);
    dev_free(&e->amp);
    dev_free(&e->node_of_p);
    dev_free(&e->p_of_node);
    dev_free(&e->coords);
    dev_free(&e->fp);
    dev_free(&e->fexp);
    dev_free(&e->part);
    dev_free(&e->mup);
    dev_free(&e->mexp);
    delete e;
    c->ee = nullptr;
}

// d_I[b] += I_ee(d_f[b]) for the B members (caller order, member-major);
// kernel launches only; members with gate[b] == 0 are skipped (gate nullable)
tg_status ee_apply_device(tg_ctx* c, const double* d_f, double* d_I, int32_t B, const double* gate) {
    EeData* e = ee_of(c);
    if (!e) return TG_OK;
    if (B < 1 || B > e->cap) return set_error(c, TG_ERR_ARG, "e-e channel: batch %d outside [1, %d]", B, e->cap);
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(ee_prep_kernel, dim3(B), dim3(kEeT), 0, st, d_f, c->n, e->P, e->node_of_p, e->w,
                       e->fp, e->fexp, gate);
    TG_CHECK_LAUNCH(c);
    const EeLattice L{e->dim, e->ns[0], e->ns[1], e->ns[2]};
    const size_t lds = sizeof(double) * ((size_t)e->M + 2 * e->J + kEeConvR);
#define TG_EE_MAIN(ACC)                                                                              \
    hipLaunchKernelGGL((ee_main_kernel<ACC, 0>), dim3(e->ntile, B), dim3(kEeT), lds, st, L, (int)e->P, e->Mh, \
                       e->M, e->J, 1.0 / e->dw, e->w, e->gtab, e->ep, e->coords, e->amp, e->fp, e->fexp,   \
                       e->ntile, gate, nullptr, e->part)
    if (e->P <= 8 * kEeT) TG_EE_MAIN(8);
    else if (e->P <= 16 * kEeT) TG_EE_MAIN(16);
    else TG_EE_MAIN(32);
#undef TG_EE_MAIN
    TG_CHECK_LAUNCH(c);
    hipLaunchKernelGGL(ee_reduce_kernel, dim3(ceil_div(c->n, 256), B), dim3(256), 0, st, e->part,
                       e->ntile, e->P, e->p_of_node, c->n, gate, d_I);
    TG_CHECK_LAUNCH(c);
    return TG_OK;
}

// d_out += J_ee(f)^T mu for one member (nodes, FP64); kernel launches only
tg_status ee_vjp_device(tg_ctx* c, const double* d_f, const double* d_mu, double* d_out) {
    EeData* e = ee_of(c);
    if (!e) return TG_OK;
    hipStream_t st = c->stream;
    if (!e->mup) {
        TG_TRY(dev_alloc(c, &e->mup, (size_t)e->P));
        TG_TRY(dev_alloc(c, &e->mexp, 1));
    }
    hipLaunchKernelGGL(ee_prep_kernel, dim3(1), dim3(kEeT), 0, st, d_f, c->n, e->P, e->node_of_p, e->w,
                       e->fp, e->fexp, nullptr);
    TG_CHECK_LAUNCH(c);
    const EeLattice L{e->dim, e->ns[0], e->ns[1], e->ns[2]};
    hipLaunchKernelGGL(ee_vjp_prep_kernel, dim3(1), dim3(kEeT), 0, st, L, e->P, d_mu, e->fp, e->w, e->mup,
                       e->mexp);
    TG_CHECK_LAUNCH(c);
    const size_t lds = sizeof(double) * ((size_t)e->M + 2 * e->J + kEeConvR);
#define TG_EE_VJP(ACC, MODE, EXP)                                                                     \
    hipLaunchKernelGGL((ee_main_kernel<ACC, MODE>), dim3(e->ntile, 1), dim3(kEeT), lds, st, L, (int)e->P, \
                       e->Mh, e->M, e->J, 1.0 / e->dw, e->w, e->gtab, e->ep, e->coords, e->amp, e->fp,   \
                       EXP, e->ntile, nullptr, e->mup, e->part)
#define TG_EE_VJP_PASS(MODE, EXP)                                                                     \
    if (e->P <= 8 * kEeT) TG_EE_VJP(8, MODE, EXP);                                                    \
    else if (e->P <= 16 * kEeT) TG_EE_VJP(16, MODE, EXP);                                             \
    else TG_EE_VJP(32, MODE, EXP);                                                                     \
    TG_CHECK_LAUNCH(c);                                                                                 \
    hipLaunchKernelGGL(ee_vjp_reduce_kernel, dim3(ceil_div(e->P, 256)), dim3(256), 0, st, e->part,      \
                       e->ntile, e->P, e->node_of_p, d_out);                                             \
    TG_CHECK_LAUNCH(c);
    TG_EE_VJP_PASS(1, e->fexp)
    TG_EE_VJP_PASS(2, e->mexp)
#undef TG_EE_VJP_PASS
#undef TG_EE_VJP
    return TG_OK;
}

}  // namespace tg
