// tg_kentry.hpp -- entries of the emission kernel K computed from the
// projection data, for the kernels that build K's bands directly (no n x n
// array is ever stored), and the fp16 hi/lo split of the batched path.
//
//   K_ij = amp sum_b R_i(b) Q_j(b),   Q_j = R_j (*) g  (tg_assemble.hip)
//
// over the overlap of node i's hat window [off_i, off_i + len_i) and node
// j's convolved window [off_j + m_lo, off_j + m_lo + len_j + nm - 1), b
// ascending: every consumer gets the same FP64 value for the same (i, j).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tg {

struct KSrc {
    const double* R;      // n x maxw projections (node order)
    const double* Q;      // n x qw convolved projections
    const int32_t* off;   // per node: first hat of its window
    const int32_t* len;   // per node: window length
    int32_t maxw, qw, m_lo, nm;
    double amp;           // g^2
};

// Ri: node i's projection window (its R row, in global memory or LDS)
__device__ inline double k_entry_w(const KSrc& k, const double* Ri, int offi, int leni, int64_t j) {
    const int qoff = k.off[j] + k.m_lo;
    const int qlen = k.len[j] + k.nm - 1;
    const int b0 = offi > qoff ? offi : qoff;
    const int b1 = (offi + leni < qoff + qlen) ? offi + leni : qoff + qlen;
    const double* Qj = k.Q + j * (int64_t)k.qw;
    double s = 0.0;
    for (int b = b0; b < b1; ++b) s += Ri[b - offi] * Qj[b - qoff];
    return s * k.amp;
}

__device__ inline double k_entry(const KSrc& k, int64_t i, int64_t j) {
    return k_entry_w(k, k.R + i * (int64_t)k.maxw, k.off[i], k.len[i], j);
}

// the windows of nodes i and j overlap (else K_ij = 0 exactly)
__device__ inline bool k_windows_meet(const KSrc& k, int offi, int leni, int64_t j) {
    const int qoff = k.off[j] + k.m_lo;
    const int qlen = k.len[j] + k.nm - 1;
    return qoff < offi + leni && offi < qoff + qlen;
}

// FP64 -> fp16 through FP32 with the two hardware conversions.  The empty
// asm statement keeps the FP32 value opaque: without it the compiler folds
// the two conversions into one correctly rounded double -> half conversion,
// which gfx950 has no instruction for and which it emulates with ~40 integer
// operations.
__device__ inline _Float16 to_fp16(double s) {
    float t = (float)s;
    __asm__ volatile("" : "+v"(t));
    return (_Float16)t;
}

// hi = s rounded to fp16 through FP32; a double rounding tie moves hi by one
// fp16 ulp, which the exact FP64 residual s - hi absorbs, so hi + lo keeps
// its ~22 significant bits
__device__ inline void split_fp16(double x, double scale, _Float16& hi, _Float16& lo) {
    const double s = x * scale;
    hi = to_fp16(s);
    lo = to_fp16(s - (double)hi);
}

// the split of x at this scale has a nonzero bit pattern (the packed band
// holds the units of such entries)
__device__ inline bool split_nonzero(double x, double scale) {
    _Float16 hi, lo;
    split_fp16(x, scale, hi, lo);
    return (__builtin_bit_cast(uint16_t, hi) | __builtin_bit_cast(uint16_t, lo)) != 0;
}

}  // namespace tg
