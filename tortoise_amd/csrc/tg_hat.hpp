// tg_hat.hpp -- hat-averaged occupation factors of the nodal W-form.
//
// The W-form's gain term at node i carries the Pauli factor (1 - x_i) and its
// loss term x_i.  With x = f (the nodal value) the P1 product of the test
// function and the occupation, int phi_i (1 - f(k)) ..., is lumped at the
// node; the reference integrates it exactly per element tuple
// (include/Physics/Scattering/IntegratorCore.hpp:298-316).  With
//   x_i = fbar_i = int phi_i f / int phi_i
// (the consistent P1 mass over the lumped one) each test function keeps its
// hat's mean occupation: on a simplex int phi_a phi_b = vol (1 + delta_ab) /
// ((d+1)(d+2)) and int phi_a = vol / (d+1), so
//   fbar_i = sum_{e ni i} sum_{b in e} (1 + delta_ib) f_b / ((d+2) |star(i)|).
// The stencil is stored with a fixed number of slots per node (7 in 2D, 15 in
// 3D), each a 32-bit word (neighbour << 6) | multiplicity, where the
// multiplicity counts the elements sharing the edge (twice the star size for
// the node itself) and is 0 for unused slots (which point at the node).  The
// denominator (d+2)|star(i)| is the sum of the multiplicities.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "tg_mesh.hpp"

namespace tg {

__host__ __device__ constexpr int hat_slots(int dim) { return dim == 2 ? 7 : 15; }
constexpr int kHatMultBits = 6;

// fbar_i from the packed stencil row of node i; load(j) returns f_j (LDS or
// global).  The products are summed in slot order, the same in every kernel.
template <int S, typename Load>
__device__ __forceinline__ double hat_value(const uint32_t* __restrict__ row, Load load) {
#pragma clang fp contract(off)
    uint32_t w[S];
#pragma unroll
    for (int k = 0; k < S; ++k) w[k] = row[k];
    double acc = 0.0;
    uint32_t den = 0;
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const uint32_t m = w[k] & ((1u << kHatMultBits) - 1u);
        acc += (double)m * load((int64_t)(w[k] >> kHatMultBits));
        den += m;
    }
    return acc / (double)den;
}

// runtime slot count (2D or 3D)
template <typename Load>
__device__ __forceinline__ double hat_value_dim(int dim, const uint32_t* __restrict__ hs, int64_t i, Load load) {
    return dim == 2 ? hat_value<7>(hs + i * 7, load) : hat_value<15>(hs + i * 15, load);
}

// the multiplicity sum (d+2)|star(i)| of node i
__device__ __forceinline__ uint32_t hat_den(int dim, const uint32_t* __restrict__ hs, int64_t i) {
    const int S = hat_slots(dim);
    uint32_t den = 0;
    for (int k = 0; k < S; ++k) den += hs[i * S + k] & ((1u << kHatMultBits) - 1u);
    return den;
}

// Number-defect weight of the hat form.  The hat factors break the pairwise
// symmetry that makes the nodal W-form conserve the particle number, so with
// TG_CONSERVE the defect d = sum_i V_i I_i is removed as I_i -= s w_i,
// s = d / sum_i V_i w_i, with w_i the gross rate node i would have at its
// own occupation x_i if every other node were half filled:
//   w_i = ((1 - x_i) (W 1)_i + x_i (W^T 1)_i) / V_i,  W = (n+1) K + n K^T.
// It is linear in the state (no kinks for the DP5), positive for x_i in
// [0,1] and well beyond, and scales with the node's rates, so every node's
// rates move by a similar small fraction.  Weights that vanish with
// x_i (1 - x_i) concentrate the removal where the state is partially
// occupied; on distributions the mesh does not resolve that makes s large
// and the DP5 stiff (steps underflow on the 128 x 128-section ensemble).
// rk = K 1, ck = K^T 1.
__host__ __device__ inline double hat_defect_weight(double x, double np1, double nb, double rk, double ck,
                                                    double V) {
#pragma clang fp contract(off)
    return ((1.0 - x) * (np1 * rk + nb * ck) + x * (np1 * ck + nb * rk)) / V;
}

// Builds the packed stencil of the structured mesh (host); false when the
// node count does not fit the 26-bit neighbour field.
inline bool hat_build(const MeshDims& md, int64_t n, int64_t nel, const std::vector<int32_t>& conn,
                      std::vector<uint32_t>& out) {
    const int dim = md.dim, nv = dim + 1, S = hat_slots(dim);
    if (n >= ((int64_t)1 << (32 - kHatMultBits))) return false;
    std::vector<int32_t> nb((size_t)n * S, -1);
    std::vector<uint32_t> mult((size_t)n * S, 0u);
    // the node itself takes slot 0
    for (int64_t i = 0; i < n; ++i) nb[(size_t)i * S] = (int32_t)i;
    for (int64_t e = 0; e < nel; ++e)
        for (int a = 0; a < nv; ++a) {
            const int32_t i = conn[e * nv + a];
            for (int b = 0; b < nv; ++b) {
                const int32_t m = conn[e * nv + b];
                int k = 0;
                while (k < S && nb[(size_t)i * S + k] != m && nb[(size_t)i * S + k] >= 0) ++k;
                if (k == S) return false;
                nb[(size_t)i * S + k] = m;
                mult[(size_t)i * S + k] += (b == a) ? 2u : 1u;
            }
        }
    out.assign((size_t)n * S, 0u);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < S; ++k) {
            const int32_t j = nb[(size_t)i * S + k];
            const uint32_t m = mult[(size_t)i * S + k];
            if (m >= (1u << kHatMultBits)) return false;
            out[(size_t)i * S + k] = ((uint32_t)(j >= 0 ? j : (int32_t)i) << kHatMultBits) | m;
        }
    return true;
}

}  // namespace tg
