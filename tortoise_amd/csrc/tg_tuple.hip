// tg_tuple.hip -- the reference's own scattering operator on the GPU: element
// tuples of the e-ph channel {el in, pn in, el out}, the Monte-Carlo
// delta-shell integral per tuple, and the per-stage propagator over DG
// (linear-form) element coefficients.
//
// Reference: ScatteringChannel<NDim,3>::precalculateIndices
// (src/Physics/Scattering/ScatteringChannel.cpp:421-538), ::propagate
// (:618-670) with scatteringIntegrationTypeB
// (include/Physics/Scattering/IntegratorCore.hpp:260-357), the MC point matrix
// (ScatteringChannel.cpp:677-688) and Function::applyInverseMass
// (src/Geometry/Structured/FunctionSpace/Function.cpp:801-804).
//
// Design
//   * Tuple list, built on the GPU in three passes over (e0, e2) element
//     pairs (one workgroup per e0, e2 in chunks of the workgroup): energy
//     window of the pair, umklapp vectors, phonon elements whose box meets
//     k2 - k0 - G u, |det D_C| and the MC screening value.  Counting, then
//     filling in (e0, e2, umklapp, e1) order, so the list is deterministic.
//   * The MC point matrix is fixed between redraws, so everything that does
//     not depend on the state is integrated once per redraw: with a phonon
//     bath of occupation n the integrand is
//         (1+n) f(k2) - n f(k0) - f(k0) f(k2),
//     bilinear in the two electron elements' coefficients c0, c2.  Per tuple
//     and output basis j of leg l the moments
//         Q_l[j][a][b] = sum_m w phi_j(x_l) phi_a(x_0) phi_b(x_2)
//     (phi_0 = 1, phi_j = x_{j-1}, w = prod(vol*dim!)/divisor) carry the whole
//     operator.  Each tuple is stored twice, once in the list of its leg-0
//     element and once in the list of its leg-2 element, with that leg's
//     moments; basis 0 of both is the same integral.
//   * Per stage one workgroup per element sums its two lists in a fixed order
//     (deterministic, no atomics), applies the inverse mass and writes the
//     element's coefficients for every member.  The moments are read once per
//     member chunk; the other element's coefficients are L2-resident gathers.
//   * The geometry (D, N, D_C^-1, constrained coordinates, inside tests) is
//     evaluated with contraction off and the same operation order as the CPU
//     restatement, so both make the same inside/outside decision per point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <cstdio>
#include <new>
#include <vector>

#include "../../include/tortoise_gpu.h"
#include "tg_dp5_kernels.hpp"
#include "tg_mesh.hpp"

struct tg_tuple_ctx {
    int dim = 2;
    int32_t nsec[3] = {1, 1, 1};
    double origin[3] = {0, 0, 0};
    double extent[3] = {1, 1, 1};
    double dG[3][3] = {}, dGinv[3][3] = {}, zone[3][3] = {};   // Geo's affine mesh and zone
    int32_t axis = 1;
    double gen_vol = 0;         // |det dG| / dim!
    int64_t n = 0, nel = 0;
    double vol = 0;
    double omega0 = 0, amplitude = 1;
    int32_t nu = 1, nmc = 50, rows = 4;
    double min_det = 1e-9, threshold = 1e-9;
    int32_t max_batch = 1;

    double* d_eps = nullptr;     // nodal electron dispersion
    double* d_pts = nullptr;     // MC point matrix [rows][nmc]
    double* d_etab = nullptr;    // element table (forms, boxes, energy ranges)
    double minv[16] = {};        // LF inverse mass of one element (row-major NC x NC)
    double* d_minv = nullptr;

    int64_t ntup = 0;            // tuples (e0, e1, e2, uid)
    int32_t* d_tup = nullptr;    // [ntup][4], ordered by (e0, e2, uid, e1)
    int64_t* d_off0 = nullptr;   // tuples of element e as leg 0: [off0[e], off0[e+1])
    int64_t* d_off2 = nullptr;   // ... as leg 2, positions in the leg-2 list
    int32_t* d_oth0 = nullptr;   // leg-0 list: the tuple's e2
    int32_t* d_oth2 = nullptr;   // leg-2 list: the tuple's e0
    int64_t* d_pos2 = nullptr;   // tuple -> position in the leg-2 list
    double* d_q0 = nullptr;      // leg-0 moments, [nq][ntup]
    double* d_q2 = nullptr;      // leg-2 moments, [nq][ntup]
    bool have_points = false;

    // dynamical phonon band: three leg lists with two other elements each and
    // three blocks of moments per entry
    bool dyn = false;
    double* d_eps_ph = nullptr;  // phonon nodal energies (dispersive band), or null
    double pmin = 0, pmax = 0;
    int64_t* d_doff[3] = {};
    int32_t* d_dothA[3] = {};
    int32_t* d_dothB[3] = {};
    int64_t* d_dpos[3] = {};
    double* d_dq[3] = {};

    double* d_f = nullptr;       // staging [max_batch][nel][dim+1]
    double* d_ft = nullptr;      // element-major copy of the states [nel][batch][dim+1]
    double* d_out = nullptr;
    double* d_bath = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.f;
    bool time_kernel = false;

    void* dp5 = nullptr;         // DP5 work buffers (tg_tuple_dp5)

    // 4-leg channel (tg_tuple4_create): tuples [ntup][5] in d_tup, the
    // element entries (tuple << 2 | leg) from d_goff[e], results per
    // (tuple, member)
    int nlegs = 3;
    double kappa = 0;
    int64_t* d_goff = nullptr;
    int32_t* d_gent = nullptr;
    double* d_res4 = nullptr;

    // time-dependent bath (tg_tuple_set_bath_schedule): batch x ns samples
    std::vector<double> h_bsched;
    double* d_bsched = nullptr;
    int32_t bs_ns = 0, bs_batch = 0;
    double bs_t0 = 0, bs_dt = 0, bs_atol = 0, bs_rtol = 0;

    hipStream_t stream = nullptr;
    int device = 0;
    char err[512] = {0};
};

namespace {

constexpr int kEnumThreads = 256;
constexpr int kApplyThreads = 256;
constexpr int kApplyMembers = 8;       // members per workgroup when batch > 1

tg_status fail(tg_tuple_ctx* c, tg_status code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
tg_status fail(tg_tuple_ctx* c, tg_status code, const char* fmt, ...) {
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof(c->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define TT_HIP(ctx, call)                                                                  \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess)                                                              \
            return fail((ctx), TG_ERR_HIP, "%s failed: %s (%s:%d)", #call,                 \
                        hipGetErrorString(_e), __FILE__, __LINE__);                        \
    } while (0)

template <typename T>
tg_status dalloc(tg_tuple_ctx* c, T** p, size_t count) {
    if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (count == 0) count = 1;
    if (hipMalloc((void**)p, sizeof(T) * count) != hipSuccess) {
        *p = nullptr;
        return fail(c, TG_ERR_NOMEM, "hipMalloc of %zu bytes failed", sizeof(T) * count);
    }
    return TG_OK;
}

// ------------------------------------------------------------- geometry
struct Geo {
    int32_t nsec[3];
    double origin[3], extent[3];
    // affine mesh: node index i at origin + sum_c dG[.][c] i_c (dG[d][c]:
    // component d of side c over nsec[c]); umklapp u at sum_c u_c zone[.][c];
    // axis: sides along the axes (the box origin + extent, dG diagonal)
    double dG[3][3], dGinv[3][3], zone[3][3];
    int32_t axis;
    double omega0, vol;
    double pmin, pmax;    // phonon energy range over the zone
    int32_t nu, nuv, nmc, rows;
    const double* eps;
    const double* eps_ph; // phonon nodal energies, or null: the constant omega0
    const double* pts;
    const double* etab;   // per element: momentum forms, energy forms, box, energy ranges
};

// Element table record: DIM*(DIM+1) momentum form coefficients, DIM+1
// electron energy form coefficients, DIM box minima, DIM box maxima, the
// electron energy min and max, DIM+1 phonon energy form coefficients, the
// phonon energy min and max.
template <int DIM>
constexpr int etab_stride() { return DIM * (DIM + 1) + (DIM + 1) + 2 * DIM + 2 + (DIM + 1) + 2; }
template <int DIM>
constexpr int etab_ph() { return DIM * (DIM + 1) + (DIM + 1) + 2 * DIM + 2; }

// Everything from here to the moment kernel decides which tuples and which MC
// points count, so it is evaluated without contraction and in the operation
// order of the CPU restatement.
#pragma clang fp contract(off)

// coordinate d of a node: origin + sum_c dG[d][c] i_c, the terms added in
// axis order (on an axis-aligned mesh the other axes add exact zeros: the
// box value origin + (extent / nsec) i)
template <int DIM>
__host__ __device__ inline double node_coord(const Geo& g, int64_t node, int d) {
    int64_t r = node;
    double x = g.origin[d];
    for (int c = 0; c < DIM; ++c) {
        const int64_t id = r % (g.nsec[c] + 1);
        r /= (g.nsec[c] + 1);
        x = x + g.dG[d][c] * (double)id;
    }
    return x;
}

// the umklapp vector's component d: sum_c u_c zone[d][c]
template <int DIM>
__host__ __device__ inline double umklapp_comp(const Geo& g, const int u[3], int d) {
    double s = 0.0;
    for (int c = 0; c < DIM; ++c) s = s + (double)u[c] * g.zone[d][c];
    return s;
}

template <int DIM>
__device__ inline void elem_nodes(const Geo& g, int64_t e, int64_t v[DIM + 1]) {
    tg::MeshDims md{DIM, {g.nsec[0], g.nsec[1], g.nsec[2]}};
    int64_t w[4];
    tg::element_nodes(md, e, w);
#pragma unroll
    for (int a = 0; a <= DIM; ++a) v[a] = w[a];
}

// Linear forms of the momentum coordinates and the energy on element e:
// kf[d][0..DIM] = [k_d(v0), k_d(v1) - k_d(v0), ...] (ScatteringChannel.cpp:285-301).
template <int DIM>
__device__ inline void elem_forms_compute(const Geo& g, int64_t e, double kf[DIM][DIM + 1],
                                          double ef[DIM + 1]) {
    int64_t v[DIM + 1];
    elem_nodes<DIM>(g, e, v);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        double k0 = node_coord<DIM>(g, v[0], d);
        kf[d][0] = k0;
#pragma unroll
        for (int a = 1; a <= DIM; ++a) kf[d][a] = node_coord<DIM>(g, v[a], d) - k0;
    }
    double e0 = g.eps[v[0]];
    ef[0] = e0;
#pragma unroll
    for (int a = 1; a <= DIM; ++a) ef[a] = g.eps[v[a]] - e0;
}

// Forms of element e read back from the table; the phonon band has the same
// momentum forms and its own energy form.
template <int DIM>
__device__ inline void elem_forms(const Geo& g, int64_t e, bool phonon, double kf[DIM][DIM + 1],
                                  double ef[DIM + 1]) {
    const double* r = g.etab + e * etab_stride<DIM>();
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int a = 0; a <= DIM; ++a) kf[d][a] = r[d * (DIM + 1) + a];
#pragma unroll
    for (int a = 0; a <= DIM; ++a) ef[a] = phonon ? r[etab_ph<DIM>() + a] : r[DIM * (DIM + 1) + a];
}

template <int DIM>
__host__ __device__ inline void umklapp_vec(int32_t nu, int32_t uid, int u[3]) {
    int span = 2 * nu + 1;
    u[0] = u[1] = u[2] = 0;
#pragma unroll
    for (int d = DIM - 1; d >= 0; --d) { u[d] = uid % span - nu; uid /= span; }
}

// D ((DIM+1) x 3DIM) and N of the tuple {e0 in, e1 in, e2 out}
// (ScatteringChannel.cpp:285-352).
template <int DIM>
__device__ inline void tuple_DN(const Geo& g, const int64_t el[3], const int u[3], double D[DIM + 1][3 * DIM],
                                double N[DIM + 1]) {
    const double dir[3] = {1.0, 1.0, -1.0};
#pragma unroll
    for (int r = 0; r <= DIM; ++r) N[r] = 0.0;
#pragma unroll
    for (int leg = 0; leg < 3; ++leg) {
        double kf[DIM][DIM + 1], ef[DIM + 1];
        elem_forms<DIM>(g, el[leg], leg == 1, kf, ef);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
#pragma unroll
            for (int a = 0; a < DIM; ++a) D[d][leg * DIM + a] = dir[leg] * kf[d][1 + a];
            N[d] -= dir[leg] * kf[d][0];
        }
#pragma unroll
        for (int a = 0; a < DIM; ++a) D[DIM][leg * DIM + a] = dir[leg] * ef[1 + a];
        N[DIM] -= dir[leg] * ef[0];
    }
    // N_d -= sum_c u_c zone[d][c], one term at a time (ScatteringChannel.cpp:330-334)
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int c = 0; c < DIM; ++c) N[d] -= (double)u[c] * g.zone[d][c];
}

// Inverse of the NC x NC matrix by Gauss-Jordan with partial pivoting;
// returns the determinant, 0 when singular.  Row swaps are done by selects
// so the arrays stay in registers.
template <int NC>
__host__ __device__ inline double invert_small(const double A[NC][NC], double inv[NC][NC]) {
    double a[NC][2 * NC];
#pragma unroll
    for (int r = 0; r < NC; ++r)
#pragma unroll
        for (int c = 0; c < 2 * NC; ++c) a[r][c] = c < NC ? A[r][c] : (c - NC == r ? 1.0 : 0.0);
    double det = 1.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        int p = c;
        double best = fabs(a[c][c]);
#pragma unroll
        for (int r = c + 1; r < NC; ++r) {
            double v = fabs(a[r][c]);
            if (v > best) { best = v; p = r; }
        }
        if (best == 0.0) return 0.0;
        if (p != c) {
#pragma unroll
            for (int r = c + 1; r < NC; ++r)
                if (r == p)
#pragma unroll
                    for (int q = 0; q < 2 * NC; ++q) { double t = a[r][q]; a[r][q] = a[c][q]; a[c][q] = t; }
            det = -det;
        }
        double piv = a[c][c];
        det *= piv;
#pragma unroll
        for (int q = 0; q < 2 * NC; ++q) a[c][q] /= piv;
#pragma unroll
        for (int r = 0; r < NC; ++r) {
            if (r == c) continue;
            double m = a[r][c];
#pragma unroll
            for (int q = 0; q < 2 * NC; ++q) a[r][q] -= m * a[c][q];
        }
    }
#pragma unroll
    for (int r = 0; r < NC; ++r)
#pragma unroll
        for (int c = 0; c < NC; ++c) inv[r][c] = a[r][NC + c];
    return det;
}

template <int DIM>
__device__ inline double dc_inverse(const double D[DIM + 1][3 * DIM], double inv[DIM + 1][DIM + 1]) {
    constexpr int NC = DIM + 1, C0 = 2 * DIM - 1;
    double DC[NC][NC];
#pragma unroll
    for (int r = 0; r < NC; ++r)
#pragma unroll
        for (int q = 0; q < NC; ++q) DC[r][q] = D[r][C0 + q];
    return invert_small<NC>(DC, inv);
}

// Legs' points of MC sample m (IntegratorCore.hpp:283-320) and the inside
// tests of the two constrained legs.
template <int DIM>
__device__ inline bool shell_point(const Geo& g, const double D[DIM + 1][3 * DIM], const double N[DIM + 1],
                                   const double inv[DIM + 1][DIM + 1], int32_t m, double K[3][DIM]) {
    constexpr int NC = DIM + 1, C0 = 2 * DIM - 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) K[0][d] = g.pts[(size_t)d * g.nmc + m];
#pragma unroll
    for (int d = 0; d < DIM - 1; ++d) K[1][d] = g.pts[(size_t)(g.rows - (DIM - 1) + d) * g.nmc + m];
    double rhs[NC];
#pragma unroll
    for (int r = 0; r < NC; ++r) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < C0; ++c) s += D[r][c] * K[c / DIM][c % DIM];
        rhs[r] = s - N[r];
    }
#pragma unroll
    for (int r = 0; r < NC; ++r) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) s += inv[r][c] * rhs[c];
        if (r == 0) K[1][DIM - 1] = -s;
        else K[2][r - 1] = -s;
    }
    bool inside = true;
#pragma unroll
    for (int leg = 1; leg < 3; ++leg) {
        double sum = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) { inside &= K[leg][d] > 0.0; sum += K[leg][d]; }
        inside &= sum < 1.0;
    }
    return inside;
}

__host__ __device__ inline double shell_divisor(int dim, double det, int32_t nmc) {
    double w = fabs(det);
    w *= (double)nmc * (double)tg::factorial(dim - 1);
    w *= (double)tg::factorial(dim);
    return w;
}

// Screening (ScatteringChannel.cpp:474-503 with the channel's points): true
// when |det D_C| >= min_det and hits * c^3 * prod(vol*dim!) / divisor exceeds
// the threshold; the value grows with the hit count, so the first passing
// count decides.
template <int DIM>
__device__ inline bool tuple_passes(const Geo& g, const double D[DIM + 1][3 * DIM], const double N[DIM + 1],
                                    double min_det, double threshold) {
    double inv[DIM + 1][DIM + 1];
    double det = dc_inverse<DIM>(D, inv);
    if (!(fabs(det) >= min_det)) return false;
    double c = sqrt(2.0);
    double cpow = c * c * c;
    double amp = 1.0;
    for (int leg = 0; leg < 3; ++leg) amp *= g.vol * (double)tg::factorial(DIM);
    double w = shell_divisor(DIM, det, g.nmc);
    int64_t hits = 0;
    for (int32_t m = 0; m < g.nmc; ++m) {
        double K[3][DIM];
        if (!shell_point<DIM>(g, D, N, inv, m, K)) continue;
        ++hits;
        if ((double)hits * cpow * amp / w > threshold) return true;
    }
    return false;
}

template <int DIM>
__device__ inline void elem_box_erange(const Geo& g, int64_t e, double lo[DIM], double hi[DIM], double* elo,
                                       double* ehi) {
    const double* r = g.etab + e * etab_stride<DIM>() + DIM * (DIM + 1) + (DIM + 1);
#pragma unroll
    for (int d = 0; d < DIM; ++d) { lo[d] = r[d]; hi[d] = r[DIM + d]; }
    *elo = r[2 * DIM];
    *ehi = r[2 * DIM + 1];
}

template <int DIM>
__device__ inline void elem_erange(const Geo& g, int64_t e, double* elo, double* ehi) {
    const double* r = g.etab + e * etab_stride<DIM>() + DIM * (DIM + 1) + (DIM + 1) + 2 * DIM;
    *elo = r[0];
    *ehi = r[1];
}

template <int DIM>
__device__ inline void elem_prange(const Geo& g, int64_t e, double* plo, double* phi) {
    const double* r = g.etab + e * etab_stride<DIM>() + etab_ph<DIM>() + DIM + 1;
    *plo = r[0];
    *phi = r[1];
}

// One thread per element: the table record (forms, bounding box, energy range).
template <int DIM>
__global__ void elem_table_kernel(Geo g, int64_t nel, double* __restrict__ tab) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nel) return;
    double kf[DIM][DIM + 1], ef[DIM + 1];
    elem_forms_compute<DIM>(g, e, kf, ef);
    double* r = tab + e * etab_stride<DIM>();
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int a = 0; a <= DIM; ++a) r[d * (DIM + 1) + a] = kf[d][a];
#pragma unroll
    for (int a = 0; a <= DIM; ++a) r[DIM * (DIM + 1) + a] = ef[a];
    int64_t v[DIM + 1];
    elem_nodes<DIM>(g, e, v);
    double lo[DIM], hi[DIM], elo = 1e300, ehi = -1e300;
#pragma unroll
    for (int d = 0; d < DIM; ++d) { lo[d] = 1e300; hi[d] = -1e300; }
#pragma unroll
    for (int a = 0; a <= DIM; ++a) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            double x = node_coord<DIM>(g, v[a], d);
            lo[d] = x < lo[d] ? x : lo[d];
            hi[d] = x > hi[d] ? x : hi[d];
        }
        double en = g.eps[v[a]];
        elo = en < elo ? en : elo;
        ehi = en > ehi ? en : ehi;
    }
    double* b = r + DIM * (DIM + 1) + (DIM + 1);
#pragma unroll
    for (int d = 0; d < DIM; ++d) { b[d] = lo[d]; b[DIM + d] = hi[d]; }
    b[2 * DIM] = elo;
    b[2 * DIM + 1] = ehi;
    // phonon energy form and range (a P1 band, or the constant omega0)
    double* q = r + etab_ph<DIM>();
    if (g.eps_ph) {
        const double p0 = g.eps_ph[v[0]];
        q[0] = p0;
        double plo = 1e300, phi = -1e300;
#pragma unroll
        for (int a = 0; a <= DIM; ++a) {
            if (a) q[a] = g.eps_ph[v[a]] - p0;
            const double en = g.eps_ph[v[a]];
            plo = en < plo ? en : plo;
            phi = en > phi ? en : phi;
        }
        q[DIM + 1] = plo;
        q[DIM + 2] = phi;
    } else {
        q[0] = g.omega0;
#pragma unroll
        for (int a = 1; a <= DIM; ++a) q[a] = 0.0;
        q[DIM + 1] = g.omega0;
        q[DIM + 2] = g.omega0;
    }
}

// Walks the tuples of the pair (e0, e2) in (uid, e1) order; calls emit(e1, uid)
// for each kept one.  Returns the count.
template <int DIM, typename Emit>
__device__ inline int32_t pair_tuples(const Geo& g, int64_t e0, int64_t e2, double min_det, double threshold,
                                      Emit emit) {
    constexpr int NES = DIM == 2 ? 2 : 6;
    double lo0[DIM], hi0[DIM], lo2[DIM], hi2[DIM], el0, eh0, el2, eh2;
    elem_erange<DIM>(g, e0, &el0, &eh0);
    elem_erange<DIM>(g, e2, &el2, &eh2);
    // energy window of the three legs (ScatteringChannel.hpp:212-222): first
    // with the phonon band's whole range, then per phonon element
    if (!(el0 + g.pmin - eh2 <= 0.0 && eh0 + g.pmax - el2 >= 0.0)) return 0;
    elem_box_erange<DIM>(g, e0, lo0, hi0, &el0, &eh0);
    elem_box_erange<DIM>(g, e2, lo2, hi2, &el2, &eh2);
    int32_t cnt = 0;
    for (int32_t uid = 0; uid < g.nuv; ++uid) {
        int u[3];
        umklapp_vec<DIM>(g.nu, uid, u);
        double lo[DIM], hi[DIM];
        int slo[3] = {0, 0, 0}, shi[3] = {0, 0, 0};
        bool empty = false;
        // k1 = k2 - k0 - G u lies in the box [lo, hi]; the sections that can
        // hold it: on an axis-aligned mesh the box's own section range, else
        // the reduced coordinates of the box's corners, one section wider
        // (a superset: the element boxes and the tuple tests decide)
        double smin[3], smax[3];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const double U = umklapp_comp<DIM>(g, u, d);
            lo[d] = lo2[d] - hi0[d] - U;
            hi[d] = hi2[d] - lo0[d] - U;
            if (g.axis) {
                double h = g.extent[d] / g.nsec[d];
                smin[d] = floor((lo[d] - g.origin[d]) / h);
                smax[d] = floor((hi[d] - g.origin[d]) / h);
            } else {
                smin[d] = 1e300;
                smax[d] = -1e300;
            }
        }
        if (!g.axis) {
            for (int cm = 0; cm < (1 << DIM); ++cm) {
                double x[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) x[d] = ((cm >> d) & 1 ? hi[d] : lo[d]) - g.origin[d];
#pragma unroll
                for (int c = 0; c < DIM; ++c) {
                    double s = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) s = s + g.dGinv[c][d] * x[d];
                    smin[c] = s < smin[c] ? s : smin[c];
                    smax[c] = s > smax[c] ? s : smax[c];
                }
            }
#pragma unroll
            for (int c = 0; c < DIM; ++c) {
                smin[c] = floor(smin[c]) - 1.0;
                smax[c] = floor(smax[c]) + 1.0;
            }
        }
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const double a = smin[d], b = smax[d];
            empty |= b < 0.0 || a > g.nsec[d] - 1;
            slo[d] = a < 0.0 ? 0 : (a > g.nsec[d] - 1 ? g.nsec[d] - 1 : (int)a);
            shi[d] = b > g.nsec[d] - 1 ? g.nsec[d] - 1 : (b < 0.0 ? 0 : (int)b);
        }
        if (empty) continue;
        int sc[3] = {slo[0], slo[1], slo[2]};
        for (;;) {
            int64_t sid = 0, stride = 1;
#pragma unroll
            for (int d = 0; d < DIM; ++d) { sid += stride * sc[d]; stride *= g.nsec[d]; }
            for (int t = 0; t < NES; ++t) {
                int64_t e1 = sid * NES + t;
                double lo1[DIM], hi1[DIM], a1, b1;
                elem_box_erange<DIM>(g, e1, lo1, hi1, &a1, &b1);
                bool meet = true;
#pragma unroll
                for (int d = 0; d < DIM; ++d) meet &= lo1[d] <= hi[d] && hi1[d] >= lo[d];
                if (!meet) continue;
                double pl1, ph1;
                elem_prange<DIM>(g, e1, &pl1, &ph1);
                if (!(el0 + pl1 - eh2 <= 0.0 && eh0 + ph1 - el2 >= 0.0)) continue;
                int64_t el[3] = {e0, e1, e2};
                double D[DIM + 1][3 * DIM], N[DIM + 1];
                tuple_DN<DIM>(g, el, u, D, N);
                if (!tuple_passes<DIM>(g, D, N, min_det, threshold)) continue;
                emit(e1, uid);
                ++cnt;
            }
            int d = 0;
            while (d < DIM && ++sc[d] > shi[d]) { sc[d] = slo[d]; ++d; }
            if (d == DIM) break;
        }
    }
    return cnt;
}

__device__ inline int64_t block_exclusive_scan(int64_t v, int64_t* sh, int64_t* total) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < kEnumThreads; off <<= 1) {
        int64_t x = tid >= off ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += x;
        __syncthreads();
    }
    int64_t incl = sh[tid];
    *total = sh[kEnumThreads - 1];
    __syncthreads();
    return incl - v;
}

// One workgroup per e0: tuple count of every e0 (write == false) or the
// tuples themselves at base[e0] in (e2, uid, e1) order (write == true).
template <int DIM>
__global__ __launch_bounds__(kEnumThreads) void enumerate_kernel(Geo g, int64_t nel, double min_det,
                                                                   double threshold, int64_t* counts,
                                                                   const int64_t* base, int32_t* tup,
                                                                   bool write) {
    __shared__ int64_t sh[kEnumThreads];
    const int64_t e0 = blockIdx.x;
    int64_t pos = write ? base[e0] : 0;
    int64_t total = 0;
    for (int64_t c0 = 0; c0 < nel; c0 += kEnumThreads) {
        const int64_t e2 = c0 + threadIdx.x;
        int32_t cnt = 0;
        if (e2 < nel) cnt = pair_tuples<DIM>(g, e0, e2, min_det, threshold, [](int64_t, int32_t) {});
        if (!write) {
            total += cnt;
            continue;
        }
        int64_t chunk_total;
        int64_t my = pos + block_exclusive_scan(cnt, sh, &chunk_total);
        if (cnt > 0)
            pair_tuples<DIM>(g, e0, e2, min_det, threshold, [&](int64_t e1, int32_t uid) {
                tup[my * 4 + 0] = (int32_t)e0;
                tup[my * 4 + 1] = (int32_t)e1;
                tup[my * 4 + 2] = (int32_t)e2;
                tup[my * 4 + 3] = uid;
                ++my;
            });
        pos += chunk_total;
    }
    if (!write) {
        int64_t sum_total;
        (void)block_exclusive_scan(total, sh, &sum_total);
        if (threadIdx.x == 0) counts[e0] = sum_total;
    }
}

// Moments of one tuple (2D: 3 bases; 3D: 4).  q0 / q2 are the leg-0 and leg-2
// lists, SoA [k][ntup] with k = (j*NC + a)*NC + b.  The MC points are the
// same for every lane, so their loads are uniform.
template <int DIM>
__global__ __launch_bounds__(256) void moments_kernel(Geo g, const int32_t* __restrict__ tup, int64_t ntup,
                                                        const int64_t* __restrict__ pos2,
                                                        double* __restrict__ q0, double* __restrict__ q2) {
    constexpr int NC = DIM + 1;
    constexpr int NQ = NC * NC * NC;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntup) return;
    const int64_t el[3] = {tup[t * 4], tup[t * 4 + 1], tup[t * 4 + 2]};
    int u[3];
    umklapp_vec<DIM>(g.nu, tup[t * 4 + 3], u);
    double D[DIM + 1][3 * DIM], N[DIM + 1], inv[DIM + 1][DIM + 1];
    tuple_DN<DIM>(g, el, u, D, N);
    double det = dc_inverse<DIM>(D, inv);
    double m0[NQ], m2[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) { m0[k] = 0.0; m2[k] = 0.0; }
    for (int32_t m = 0; m < g.nmc; ++m) {
        double K[3][DIM];
        if (!shell_point<DIM>(g, D, N, inv, m, K)) continue;
        double p0[NC], p2[NC];
        p0[0] = 1.0;
        p2[0] = 1.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) { p0[1 + d] = K[0][d]; p2[1 + d] = K[2][d]; }
#pragma unroll
        for (int a = 0; a < NC; ++a)
#pragma unroll
            for (int b = 0; b < NC; ++b) {
                double pab = p0[a] * p2[b];
#pragma unroll
                for (int j = 0; j < NC; ++j) {
                    m0[(j * NC + a) * NC + b] += p0[j] * pab;
                    m2[(j * NC + a) * NC + b] += p2[j] * pab;
                }
            }
    }
    double amp = 1.0;
    for (int leg = 0; leg < 3; ++leg) amp *= g.vol * (double)tg::factorial(DIM);
    double w = amp / shell_divisor(DIM, det, g.nmc);
    const int64_t p = pos2[t];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        q0[(int64_t)k * ntup + t] = m0[k] * w;
        q2[(int64_t)k * ntup + p] = m2[k] * w;
    }
}

// Dynamical phonons: the integrand F2 + P F2 - F0 P - F0 F2 (fermion
// electrons, boson phonons; the cubic terms of the two products cancel) is a
// sum of pairwise products, so per leg l three blocks of moments carry it:
//   block 0: phi_j(x_l) phi_a(x0) phi_b(x2)
//   block 1: phi_j(x_l) phi_a(x1) phi_b(x2)
//   block 2: phi_j(x_l) phi_a(x0) phi_b(x1)
// at k = ((block*NC + j)*NC + a)*NC + b of the leg's list, entry pos[t].
template <int DIM, int LEG>
__global__ __launch_bounds__(256) void moments_dyn_kernel(Geo g, const int32_t* __restrict__ tup, int64_t ntup,
                                                            const int64_t* __restrict__ pos, double* __restrict__ q) {
    constexpr int NC = DIM + 1;
    constexpr int NQ = NC * NC * NC;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntup) return;
    const int64_t el[3] = {tup[t * 4], tup[t * 4 + 1], tup[t * 4 + 2]};
    int u[3];
    umklapp_vec<DIM>(g.nu, tup[t * 4 + 3], u);
    double D[DIM + 1][3 * DIM], N[DIM + 1], inv[DIM + 1][DIM + 1];
    tuple_DN<DIM>(g, el, u, D, N);
    double det = dc_inverse<DIM>(D, inv);
    double m[3 * NQ];
#pragma unroll
    for (int k = 0; k < 3 * NQ; ++k) m[k] = 0.0;
    for (int32_t mm = 0; mm < g.nmc; ++mm) {
        double K[3][DIM];
        if (!shell_point<DIM>(g, D, N, inv, mm, K)) continue;
        double p[3][NC];
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            p[l][0] = 1.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) p[l][1 + d] = K[l][d];
        }
        constexpr int X[3] = {0, 1, 0}, Y[3] = {2, 2, 1};
#pragma unroll
        for (int blk = 0; blk < 3; ++blk)
#pragma unroll
            for (int a = 0; a < NC; ++a)
#pragma unroll
                for (int b = 0; b < NC; ++b) {
                    const double pab = p[X[blk]][a] * p[Y[blk]][b];
#pragma unroll
                    for (int j = 0; j < NC; ++j) m[((blk * NC + j) * NC + a) * NC + b] += p[LEG][j] * pab;
                }
    }
    double amp = 1.0;
    for (int leg = 0; leg < 3; ++leg) amp *= g.vol * (double)tg::factorial(DIM);
    const double w = amp / shell_divisor(DIM, det, g.nmc);
    const int64_t P = LEG == 0 ? t : pos[t];
#pragma unroll
    for (int k = 0; k < 3 * NQ; ++k) q[(int64_t)k * ntup + P] = m[k] * w;
}

#pragma clang fp contract(on)

// Leg-2 list: position of every tuple (stable counting sort on e2) and the
// other element of each entry.
__global__ void fill_leg_lists(const int32_t* __restrict__ tup, int64_t ntup, const int64_t* __restrict__ pos2,
                               int32_t* __restrict__ oth0, int32_t* __restrict__ oth2) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntup) return;
    oth0[t] = tup[t * 4 + 2];
    oth2[pos2[t]] = tup[t * 4];
}

// Propagator: one workgroup per (element, chunk of MB members).
//
// An element's entries (its leg-0 list, then its leg-2 list) are dealt to
// SL slots in rounds: entry i of a list goes to slot (i - first) % SL, thread
// h owns slot h and sums its entries in order for each member of the chunk.
// The element's result is the sum of the slots in slot order, then
// amplitude * Minv^T.  The order depends only on the lists, so a member of a
// batch is bitwise equal to its single evaluation.  A round's moments are
// staged in LDS (coalesced SoA reads, the next round's loads in flight during
// the current round's arithmetic); in 2D each thread then keeps its entry's
// moments in registers for all members of the chunk.  States are read from
// an element-major copy ft[e][member][a], so the chunk's members of one
// element are contiguous.  For an entry of the leg-0 list (sign +1)
// c0 = own coefficients, c2 = the other element's; for the leg-2 list (sign
// -1) the roles swap.  With
//   G[a][b] = (1+n) d_a0 c2_b - n c0_a d_b0 - c0_a c2_b
// the entry adds sign * sum_ab Q[j][a][b] G[a][b] to output basis j.
// Explicit fma with contraction off: every chunk size rounds identically.
#pragma clang fp contract(off)

template <int DIM>
constexpr int prop_slots() { return DIM == 2 ? 256 : 128; }

template <int DIM>
constexpr size_t prop_lds_bytes() {
    constexpr int NC = DIM + 1, NQ = NC * NC * NC, SL = prop_slots<DIM>();
    return sizeof(double) * (size_t)NQ * SL + sizeof(int32_t) * SL;
}

// The bilinear weight of an entry.  Propagator (RATES = false):
//   G[a][b] = (1+n) d_a0 c2_b - n c0_a d_b0 - c0_a c2_b.
// Scattering rates (RATES = true, ScatteringChannel.cpp:544-612 with the
// rate leg's forms replaced by constants; the factor 2 is applied at the end):
//   leg 0 (own = c0):  G[a][b] = d_a0 (n d_b0 + c2_b)
//   leg 2 (own = c2):  G[a][b] = d_b0 ((1+n) d_a0 - c0_a)
template <int DIM, bool RATES, typename QAt>
__device__ inline void entry_terms(QAt qat, bool leg0, const double c0[DIM + 1], const double c2[DIM + 1], double n,
                                   double np1, double out[DIM + 1]) {
    constexpr int NC = DIM + 1;
    double G[NC][NC];
    if constexpr (!RATES) {
#pragma unroll
        for (int a = 0; a < NC; ++a)
#pragma unroll
            for (int b = 0; b < NC; ++b) G[a][b] = -(c0[a] * c2[b]);
#pragma unroll
        for (int b = 0; b < NC; ++b) G[0][b] = fma(np1, c2[b], G[0][b]);
#pragma unroll
        for (int a = 0; a < NC; ++a) G[a][0] = fma(-n, c0[a], G[a][0]);
    } else {
#pragma unroll
        for (int a = 0; a < NC; ++a)
#pragma unroll
            for (int b = 0; b < NC; ++b) G[a][b] = 0.0;
        if (leg0) {
#pragma unroll
            for (int b = 0; b < NC; ++b) G[0][b] = c2[b];
            G[0][0] += n;
        } else {
#pragma unroll
            for (int a = 0; a < NC; ++a) G[a][0] = -c0[a];
            G[0][0] += np1;
        }
    }
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < NC; ++a)
#pragma unroll
            for (int b = 0; b < NC; ++b) s = fma(qat((j * NC + a) * NC + b), G[a][b], s);
        out[j] = s;
    }
}

template <int DIM, int MB, bool RATES>
__global__ __launch_bounds__(kApplyThreads) void propagate_kernel(
    int64_t nel, int64_t ntup, const int64_t* __restrict__ off0, const int64_t* __restrict__ off2,
    const int32_t* __restrict__ oth0, const int32_t* __restrict__ oth2, const double* __restrict__ q0,
    const double* __restrict__ q2, const double* __restrict__ ft, const double* __restrict__ bath,
    int32_t batch, double amplitude, const double* __restrict__ minv_g, double* __restrict__ out) {
    constexpr int NC = DIM + 1;
    constexpr int NQ = NC * NC * NC;
    constexpr int SL = prop_slots<DIM>();
    constexpr bool QREG = DIM == 2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* sQ = reinterpret_cast<double*>(smem_raw);                 // [NQ][SL]
    int32_t* sO = reinterpret_cast<int32_t*>(sQ + (size_t)NQ * SL);   // [SL]
    __shared__ double sOwn[MB][NC];
    __shared__ double sN[MB];
    __shared__ double sTot[MB][NC];
    const int64_t e = blockIdx.x;
    const int mb = blockIdx.y * MB;
    const int nm = min(MB, batch - mb);
    const int h = threadIdx.x;
    const bool slot_thread = h < SL;
    if (threadIdx.x < MB * NC) {
        const int b = threadIdx.x / NC, a = threadIdx.x % NC;
        sOwn[b][a] = b < nm ? ft[((int64_t)e * batch + mb + b) * NC + a] : 0.0;
    }
    if (threadIdx.x < MB) sN[threadIdx.x] = (int)threadIdx.x < nm ? bath[mb + threadIdx.x] : 0.0;
    double acc[MB][NC];
#pragma unroll
    for (int b = 0; b < MB; ++b)
#pragma unroll
        for (int a = 0; a < NC; ++a) acc[b][a] = 0.0;
    // rounds of SL entries: the leg-0 list, then the leg-2 list
    constexpr int PER = NQ * SL / kApplyThreads;
    static_assert(NQ * SL % kApplyThreads == 0, "staging must tile the workgroup");
    const int64_t a0 = off0[e], len0 = off0[e + 1] - a0;
    const int64_t a2 = off2[e], len2 = off2[e + 1] - a2;
    const int64_t nr0 = (len0 + SL - 1) / SL, nr = nr0 + (len2 + SL - 1) / SL;
    double stage[PER];
    int32_t stage_o = 0;
    auto fetch = [&](int64_t r) {
        const bool l0 = r < nr0;
        const int64_t base = l0 ? a0 + r * SL : a2 + (r - nr0) * SL;
        const int cnt = (int)min((int64_t)SL, (l0 ? a0 + len0 : a2 + len2) - base);
        const double* q = l0 ? q0 : q2;
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int idx = threadIdx.x + p * kApplyThreads;
            const int k = idx / SL, hh = idx % SL;
            stage[p] = hh < cnt ? q[(int64_t)k * ntup + base + hh] : 0.0;
        }
        const int32_t* oth = l0 ? oth0 : oth2;
        stage_o = (int)threadIdx.x < cnt ? oth[base + threadIdx.x] : 0;
    };
    if (nr > 0) fetch(0);
    for (int64_t r = 0; r < nr; ++r) {
        const bool l0 = r < nr0;
        const int64_t base = l0 ? a0 + r * SL : a2 + (r - nr0) * SL;
        const int cnt = (int)min((int64_t)SL, (l0 ? a0 + len0 : a2 + len2) - base);
#pragma unroll
        for (int p = 0; p < PER; ++p) sQ[threadIdx.x + p * kApplyThreads] = stage[p];
        if (slot_thread) sO[threadIdx.x] = stage_o;
        __syncthreads();
        if (r + 1 < nr) fetch(r + 1);
        if (slot_thread && h < cnt) {
            const int64_t o = sO[h];
            double Qr[QREG ? NQ : 1];
            if constexpr (QREG) {
#pragma unroll
                for (int k = 0; k < NQ; ++k) Qr[k] = sQ[k * SL + h];
            }
#pragma unroll
            for (int b = 0; b < MB; ++b) {
                if (b < nm) {
                    double own[NC], other[NC], c0[NC], c2[NC], t[NC];
#pragma unroll
                    for (int a = 0; a < NC; ++a) {
                        own[a] = sOwn[b][a];
                        other[a] = ft[((int64_t)o * batch + mb + b) * NC + a];
                    }
#pragma unroll
                    for (int a = 0; a < NC; ++a) {
                        c0[a] = l0 ? own[a] : other[a];
                        c2[a] = l0 ? other[a] : own[a];
                    }
                    if constexpr (QREG)
                        entry_terms<DIM, RATES>([&](int k) { return Qr[k]; }, l0, c0, c2, sN[b], 1.0 + sN[b], t);
                    else
                        entry_terms<DIM, RATES>([&](int k) { return sQ[k * SL + h]; }, l0, c0, c2, sN[b],
                                                1.0 + sN[b], t);
#pragma unroll
                    for (int j = 0; j < NC; ++j) acc[b][j] = (l0 || RATES) ? acc[b][j] + t[j] : acc[b][j] - t[j];
                }
            }
        }
        __syncthreads();
    }
    // slot-ordered sum per (member, basis) through the (now free) staging area
    double* sRed = sQ;   // [SL][NC]
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        if (b < nm) {
            if (slot_thread)
#pragma unroll
                for (int j = 0; j < NC; ++j) sRed[h * NC + j] = acc[b][j];
            __syncthreads();
            if (threadIdx.x < NC) {
                double total = 0.0;
                for (int hh = 0; hh < SL; ++hh) total += sRed[hh * NC + threadIdx.x];
                sTot[b][threadIdx.x] = total;
            }
            __syncthreads();
        }
    }
    if (threadIdx.x < nm * NC) {
        const int b = threadIdx.x / NC, j = threadIdx.x % NC;
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < NC; ++a) v = fma(sTot[b][a], minv_g[a * NC + j], v);
        out[((int64_t)(mb + b) * nel + e) * NC + j] = RATES ? 2.0 * (amplitude * v) : amplitude * v;
    }
}

// ft[band][e][b][a] = f[b][band][e][a]: the members of one element are
// contiguous (nbands = 1 for the phonon bath, 2 with a dynamical phonon band)
__global__ void transpose_state_kernel(const double* __restrict__ f, double* __restrict__ ft, int64_t nel,
                                       int32_t batch, int nc, int nbands) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)batch * nbands * nel * nc;
    if (i >= total) return;
    const int a = (int)(i % nc);
    const int64_t r = i / nc;
    const int64_t e = r % nel;
    const int64_t bb = r / nel;
    const int band = (int)(bb % nbands);
    const int64_t b = bb / nbands;
    ft[(((int64_t)band * nel + e) * batch + b) * nc + a] = f[i];
}

// The dynamical-phonon propagator: one workgroup per (element, chunk of MB
// members) sums the element's leg-0 and leg-2 lists into its electron
// coefficients and its leg-1 list into its phonon coefficients, slot-ordered
// as the bath kernel (SL slots per round, thread h = slot h, members in the
// inner loop, so a batch member is bitwise equal to its single evaluation).
// Entry of leg l with coefficients c0 (leg 0), p (leg 1), c2 (leg 2):
//   t_j = sum_ab Q0[j][a][b] (d_a0 c2_b - c0_a c2_b) + Q1[j][a][b] p_a c2_b - Q2[j][a][b] c0_a p_b
// added with sign +, +, - for l = 0, 1, 2.  States are read from the
// element-major copy ft[band][e][b][a]; out is [b][band][e][a].
struct DynLists {
    const int64_t* off[3];
    const int32_t* othA[3];
    const int32_t* othB[3];
    const double* q[3];
};

template <int DIM>
constexpr int dyn_slots() { return DIM == 2 ? 128 : 64; }

template <int DIM>
constexpr size_t dyn_lds_bytes() {
    constexpr int NC = DIM + 1, NQ = 3 * NC * NC * NC, SL = dyn_slots<DIM>();
    return sizeof(double) * (size_t)NQ * SL + 2 * sizeof(int32_t) * SL;
}

// RATES: ScatteringChannel::scatteringRates of the electron band with the
// rate leg's forms replaced by constants (factor 2 at the end):
//   leg-0 entries:  G0[a][b] = d_a0 c2_b, G2[a][b] = d_a0 p_b        (2 (P + F2))
//   leg-2 entries:  G0[a][b] = -c0_a d_b0, G1[a][b] = (d_a0 + p_a) d_b0  (2 (1 + P - F0))
// the leg-1 list is skipped and out is the electron band only, [b][e][a].
template <int DIM, int MB, bool RATES>
__global__ __launch_bounds__(kApplyThreads) void propagate_dyn_kernel(int64_t nel, int64_t ntup, DynLists L,
                                                                        const double* __restrict__ ft, int32_t batch,
                                                                        double amplitude, const double* __restrict__ minv_g,
                                                                        double* __restrict__ out) {
    constexpr int NC = DIM + 1;
    constexpr int NQ = 3 * NC * NC * NC;
    constexpr int SL = dyn_slots<DIM>();
    constexpr int PER = (NQ * SL + kApplyThreads - 1) / kApplyThreads;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* sQ = reinterpret_cast<double*>(smem_raw);                  // [NQ][SL]
    int32_t* sA = reinterpret_cast<int32_t*>(sQ + (size_t)NQ * SL);    // [SL]
    int32_t* sB = sA + SL;                                             // [SL]
    __shared__ double sOwn[2][MB][NC];
    __shared__ double sTot[2][MB][NC];
    const int64_t e = blockIdx.x;
    const int mb = blockIdx.y * MB;
    const int nm = min(MB, batch - mb);
    const int h = threadIdx.x;
    const bool slot_thread = h < SL;
    const double* fe = ft;                                  // electrons [e][b][a]
    const double* fp = ft + (int64_t)nel * batch * NC;      // phonons
    if (threadIdx.x < 2 * MB * NC) {
        const int band = threadIdx.x / (MB * NC), r = threadIdx.x % (MB * NC);
        const int b = r / NC, a = r % NC;
        sOwn[band][b][a] = b < nm ? (band ? fp : fe)[((int64_t)e * batch + mb + b) * NC + a] : 0.0;
    }
    double acc[2][MB][NC];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int b = 0; b < MB; ++b)
#pragma unroll
            for (int a = 0; a < NC; ++a) acc[q][b][a] = 0.0;
    // rounds: the leg-0 list, the leg-2 list, then the leg-1 list
    constexpr int ORDER[3] = {0, 2, 1};
    int64_t first[3], len[3], nrl[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        first[i] = L.off[ORDER[i]][e];
        len[i] = L.off[ORDER[i]][e + 1] - first[i];
        nrl[i] = (len[i] + SL - 1) / SL;
    }
    if (RATES) nrl[2] = 0;
    const int64_t nr = nrl[0] + nrl[1] + nrl[2];
    double stage[PER];
    int32_t stage_a = 0, stage_b = 0;
    auto locate = [&](int64_t r, int& li, int64_t& base, int& cnt) {
        li = r < nrl[0] ? 0 : (r < nrl[0] + nrl[1] ? 1 : 2);
        const int64_t rr = li == 0 ? r : (li == 1 ? r - nrl[0] : r - nrl[0] - nrl[1]);
        base = first[li] + rr * SL;
        cnt = (int)min((int64_t)SL, first[li] + len[li] - base);
    };
    auto fetch = [&](int64_t r) {
        int li, cnt;
        int64_t base;
        locate(r, li, base, cnt);
        const int leg = ORDER[li];
        const double* q = L.q[leg];
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int idx = threadIdx.x + p * kApplyThreads;
            const int k = idx / SL, hh = idx % SL;
            stage[p] = (idx < NQ * SL && hh < cnt) ? q[(int64_t)k * ntup + base + hh] : 0.0;
        }
        stage_a = (int)threadIdx.x < cnt ? L.othA[leg][base + threadIdx.x] : 0;
        stage_b = (int)threadIdx.x < cnt ? L.othB[leg][base + threadIdx.x] : 0;
    };
    if (nr > 0) fetch(0);
    for (int64_t r = 0; r < nr; ++r) {
        int li, cnt;
        int64_t base;
        locate(r, li, base, cnt);
        const int leg = ORDER[li];
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int idx = threadIdx.x + p * kApplyThreads;
            if (idx < NQ * SL) sQ[idx] = stage[p];
        }
        if (slot_thread) {
            sA[threadIdx.x] = stage_a;
            sB[threadIdx.x] = stage_b;
        }
        __syncthreads();
        if (r + 1 < nr) fetch(r + 1);
        if (slot_thread && h < cnt) {
            const int64_t oa = sA[h], ob = sB[h];
#pragma unroll
            for (int b = 0; b < MB; ++b) {
                if (b < nm) {
                    double c0[NC], pp[NC], c2[NC];
#pragma unroll
                    for (int a = 0; a < NC; ++a) {
                        const int64_t ia = (oa * batch + mb + b) * NC + a, ib = (ob * batch + mb + b) * NC + a;
                        if (leg == 0) { c0[a] = sOwn[0][b][a]; pp[a] = fp[ia]; c2[a] = fe[ib]; }
                        else if (leg == 1) { c0[a] = fe[ia]; pp[a] = sOwn[1][b][a]; c2[a] = fe[ib]; }
                        else { c0[a] = fe[ia]; pp[a] = fp[ib]; c2[a] = sOwn[0][b][a]; }
                    }
                    double G0[NC][NC], G1[NC][NC], G2[NC][NC];
                    if constexpr (!RATES) {
#pragma unroll
                        for (int a = 0; a < NC; ++a)
#pragma unroll
                            for (int bb = 0; bb < NC; ++bb) {
                                G0[a][bb] = -(c0[a] * c2[bb]);
                                G1[a][bb] = pp[a] * c2[bb];
                                G2[a][bb] = -(c0[a] * pp[bb]);
                            }
#pragma unroll
                        for (int bb = 0; bb < NC; ++bb) G0[0][bb] += c2[bb];
                    } else {
#pragma unroll
                        for (int a = 0; a < NC; ++a)
#pragma unroll
                            for (int bb = 0; bb < NC; ++bb) G0[a][bb] = G1[a][bb] = G2[a][bb] = 0.0;
                        if (leg == 0) {
#pragma unroll
                            for (int bb = 0; bb < NC; ++bb) {
                                G0[0][bb] = c2[bb];
                                G2[0][bb] = pp[bb];
                            }
                        } else {
#pragma unroll
                            for (int a = 0; a < NC; ++a) {
                                G0[a][0] = -c0[a];
                                G1[a][0] = pp[a];
                            }
                            G1[0][0] += 1.0;
                        }
                    }
#pragma unroll
                    for (int j = 0; j < NC; ++j) {
                        double t = 0.0;
#pragma unroll
                        for (int a = 0; a < NC; ++a)
#pragma unroll
                            for (int bb = 0; bb < NC; ++bb) {
                                t = fma(sQ[(((0 * NC + j) * NC + a) * NC + bb) * SL + h], G0[a][bb], t);
                                t = fma(sQ[(((1 * NC + j) * NC + a) * NC + bb) * SL + h], G1[a][bb], t);
                                t = fma(sQ[(((2 * NC + j) * NC + a) * NC + bb) * SL + h], G2[a][bb], t);
                            }
                        if (leg == 1) acc[1][b][j] = acc[1][b][j] + t;
                        else if (leg == 0 || RATES) acc[0][b][j] = acc[0][b][j] + t;
                        else acc[0][b][j] = acc[0][b][j] - t;
                    }
                }
            }
        }
        __syncthreads();
    }
    // slot-ordered sums per (band, member, basis) through the staging area
    double* sRed = sQ;   // [SL][NC]
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            if (b < nm) {
                if (slot_thread)
#pragma unroll
                    for (int j = 0; j < NC; ++j) sRed[h * NC + j] = acc[q][b][j];
                __syncthreads();
                if (threadIdx.x < NC) {
                    double total = 0.0;
                    for (int hh = 0; hh < SL; ++hh) total += sRed[hh * NC + threadIdx.x];
                    sTot[q][b][threadIdx.x] = total;
                }
                __syncthreads();
            }
        }
    if (threadIdx.x < (RATES ? 1 : 2) * nm * NC) {
        const int q = threadIdx.x / (nm * NC), r = threadIdx.x % (nm * NC);
        const int b = r / NC, j = r % NC;
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < NC; ++a) v = fma(sTot[q][b][a], minv_g[a * NC + j], v);
        if (RATES) out[((int64_t)(mb + b) * nel + e) * NC + j] = 2.0 * (amplitude * v);
        else out[(((int64_t)(mb + b) * 2 + q) * nel + e) * NC + j] = amplitude * v;
    }
}


// ------------------------------------------------------ 4-leg channel
// The channel {el in, el in, el out, el out} of the electron band
// (ScatteringChannel<NDim,4>, ScatteringChannel.cpp:716-722): legs 0 and 1
// free, leg 2 partially free, leg 3 constrained (the NLEGS == 4 branch of
// IntegratorCore.hpp:283-357).  The integrand is cubic in the states of four
// elements, so per-tuple moments would be (DIM+1)^4 per output; instead every
// evaluation sums the MC points directly, one thread per (tuple, member),
// recomputing the status-independent geometry as the reference does per tuple
// (ScatteringChannel.cpp:618-650).  Each element then adds its (tuple, leg)
// entries in tuple order and applies the inverse mass.  Same operation order
// and no contraction, as the CPU restatement.
#pragma clang fp contract(off)

__device__ inline double dir4(int leg) { return leg < 2 ? 1.0 : -1.0; }

// D ((DIM+1) x 4DIM) and N of the tuple (ScatteringChannel.cpp:285-352)
template <int DIM>
__device__ inline void tuple4_DN(const Geo& g, const int64_t el[4], const int u[3], double D[DIM + 1][4 * DIM],
                                 double N[DIM + 1]) {
#pragma unroll
    for (int r = 0; r <= DIM; ++r) N[r] = 0.0;
#pragma unroll
    for (int leg = 0; leg < 4; ++leg) {
        double kf[DIM][DIM + 1], ef[DIM + 1];
        elem_forms<DIM>(g, el[leg], false, kf, ef);
        const double s = dir4(leg);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
#pragma unroll
            for (int a = 0; a < DIM; ++a) D[d][leg * DIM + a] = s * kf[d][1 + a];
            N[d] -= s * kf[d][0];
        }
#pragma unroll
        for (int a = 0; a < DIM; ++a) D[DIM][leg * DIM + a] = s * ef[1 + a];
        N[DIM] -= s * ef[0];
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int c = 0; c < DIM; ++c) N[d] -= (double)u[c] * g.zone[d][c];
}

template <int DIM>
__device__ inline double dc4_inverse(const double D[DIM + 1][4 * DIM], double inv[DIM + 1][DIM + 1]) {
    constexpr int NC = DIM + 1, C0 = 3 * DIM - 1;
    double DC[NC][NC];
#pragma unroll
    for (int r = 0; r < NC; ++r)
#pragma unroll
        for (int q = 0; q < NC; ++q) DC[r][q] = D[r][C0 + q];
    return invert_small<NC>(DC, inv);
}

// Legs' points of MC sample m: legs 0 and 1 from the top 2 DIM rows, leg 2's
// free coordinates from the bottom DIM - 1 rows, then
// alpha = -D_C^-1 (D_F K_F - N) gives leg 2's last coordinate and leg 3;
// true when both constrained legs lie inside their simplices.
template <int DIM>
__device__ inline bool shell4_point(const Geo& g, const double D[DIM + 1][4 * DIM], const double N[DIM + 1],
                                    const double inv[DIM + 1][DIM + 1], int32_t m, double K[4][DIM]) {
    constexpr int NC = DIM + 1, C0 = 3 * DIM - 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        K[0][d] = g.pts[(size_t)d * g.nmc + m];
        K[1][d] = g.pts[(size_t)(DIM + d) * g.nmc + m];
    }
#pragma unroll
    for (int d = 0; d < DIM - 1; ++d) K[2][d] = g.pts[(size_t)(g.rows - (DIM - 1) + d) * g.nmc + m];
    double rhs[NC];
#pragma unroll
    for (int r = 0; r < NC; ++r) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < C0; ++c) s += D[r][c] * K[c / DIM][c % DIM];
        rhs[r] = s - N[r];
    }
#pragma unroll
    for (int r = 0; r < NC; ++r) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) s += inv[r][c] * rhs[c];
        if (r == 0) K[2][DIM - 1] = -s;
        else K[3][r - 1] = -s;
    }
    bool inside = true;
#pragma unroll
    for (int leg = 2; leg < 4; ++leg) {
        double sum = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) { inside &= K[leg][d] > 0.0; sum += K[leg][d]; }
        inside &= sum < 1.0;
    }
    return inside;
}

// |det D_C| NMC (DIM-1)! (DIM!)^2 (IntegratorCore.hpp:350-356)
template <int DIM>
__device__ inline double shell4_divisor(double det, int32_t nmc) {
    double w = fabs(det);
    w *= (double)nmc * (double)tg::factorial(DIM - 1);
    w *= (double)tg::factorial(DIM);
    w *= (double)tg::factorial(DIM);
    return w;
}

// Screening of a 4-leg tuple: |det D_C| >= min_det and
// hits c^4 prod(vol DIM!) / divisor > threshold (c = sqrt 2, the modal
// order-0 form, ScatteringChannel.cpp:465-503)
template <int DIM>
__device__ inline bool tuple4_passes(const Geo& g, const double D[DIM + 1][4 * DIM], const double N[DIM + 1],
                                     double min_det, double threshold) {
    double inv[DIM + 1][DIM + 1];
    const double det = dc4_inverse<DIM>(D, inv);
    if (!(fabs(det) >= min_det)) return false;
    int64_t hits = 0;
    for (int32_t m = 0; m < g.nmc; ++m) {
        double K[4][DIM];
        hits += shell4_point<DIM>(g, D, N, inv, m, K) ? 1 : 0;
    }
    const double c = sqrt(2.0);
    const double cpow = c * c * c * c;
    double amp = 1.0;
    for (int leg = 0; leg < 4; ++leg) amp *= g.vol * (double)tg::factorial(DIM);
    return (double)hits * cpow * amp / shell4_divisor<DIM>(det, g.nmc) > threshold;
}

// The tuples of the element pair (e0, e1) in (e2, uid, e3) order: the energy
// window of the four legs (ScatteringChannel.hpp:212-222; first with leg 3
// over the band's range [emin, emax]) and the leg-3 elements whose box meets
// k0 + k1 - k2 + G u.  Calls emit(e2, e3, uid) for each kept one.
template <int DIM, typename Emit>
__device__ inline int32_t pair4_tuples(const Geo& g, int64_t nel, int64_t e0, int64_t e1, double emin, double emax,
                                       double min_det, double threshold, Emit emit) {
    constexpr int NES = DIM == 2 ? 2 : 6;
    double lo0[DIM], hi0[DIM], lo1[DIM], hi1[DIM], el0, eh0, el1, eh1;
    elem_box_erange<DIM>(g, e0, lo0, hi0, &el0, &eh0);
    elem_box_erange<DIM>(g, e1, lo1, hi1, &el1, &eh1);
    int32_t cnt = 0;
    for (int64_t e2 = 0; e2 < nel; ++e2) {
        double lo2[DIM], hi2[DIM], el2, eh2;
        elem_box_erange<DIM>(g, e2, lo2, hi2, &el2, &eh2);
        if (!(el0 + el1 - eh2 - emax <= 0.0 && eh0 + eh1 - el2 - emin >= 0.0)) continue;
        for (int32_t uid = 0; uid < g.nuv; ++uid) {
            int u[3];
            umklapp_vec<DIM>(g.nu, uid, u);
            double lo[DIM], hi[DIM], smin[3], smax[3];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                const double U = umklapp_comp<DIM>(g, u, d);
                lo[d] = lo0[d] + lo1[d] - hi2[d] + U;
                hi[d] = hi0[d] + hi1[d] - lo2[d] + U;
                if (g.axis) {
                    const double h = g.extent[d] / g.nsec[d];
                    smin[d] = floor((lo[d] - g.origin[d]) / h);
                    smax[d] = floor((hi[d] - g.origin[d]) / h);
                } else {
                    smin[d] = 1e300;
                    smax[d] = -1e300;
                }
            }
            if (!g.axis) {
                for (int cm = 0; cm < (1 << DIM); ++cm) {
                    double x[DIM];
#pragma unroll
                    for (int d = 0; d < DIM; ++d) x[d] = ((cm >> d) & 1 ? hi[d] : lo[d]) - g.origin[d];
#pragma unroll
                    for (int c = 0; c < DIM; ++c) {
                        double s = 0.0;
#pragma unroll
                        for (int d = 0; d < DIM; ++d) s = s + g.dGinv[c][d] * x[d];
                        smin[c] = s < smin[c] ? s : smin[c];
                        smax[c] = s > smax[c] ? s : smax[c];
                    }
                }
#pragma unroll
                for (int c = 0; c < DIM; ++c) {
                    smin[c] = floor(smin[c]) - 1.0;
                    smax[c] = floor(smax[c]) + 1.0;
                }
            }
            int slo[3] = {0, 0, 0}, shi[3] = {0, 0, 0};
            bool empty = false;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                const double a = smin[d], b = smax[d];
                empty |= b < 0.0 || a > g.nsec[d] - 1;
                slo[d] = a < 0.0 ? 0 : (a > g.nsec[d] - 1 ? g.nsec[d] - 1 : (int)a);
                shi[d] = b > g.nsec[d] - 1 ? g.nsec[d] - 1 : (b < 0.0 ? 0 : (int)b);
            }
            if (empty) continue;
            int sc[3] = {slo[0], slo[1], slo[2]};
            for (;;) {
                int64_t sid = 0, stride = 1;
#pragma unroll
                for (int d = 0; d < DIM; ++d) { sid += stride * sc[d]; stride *= g.nsec[d]; }
                for (int t = 0; t < NES; ++t) {
                    const int64_t e3 = sid * NES + t;
                    double lo3[DIM], hi3[DIM], el3, eh3;
                    elem_box_erange<DIM>(g, e3, lo3, hi3, &el3, &eh3);
                    bool meet = true;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) meet &= lo3[d] <= hi[d] && hi3[d] >= lo[d];
                    if (!meet) continue;
                    if (!(el0 + el1 - eh2 - eh3 <= 0.0 && eh0 + eh1 - el2 - el3 >= 0.0)) continue;
                    const int64_t el[4] = {e0, e1, e2, e3};
                    double D[DIM + 1][4 * DIM], N[DIM + 1];
                    tuple4_DN<DIM>(g, el, u, D, N);
                    if (!tuple4_passes<DIM>(g, D, N, min_det, threshold)) continue;
                    emit(e2, e3, uid);
                    ++cnt;
                }
                int d = 0;
                while (d < DIM && ++sc[d] > shi[d]) { sc[d] = slo[d]; ++d; }
                if (d == DIM) break;
            }
        }
    }
    return cnt;
}

// One thread per element pair p = e0 nel + e1: pass 0 counts its tuples into
// cnt[p], pass 1 writes them from off[p] as [e0, e1, e2, e3, uid].
template <int DIM>
__global__ void __launch_bounds__(kEnumThreads)
enum4_kernel(Geo g, int64_t nel, double emin, double emax, double min_det, double threshold, int pass,
             int64_t* __restrict__ cnt, const int64_t* __restrict__ off, int32_t* __restrict__ tup) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nel * nel) return;
    const int64_t e0 = p / nel, e1 = p % nel;
    if (pass == 0) {
        cnt[p] = pair4_tuples<DIM>(g, nel, e0, e1, emin, emax, min_det, threshold, [](int64_t, int64_t, int32_t) {});
    } else {
        int64_t w = off[p];
        pair4_tuples<DIM>(g, nel, e0, e1, emin, emax, min_det, threshold, [&](int64_t e2, int64_t e3, int32_t uid) {
            int32_t* o = tup + w * 5;
            o[0] = (int32_t)e0;
            o[1] = (int32_t)e1;
            o[2] = (int32_t)e2;
            o[3] = (int32_t)e3;
            o[4] = uid;
            ++w;
        });
    }
}

// One thread per (tuple, member), member fastest: the tuple's MC integral
// (scatteringIntegrationTypeB) times constructAmplitude
// (ScatteringChannel.cpp:399-415) and the leg's sign, written as
// res[(t batch + b)][leg][DIM+1]; zero when |det D_C| < min_det (:621-622).
// functAmplitude: 1 for kappa <= 0, else 1/(|q|^2 + kappa^2)^2 with q the
// barycentre of leg 2 minus that of leg 0.
template <int DIM>
__global__ void __launch_bounds__(kApplyThreads)
tuple4_kernel(Geo g, int64_t ntup, int64_t nel, int32_t batch, const int32_t* __restrict__ tup,
              const double* __restrict__ f, double amplitude, double kappa, double min_det,
              double* __restrict__ res) {
    constexpr int NC = DIM + 1;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntup * batch) return;
    const int64_t t = i / batch;
    const int32_t b = (int32_t)(i % batch);
    const int32_t* tp = tup + t * 5;
    const int64_t el[4] = {tp[0], tp[1], tp[2], tp[3]};
    int u[3];
    umklapp_vec<DIM>(g.nu, tp[4], u);
    double D[DIM + 1][4 * DIM], N[DIM + 1], inv[NC][NC];
    tuple4_DN<DIM>(g, el, u, D, N);
    const double det = dc4_inverse<DIM>(D, inv);
    double acc[4][NC];
#pragma unroll
    for (int l = 0; l < 4; ++l)
#pragma unroll
        for (int a = 0; a < NC; ++a) acc[l][a] = 0.0;
    double* r = res + (size_t)i * 4 * NC;
    if (!(fabs(det) >= min_det)) {
#pragma unroll
        for (int l = 0; l < 4; ++l)
#pragma unroll
            for (int a = 0; a < NC; ++a) r[l * NC + a] = 0.0;
        return;
    }
    // the forms (ScatteringChannel.cpp:628-640, fermion statistics): direct
    // -f0 f1 (1 - f2)(1 - f3), time reversed (1 - f0)(1 - f1) f2 f3
    double fa[4][NC], fb[4][NC];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const double* c = f + ((size_t)b * nel + el[l]) * NC;
#pragma unroll
        for (int a = 0; a < NC; ++a) {
            const double one = a == 0 ? 1.0 : 0.0, x = c[a];
            if (l < 2) {
                fa[l][a] = l == 0 ? -x : x;
                fb[l][a] = one - x;
            } else {
                fa[l][a] = one - x;
                fb[l][a] = x;
            }
        }
    }
    for (int32_t m = 0; m < g.nmc; ++m) {
        double K[4][DIM];
        if (!shell4_point<DIM>(g, D, N, inv, m, K)) continue;
        double pa = 1.0, pb = 1.0;
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            double va = fa[l][0], vb = fb[l][0];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                va += fa[l][1 + d] * K[l][d];
                vb += fb[l][1 + d] * K[l][d];
            }
            pa *= va;
            pb *= vb;
        }
        const double val = pa + pb;
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            acc[l][0] += val;
#pragma unroll
            for (int d = 0; d < DIM; ++d) acc[l][1 + d] += val * K[l][d];
        }
    }
    const double w = shell4_divisor<DIM>(det, g.nmc);
    double amp = amplitude;
    for (int l = 0; l < 4; ++l) amp *= g.vol * (double)tg::factorial(DIM);
    if (kappa > 0.0) {
        int64_t v0[DIM + 1], v2[DIM + 1];
        elem_nodes<DIM>(g, el[0], v0);
        elem_nodes<DIM>(g, el[2], v2);
        double q2 = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            double c0 = 0.0, c2 = 0.0;
#pragma unroll
            for (int a = 0; a <= DIM; ++a) {
                c0 += node_coord<DIM>(g, v0[a], d);
                c2 += node_coord<DIM>(g, v2[a], d);
            }
            const double q = (c2 - c0) / (double)(DIM + 1);
            q2 += q * q;
        }
        const double vq = 1.0 / (q2 + kappa * kappa);
        amp *= vq * vq;
    }
#pragma unroll
    for (int l = 0; l < 4; ++l)
#pragma unroll
        for (int a = 0; a < NC; ++a) r[l * NC + a] = dir4(l) * (amp * (acc[l][a] / w));
}

// One 64-lane workgroup per (element, member), b = blockIdx % batch: the
// element's entries (tuple << 2 | leg, in tuple order) are loaded 256 at a
// time into LDS by all lanes, then lane a < DIM+1 adds coefficient a of them
// in entry order (the order of the sequential sum), then the inverse mass
// (Function.cpp:801-804) into out [batch][nel][DIM+1].
constexpr int kGather4Chunk = 256;
template <int DIM>
__global__ void __launch_bounds__(64)
gather4_kernel(int64_t nel, int32_t batch, const int64_t* __restrict__ goff, const int32_t* __restrict__ gent,
               const double* __restrict__ res, const double* __restrict__ minv_g, double* __restrict__ out) {
    constexpr int NC = DIM + 1;
    __shared__ double sh[kGather4Chunk][NC];
    const int64_t e = blockIdx.x / batch;
    const int32_t b = (int32_t)(blockIdx.x % batch);
    const int lane = threadIdx.x;
    const int64_t k_end = goff[e + 1];
    constexpr int U = kGather4Chunk / 64;
    // this lane's entries of the next chunk, loaded while the current one is summed
    double nxt[U][NC];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t k = k0 + u * 64 + lane;
            if (k < k_end) {
                const int32_t ent = gent[k];
                const double* r = res + ((size_t)(ent >> 2) * batch + b) * 4 * NC + (ent & 3) * NC;
#pragma unroll
                for (int a = 0; a < NC; ++a) nxt[u][a] = r[a];
            }
        }
    };
    double acc = 0.0;
    int64_t k0 = goff[e];
    if (k0 < k_end) load(k0);
    for (; k0 < k_end; k0 += kGather4Chunk) {
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u * 64 + lane < k_end)
#pragma unroll
                for (int a = 0; a < NC; ++a) sh[u * 64 + lane][a] = nxt[u][a];
        __syncthreads();
        if (k0 + kGather4Chunk < k_end) load(k0 + kGather4Chunk);
        if (lane < NC) {
            const int n = (int)min<int64_t>(kGather4Chunk, k_end - k0);
            int j = 0;
            // reads issued ahead, adds in entry order
            for (; j + 16 <= n; j += 16) {
                double v[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) v[q] = sh[j + q][lane];
#pragma unroll
                for (int q = 0; q < 16; ++q) acc += v[q];
            }
            for (; j < n; ++j) acc += sh[j][lane];
        }
        __syncthreads();
    }
    if (lane < NC) sh[0][lane] = acc;
    __syncthreads();
    if (lane < NC) {
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < NC; ++a) v += sh[0][a] * minv_g[a * NC + lane];
        out[((size_t)b * nel + e) * NC + lane] = v;
    }
}

#pragma clang fp contract(on)

Geo make_geo(const tg_tuple_ctx* c) {
    Geo g{};
    for (int d = 0; d < 3; ++d) {
        g.nsec[d] = c->nsec[d];
        g.origin[d] = c->origin[d];
        g.extent[d] = c->extent[d];
        for (int e = 0; e < 3; ++e) {
            g.dG[d][e] = c->dG[d][e];
            g.dGinv[d][e] = c->dGinv[d][e];
            g.zone[d][e] = c->zone[d][e];
        }
    }
    g.axis = c->axis;
    g.omega0 = c->omega0;
    g.vol = c->vol;
    g.nu = c->nu;
    g.nuv = 1;
    for (int d = 0; d < c->dim; ++d) g.nuv *= 2 * c->nu + 1;
    g.nmc = c->nmc;
    g.rows = c->rows;
    g.eps = c->d_eps;
    g.eps_ph = c->d_eps_ph;
    g.pmin = c->pmin;
    g.pmax = c->pmax;
    g.pts = c->d_pts;
    g.etab = c->d_etab;
    return g;
}

tg_status build_lists(tg_tuple_ctx* c) {
    // leg-0 offsets from the (e0-ordered) tuples, leg-2 positions by a stable
    // counting sort on e2 (host: the lists are built once per tuple list)
    const int64_t nel = c->nel, T = c->ntup;
    std::vector<int32_t> tup((size_t)T * 4);
    if (T) TT_HIP(c, hipMemcpy(tup.data(), c->d_tup, sizeof(int32_t) * 4 * T, hipMemcpyDeviceToHost));
    std::vector<int64_t> off0(nel + 1, 0), off2(nel + 1, 0), pos2(T > 0 ? T : 1);
    for (int64_t t = 0; t < T; ++t) {
        ++off0[tup[t * 4] + 1];
        ++off2[tup[t * 4 + 2] + 1];
    }
    for (int64_t e = 0; e < nel; ++e) { off0[e + 1] += off0[e]; off2[e + 1] += off2[e]; }
    std::vector<int64_t> fillp(off2.begin(), off2.end() - 1);
    for (int64_t t = 0; t < T; ++t) pos2[t] = fillp[tup[t * 4 + 2]]++;
    tg_status s;
    if ((s = dalloc(c, &c->d_off0, nel + 1)) != TG_OK) return s;
    if ((s = dalloc(c, &c->d_off2, nel + 1)) != TG_OK) return s;
    if ((s = dalloc(c, &c->d_pos2, T)) != TG_OK) return s;
    if ((s = dalloc(c, &c->d_oth0, T)) != TG_OK) return s;
    if ((s = dalloc(c, &c->d_oth2, T)) != TG_OK) return s;
    const int nq = (c->dim + 1) * (c->dim + 1) * (c->dim + 1);
    if ((s = dalloc(c, &c->d_q0, (size_t)nq * T)) != TG_OK) return s;
    if ((s = dalloc(c, &c->d_q2, (size_t)nq * T)) != TG_OK) return s;
    TT_HIP(c, hipMemcpy(c->d_off0, off0.data(), sizeof(int64_t) * (nel + 1), hipMemcpyHostToDevice));
    TT_HIP(c, hipMemcpy(c->d_off2, off2.data(), sizeof(int64_t) * (nel + 1), hipMemcpyHostToDevice));
    if (T) {
        TT_HIP(c, hipMemcpy(c->d_pos2, pos2.data(), sizeof(int64_t) * T, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(fill_leg_lists, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, c->stream,
                           c->d_tup, T, c->d_pos2, c->d_oth0, c->d_oth2);
        TT_HIP(c, hipGetLastError());
    }
    if (!c->dyn) return TG_OK;
    // dynamical phonons: the leg-1 list too, and per list the two other
    // elements in leg order (leg 0: e1, e2; leg 1: e0, e2; leg 2: e0, e1)
    std::vector<int64_t> off1(nel + 1, 0), pos1(T > 0 ? T : 1), pos0(T > 0 ? T : 1);
    for (int64_t t = 0; t < T; ++t) ++off1[tup[t * 4 + 1] + 1];
    for (int64_t e = 0; e < nel; ++e) off1[e + 1] += off1[e];
    std::vector<int64_t> fill1(off1.begin(), off1.end() - 1);
    for (int64_t t = 0; t < T; ++t) {
        pos1[t] = fill1[tup[t * 4 + 1]]++;
        pos0[t] = t;
    }
    const std::vector<int64_t>* offs[3] = {&off0, &off1, &off2};
    const std::vector<int64_t>* poss[3] = {&pos0, &pos1, &pos2};
    const int nqd = 3 * nq;
    for (int l = 0; l < 3; ++l) {
        std::vector<int32_t> a((size_t)(T > 0 ? T : 1)), b((size_t)(T > 0 ? T : 1));
        const int la = l == 0 ? 1 : 0, lb = l == 2 ? 1 : 2;
        for (int64_t t = 0; t < T; ++t) {
            const int64_t q = (*poss[l])[t];
            a[q] = tup[t * 4 + la];
            b[q] = tup[t * 4 + lb];
        }
        if ((s = dalloc(c, &c->d_doff[l], nel + 1)) != TG_OK) return s;
        if ((s = dalloc(c, &c->d_dothA[l], T)) != TG_OK) return s;
        if ((s = dalloc(c, &c->d_dothB[l], T)) != TG_OK) return s;
        if ((s = dalloc(c, &c->d_dpos[l], T)) != TG_OK) return s;
        if ((s = dalloc(c, &c->d_dq[l], (size_t)nqd * T)) != TG_OK) return s;
        TT_HIP(c, hipMemcpy(c->d_doff[l], offs[l]->data(), sizeof(int64_t) * (nel + 1), hipMemcpyHostToDevice));
        if (T) {
            TT_HIP(c, hipMemcpy(c->d_dothA[l], a.data(), sizeof(int32_t) * T, hipMemcpyHostToDevice));
            TT_HIP(c, hipMemcpy(c->d_dothB[l], b.data(), sizeof(int32_t) * T, hipMemcpyHostToDevice));
            TT_HIP(c, hipMemcpy(c->d_dpos[l], poss[l]->data(), sizeof(int64_t) * T, hipMemcpyHostToDevice));
        }
    }
    return TG_OK;
}

tg_status enumerate(tg_tuple_ctx* c) {
    const int64_t nel = c->nel;
    if (nel > 0x7fffffff) return fail(c, TG_ERR_ARG, "element count %lld exceeds int32", (long long)nel);
    Geo g = make_geo(c);
    int64_t* d_counts = nullptr;
    int64_t* d_base = nullptr;
    tg_status s;
    if ((s = dalloc(c, &d_counts, nel)) != TG_OK) return s;
    if ((s = dalloc(c, &d_base, nel)) != TG_OK) { (void)hipFree(d_counts); return s; }
    auto cleanup = [&]() { (void)hipFree(d_counts); (void)hipFree(d_base); };
    auto launch_enum = [&](const int64_t* b, int32_t* tp, bool write) {
        if (c->dim == 2)
            hipLaunchKernelGGL(enumerate_kernel<2>, dim3((unsigned)nel), dim3(kEnumThreads), 0, c->stream, g,
                               nel, c->min_det, c->threshold, d_counts, b, tp, write);
        else
            hipLaunchKernelGGL(enumerate_kernel<3>, dim3((unsigned)nel), dim3(kEnumThreads), 0, c->stream, g,
                               nel, c->min_det, c->threshold, d_counts, b, tp, write);
    };
    launch_enum(nullptr, nullptr, false);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { cleanup(); return fail(c, TG_ERR_HIP, "tuple count: %s", hipGetErrorString(e)); }
    std::vector<int64_t> counts(nel), base(nel);
    e = hipMemcpy(counts.data(), d_counts, sizeof(int64_t) * nel, hipMemcpyDeviceToHost);
    if (e != hipSuccess) { cleanup(); return fail(c, TG_ERR_HIP, "tuple count copy: %s", hipGetErrorString(e)); }
    int64_t total = 0;
    for (int64_t i = 0; i < nel; ++i) { base[i] = total; total += counts[i]; }
    if (total > ((int64_t)1 << 40)) { cleanup(); return fail(c, TG_ERR_ARG, "%lld tuples", (long long)total); }
    if ((s = dalloc(c, &c->d_tup, (size_t)total * 4)) != TG_OK) { cleanup(); return s; }
    e = hipMemcpy(d_base, base.data(), sizeof(int64_t) * nel, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_enum(d_base, c->d_tup, true);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    cleanup();
    if (e != hipSuccess) return fail(c, TG_ERR_HIP, "tuple fill: %s", hipGetErrorString(e));
    c->ntup = total;
    return build_lists(c);
}

// The element entries of a 4-leg list: for element e the (tuple << 2 | leg)
// of every leg that lands on it, in tuple order then leg order, and the
// per-(tuple, member) result buffer.
tg_status build_lists4(tg_tuple_ctx* c, const int32_t* idx) {
    const int64_t nel = c->nel, ntup = c->ntup;
    if (ntup >= ((int64_t)1 << 29)) return fail(c, TG_ERR_ARG, "%lld tuples: at most 2^29 - 1", (long long)ntup);
    std::vector<int64_t> off((size_t)nel + 1, 0);
    for (int64_t t = 0; t < ntup; ++t)
        for (int l = 0; l < 4; ++l) ++off[(size_t)idx[t * 5 + l] + 1];
    for (int64_t e = 0; e < nel; ++e) off[(size_t)e + 1] += off[(size_t)e];
    std::vector<int32_t> ent((size_t)std::max<int64_t>(4 * ntup, 1));
    std::vector<int64_t> pos(off.begin(), off.end() - 1);
    for (int64_t t = 0; t < ntup; ++t)
        for (int l = 0; l < 4; ++l) ent[(size_t)pos[(size_t)idx[t * 5 + l]]++] = (int32_t)((t << 2) | l);
    tg_status s;
    if ((s = dalloc(c, &c->d_goff, (size_t)nel + 1)) != TG_OK) return s;
    if ((s = dalloc(c, &c->d_gent, ent.size())) != TG_OK) return s;
    if ((s = dalloc(c, &c->d_res4, (size_t)std::max<int64_t>(ntup, 1) * c->max_batch * 4 * (c->dim + 1))) != TG_OK)
        return s;
    TT_HIP(c, hipMemcpy(c->d_goff, off.data(), sizeof(int64_t) * (nel + 1), hipMemcpyHostToDevice));
    TT_HIP(c, hipMemcpy(c->d_gent, ent.data(), sizeof(int32_t) * ent.size(), hipMemcpyHostToDevice));
    return TG_OK;
}

// The 4-leg tuple list on the GPU: one thread per element pair (e0, e1),
// counting then filling, so the list is in (e0, e1, e2, uid, e3) order.
tg_status enumerate4(tg_tuple_ctx* c) {
    const int64_t nel = c->nel;
    if (nel > 46340) return fail(c, TG_ERR_ARG, "%lld elements: the 4-leg pair grid needs nel^2 < 2^31", (long long)nel);
    Geo g = make_geo(c);
    // the band's energy range from the element table
    const int stride = c->dim == 2 ? etab_stride<2>() : etab_stride<3>();
    std::vector<double> tab((size_t)nel * stride);
    TT_HIP(c, hipMemcpy(tab.data(), c->d_etab, sizeof(double) * tab.size(), hipMemcpyDeviceToHost));
    const int eoff = c->dim * (c->dim + 1) + (c->dim + 1) + 2 * c->dim;
    double emin = 0.0, emax = 0.0;
    for (int64_t e = 0; e < nel; ++e) {
        const double lo = tab[(size_t)e * stride + eoff], hi = tab[(size_t)e * stride + eoff + 1];
        if (e == 0 || lo < emin) emin = lo;
        if (e == 0 || hi > emax) emax = hi;
    }
    const int64_t np = nel * nel;
    int64_t* d_cnt = nullptr;
    int64_t* d_off = nullptr;
    tg_status s;
    if ((s = dalloc(c, &d_cnt, np)) != TG_OK) return s;
    if ((s = dalloc(c, &d_off, np)) != TG_OK) { (void)hipFree(d_cnt); return s; }
    auto cleanup = [&]() { (void)hipFree(d_cnt); (void)hipFree(d_off); };
    const dim3 grid((unsigned)((np + kEnumThreads - 1) / kEnumThreads));
    auto launch = [&](int pass, int32_t* tp) {
        if (c->dim == 2)
            hipLaunchKernelGGL(enum4_kernel<2>, grid, dim3(kEnumThreads), 0, c->stream, g, nel, emin, emax, c->min_det,
                               c->threshold, pass, d_cnt, d_off, tp);
        else
            hipLaunchKernelGGL(enum4_kernel<3>, grid, dim3(kEnumThreads), 0, c->stream, g, nel, emin, emax, c->min_det,
                               c->threshold, pass, d_cnt, d_off, tp);
    };
    launch(0, nullptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { cleanup(); return fail(c, TG_ERR_HIP, "4-leg tuple count: %s", hipGetErrorString(e)); }
    std::vector<int64_t> cnt((size_t)np), off((size_t)np);
    e = hipMemcpy(cnt.data(), d_cnt, sizeof(int64_t) * np, hipMemcpyDeviceToHost);
    if (e != hipSuccess) { cleanup(); return fail(c, TG_ERR_HIP, "4-leg count copy: %s", hipGetErrorString(e)); }
    int64_t total = 0;
    for (int64_t p = 0; p < np; ++p) { off[(size_t)p] = total; total += cnt[(size_t)p]; }
    if (total >= ((int64_t)1 << 29)) { cleanup(); return fail(c, TG_ERR_ARG, "%lld tuples: at most 2^29 - 1", (long long)total); }
    if ((s = dalloc(c, &c->d_tup, (size_t)total * 5)) != TG_OK) { cleanup(); return s; }
    e = hipMemcpy(d_off, off.data(), sizeof(int64_t) * np, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch(1, c->d_tup);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    cleanup();
    if (e != hipSuccess) return fail(c, TG_ERR_HIP, "4-leg tuple fill: %s", hipGetErrorString(e));
    c->ntup = total;
    std::vector<int32_t> idx((size_t)std::max<int64_t>(total, 1) * 5);
    if (total) TT_HIP(c, hipMemcpy(idx.data(), c->d_tup, sizeof(int32_t) * 5 * total, hipMemcpyDeviceToHost));
    return build_lists4(c, idx.data());
}

// One 4-leg evaluation of `batch` members from device states d_f into d_out.
tg_status propagate4_device(tg_tuple_ctx* c, const double* d_f, double* d_out, int32_t batch) {
    Geo g = make_geo(c);
    if (c->time_kernel) TT_HIP(c, hipEventRecord(c->ev0, c->stream));
    const int64_t nt = c->ntup * batch, ne = c->nel * batch;
    const dim3 gt((unsigned)std::max<int64_t>((nt + kApplyThreads - 1) / kApplyThreads, 1));
    const dim3 ge((unsigned)ne);   // one 64-lane workgroup per (element, member)
    if (c->dim == 2) {
        if (nt)
            hipLaunchKernelGGL(tuple4_kernel<2>, gt, dim3(kApplyThreads), 0, c->stream, g, c->ntup, c->nel, batch,
                               c->d_tup, d_f, c->amplitude, c->kappa, c->min_det, c->d_res4);
        hipLaunchKernelGGL(gather4_kernel<2>, ge, dim3(64), 0, c->stream, c->nel, batch, c->d_goff,
                           c->d_gent, c->d_res4, c->d_minv, d_out);
    } else {
        if (nt)
            hipLaunchKernelGGL(tuple4_kernel<3>, gt, dim3(kApplyThreads), 0, c->stream, g, c->ntup, c->nel, batch,
                               c->d_tup, d_f, c->amplitude, c->kappa, c->min_det, c->d_res4);
        hipLaunchKernelGGL(gather4_kernel<3>, ge, dim3(64), 0, c->stream, c->nel, batch, c->d_goff,
                           c->d_gent, c->d_res4, c->d_minv, d_out);
    }
    TT_HIP(c, hipGetLastError());
    if (c->time_kernel) TT_HIP(c, hipEventRecord(c->ev1, c->stream));
    return TG_OK;
}

tg_status moments(tg_tuple_ctx* c) {
    if (c->ntup == 0) return TG_OK;
    Geo g = make_geo(c);
    dim3 grid((unsigned)((c->ntup + 255) / 256));
    if (c->dyn) {
#define TT_MDYN(D, L) hipLaunchKernelGGL((moments_dyn_kernel<D, L>), grid, dim3(256), 0, c->stream, g, c->d_tup, \
                                         c->ntup, c->d_dpos[L], c->d_dq[L])
        if (c->dim == 2) { TT_MDYN(2, 0); TT_MDYN(2, 1); TT_MDYN(2, 2); }
        else { TT_MDYN(3, 0); TT_MDYN(3, 1); TT_MDYN(3, 2); }
#undef TT_MDYN
        TT_HIP(c, hipGetLastError());
        TT_HIP(c, hipStreamSynchronize(c->stream));
        return TG_OK;
    }
    if (c->dim == 2)
        hipLaunchKernelGGL(moments_kernel<2>, grid, dim3(256), 0, c->stream, g, c->d_tup, c->ntup, c->d_pos2,
                           c->d_q0, c->d_q2);
    else
        hipLaunchKernelGGL(moments_kernel<3>, grid, dim3(256), 0, c->stream, g, c->d_tup, c->ntup, c->d_pos2,
                           c->d_q0, c->d_q2);
    TT_HIP(c, hipGetLastError());
    TT_HIP(c, hipStreamSynchronize(c->stream));
    return TG_OK;
}

template <int DIM, int MB, bool RATES>
tg_status launch_propagate(tg_tuple_ctx* c, const double* d_ft, const double* d_bath, double* d_out, int32_t batch) {
    constexpr size_t lds = prop_lds_bytes<DIM>();
    static bool attr_set = false;   // per process; the attribute belongs to the function
    if (!attr_set) {
        TT_HIP(c, hipFuncSetAttribute(reinterpret_cast<const void*>(&propagate_kernel<DIM, MB, RATES>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr_set = true;
    }
    dim3 grid((unsigned)c->nel, (unsigned)((batch + MB - 1) / MB));
    hipLaunchKernelGGL((propagate_kernel<DIM, MB, RATES>), grid, dim3(kApplyThreads), lds, c->stream, c->nel, c->ntup,
                       c->d_off0, c->d_off2, c->d_oth0, c->d_oth2, c->d_q0, c->d_q2, d_ft, d_bath, batch,
                       c->amplitude, c->d_minv, d_out);
    TT_HIP(c, hipGetLastError());
    return TG_OK;
}

template <int DIM, int MB, bool RATES>
tg_status launch_dyn(tg_tuple_ctx* c, const double* d_ft, double* d_out, int32_t batch) {
    constexpr size_t lds = dyn_lds_bytes<DIM>();
    static bool attr_set = false;
    if (!attr_set) {
        TT_HIP(c, hipFuncSetAttribute(reinterpret_cast<const void*>(&propagate_dyn_kernel<DIM, MB, RATES>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr_set = true;
    }
    DynLists L{};
    for (int l = 0; l < 3; ++l) {
        L.off[l] = c->d_doff[l];
        L.othA[l] = c->d_dothA[l];
        L.othB[l] = c->d_dothB[l];
        L.q[l] = c->d_dq[l];
    }
    dim3 grid((unsigned)c->nel, (unsigned)((batch + MB - 1) / MB));
    hipLaunchKernelGGL((propagate_dyn_kernel<DIM, MB, RATES>), grid, dim3(kApplyThreads), lds, c->stream, c->nel, c->ntup,
                       L, d_ft, batch, c->amplitude, c->d_minv, d_out);
    TT_HIP(c, hipGetLastError());
    return TG_OK;
}

template <bool RATES = false>
tg_status propagate_device(tg_tuple_ctx* c, const double* d_f, const double* d_bath, double* d_out,
                           int32_t batch) {
    if (!RATES && c->nlegs == 4) return propagate4_device(c, d_f, d_out, batch);
    const int nc = c->dim + 1;
    const int nbands = c->dyn ? 2 : 1;
    const int64_t total = (int64_t)batch * nbands * c->nel * nc;
    if (c->time_kernel) TT_HIP(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(transpose_state_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream, d_f,
                       c->d_ft, c->nel, batch, nc, nbands);
    TT_HIP(c, hipGetLastError());
    tg_status s;
    if (c->dyn) {
        if (c->dim == 2)
            s = batch == 1 ? launch_dyn<2, 1, RATES>(c, c->d_ft, d_out, batch)
                           : launch_dyn<2, kApplyMembers, RATES>(c, c->d_ft, d_out, batch);
        else
            s = batch == 1 ? launch_dyn<3, 1, RATES>(c, c->d_ft, d_out, batch)
                           : launch_dyn<3, kApplyMembers, RATES>(c, c->d_ft, d_out, batch);
        if (s != TG_OK) return s;
        if (c->time_kernel) TT_HIP(c, hipEventRecord(c->ev1, c->stream));
        return TG_OK;
    }
    if (c->dim == 2)
        s = batch == 1 ? launch_propagate<2, 1, RATES>(c, c->d_ft, d_bath, d_out, batch)
                       : launch_propagate<2, kApplyMembers, RATES>(c, c->d_ft, d_bath, d_out, batch);
    else
        s = batch == 1 ? launch_propagate<3, 1, RATES>(c, c->d_ft, d_bath, d_out, batch)
                       : launch_propagate<3, kApplyMembers, RATES>(c, c->d_ft, d_bath, d_out, batch);
    if (s != TG_OK) return s;
    if (c->time_kernel) TT_HIP(c, hipEventRecord(c->ev1, c->stream));
    return TG_OK;
}

// ------------------------------------------------------------------ DP5
// Band::relativeError (Band.cpp:87-101) of a fermion band on DG coefficients:
// Function's pointwise operations act on the element-vertex values c0 (+ cj),
// integrate(rhs) of the P1 result is exact per element:
//   sum_e (sum_v r_v^2 + (sum_v r_v)^2),  r_v = err_v / (rtol min(|p_v|,|1-p_v|) + atol)
// (penalty 0, the Band default); the controller divides by (dim+1)(dim+2) nel.
constexpr int kRelThreads = 256;

template <int NC, bool BOSON>
__global__ __launch_bounds__(kRelThreads) void dg_relerr_partial_kernel(int64_t nel, const double* __restrict__ err,
                                                                          const double* __restrict__ pop,
                                                                          int64_t stride, int64_t offset, double rtol,
                                                                          double atol, int nblk,
                                                                          double* __restrict__ part) {
    __shared__ double sw[kRelThreads / 64];
    const int b = blockIdx.y;
    const double* ev = err + (int64_t)b * stride + offset;
    const double* pv = pop + (int64_t)b * stride + offset;
    double s = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * kRelThreads + threadIdx.x; e < nel; e += (int64_t)nblk * kRelThreads) {
        double sr = 0.0, sr2 = 0.0;
#pragma unroll
        for (int a = 0; a < NC; ++a) {
            const double y = pv[e * NC] + (a ? pv[e * NC + a] : 0.0);
            const double x = ev[e * NC] + (a ? ev[e * NC + a] : 0.0);
            double scale;
            if constexpr (BOSON) {
                scale = rtol * y + atol;
            } else {
                const double m = fabs(y) < fabs(1.0 - y) ? fabs(y) : fabs(1.0 - y);
                scale = rtol * m + atol;
            }
            const double r = x / scale;
            sr += r;
            sr2 += r * r;
        }
        s += sr2 + sr * sr;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kRelThreads / 64; ++w) t += sw[w];
        part[(int64_t)b * nblk + blockIdx.x] = t;
    }
}

// per member: the error-norm sum, or the larger of the two bands' sums
// (MaterialStatus::relativeError takes the max over bands, MaterialStatus.cpp:189-196)
__global__ void dg_relerr_final_kernel(const double* __restrict__ part, const double* __restrict__ part2, int nblk,
                                       int B, double* __restrict__ norms) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double s = 0.0;
    for (int i = 0; i < nblk; ++i) s += part[(int64_t)b * nblk + i];
    if (part2) {
        double q = 0.0;
        for (int i = 0; i < nblk; ++i) q += part2[(int64_t)b * nblk + i];
        if (q >= s) s = q;
    }
    norms[b] = s;
}

// Electron energy flux of DG derivatives I (LF coefficients): per member
// sum_e int_e eps I = vol / ((dim+1)(dim+2)) sum_e (sum_v eps_v I_v + sum_v eps_v sum_v I_v),
// the exact P1 x P1 integral on each simplex with the element-vertex values
// (eps from the element table's energy form).  Block partials in a fixed
// order, then the final sum times the volume factor.
template <int DIM>
__global__ __launch_bounds__(kRelThreads) void dg_flux_partial_kernel(int64_t nel, const double* __restrict__ I,
                                                                        int64_t stride, const double* __restrict__ etab,
                                                                        int nblk, double* __restrict__ part) {
    constexpr int NC = DIM + 1;
    __shared__ double sw[kRelThreads / 64];
    const int b = blockIdx.y;
    const double* iv = I + (int64_t)b * stride;
    double s = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * kRelThreads + threadIdx.x; e < nel; e += (int64_t)nblk * kRelThreads) {
        const double* ef = etab + e * etab_stride<DIM>() + DIM * (DIM + 1);
        double sei = 0.0, se = 0.0, si = 0.0;
#pragma unroll
        for (int a = 0; a < NC; ++a) {
            const double ev = ef[0] + (a ? ef[a] : 0.0);
            const double x = iv[e * NC] + (a ? iv[e * NC + a] : 0.0);
            sei += ev * x;
            se += ev;
            si += x;
        }
        s += sei + se * si;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kRelThreads / 64; ++w) t += sw[w];
        part[(int64_t)b * nblk + blockIdx.x] = t;
    }
}

__global__ void dg_flux_final_kernel(const double* __restrict__ part, int nblk, int B, double scale,
                                     const double* __restrict__ h, double* __restrict__ flux) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || (h && h[b] == 0.0)) return;
    double s = 0.0;
    for (int i = 0; i < nblk; ++i) s += part[(int64_t)b * nblk + i];
    flux[b] = scale * s;
}

struct TupleDp5 {
    int32_t cap = 0;
    int64_t n = 0;
    int nblk = 1;
    double* k[7] = {};
    double* ys = nullptr;
    double* y5 = nullptr;
    double* errv = nullptr;
    double* y = nullptr;
    double* h = nullptr;
    int8_t* acc = nullptr;
    tg::MemberDev* md = nullptr;
    double* norms = nullptr;
    double* part = nullptr;
    double* part2 = nullptr;     // phonon band's partials (dynamical phonons)
    double* fpart = nullptr;     // energy-flux partials (cap x nblk)
    double* flux = nullptr;      // electron energy flux per stage and member (7 x cap)
    double* hlog = nullptr;      // accepted step sizes (cap x hlog_cap)
    size_t hlog_cap = 0;
    int32_t* running = nullptr;
    int32_t* running_host = nullptr;
    double* dense = nullptr;
    size_t dense_cap = 0;
};

void tuple_dp5_free(tg_tuple_ctx* c) {
    TupleDp5* w = (TupleDp5*)c->dp5;
    if (!w) return;
    void* bufs[] = {w->ys,    w->y5,   w->errv,  w->y,    w->h,       w->acc,  w->md,
                    w->norms, w->part, w->part2, w->fpart, w->flux, w->running, w->dense, w->hlog};
    for (int j = 0; j < 7; ++j)
        if (w->k[j]) (void)hipFree(w->k[j]);
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (w->running_host) (void)hipHostFree(w->running_host);
    delete w;
    c->dp5 = nullptr;
}

tg_status tuple_dp5_ensure(tg_tuple_ctx* c) {
    if (c->dp5) return TG_OK;
    TupleDp5* w = new (std::nothrow) TupleDp5();
    if (!w) return fail(c, TG_ERR_NOMEM, "DP5 work");
    c->dp5 = w;
    w->cap = c->max_batch;
    w->n = c->nel * (c->dim + 1) * (c->dyn ? 2 : 1);
    w->nblk = (int)std::min<int64_t>(64, (c->nel + kRelThreads - 1) / kRelThreads);
    const size_t sz = (size_t)w->n * w->cap;
    tg_status s;
    for (int j = 0; j < 7; ++j)
        if ((s = dalloc(c, &w->k[j], sz)) != TG_OK) return s;
    if ((s = dalloc(c, &w->ys, sz)) != TG_OK || (s = dalloc(c, &w->y5, sz)) != TG_OK ||
        (s = dalloc(c, &w->errv, sz)) != TG_OK || (s = dalloc(c, &w->y, sz)) != TG_OK ||
        (s = dalloc(c, &w->h, (size_t)w->cap)) != TG_OK || (s = dalloc(c, &w->acc, (size_t)w->cap)) != TG_OK ||
        (s = dalloc(c, &w->md, (size_t)w->cap)) != TG_OK || (s = dalloc(c, &w->norms, (size_t)w->cap)) != TG_OK ||
        (s = dalloc(c, &w->part, (size_t)w->cap * w->nblk)) != TG_OK ||
        (s = dalloc(c, &w->part2, (size_t)w->cap * w->nblk)) != TG_OK ||
        (s = dalloc(c, &w->fpart, (size_t)w->cap * w->nblk)) != TG_OK ||
        (s = dalloc(c, &w->flux, (size_t)7 * w->cap)) != TG_OK || (s = dalloc(c, &w->running, 1)) != TG_OK)
        return s;
    TT_HIP(c, hipHostMalloc((void**)&w->running_host, sizeof(int32_t), hipHostMallocDefault));
    return TG_OK;
}

// the electron energy flux of the derivative k (electron band) of every
// running member into flux[0..B) (kernel launches only)
tg_status tuple_flux(tg_tuple_ctx* c, TupleDp5* w, const double* k, int32_t B, const double* h, double* flux) {
    const dim3 grid(w->nblk, B);
    if (c->dim == 2)
        hipLaunchKernelGGL(dg_flux_partial_kernel<2>, grid, dim3(kRelThreads), 0, c->stream, c->nel, k, w->n,
                           c->d_etab, w->nblk, w->fpart);
    else
        hipLaunchKernelGGL(dg_flux_partial_kernel<3>, grid, dim3(kRelThreads), 0, c->stream, c->nel, k, w->n,
                           c->d_etab, w->nblk, w->fpart);
    const double scale = c->vol / (double)((c->dim + 1) * (c->dim + 2));
    hipLaunchKernelGGL(dg_flux_final_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, c->stream, w->fpart,
                       w->nblk, B, scale, h, flux);
    TT_HIP(c, hipGetLastError());
    return TG_OK;
}

// one attempted step for all members (kernel launches only; captured)
tg_status tuple_attempt(tg_tuple_ctx* c, TupleDp5* w, int32_t B, const tg::Ctl& ctl, double rtol, double atol,
                        double* d_dense, double* d_hlog) {
    using namespace tg;
    hipStream_t st = c->stream;
    const int64_t n = w->n;
    double** k = w->k;
    const dim3 blk(256);
    const dim3 grd((unsigned)std::max<int64_t>(1, (n + 1023) / 1024), B);
    // stage state at t + cs h (the bath schedule, if any, is evaluated there)
    auto stage = [&](std::initializer_list<double> coefs, int nk, double* out, double cs) -> tg_status {
        StageArgs a{};
        int j = 0;
        for (double v : coefs) a.a[j++] = v;
        for (int q = 0; q < nk; ++q) a.k[q] = k[q];
        a.nk = nk;
        hipLaunchKernelGGL(stage_kernel, grd, blk, 0, st, n, B, w->y, a, w->h, out, ctl.bath, w->md, cs,
                           c->d_bath);
        TT_HIP(c, hipGetLastError());
        return TG_OK;
    };
    auto eval = [&](const double* in, int q) {
        tg_status r = propagate_device(c, in, c->d_bath, k[q], B);
        if (r == TG_OK && ctl.flux) r = tuple_flux(c, w, k[q], B, w->h, w->flux + (int64_t)q * w->cap);
        return r;
    };
    const dim3 cgrd((unsigned)((B + 255) / 256));
    hipLaunchKernelGGL(prep_kernel, cgrd, dim3(256), 0, st, w->md, ctl, w->h, w->running);
    TT_HIP(c, hipGetLastError());
    tg_status s;
    // stages k2..k6, y5 and k7 (DormandPrince54.hpp:565-578)
    if ((s = stage({A21}, 1, w->ys, C2)) != TG_OK || (s = eval(w->ys, 1)) != TG_OK) return s;
    if ((s = stage({A31, A32}, 2, w->ys, C3)) != TG_OK || (s = eval(w->ys, 2)) != TG_OK) return s;
    if ((s = stage({A41, A42, A43}, 3, w->ys, C4)) != TG_OK || (s = eval(w->ys, 3)) != TG_OK) return s;
    if ((s = stage({A51, A52, A53, A54}, 4, w->ys, C5)) != TG_OK || (s = eval(w->ys, 4)) != TG_OK) return s;
    if ((s = stage({A61, A62, A63, A64, A65}, 5, w->ys, 1.0)) != TG_OK || (s = eval(w->ys, 5)) != TG_OK) return s;
    if ((s = stage({B1, 0.0, B3, B4, B5, B6}, 6, w->y5, 1.0)) != TG_OK || (s = eval(w->y5, 6)) != TG_OK) return s;
    {   // err = y5 - y4 (:580)
        StageArgs a{};
        const double bh[7] = {BH1, 0.0, BH3, BH4, BH5, BH6, BH7};
        for (int q = 0; q < 7; ++q) {
            a.a[q] = bh[q];
            a.k[q] = k[q];
        }
        a.nk = 7;
        hipLaunchKernelGGL(error_kernel, grd, blk, 0, st, n, B, w->y, a, w->h, w->y5, w->errv);
        TT_HIP(c, hipGetLastError());
    }
    {
        const int64_t band = c->nel * (c->dim + 1);
#define TT_REL(NC, BOS, OFF, PART)                                                                              \
    hipLaunchKernelGGL((dg_relerr_partial_kernel<NC, BOS>), dim3(w->nblk, B), dim3(kRelThreads), 0, st, c->nel, \
                       w->errv, w->y5, w->n, OFF, rtol, atol, w->nblk, PART)
        if (c->dim == 2) {
            TT_REL(3, false, 0, w->part);
            if (c->dyn) TT_REL(3, true, band, w->part2);
        } else {
            TT_REL(4, false, 0, w->part);
            if (c->dyn) TT_REL(4, true, band, w->part2);
        }
#undef TT_REL
    }
    hipLaunchKernelGGL(dg_relerr_final_kernel, cgrd, dim3(256), 0, st, w->part, c->dyn ? w->part2 : (double*)nullptr,
                       w->nblk, B, w->norms);
    hipLaunchKernelGGL(controller_kernel, cgrd, dim3(256), 0, st, w->md, ctl, w->norms, w->acc, d_hlog,
                       w->running);
    TT_HIP(c, hipGetLastError());
    if (d_dense) {
        hipLaunchKernelGGL(dense_kernel, dim3((unsigned)std::max<int64_t>(1, (n + 1023) / 1024), B), blk, 0, st, n,
                           w->md, ctl, w->y, k[0], k[2], k[3], k[4], k[5], k[6], d_dense);
        TT_HIP(c, hipGetLastError());
    }
    hipLaunchKernelGGL(accept_kernel, grd, blk, 0, st, n, w->acc, w->y, w->y5, k[0], k[6]);
    TT_HIP(c, hipGetLastError());
    TT_HIP(c, hipMemcpyAsync(w->running_host, w->running, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    return TG_OK;
}

}  // namespace

// ================================================================= C ABI
extern "C" const char* tg_tuple_last_error(const tg_tuple_ctx* c) { return c ? c->err : "null context"; }

extern "C" void tg_tuple_destroy(tg_tuple_ctx* c) {
    if (!c) return;
    tuple_dp5_free(c);
    for (int l = 0; l < 3; ++l) {
        void* dl[] = {c->d_doff[l], c->d_dothA[l], c->d_dothB[l], c->d_dpos[l], c->d_dq[l]};
        for (void* p : dl)
            if (p) (void)hipFree(p);
    }
    if (c->d_eps_ph) (void)hipFree(c->d_eps_ph);
    if (c->d_bsched) (void)hipFree(c->d_bsched);
    void* bufs[] = {c->d_eps, c->d_pts, c->d_tup, c->d_off0, c->d_off2, c->d_oth0, c->d_oth2, c->d_pos2,
                    c->d_q0,  c->d_q2,  c->d_f,   c->d_out,  c->d_bath, c->d_minv, c->d_etab, c->d_ft,
                    c->d_goff, c->d_gent, c->d_res4};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" tg_status tg_tuple_create(const tg_tuple_desc* d, tg_tuple_ctx** out) {
    if (!out) return TG_ERR_ARG;
    *out = nullptr;
    if (!d) return TG_ERR_ARG;
    tg_tuple_ctx* c = new (std::nothrow) tg_tuple_ctx();
    if (!c) return TG_ERR_NOMEM;
    auto bail = [&](tg_status s) { *out = c; return s; };   // caller reads the message, then destroys
    if (d->dim != 2 && d->dim != 3) return bail(fail(c, TG_ERR_ARG, "dim %d: the tuple operator covers 2D and 3D", d->dim));
    c->dim = d->dim;
    c->n = 1;
    c->nel = tg::factorial(d->dim);
    // the zone: an axis-aligned box (all-zero sides) or a general
    // parallelepiped mesh with its own zone sides
    bool general = false;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) general |= d->sides[a][b] != 0.0;
    for (int q = 0; q < 3; ++q) {
        c->nsec[q] = q < d->dim ? d->nsec[q] : 1;
        c->origin[q] = q < d->dim ? d->origin[q] : 0.0;
        c->extent[q] = q < d->dim ? d->extent[q] : 1.0;
    }
    for (int q = 0; q < d->dim; ++q) {
        if (d->nsec[q] < 1 || (!general && !(d->extent[q] > 0.0)))
            return bail(fail(c, TG_ERR_ARG, "axis %d: nsec %d, extent %g", q, d->nsec[q], d->extent[q]));
        c->n *= d->nsec[q] + 1;
        c->nel *= d->nsec[q];
    }
    {
        // dG[d][c] = component d of side c / nsec[c]; zone[d][c] likewise of zone side c
        const int D = d->dim;
        double S[3][3] = {}, Z[3][3] = {};
        for (int cc = 0; cc < D; ++cc)
            for (int q = 0; q < D; ++q) {
                S[q][cc] = general ? d->sides[cc][q] : (q == cc ? d->extent[q] : 0.0);
                Z[q][cc] = general ? d->zone[cc][q] : (q == cc ? d->extent[q] : 0.0);
            }
        c->axis = 1;
        for (int q = 0; q < D; ++q)
            for (int cc = 0; cc < D; ++cc) {
                c->dG[q][cc] = S[q][cc] / (double)c->nsec[cc];
                c->zone[q][cc] = Z[q][cc];
                if (q != cc && S[q][cc] != 0.0) c->axis = 0;
            }
        for (int q = 0; q < D; ++q)
            if (c->axis && !(S[q][q] > 0.0)) c->axis = 0;
        if (c->axis)
            for (int q = 0; q < D; ++q) c->extent[q] = S[q][q];
        // determinant and inverse of dG (section volume, reduced coordinates)
        double det;
        if (D == 2) {
            det = c->dG[0][0] * c->dG[1][1] - c->dG[0][1] * c->dG[1][0];
        } else {
            det = c->dG[0][0] * (c->dG[1][1] * c->dG[2][2] - c->dG[1][2] * c->dG[2][1]) -
                  c->dG[0][1] * (c->dG[1][0] * c->dG[2][2] - c->dG[1][2] * c->dG[2][0]) +
                  c->dG[0][2] * (c->dG[1][0] * c->dG[2][1] - c->dG[1][1] * c->dG[2][0]);
        }
        if (!(std::fabs(det) > 0.0) || !std::isfinite(det))
            return bail(fail(c, TG_ERR_ARG, "the mesh sides are degenerate (determinant %g)", det));
        // umklapp vectors are combinations of the zone's sides: a degenerate
        // zone would repeat tuples under different umklapp ids
        if (d->n_umklapp > 0) {
            const double zdet = D == 2 ? Z[0][0] * Z[1][1] - Z[0][1] * Z[1][0]
                                       : Z[0][0] * (Z[1][1] * Z[2][2] - Z[1][2] * Z[2][1]) -
                                             Z[0][1] * (Z[1][0] * Z[2][2] - Z[1][2] * Z[2][0]) +
                                             Z[0][2] * (Z[1][0] * Z[2][1] - Z[1][1] * Z[2][0]);
            if (!(std::fabs(zdet) > 0.0) || !std::isfinite(zdet))
                return bail(fail(c, TG_ERR_ARG, "the zone sides are degenerate (determinant %g) with n_umklapp %d",
                                 zdet, d->n_umklapp));
        }
        double A[3][3] = {}, Ai[3][3] = {};
        for (int q = 0; q < D; ++q)
            for (int cc = 0; cc < D; ++cc) A[q][cc] = c->dG[q][cc];
        if (D == 2) {
            double in2[2][2], a2[2][2] = {{A[0][0], A[0][1]}, {A[1][0], A[1][1]}};
            invert_small<2>(a2, in2);
            for (int q = 0; q < 2; ++q)
                for (int cc = 0; cc < 2; ++cc) Ai[q][cc] = in2[q][cc];
        } else {
            invert_small<3>(A, Ai);
        }
        for (int q = 0; q < D; ++q)
            for (int cc = 0; cc < D; ++cc) c->dGinv[q][cc] = Ai[q][cc];
        c->gen_vol = std::fabs(det) / (double)tg::factorial(D);
    }
    if (!d->eps) return bail(fail(c, TG_ERR_ARG, "eps is null"));
    if (d->num_mc < 1 || d->num_mc > (1 << 20)) return bail(fail(c, TG_ERR_ARG, "num_mc %d outside [1, 2^20]", d->num_mc));
    if (d->n_umklapp < 0 || d->n_umklapp > 3) return bail(fail(c, TG_ERR_ARG, "n_umklapp %d outside [0, 3]", d->n_umklapp));
    if (d->max_batch < 1) return bail(fail(c, TG_ERR_ARG, "max_batch %d < 1", d->max_batch));
    c->omega0 = d->omega0;
    c->amplitude = d->amplitude;
    c->nu = d->n_umklapp;
    c->nmc = d->num_mc;
    c->rows = std::max(d->dim * 2 - 1, 2 * d->dim);
    c->min_det = d->min_det;
    c->threshold = d->threshold;
    c->max_batch = d->max_batch;
    c->dyn = d->dynamic_phonons != 0;
    if (c->axis) {
        double v = 1.0;
        for (int q = 0; q < d->dim; ++q) v *= c->extent[q] / (double)d->nsec[q];
        c->vol = std::fabs(v / (double)tg::factorial(d->dim));
    } else {
        c->vol = c->gen_vol;   // |det(section sides)| / dim!
    }
    // LF inverse mass (Mesh.cpp:95-123): M_ab = vol * dim! * int_ref phi_a phi_b
    {
        const int nc = d->dim + 1;
        double M[4][4] = {}, Mi[4][4] = {};
        const double f1 = 1.0 / tg::factorial(d->dim + 1), f2 = 1.0 / tg::factorial(d->dim + 2);
        const double scale = c->vol * tg::factorial(d->dim);
        for (int a = 0; a < nc; ++a)
            for (int b = 0; b < nc; ++b) {
                double m = (a == 0 && b == 0) ? 1.0 / tg::factorial(d->dim) : (a == 0 || b == 0) ? f1 : (a == b ? 2.0 * f2 : f2);
                M[a][b] = scale * m;
            }
        if (nc == 3) {
            double A[3][3], I3[3][3];
            for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) A[a][b] = M[a][b];
            invert_small<3>(A, I3);
            for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) Mi[a][b] = I3[a][b];
        } else {
            invert_small<4>(M, Mi);
        }
        for (int a = 0; a < nc; ++a)
            for (int b = 0; b < nc; ++b) c->minv[a * nc + b] = Mi[a][b];
    }
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipStreamCreate(&c->stream);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) return bail(fail(c, TG_ERR_HIP, "stream/event: %s", hipGetErrorString(e)));
    for (int64_t i = 0; i < c->n; ++i)
        if (!std::isfinite(d->eps[i])) return bail(fail(c, TG_ERR_NONFINITE, "eps[%lld] is not finite", (long long)i));
    c->pmin = c->pmax = d->omega0;
    if (d->eps_ph) {
        for (int64_t i = 0; i < c->n; ++i) {
            if (!std::isfinite(d->eps_ph[i]))
                return bail(fail(c, TG_ERR_NONFINITE, "eps_ph[%lld] is not finite", (long long)i));
            if (i == 0 || d->eps_ph[i] < c->pmin) c->pmin = d->eps_ph[i];
            if (i == 0 || d->eps_ph[i] > c->pmax) c->pmax = d->eps_ph[i];
        }
    }
    const int nbands = c->dyn ? 2 : 1;
    const int nc = d->dim + 1;
    tg_status s;
    if ((s = dalloc(c, &c->d_eps, c->n)) != TG_OK) return bail(s);
    if ((s = dalloc(c, &c->d_pts, (size_t)c->rows * c->nmc)) != TG_OK) return bail(s);
    if ((s = dalloc(c, &c->d_f, (size_t)c->max_batch * nbands * c->nel * nc)) != TG_OK) return bail(s);
    if ((s = dalloc(c, &c->d_out, (size_t)c->max_batch * nbands * c->nel * nc)) != TG_OK) return bail(s);
    if ((s = dalloc(c, &c->d_ft, (size_t)c->max_batch * nbands * c->nel * nc)) != TG_OK) return bail(s);
    if ((s = dalloc(c, &c->d_minv, 16)) != TG_OK) return bail(s);
    if ((s = dalloc(c, &c->d_bath, c->max_batch)) != TG_OK) return bail(s);
    if (d->eps_ph) {
        if ((s = dalloc(c, &c->d_eps_ph, c->n)) != TG_OK) return bail(s);
        e = hipMemcpy(c->d_eps_ph, d->eps_ph, sizeof(double) * c->n, hipMemcpyHostToDevice);
        if (e != hipSuccess) return bail(fail(c, TG_ERR_HIP, "upload: %s", hipGetErrorString(e)));
    }
    e = hipMemcpy(c->d_eps, d->eps, sizeof(double) * c->n, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(c->d_minv, c->minv, sizeof(double) * 16, hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(fail(c, TG_ERR_HIP, "upload: %s", hipGetErrorString(e)));
    {
        const int stride = d->dim == 2 ? etab_stride<2>() : etab_stride<3>();
        if ((s = dalloc(c, &c->d_etab, (size_t)c->nel * stride)) != TG_OK) return bail(s);
        Geo g = make_geo(c);
        dim3 grid((unsigned)((c->nel + 255) / 256));
        if (d->dim == 2)
            hipLaunchKernelGGL(elem_table_kernel<2>, grid, dim3(256), 0, c->stream, g, c->nel, c->d_etab);
        else
            hipLaunchKernelGGL(elem_table_kernel<3>, grid, dim3(256), 0, c->stream, g, c->nel, c->d_etab);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return bail(fail(c, TG_ERR_HIP, "element table: %s", hipGetErrorString(e)));
    }
    *out = c;
    return TG_OK;
}

extern "C" int32_t tg_tuple_point_rows(int32_t dim, int32_t nlegs) {
    return std::max(dim * (nlegs - 1) - 1, 2 * dim);
}

extern "C" tg_status tg_tuple4_create(const tg_tuple_desc* d, double kappa, tg_tuple_ctx** out) {
    if (!out) return TG_ERR_ARG;
    *out = nullptr;
    if (!d) return TG_ERR_ARG;
    tg_status s = tg_tuple_create(d, out);
    if (s != TG_OK) return s;
    tg_tuple_ctx* c = *out;
    if (d->dynamic_phonons) return fail(c, TG_ERR_ARG, "the 4-leg channel has no phonon band (dynamic_phonons must be 0)");
    if (!std::isfinite(kappa)) return fail(c, TG_ERR_ARG, "kappa is not finite");
    c->nlegs = 4;
    c->kappa = kappa;
    c->rows = tg_tuple_point_rows(c->dim, 4);
    return dalloc(c, &c->d_pts, (size_t)c->rows * c->nmc);
}

extern "C" tg_status tg_tuple_set_points(tg_tuple_ctx* c, const double* pts, int32_t flags) {
    if (!c || !pts) return TG_ERR_ARG;
    (void)flags;
    for (int64_t i = 0; i < (int64_t)c->rows * c->nmc; ++i)
        if (!std::isfinite(pts[i])) return fail(c, TG_ERR_NONFINITE, "pts[%lld] is not finite", (long long)i);
    TT_HIP(c, hipMemcpy(c->d_pts, pts, sizeof(double) * c->rows * c->nmc, hipMemcpyHostToDevice));
    tg_status s;
    if (c->nlegs == 4) {
        // no moments: every evaluation integrates the points directly
        if (!c->d_tup && (s = enumerate4(c)) != TG_OK) return s;
        c->have_points = true;
        return TG_OK;
    }
    if (!c->d_tup && (s = enumerate(c)) != TG_OK) return s;
    if ((s = moments(c)) != TG_OK) return s;
    c->have_points = true;
    return TG_OK;
}

extern "C" tg_status tg_tuple_set_indices(tg_tuple_ctx* c, const int32_t* idx, int64_t ntup) {
    if (!c || (ntup > 0 && !idx) || ntup < 0) return TG_ERR_ARG;
    int32_t nuv = 1;
    for (int d = 0; d < c->dim; ++d) nuv *= 2 * c->nu + 1;
    if (c->nlegs == 4) {
        // [ntup][5] in any order: each element sums its entries in list order
        for (int64_t t = 0; t < ntup; ++t) {
            for (int l = 0; l < 4; ++l)
                if (idx[t * 5 + l] < 0 || idx[t * 5 + l] >= c->nel)
                    return fail(c, TG_ERR_ARG, "tuple %lld: element %d outside [0, %lld)", (long long)t, idx[t * 5 + l],
                                (long long)c->nel);
            if (idx[t * 5 + 4] < 0 || idx[t * 5 + 4] >= nuv)
                return fail(c, TG_ERR_ARG, "tuple %lld: umklapp id %d outside [0, %d)", (long long)t, idx[t * 5 + 4], nuv);
        }
        if (ntup >= ((int64_t)1 << 29)) return fail(c, TG_ERR_ARG, "%lld tuples: at most 2^29 - 1", (long long)ntup);
        tg_status s;
        if ((s = dalloc(c, &c->d_tup, (size_t)ntup * 5)) != TG_OK) return s;
        if (ntup) TT_HIP(c, hipMemcpy(c->d_tup, idx, sizeof(int32_t) * 5 * ntup, hipMemcpyHostToDevice));
        c->ntup = ntup;
        return build_lists4(c, idx);
    }
    for (int64_t t = 0; t < ntup; ++t) {
        for (int l = 0; l < 3; ++l)
            if (idx[t * 4 + l] < 0 || idx[t * 4 + l] >= c->nel)
                return fail(c, TG_ERR_ARG, "tuple %lld: element %d outside [0, %lld)", (long long)t, idx[t * 4 + l],
                            (long long)c->nel);
        if (idx[t * 4 + 3] < 0 || idx[t * 4 + 3] >= nuv)
            return fail(c, TG_ERR_ARG, "tuple %lld: umklapp id %d outside [0, %d)", (long long)t, idx[t * 4 + 3], nuv);
    }
    // The leg lists are built from a list ordered by the leg-0 element.  The
    // reference writes one full sweep per umklapp vector
    // (ScatteringChannel.cpp:473-532), so a list in any other order is put in
    // leg-0 order by a stable counting sort (tuples of one e0 keep their
    // order; the propagate sums depend on the order only through rounding).
    std::vector<int32_t> sorted;
    bool ordered = true;
    for (int64_t t = 1; t < ntup && ordered; ++t) ordered = idx[t * 4] >= idx[(t - 1) * 4];
    if (!ordered) {
        std::vector<int64_t> start((size_t)c->nel + 1, 0);
        for (int64_t t = 0; t < ntup; ++t) ++start[(size_t)idx[t * 4] + 1];
        for (int64_t e = 0; e < c->nel; ++e) start[(size_t)e + 1] += start[(size_t)e];
        sorted.resize((size_t)ntup * 4);
        for (int64_t t = 0; t < ntup; ++t) {
            const int64_t d = start[(size_t)idx[t * 4]]++;
            for (int q = 0; q < 4; ++q) sorted[(size_t)d * 4 + q] = idx[t * 4 + q];
        }
        idx = sorted.data();
    }
    tg_status s;
    if ((s = dalloc(c, &c->d_tup, (size_t)ntup * 4)) != TG_OK) return s;
    if (ntup) TT_HIP(c, hipMemcpy(c->d_tup, idx, sizeof(int32_t) * 4 * ntup, hipMemcpyHostToDevice));
    c->ntup = ntup;
    if ((s = build_lists(c)) != TG_OK) return s;
    if (c->have_points && (s = moments(c)) != TG_OK) return s;
    return TG_OK;
}

extern "C" tg_status tg_tuple_count(const tg_tuple_ctx* c, int64_t* ntup) {
    if (!c || !ntup) return TG_ERR_ARG;
    *ntup = c->d_tup ? c->ntup : -1;
    return TG_OK;
}

extern "C" tg_status tg_tuple_copy_indices(tg_tuple_ctx* c, int32_t* idx, int64_t cap) {
    if (!c || !idx) return TG_ERR_ARG;
    if (!c->d_tup) return fail(c, TG_ERR_STATE, "no tuple list yet (call tg_tuple_set_points)");
    if (cap < c->ntup) return fail(c, TG_ERR_ARG, "cap %lld < %lld tuples", (long long)cap, (long long)c->ntup);
    const size_t rec = (size_t)c->nlegs + 1;
    if (c->ntup) TT_HIP(c, hipMemcpy(idx, c->d_tup, sizeof(int32_t) * rec * c->ntup, hipMemcpyDeviceToHost));
    return TG_OK;
}

extern "C" tg_status tg_tuple_set_bath_schedule(tg_tuple_ctx* c, const tg_bath_schedule* s) {
    if (!c) return TG_ERR_ARG;
    if (c->nlegs == 4) return fail(c, TG_ERR_STATE, "the 4-leg channel has no bath");
    if (!s) {
        c->h_bsched.clear();
        c->bs_ns = c->bs_batch = 0;
        return TG_OK;
    }
    if (c->dyn) return fail(c, TG_ERR_STATE, "the phonon band evolves (dynamic_phonons): it has no bath");
    if (!s->samples || s->n_samples < 2 || s->batch < 1 || s->batch > c->max_batch)
        return fail(c, TG_ERR_ARG, "bath schedule needs samples, n_samples >= 2 and batch in [1, max_batch]");
    if (!(s->dt > 0.0) || !std::isfinite(s->t_start) || !std::isfinite(s->dt) || !(s->atol >= 0.0) ||
        !(s->rtol >= 0.0) || (!(s->atol > 0.0) && !(s->rtol > 0.0)))
        return fail(c, TG_ERR_ARG, "bath schedule needs dt > 0, finite t_start, atol, rtol >= 0 not both 0");
    const size_t cnt = (size_t)s->batch * s->n_samples;
    for (size_t i = 0; i < cnt; ++i)
        if (!std::isfinite(s->samples[i])) return fail(c, TG_ERR_NONFINITE, "bath schedule sample %zu is not finite", i);
    tg_status r;
    if ((r = dalloc(c, &c->d_bsched, cnt)) != TG_OK) return r;
    TT_HIP(c, hipMemcpy(c->d_bsched, s->samples, sizeof(double) * cnt, hipMemcpyHostToDevice));
    c->h_bsched.assign(s->samples, s->samples + cnt);
    c->bs_ns = s->n_samples;
    c->bs_batch = s->batch;
    c->bs_t0 = s->t_start;
    c->bs_dt = s->dt;
    c->bs_atol = s->atol;
    c->bs_rtol = s->rtol;
    return TG_OK;
}

extern "C" tg_status tg_tuple_set_amplitude(tg_tuple_ctx* c, double amplitude) {
    if (!c || !std::isfinite(amplitude)) return TG_ERR_ARG;
    c->amplitude = amplitude;
    return TG_OK;
}

template <bool RATES>
tg_status tuple_apply(tg_tuple_ctx* c, const double* f, const double* bath_n, double* out, int64_t nel, int32_t batch,
                      int32_t flags) {
    if (!c || !f || (!bath_n && !c->dyn && c->nlegs == 3) || !out) return TG_ERR_ARG;
    if (RATES && c->nlegs == 4) return fail(c, TG_ERR_STATE, "scattering rates: 3-leg channels only");
    if (!c->have_points) return fail(c, TG_ERR_STATE, "no MC points yet (call tg_tuple_set_points)");
    if (nel != c->nel) return fail(c, TG_ERR_ARG, "nel=%lld, mesh has %lld elements", (long long)nel, (long long)c->nel);
    if (batch < 1 || batch > c->max_batch)
        return fail(c, TG_ERR_ARG, "batch %d outside [1, max_batch=%d]", batch, c->max_batch);
    c->time_kernel = (flags & TG_TIME_KERNEL) != 0;
    const size_t len = (size_t)c->nel * (c->dim + 1) * batch * (c->dyn ? 2 : 1);
    const size_t olen = RATES ? (size_t)c->nel * (c->dim + 1) * batch : len;   // rates: the electron band
    if (c->nlegs == 4) {
        if (flags & TG_PTR_DEVICE) {
            tg_status s = propagate4_device(c, f, out, batch);
            if (s != TG_OK) return s;
        } else {
            TT_HIP(c, hipMemcpyAsync(c->d_f, f, sizeof(double) * len, hipMemcpyHostToDevice, c->stream));
            tg_status s = propagate4_device(c, c->d_f, c->d_out, batch);
            if (s != TG_OK) return s;
            TT_HIP(c, hipMemcpyAsync(out, c->d_out, sizeof(double) * len, hipMemcpyDeviceToHost, c->stream));
        }
    } else if (flags & TG_PTR_DEVICE) {
        tg_status s = propagate_device<RATES>(c, f, bath_n, out, batch);
        if (s != TG_OK) return s;
    } else {
        TT_HIP(c, hipMemcpyAsync(c->d_f, f, sizeof(double) * len, hipMemcpyHostToDevice, c->stream));
        if (bath_n)
            TT_HIP(c, hipMemcpyAsync(c->d_bath, bath_n, sizeof(double) * batch, hipMemcpyHostToDevice, c->stream));
        tg_status s = propagate_device<RATES>(c, c->d_f, c->d_bath, c->d_out, batch);
        if (s != TG_OK) return s;
        TT_HIP(c, hipMemcpyAsync(out, c->d_out, sizeof(double) * olen, hipMemcpyDeviceToHost, c->stream));
    }
    TT_HIP(c, hipStreamSynchronize(c->stream));
    if (c->time_kernel) TT_HIP(c, hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
    return TG_OK;
}

extern "C" tg_status tg_tuple_propagate(tg_tuple_ctx* c, const double* f, const double* bath_n, double* out,
                                        int64_t nel, int32_t batch, int32_t flags) {
    return tuple_apply<false>(c, f, bath_n, out, nel, batch, flags);
}

extern "C" tg_status tg_tuple_rates(tg_tuple_ctx* c, const double* f, const double* bath_n, double* out,
                                    int64_t nel, int32_t batch, int32_t flags) {
    return tuple_apply<true>(c, f, bath_n, out, nel, batch, flags);
}

extern "C" tg_status tg_tuple_last_kernel_ms(tg_tuple_ctx* c, double* ms) {
    if (!c || !ms) return TG_ERR_ARG;
    *ms = c->last_ms;
    return TG_OK;
}

extern "C" void* tg_tuple_stream(const tg_tuple_ctx* c) { return c ? (void*)c->stream : nullptr; }

// .scattIndices file of a channel (ScatteringChannel.cpp:422-459,520-537):
// for each leg and axis the int nSecSplits, an unsigned long tuple count,
// then per tuple the legs' element ids and the umklapp id, all native-endian.
extern "C" tg_status tg_tuple_save_indices(tg_tuple_ctx* c, const char* path) {
    if (!c || !path) return TG_ERR_ARG;
    if (!c->d_tup) return fail(c, TG_ERR_STATE, "no tuple list yet");
    const size_t rec = (size_t)c->nlegs + 1;
    std::vector<int32_t> idx((size_t)c->ntup * rec);
    if (c->ntup) TT_HIP(c, hipMemcpy(idx.data(), c->d_tup, sizeof(int32_t) * rec * c->ntup, hipMemcpyDeviceToHost));
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return fail(c, TG_ERR_ARG, "cannot open %s for writing", path);
    bool ok = true;
    for (int l = 0; l < c->nlegs; ++l)
        for (int d = 0; d < c->dim; ++d) ok &= std::fwrite(&c->nsec[d], sizeof(int32_t), 1, fp) == 1;
    unsigned long total = (unsigned long)c->ntup;
    ok &= std::fwrite(&total, sizeof(unsigned long), 1, fp) == 1;
    if (c->ntup) ok &= std::fwrite(idx.data(), sizeof(int32_t) * rec, (size_t)c->ntup, fp) == (size_t)c->ntup;
    ok &= std::fclose(fp) == 0;
    return ok ? TG_OK : fail(c, TG_ERR_ARG, "short write to %s", path);
}

extern "C" tg_status tg_tuple_load_indices(tg_tuple_ctx* c, const char* path) {
    if (!c || !path) return TG_ERR_ARG;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(c, TG_ERR_ARG, "cannot open %s", path);
    bool same = true;
    for (int l = 0; l < c->nlegs; ++l)
        for (int d = 0; d < c->dim; ++d) {
            int32_t ns = 0;
            same &= std::fread(&ns, sizeof(int32_t), 1, fp) == 1 && ns == c->nsec[d];
        }
    unsigned long total = 0;
    if (!same || std::fread(&total, sizeof(unsigned long), 1, fp) != 1) {
        std::fclose(fp);
        return fail(c, TG_ERR_ARG, "%s: mesh sections differ from this channel's", path);
    }
    const size_t rec = (size_t)c->nlegs + 1;
    std::vector<int32_t> idx((size_t)total * rec);
    bool ok = total == 0 || std::fread(idx.data(), sizeof(int32_t) * rec, total, fp) == total;
    std::fclose(fp);
    if (!ok) return fail(c, TG_ERR_ARG, "%s: truncated tuple list", path);
    return tg_tuple_set_indices(c, idx.data(), (int64_t)total);
}

// MaterialTimeStatus::propagate (MaterialTimeStatus.cpp:102-137) with the
// tuple operator: DP5(4) (DormandPrince54.hpp:479-665) of the electron band's
// DG coefficients for `batch` members, each with its own steps; the phonon
// band is a bath (its constraint makes its error zero).  One attempted step
// is captured as a HIP graph and replayed one attempt ahead of the "any
// member running" flag it reads, as in tg_dp5_integrate.
extern "C" tg_status tg_tuple_dp5(tg_tuple_ctx* c, double* f, int64_t nel, int32_t batch, double t0, double t1,
                                  double dense_dt, double rtol, double atol, const double* bath_n, int32_t flags,
                                  double* dense_out, int32_t max_dense, tg_step_stats* stats,
                                  double* h_accepted, int32_t h_cap) {
    using namespace tg;
    const bool sched_on = (flags & TG_BATH_SCHEDULE) != 0;
    if (h_accepted && h_cap < 1) return TG_ERR_ARG;
    if (!c || !f || (!bath_n && !c->dyn && !sched_on && c->nlegs == 3)) return TG_ERR_ARG;
    if (!c->have_points) return fail(c, TG_ERR_STATE, "no MC points yet (call tg_tuple_set_points)");
    if (nel != c->nel) return fail(c, TG_ERR_ARG, "nel=%lld, mesh has %lld elements", (long long)nel, (long long)c->nel);
    if (batch < 1 || batch > c->max_batch)
        return fail(c, TG_ERR_ARG, "batch %d outside [1, max_batch=%d]", batch, c->max_batch);
    if (!(dense_dt > 0.0)) return fail(c, TG_ERR_ARG, "dense_dt must be > 0");
    if (dense_out && max_dense < 1) return fail(c, TG_ERR_ARG, "max_dense < 1");
    if (sched_on && (c->dyn || c->bs_ns < 2))
        return fail(c, TG_ERR_STATE, "TG_BATH_SCHEDULE needs a bath channel and tg_tuple_set_bath_schedule");
    if (sched_on && batch > c->bs_batch)
        return fail(c, TG_ERR_ARG, "batch %d exceeds the %d members of the bath schedule", batch, c->bs_batch);
    tg_status s;
    if ((s = tuple_dp5_ensure(c)) != TG_OK) return s;
    TupleDp5* w = (TupleDp5*)c->dp5;
    const bool dev = flags & TG_PTR_DEVICE;
    hipStream_t st = c->stream;
    const size_t vec = (size_t)w->n * batch;
    TT_HIP(c, hipMemcpyAsync(w->y, f, sizeof(double) * vec, dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    // the bath: constant per member, or the schedule's n_b(t0) for k1
    BathSched bs{};
    std::vector<double> bath0;
    if (sched_on) {
        bs.tab = c->h_bsched.data();
        bs.ns = c->bs_ns;
        bs.t_start = c->bs_t0;
        bs.dt = c->bs_dt;
        bs.atol = c->bs_atol;
        bs.rtol = c->bs_rtol;
        bath0.resize(batch);
        for (int b = 0; b < batch; ++b) bath0[b] = sched_value(bs.tab + (size_t)b * bs.ns, bs.ns, bs.t_start, bs.dt, t0);
        TT_HIP(c, hipMemcpyAsync(c->d_bath, bath0.data(), sizeof(double) * batch, hipMemcpyHostToDevice, st));
    } else if (bath_n) {
        TT_HIP(c, hipMemcpyAsync(c->d_bath, bath_n, sizeof(double) * batch,
                                 dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    }
    double* d_dense = nullptr;
    if (dense_out) {
        if (dev) {
            d_dense = dense_out;
        } else {
            const size_t need = (size_t)max_dense * vec;
            if (w->dense_cap < need) {
                if (w->dense) (void)hipFree(w->dense);
                w->dense = nullptr;
                w->dense_cap = 0;
                if ((s = dalloc(c, &w->dense, need)) != TG_OK) return s;
                w->dense_cap = need;
            }
            d_dense = w->dense;
        }
    }
    Ctl ctl{};
    ctl.t1 = t1;
    ctl.dense_dt = dense_dt;
    ctl.min_h = dense_dt / 10000.;
    ctl.denom = (double)((c->dim + 1) * (c->dim + 2)) * (double)c->nel;
    ctl.B = batch;
    ctl.replay = 0;
    ctl.h_cap = h_accepted ? h_cap : 1;
    ctl.max_dense = d_dense ? max_dense : 0;
    if (h_accepted) {
        const size_t need = (size_t)batch * h_cap;
        if (w->hlog_cap < need) {
            if ((s = dalloc(c, &w->hlog, need)) != TG_OK) return s;
            w->hlog_cap = need;
        }
        // entries past a member's accepted count keep the caller's values
        TT_HIP(c, hipMemcpyAsync(w->hlog, h_accepted, sizeof(double) * need, hipMemcpyHostToDevice, st));
    }
    ctl.sched = nullptr;
    ctl.nsched = nullptr;
    ctl.bath = bs;
    ctl.bath.tab = sched_on ? c->d_bsched : nullptr;
    // with dynamical phonons or four legs there is no bath
    ctl.flux = (c->dyn || c->nlegs == 4) ? nullptr : w->flux;
    ctl.flux_stride = w->cap;
    std::vector<MemberDev> ms(batch);
    for (int b = 0; b < batch; ++b) {
        MemberDev& m = ms[b];
        std::memset(&m, 0, sizeof(m));
        m.t = t0;
        // the first step is vetted too (DormandPrince54.hpp:562-564)
        m.h = std::fmin(dense_dt, step_vetoer(bs, sched_on ? bs.tab + (size_t)b * bs.ns : nullptr, t0, dense_dt,
                                              dense_dt));
        m.dense_time = t0 + dense_dt;
        m.olderror = 1.0;
        m.st.evals = 1;
    }
    TT_HIP(c, hipMemcpyAsync(w->md, ms.data(), sizeof(MemberDev) * batch, hipMemcpyHostToDevice, st));
    const bool timing = c->time_kernel;
    c->time_kernel = false;   // no event records inside the graph
    s = propagate_device(c, w->y, c->d_bath, w->k[0], batch);   // k1 = timeOp(y0)
    if (s == TG_OK && ctl.flux) s = tuple_flux(c, w, w->k[0], batch, nullptr, w->flux);
    if (s != TG_OK) { c->time_kernel = timing; return s; }
    TT_HIP(c, hipStreamSynchronize(st));

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    TT_HIP(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    tg_status cap = tuple_attempt(c, w, batch, ctl, rtol, atol, d_dense, h_accepted ? w->hlog : nullptr);
    c->time_kernel = timing;
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    if (cap != TG_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return cap;
    }
    if (ec != hipSuccess) return fail(c, TG_ERR_HIP, "graph capture of a DP5 attempt: %s", hipGetErrorString(ec));
    hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(c, TG_ERR_HIP, "graph instantiation: %s", hipGetErrorString(ei));

    tg_status rc = TG_OK;
    hipEvent_t ev[2] = {nullptr, nullptr};
    for (auto& e : ev)
        if (rc == TG_OK && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            rc = fail(c, TG_ERR_HIP, "event creation");
    const int64_t max_attempts = 100000000;   // guards a controller that never ends
    for (int64_t a = 0; rc == TG_OK; ++a) {
        if (a >= max_attempts) {
            rc = fail(c, TG_ERR_STATE, "more than %lld DP5 attempts", (long long)max_attempts);
            break;
        }
        if (hipGraphLaunch(exec, st) != hipSuccess || hipEventRecord(ev[a & 1], st) != hipSuccess) {
            rc = fail(c, TG_ERR_HIP, "graph launch of a DP5 attempt");
            break;
        }
        if (a == 0) continue;
        if (hipEventSynchronize(ev[(a - 1) & 1]) != hipSuccess) {
            rc = fail(c, TG_ERR_HIP, "DP5 attempt failed on the device");
            break;
        }
        if (*(volatile int32_t*)w->running_host == 0) break;
    }
    const hipError_t es = hipStreamSynchronize(st);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipGraphExecDestroy(exec);
    if (rc != TG_OK) return rc;
    if (es != hipSuccess) return fail(c, TG_ERR_HIP, "DP5 attempts: %s", hipGetErrorString(es));
    TT_HIP(c, hipMemcpyAsync(ms.data(), w->md, sizeof(MemberDev) * batch, hipMemcpyDeviceToHost, st));
    TT_HIP(c, hipMemcpyAsync(f, w->y, sizeof(double) * vec, dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    if (dense_out && !dev)
        TT_HIP(c, hipMemcpyAsync(dense_out, d_dense, sizeof(double) * (size_t)max_dense * vec, hipMemcpyDeviceToHost, st));
    if (h_accepted)
        TT_HIP(c, hipMemcpyAsync(h_accepted, w->hlog, sizeof(double) * (size_t)batch * h_cap, hipMemcpyDeviceToHost, st));
    TT_HIP(c, hipStreamSynchronize(st));
    if (!(flags & TG_QUIET) && batch == 1) printf("Number of adaptive steps: %lld\n", (long long)ms[0].st.attempted);
    for (int b = 0; b < batch; ++b) {
        ms[b].st.failed = ms[b].failed ? 1 : 0;
        ms[b].st.bath_energy = ms[b].ledger;
        if (stats) stats[b] = ms[b].st;
        if (ms[b].nonfinite) rc = TG_ERR_NONFINITE;
        else if (ms[b].failed && rc == TG_OK) rc = TG_ERR_STEP_UNDERFLOW;
    }
    if (rc != TG_OK) fail(c, rc, rc == TG_ERR_NONFINITE ? "non-finite error norm" : "step size underflow");
    return rc;
}
