/*
 * mc_oracle.c -- CPU FP64 restatement of the reference's own scattering
 * operator: element tuples, the Monte-Carlo delta-shell integrator, the
 * per-tuple propagator and the inverse mass.
 *
 * TEST INFRASTRUCTURE ONLY (see tortoise_oracle.h): the checker of the HIP
 * tuple operator, and the "reference algorithm" CPU timing.
 *
 * Reference files restated (paths relative to the TORTOISE tree):
 *   include/Physics/Scattering/IntegratorCore.hpp:153-155   det of D_C
 *   include/Physics/Scattering/IntegratorCore.hpp:260-357   scatteringIntegrationTypeB
 *   src/Physics/Scattering/ScatteringChannel.cpp:285-352    D and N of a tuple
 *   src/Physics/Scattering/ScatteringChannel.cpp:355-381    output forms
 *   src/Physics/Scattering/ScatteringChannel.cpp:382-398    amplitude factor
 *   src/Physics/Scattering/ScatteringChannel.cpp:421-538    tuple screening
 *   src/Physics/Scattering/ScatteringChannel.cpp:618-670    propagator + sum
 *   src/Physics/Scattering/ScatteringChannel.cpp:677-688    MC point rows
 *   include/Geometry/GeometryCore/Point.hpp:44-68           simplex sampling
 *   include/Physics/Scattering/ScatteringChannel.hpp:205-255 energy window
 *   src/Geometry/Structured/DirectSpace/Mesh.cpp:95-123     mass matrix
 *   Function.cpp:801-804                                     applyInverseMass
 *
 * Spec decisions (the reference leaves them to Eigen / std::rand / TBB):
 *   - random numbers: splitmix64 (Steele, Lea, Flood 2014), seeded by the
 *     caller; the reference draws from Eigen's Random (std::rand);
 *   - D_C^-1: Gauss-Jordan elimination with partial pivoting in a fixed
 *     operation order (the HIP path uses the same steps, so the inside-simplex
 *     decisions of the two agree point by point);
 *   - tuple screening uses the operator's own MC points (the reference draws
 *     fresh ones for the screening integral);
 *   - mass matrix: exact rational moments of the reference simplex (the
 *     reference sums a 14-digit order-2 Gauss rule: 1e-14 relative in 2D);
 *   - candidate pruning: element energy windows and momentum bounding boxes
 *     (the reference prunes by section boxes first); both keep every tuple
 *     whose delta shell meets the simplices.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "tortoise_oracle.h"

static int mc_fact(int d) { int r = 1; for (int i = 2; i <= d; ++i) r *= i; return r; }

/* ================================================================ points */
static uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double unit_draw(uint64_t* s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

/* Uniform point in the reference simplex of dimension d (1..3) from the unit
 * cube by folding (Point.hpp:44-68; 3D: Rocchini & Cignoni 2000). */
static void simplex_point(int d, uint64_t* s, double* x) {
    for (int i = 0; i < d; ++i) x[i] = unit_draw(s);
    if (d == 2) {
        if (x[0] + x[1] > 1.0) { x[0] = 1.0 - x[0]; x[1] = 1.0 - x[1]; }
    } else if (d == 3) {
        if (x[0] + x[1] > 1.0) { x[0] = 1.0 - x[0]; x[1] = 1.0 - x[1]; }
        if (x[1] + x[2] > 1.0) {
            double t = x[2];
            x[2] = 1.0 - x[0] - x[1];
            x[1] = 1.0 - t;
        } else if (x[0] + x[1] + x[2] > 1.0) {
            double t = x[2];
            x[2] = x[0] + x[1] + x[2] - 1.0;
            x[0] = 1.0 - x[1] - t;
        }
    }
}

int32_t or_mc_rows(int dim, int nlegs) {
    int a = dim * (nlegs - 1) - 1, b = 2 * dim;
    return a > b ? a : b;
}

/* MC point matrix of a channel (ScatteringChannel.cpp:677-688), row-major
 * [rows][nmc]: legs 0..nlegs-3 fill rows leg*dim.., the partially free leg
 * fills the bottom dim-1 rows; rows no leg uses are zero. */
void or_mc_points(int dim, int nlegs, int32_t nmc, uint64_t seed, double* pts) {
    int rows = or_mc_rows(dim, nlegs);
    memset(pts, 0, sizeof(double) * (size_t)rows * nmc);
    uint64_t s = seed;
    double x[3];
    for (int leg = 0; leg < nlegs - 2; ++leg)
        for (int32_t m = 0; m < nmc; ++m) {
            simplex_point(dim, &s, x);
            for (int d = 0; d < dim; ++d) pts[(size_t)(leg * dim + d) * nmc + m] = x[d];
        }
    if (dim > 1)
        for (int32_t m = 0; m < nmc; ++m) {
            simplex_point(dim - 1, &s, x);
            for (int d = 0; d < dim - 1; ++d) pts[(size_t)(rows - (dim - 1) + d) * nmc + m] = x[d];
        }
}

/* ============================================================ small LA */
/* Inverse of the n x n matrix A (row-major, n <= 4) by Gauss-Jordan with
 * partial pivoting; returns the determinant (0 when singular). */
double or_mc_invert(int n, const double* A, double* inv) {
    double a[4][8];
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < 2 * n; ++c) a[r][c] = c < n ? A[r * n + c] : (c - n == r ? 1.0 : 0.0);
    double det = 1.0;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(a[r][c]) > fabs(a[p][c])) p = r;
        if (a[p][c] == 0.0) return 0.0;
        if (p != c) {
            for (int q = 0; q < 2 * n; ++q) { double t = a[p][q]; a[p][q] = a[c][q]; a[c][q] = t; }
            det = -det;
        }
        double piv = a[c][c];
        det *= piv;
        for (int q = 0; q < 2 * n; ++q) a[c][q] /= piv;
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double m = a[r][c];
            for (int q = 0; q < 2 * n; ++q) a[r][q] -= m * a[c][q];
        }
    }
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) inv[r * n + c] = a[r][n + c];
    return det;
}

/* ====================================================== integrator core */
static double lf_apply(int dim, const double* form, const double* x) {
    double v = form[0];
    for (int d = 0; d < dim; ++d) v += form[1 + d] * x[d];
    return v;
}

/* The legs' points of MC sample m (IntegratorCore.hpp:283-320): free
 * coordinates from the point matrix, the D_C-constrained ones from
 *   alpha = -D_C^-1 (D_F K_F - N),
 * then the inside tests of the two constrained legs.  Returns 1 inside. */
static int shell_point(int dim, int nlegs, const double* D, const double* N, const double* inv,
                       const double* pts, int32_t nmc, int32_t m, double K[4][3]) {
    const int nc = dim + 1, ncol = nlegs * dim, rows = or_mc_rows(dim, nlegs);
    const int c0 = (nlegs - 1) * dim - 1;
    for (int leg = 0; leg < nlegs - 2; ++leg)
        for (int d = 0; d < dim; ++d) K[leg][d] = pts[(size_t)(leg * dim + d) * nmc + m];
    for (int d = 0; d < dim - 1; ++d) K[nlegs - 2][d] = pts[(size_t)(rows - (dim - 1) + d) * nmc + m];
    double rhs[4];
    for (int r = 0; r < nc; ++r) {
        double s = 0.0;
        for (int c = 0; c < c0; ++c) s += D[r * ncol + c] * K[c / dim][c % dim];
        rhs[r] = s - N[r];
    }
    for (int r = 0; r < nc; ++r) {
        double s = 0.0;
        for (int c = 0; c < nc; ++c) s += inv[r * nc + c] * rhs[c];
        if (r == 0) K[nlegs - 2][dim - 1] = -s;
        else K[nlegs - 1][r - 1] = -s;
    }
    int inside = 1;
    for (int leg = nlegs - 2; leg < nlegs; ++leg) {
        double sum = 0.0;
        for (int d = 0; d < dim; ++d) { inside &= K[leg][d] > 0.0; sum += K[leg][d]; }
        inside &= sum < 1.0;
    }
    return inside;
}

/* Divisor of the MC estimate (IntegratorCore.hpp:336-356):
 * |det D_C| * NMC * (dim-1)! * (dim!)^(nlegs-2); just |det D_C| for 1D two legs. */
static double shell_divisor(int dim, int nlegs, double det, int32_t nmc) {
    double w = fabs(det);
    if (nlegs == 2 && dim == 1) return w;
    w *= (double)nmc * (double)mc_fact(dim - 1);
    for (int k = 0; k < nlegs - 2; ++k) w *= (double)mc_fact(dim);
    return w;
}

/* scatteringIntegrationTypeB (IntegratorCore.hpp:260-357) with the output
 * forms of constructOutputformsLinForm (ScatteringChannel.cpp:355-381): the
 * output of leg l, basis j, is the integrand times phi_j at leg l's point,
 * phi_0 = 1, phi_j = x_{j-1}.
 *   D  (dim+1) x (nlegs*dim) row-major,  N  dim+1
 *   fa, fb: nlegs blocks of dim+1 linear-form coefficients (the "direct" and
 *   "time-reversed" products)
 *   pts: or_mc_points layout; out: nlegs*(dim+1).
 * Returns 0, or -1 if D_C is singular. */
int or_mc_integrate(int dim, int nlegs, const double* D, const double* N, const double* fa,
                    const double* fb, const double* pts, int32_t nmc, double* out) {
    const int nc = dim + 1, ncol = nlegs * dim;
    const int c0 = (nlegs - 1) * dim - 1;           /* first column of D_C */
    double DC[16], inv[16];
    for (int r = 0; r < nc; ++r)
        for (int c = 0; c < nc; ++c) DC[r * nc + c] = D[r * ncol + c0 + c];
    double det = or_mc_invert(nc, DC, inv);
    if (det == 0.0) return -1;
    for (int i = 0; i < nlegs * nc; ++i) out[i] = 0.0;
    const int32_t npt = (nlegs == 2 && dim == 1) ? 1 : nmc;
    for (int32_t m = 0; m < npt; ++m) {
        double K[4][3];                  /* point of every leg, reference coords */
        if (!shell_point(dim, nlegs, D, N, inv, pts, nmc, m, K)) continue;
        double pa = 1.0, pb = 1.0;
        for (int leg = 0; leg < nlegs; ++leg) {
            pa *= lf_apply(dim, fa + leg * nc, K[leg]);
            pb *= lf_apply(dim, fb + leg * nc, K[leg]);
        }
        double val = pa + pb;
        for (int leg = 0; leg < nlegs; ++leg) {
            out[leg * nc] += val;
            for (int d = 0; d < dim; ++d) out[leg * nc + 1 + d] += val * K[leg][d];
        }
    }
    double w = shell_divisor(dim, nlegs, det, nmc);
    for (int i = 0; i < nlegs * nc; ++i) out[i] /= w;
    return 0;
}

/* ==================================================== e-ph channel geometry */
/* The channel {el in, pn in, el out} with every band on one structured mesh:
 * leg 0 and leg 2 on the electron band (P1 dispersion from nodal eps), leg 1
 * on the phonon band (P1 dispersion from nodal eps_ph, or the constant omega0
 * when eps_ph is NULL). */
typedef struct mc_geo {
    int dim;
    const int32_t* nsec;
    double origin[3];
    double extent[3];        /* axis-aligned mesh: its box sides */
    double dG[3][3];         /* dG[d][c]: component d of mesh side c over nsec[c] */
    double dGinv[3][3];
    double zone[3][3];       /* zone[d][c]: component d of zone side c (umklapp basis) */
    int axis;                /* the mesh sides lie along the axes */
    const double* eps;
    const double* eps_ph;    /* phonon nodal energies, or NULL: constant omega0 */
    double omega0;
    int64_t nel;
    int32_t* conn;
    double* nodes;
    double vol;
} mc_geo;

/* Geometry block geo[21]: origin (3), then the mesh sides (side c at
 * geo[3 + 3c], Mesh::gVec columns, Mesh.cpp:268-281), then the zone sides
 * (zone side c at geo[12 + 3c], Region::gVec columns: the umklapp basis,
 * ScatteringChannel.cpp:330-334).  An axis-aligned box has side c =
 * extent_c e_c.  Returns the element volume. */
static double geo_setup(mc_geo* g, int dim, const int32_t* nsec, const double* geo) {
    g->axis = 1;
    for (int d = 0; d < 3; ++d) {
        g->origin[d] = d < dim ? geo[d] : 0.0;
        g->extent[d] = 1.0;
        for (int c = 0; c < 3; ++c) {
            const double sd = (d < dim && c < dim) ? geo[3 + 3 * c + d] : 0.0;
            g->dG[d][c] = (d < dim && c < dim) ? sd / (double)nsec[c] : 0.0;
            g->zone[d][c] = (d < dim && c < dim) ? geo[12 + 3 * c + d] : 0.0;
            g->dGinv[d][c] = 0.0;
            if (d != c && sd != 0.0) g->axis = 0;
        }
    }
    for (int d = 0; d < dim; ++d) {
        const double sd = geo[3 + 3 * d + d];
        if (!(sd > 0.0)) g->axis = 0;
        g->extent[d] = sd;
    }
    double det;
    if (dim == 1) {
        det = g->dG[0][0];
        g->dGinv[0][0] = 1.0 / det;
    } else if (dim == 2) {
        det = g->dG[0][0] * g->dG[1][1] - g->dG[0][1] * g->dG[1][0];
        g->dGinv[0][0] = g->dG[1][1] / det;
        g->dGinv[0][1] = -g->dG[0][1] / det;
        g->dGinv[1][0] = -g->dG[1][0] / det;
        g->dGinv[1][1] = g->dG[0][0] / det;
    } else {
        double A[9], Ai[9];
        for (int d = 0; d < 3; ++d)
            for (int c = 0; c < 3; ++c) A[3 * d + c] = g->dG[d][c];
        det = or_mc_invert(3, A, Ai);
        for (int d = 0; d < 3; ++d)
            for (int c = 0; c < 3; ++c) g->dGinv[d][c] = Ai[3 * d + c];
    }
    /* element volume: the box product, else |det(section sides)| / dim! */
    if (g->axis) return or_element_volume(dim, nsec, g->extent);
    return fabs(det) / (double)mc_fact(dim);
}

static void geo_init(mc_geo* g, int dim, const int32_t* nsec, const double* geo, const double* eps,
                     const double* eps_ph, double omega0) {
    g->dim = dim; g->nsec = nsec; g->eps = eps;
    g->eps_ph = eps_ph;
    g->omega0 = omega0;
    g->vol = geo_setup(g, dim, nsec, geo);
    g->nel = or_num_elements(dim, nsec);
    int64_t n = or_num_nodes(dim, nsec);
    g->conn = (int32_t*)malloc(sizeof(int32_t) * g->nel * (dim + 1));
    g->nodes = (double*)malloc(sizeof(double) * n * dim);
    or_mesh_elements(dim, nsec, g->conn);
    /* node coordinate d: origin + sum_c dG[d][c] i_c in axis order (on an
     * axis-aligned mesh the other axes add exact zeros) */
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < dim; ++d) {
            int64_t r = i;
            double x = g->origin[d];
            for (int c = 0; c < dim; ++c) {
                const int64_t id = r % (nsec[c] + 1);
                r /= (nsec[c] + 1);
                x = x + g->dG[d][c] * (double)id;
            }
            g->nodes[i * dim + d] = x;
        }
    }
}
static void geo_free(mc_geo* g) { free(g->conn); free(g->nodes); }

/* k-space linear form of coordinate d and of the energy on element e:
 * kf[d][0..dim] = [k_d(v0), k_d(v1)-k_d(v0), ...], ef likewise. */
static void elem_forms(const mc_geo* g, int64_t e, int phonon, double kf[3][4], double ef[4]) {
    const int dim = g->dim;
    const int32_t* v = g->conn + e * (dim + 1);
    for (int d = 0; d < dim; ++d) {
        double k0 = g->nodes[(int64_t)v[0] * dim + d];
        kf[d][0] = k0;
        for (int a = 1; a <= dim; ++a) kf[d][a] = g->nodes[(int64_t)v[a] * dim + d] - k0;
    }
    if (phonon && !g->eps_ph) {
        ef[0] = g->omega0;
        for (int a = 1; a <= dim; ++a) ef[a] = 0.0;
    } else {
        const double* en = phonon ? g->eps_ph : g->eps;
        double e0 = en[v[0]];
        ef[0] = e0;
        for (int a = 1; a <= dim; ++a) ef[a] = en[v[a]] - e0;
    }
}

/* umklapp vector u of id `uid` in the reference order (first axis outermost) */
static void umklapp_vec(int dim, int32_t nu, int32_t uid, int* u) {
    int span = 2 * nu + 1;
    for (int d = dim - 1; d >= 0; --d) { u[d] = uid % span - nu; uid /= span; }
}

/* D and N of a tuple (ScatteringChannel.cpp:285-352): legs {e0 in, e1 in, e2 out}. */
static void tuple_DN(const mc_geo* g, const int64_t* el, const int* u, double* D, double* N) {
    const int dim = g->dim, ncol = 3 * dim;
    static const double dir[3] = {1.0, 1.0, -1.0};
    for (int r = 0; r <= dim; ++r) N[r] = 0.0;
    for (int leg = 0; leg < 3; ++leg) {
        double kf[3][4], ef[4];
        elem_forms(g, el[leg], leg == 1, kf, ef);
        for (int d = 0; d < dim; ++d) {
            for (int a = 0; a < dim; ++a) D[d * ncol + leg * dim + a] = dir[leg] * kf[d][1 + a];
            N[d] -= dir[leg] * kf[d][0];
        }
        for (int a = 0; a < dim; ++a) D[dim * ncol + leg * dim + a] = dir[leg] * ef[1 + a];
        N[dim] -= dir[leg] * ef[0];
    }
    /* N_d -= sum_c u_c zone[d][c], one term at a time (ScatteringChannel.cpp:330-334) */
    for (int d = 0; d < dim; ++d)
        for (int c = 0; c < dim; ++c) N[d] -= (double)u[c] * g->zone[d][c];
}

static void elem_box(const mc_geo* g, int64_t e, double* lo, double* hi) {
    const int dim = g->dim;
    const int32_t* v = g->conn + e * (dim + 1);
    for (int d = 0; d < dim; ++d) { lo[d] = 1e300; hi[d] = -1e300; }
    for (int a = 0; a <= dim; ++a)
        for (int d = 0; d < dim; ++d) {
            double x = g->nodes[(int64_t)v[a] * dim + d];
            if (x < lo[d]) lo[d] = x;
            if (x > hi[d]) hi[d] = x;
        }
}

static void elem_erange(const mc_geo* g, const double* en, int64_t e, double* lo, double* hi) {
    const int32_t* v = g->conn + e * (g->dim + 1);
    *lo = 1e300; *hi = -1e300;
    for (int a = 0; a <= g->dim; ++a) {
        double x = en[v[a]];
        if (x < *lo) *lo = x;
        if (x > *hi) *hi = x;
    }
}

/* Screening value of a tuple: scatteringIntegrationTypeC with the modal
 * order-0 form sqrt(2) (1 in 1D) on every leg (ScatteringChannel.cpp:465-489),
 * amplitude 1 and functAmplitude 1 (screening runs at channel creation,
 * before either is set).  Evaluated as hits * c^3 * prod(vol*dim!) / divisor,
 * which grows with the hit count, so the comparison with the threshold can
 * stop at the first count that passes.  Returns -1 when |det D_C| < min_det. */
static double tuple_screen(const mc_geo* g, const double* D, const double* N, const double* pts,
                           int32_t nmc, double min_det) {
    const int dim = g->dim, nc = dim + 1;
    double DC[16], inv[16];
    for (int r = 0; r < nc; ++r)
        for (int q = 0; q < nc; ++q) DC[r * nc + q] = D[r * 3 * dim + 2 * dim - 1 + q];
    double det = or_mc_invert(nc, DC, inv);
    if (!(fabs(det) >= min_det)) return -1.0;
    int64_t hits = 0;
    for (int32_t m = 0; m < nmc; ++m) {
        double K[4][3];
        hits += shell_point(dim, 3, D, N, inv, pts, nmc, m, K);
    }
    double c = dim == 1 ? 1.0 : sqrt(2.0);
    double cpow = c * c * c;
    double amp = 1.0;
    for (int leg = 0; leg < 3; ++leg) amp *= g->vol * (double)mc_fact(dim);
    return (double)hits * cpow * amp / shell_divisor(dim, 3, det, nmc);
}

/* Tuples of the e-ph channel, ordered by (e0, e2, umklapp id, e1): every
 * (e0, e2) pair whose energy window contains omega0-conservation
 * (ScatteringChannel.hpp:212-222), every umklapp vector in {-nu..nu}^dim and
 * every phonon element whose bounding box meets k2 - k0 - G u, kept when
 * |det D_C| >= min_det and the screening value exceeds `threshold`
 * (ScatteringChannel.cpp:474-503).  idx[t*4 + {0,1,2,3}] = e0, e1, e2, uid.
 * Returns the tuple count (idx may be NULL to count; at most cap written). */
int64_t or_mc_tuples_eph(int dim, const int32_t* nsec, const double* geo,
                         const double* eps, const double* eps_ph, double omega0, int32_t nu,
                         const double* pts, int32_t nmc, double min_det, double threshold,
                         int32_t* idx, int64_t cap) {
    mc_geo g;
    geo_init(&g, dim, nsec, geo, eps, eps_ph, omega0);
    const int64_t nel = g.nel;
    const int nes = mc_fact(dim);
    int32_t nuv = 1;
    for (int d = 0; d < dim; ++d) nuv *= 2 * nu + 1;
    double* box = (double*)malloc(sizeof(double) * nel * 2 * dim);
    double* er = (double*)malloc(sizeof(double) * nel * 2);
    double* pr = (double*)malloc(sizeof(double) * nel * 2);   /* phonon energy ranges */
    double pmin = omega0, pmax = omega0;
    for (int64_t e = 0; e < nel; ++e) {
        elem_box(&g, e, box + e * 2 * dim, box + e * 2 * dim + dim);
        elem_erange(&g, eps, e, er + 2 * e, er + 2 * e + 1);
        if (eps_ph) {
            elem_erange(&g, eps_ph, e, pr + 2 * e, pr + 2 * e + 1);
            if (e == 0 || pr[2 * e] < pmin) pmin = pr[2 * e];
            if (e == 0 || pr[2 * e + 1] > pmax) pmax = pr[2 * e + 1];
        } else {
            pr[2 * e] = omega0;
            pr[2 * e + 1] = omega0;
        }
    }
    int64_t* counts = (int64_t*)calloc((size_t)nel + 1, sizeof(int64_t));
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            if (!idx) break;
            for (int64_t e = 0; e < nel; ++e) counts[e + 1] += counts[e];
        }
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t e0 = 0; e0 < nel; ++e0) {
            int64_t w = pass == 1 ? counts[e0] : 0, cnt = 0;
            const double* b0 = box + e0 * 2 * dim;
            for (int64_t e2 = 0; e2 < nel; ++e2) {
                /* energy window of the three legs (ScatteringChannel.hpp:212-222):
                 * first with the phonon band's whole range, then per phonon element */
                if (!(er[2 * e0] + pmin - er[2 * e2 + 1] <= 0.0 &&
                      er[2 * e0 + 1] + pmax - er[2 * e2] >= 0.0))
                    continue;
                const double* b2 = box + e2 * 2 * dim;
                for (int32_t uid = 0; uid < nuv; ++uid) {
                    int u[3] = {0, 0, 0};
                    umklapp_vec(dim, nu, uid, u);
                    /* k1 = k2 - k0 - G u */
                    double lo[3], hi[3];
                    int slo[3] = {0, 0, 0}, shi[3] = {0, 0, 0}, empty = 0;
                    /* sections that can hold k1: on an axis-aligned mesh the
                     * box's own section range, else the reduced coordinates of
                     * its corners, one section wider (a superset) */
                    double smin[3], smax[3];
                    for (int d = 0; d < dim; ++d) {
                        double U = 0.0;
                        for (int c = 0; c < dim; ++c) U = U + (double)u[c] * g.zone[d][c];
                        lo[d] = b2[d] - b0[dim + d] - U;
                        hi[d] = b2[dim + d] - b0[d] - U;
                        if (g.axis) {
                            double h = g.extent[d] / nsec[d];
                            smin[d] = floor((lo[d] - g.origin[d]) / h);
                            smax[d] = floor((hi[d] - g.origin[d]) / h);
                        } else {
                            smin[d] = 1e300;
                            smax[d] = -1e300;
                        }
                    }
                    if (!g.axis) {
                        for (int cm = 0; cm < (1 << dim); ++cm) {
                            double x[3];
                            for (int d = 0; d < dim; ++d) x[d] = ((cm >> d) & 1 ? hi[d] : lo[d]) - g.origin[d];
                            for (int c = 0; c < dim; ++c) {
                                double sr = 0.0;
                                for (int d = 0; d < dim; ++d) sr = sr + g.dGinv[c][d] * x[d];
                                if (sr < smin[c]) smin[c] = sr;
                                if (sr > smax[c]) smax[c] = sr;
                            }
                        }
                        for (int c = 0; c < dim; ++c) {
                            smin[c] = floor(smin[c]) - 1.0;
                            smax[c] = floor(smax[c]) + 1.0;
                        }
                    }
                    for (int d = 0; d < dim; ++d) {
                        double a = smin[d], b = smax[d];
                        if (b < 0.0 || a > nsec[d] - 1) { empty = 1; break; }
                        slo[d] = a < 0.0 ? 0 : (int)a;
                        shi[d] = b > nsec[d] - 1 ? nsec[d] - 1 : (int)b;
                    }
                    if (empty) continue;
                    int sc[3] = {slo[0], dim > 1 ? slo[1] : 0, dim > 2 ? slo[2] : 0};
                    for (;;) {
                        int64_t sid = 0, stride = 1;
                        for (int d = 0; d < dim; ++d) { sid += stride * sc[d]; stride *= nsec[d]; }
                        for (int t = 0; t < nes; ++t) {
                            int64_t e1 = sid * nes + t;
                            const double* b1 = box + e1 * 2 * dim;
                            int meet = 1;
                            for (int d = 0; d < dim; ++d)
                                meet &= b1[d] <= hi[d] && b1[dim + d] >= lo[d];
                            if (!meet) continue;
                            if (!(er[2 * e0] + pr[2 * e1] - er[2 * e2 + 1] <= 0.0 &&
                                  er[2 * e0 + 1] + pr[2 * e1 + 1] - er[2 * e2] >= 0.0))
                                continue;
                            int64_t el[3] = {e0, e1, e2};
                            double D[36], N[4];
                            tuple_DN(&g, el, u, D, N);
                            double s = tuple_screen(&g, D, N, pts, nmc, min_det);
                            if (!(s > threshold)) continue;
                            if (pass == 1 && w < cap) {
                                idx[w * 4 + 0] = (int32_t)e0;
                                idx[w * 4 + 1] = (int32_t)e1;
                                idx[w * 4 + 2] = (int32_t)e2;
                                idx[w * 4 + 3] = uid;
                            }
                            ++w;
                            ++cnt;
                        }
                        int d = 0;
                        while (d < dim && ++sc[d] > shi[d]) { sc[d] = slo[d]; ++d; }
                        if (d == dim) break;
                    }
                }
            }
            if (pass == 0) counts[e0 + 1] = cnt;
        }
    }
    int64_t total = 0;
    if (idx) total = counts[nel];
    else for (int64_t e = 0; e < nel; ++e) total += counts[e + 1];
    free(counts); free(box); free(er); free(pr);
    geo_free(&g);
    return total;
}

/* ================================================================ mass */
/* Inverse of the LF-basis mass matrix of one element (Mesh.cpp:95-123):
 * M_ab = vol * d! * int_ref phi_a phi_b, phi = {1, x_1..x_d}; exact moments
 * int x_i = 1/(d+1)!, int x_i^2 = 2/(d+2)!, int x_i x_j = 1/(d+2)!. */
void or_mc_inverse_mass(int dim, double vol, double* minv) {
    const int nc = dim + 1;
    double M[16];
    double f1 = 1.0 / (double)mc_fact(dim + 1), f2 = 1.0 / (double)mc_fact(dim + 2);
    double scale = vol * (double)mc_fact(dim);
    for (int a = 0; a < nc; ++a)
        for (int b = 0; b < nc; ++b) {
            double m;
            if (a == 0 && b == 0) m = 1.0 / (double)mc_fact(dim);
            else if (a == 0 || b == 0) m = f1;
            else m = a == b ? 2.0 * f2 : f2;
            M[a * nc + b] = scale * m;
        }
    or_mc_invert(nc, M, minv);
}

/* ============================================================ propagator */
/* ScatteringChannel::propagate (ScatteringChannel.cpp:618-670) of the e-ph
 * channel over a tuple list, then applyInverseMass (Function.cpp:801-804).
 *   f_el, f_ph: DG linear-form coefficients [nel][dim+1] of the electron and
 *   phonon bands; out_el, out_ph likewise (out_ph may be NULL).
 *   amplitude: g^2 (functAmplitude = 1).
 * Per tuple the forms are (fermion electrons, boson phonons):
 *   direct        -f(k0) * n(k1) * (1 - f(k2))
 *   time-reversed (1 - f(k0)) * (1 + n(k1)) * f(k2)
 * and leg l receives dir_l * result_l.  Threads accumulate disjoint tuple
 * ranges; their sums are added in thread order. */
void or_mc_apply_eph(int dim, const int32_t* nsec, const double* geo,
                     const double* eps, const double* eps_ph, double omega0, double amplitude,
                     int32_t nu, const double* pts, int32_t nmc, const int32_t* idx, int64_t ntup,
                     const double* f_el, const double* f_ph, double* out_el, double* out_ph,
                     int32_t nthreads) {
    mc_geo g;
    geo_init(&g, dim, nsec, geo, eps, eps_ph, omega0);
    const int nc = dim + 1;
    const int64_t nel = g.nel, len = nel * nc;
    int nt = nthreads > 0 ? nthreads : 1;
    double* acc = (double*)calloc((size_t)nt * 2 * len, sizeof(double));
    double fullamp = amplitude;
    for (int leg = 0; leg < 3; ++leg) fullamp *= g.vol * (double)mc_fact(dim);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int th = 0; th < nt; ++th) {
        double* ae = acc + (size_t)th * 2 * len;
        double* ap = ae + len;
        int64_t t0 = ntup * th / nt, t1 = ntup * (th + 1) / nt;
        for (int64_t t = t0; t < t1; ++t) {
            int64_t el[3] = {idx[t * 4], idx[t * 4 + 1], idx[t * 4 + 2]};
            int u[3] = {0, 0, 0};
            umklapp_vec(dim, nu, idx[t * 4 + 3], u);
            double D[36], N[4];
            tuple_DN(&g, el, u, D, N);
            double DC[16], inv[16];
            for (int r = 0; r < nc; ++r)
                for (int q = 0; q < nc; ++q) DC[r * nc + q] = D[r * 3 * dim + 2 * dim - 1 + q];
            if (fabs(or_mc_invert(nc, DC, inv)) < 1e-300) continue;
            double fa[12], fb[12], res[12];
            const double* c0 = f_el + el[0] * nc;
            const double* c1 = f_ph + el[1] * nc;
            const double* c2 = f_el + el[2] * nc;
            for (int a = 0; a < nc; ++a) {
                double one = a == 0 ? 1.0 : 0.0;
                fa[a] = -c0[a];                 /* leg 0 in, sign -1 of the direct product */
                fa[nc + a] = c1[a];             /* leg 1 in: n */
                fa[2 * nc + a] = one - c2[a];   /* leg 2 out, fermion: 1 - f */
                fb[a] = one - c0[a];            /* time reversed: 1 - f */
                fb[nc + a] = one + c1[a];       /* 1 + n */
                fb[2 * nc + a] = c2[a];         /* f */
            }
            or_mc_integrate(dim, 3, D, N, fa, fb, pts, nmc, res);
            for (int a = 0; a < nc; ++a) {
                ae[el[0] * nc + a] += fullamp * res[a];
                ap[el[1] * nc + a] += fullamp * res[nc + a];
                ae[el[2] * nc + a] -= fullamp * res[2 * nc + a];
            }
        }
    }
    double minv[16];
    or_mc_inverse_mass(dim, g.vol, minv);
    for (int64_t e = 0; e < nel; ++e) {
        double se[4] = {0, 0, 0, 0}, sp[4] = {0, 0, 0, 0};
        for (int th = 0; th < nt; ++th)
            for (int a = 0; a < nc; ++a) {
                se[a] += acc[(size_t)th * 2 * len + e * nc + a];
                sp[a] += acc[(size_t)th * 2 * len + len + e * nc + a];
            }
        for (int b = 0; b < nc; ++b) {
            double ve = 0.0, vp = 0.0;
            for (int a = 0; a < nc; ++a) { ve += se[a] * minv[a * nc + b]; vp += sp[a] * minv[a * nc + b]; }
            out_el[e * nc + b] = ve;
            if (out_ph) out_ph[e * nc + b] = vp;
        }
    }
    free(acc);
    geo_free(&g);
}

/* Nodal values -> DG linear-form coefficients per element (the form that
 * interpolates the nodal values at the element's vertices, LinearForm.hpp:34-40). */
void or_mc_nodal_to_dg(int dim, const int32_t* nsec, const double* f, double* dg) {
    int64_t nel = or_num_elements(dim, nsec);
    int32_t* conn = (int32_t*)malloc(sizeof(int32_t) * nel * (dim + 1));
    or_mesh_elements(dim, nsec, conn);
    for (int64_t e = 0; e < nel; ++e) {
        const int32_t* v = conn + e * (dim + 1);
        double f0 = f[v[0]];
        dg[e * (dim + 1)] = f0;
        for (int a = 1; a <= dim; ++a) dg[e * (dim + 1) + a] = f[v[a]] - f0;
    }
    free(conn);
}

/* ====================================================== DG error norm + DP5 */
/* Band::relativeError (Band.cpp:87-101) on DG linear-form coefficients
 * [nel][dim+1]: Function's pointwise operations act on the element-vertex
 * values c0 (+ cj), and integrate(rhs) of the P1 result is exact per element.
 * Fermion (stat = -1) or boson (+1) scale; the penalty term of states outside
 * the physical range as in the reference. */
double or_relative_error_dg(int dim, int64_t nel, const double* err, const double* pop, double stat,
                            double rtol, double atol, double penalty) {
    const int nc = dim + 1;
    double s_r = 0.0, s_p = 0.0;
    for (int64_t e = 0; e < nel; ++e) {
        double sr = 0.0, sr2 = 0.0, sp = 0.0, sp2 = 0.0;
        for (int a = 0; a < nc; ++a) {
            double y = pop[e * nc] + (a ? pop[e * nc + a] : 0.0);
            double x = err[e * nc] + (a ? err[e * nc + a] : 0.0);
            double scale, pen;
            if (stat == -1.0) {
                double m = fabs(y) < fabs(1.0 - y) ? fabs(y) : fabs(1.0 - y);
                scale = rtol * m + atol;
                pen = ((y < 0.0 ? -y : 0.0) + (y - 1.0 > 0.0 ? y - 1.0 : 0.0)) / scale;
            } else {
                scale = rtol * y + atol;
                pen = (y < 0.0 ? -y : 0.0) / scale;
            }
            double r = x / scale;
            sr += r; sr2 += r * r; sp += pen; sp2 += pen * pen;
        }
        s_r += sr2 + sr * sr;
        s_p += sp2 + sp * sp;
    }
    double denom = (double)((dim + 1) * (dim + 2)) * (double)nel;
    return sqrt((s_r + penalty * s_p) / denom);
}

typedef struct mc_dp5_user {
    int dim;
    int dynamic;             /* 1: the state holds the phonon band after the electron band */
    const int32_t* nsec;
    const double *geo, *eps, *eps_ph, *pts;
    double omega0, amplitude, rtol, atol;
    int32_t nu, nmc, nthreads;
    const int32_t* idx;
    int64_t ntup, nel;
    double* fph;
    const or_bath_sched* bs; /* non-NULL: the bath at the stage time */
    int legs4;               /* 1: the 4-leg channel (idx [ntup][5]) */
    double min_det, kappa;
} mc_dp5_user;

/* int eps f over the mesh for DG coefficients f: per element the exact
 * P1 x P1 integral vol/((d+1)(d+2)) (sum_v eps_v f_v + sum_v eps_v sum_v f_v)
 * with the element-vertex values (eps at the element's nodes) */
double or_energy_dg(int dim, const int32_t* nsec, const double* geo, const double* eps,
                    const double* f) {
    const int64_t nel = or_num_elements(dim, nsec);
    const int nc = dim + 1;
    int32_t* conn = (int32_t*)malloc(sizeof(int32_t) * nel * nc);
    or_mesh_elements(dim, nsec, conn);
    double tot = 0.0;
    for (int64_t e = 0; e < nel; ++e) {
        double sei = 0.0, se = 0.0, si = 0.0;
        for (int a = 0; a < nc; ++a) {
            const double ev = eps[conn[e * nc + a]];
            const double x = f[e * nc] + (a ? f[e * nc + a] : 0.0);
            sei += ev * x;
            se += ev;
            si += x;
        }
        tot += sei + se * si;
    }
    free(conn);
    mc_geo g;
    return tot * geo_setup(&g, dim, nsec, geo) / (double)((dim + 1) * (dim + 2));
}

static double mc_dp5_flux(void* u, const double* dydt) {
    const mc_dp5_user* p = (const mc_dp5_user*)u;
    return or_energy_dg(p->dim, p->nsec, p->geo, p->eps, dydt);
}

static void mc_dp5_rhs(void* u, double t, const double* y, double* dydt) {
    const mc_dp5_user* p = (const mc_dp5_user*)u;
    const int64_t len = p->nel * (p->dim + 1);
    if (p->legs4) {
        or_mc_apply_ee(p->dim, p->nsec, p->geo, p->eps, p->nu, p->pts, p->nmc, p->min_det, p->amplitude,
                       p->kappa, p->idx, p->ntup, y, 1, dydt);
        return;
    }
    if (p->bs) {
        const double nb = or_sched_value(p->bs->s, p->bs->ns, p->bs->t_start, p->bs->dt, t);
        for (int64_t e = 0; e < p->nel; ++e) p->fph[e * (p->dim + 1)] = nb;
    }
    if (p->dynamic)
        or_mc_apply_eph(p->dim, p->nsec, p->geo, p->eps, p->eps_ph, p->omega0,
                        p->amplitude, p->nu, p->pts, p->nmc, p->idx, p->ntup, y, y + len, dydt,
                        dydt + len, p->nthreads);
    else
        or_mc_apply_eph(p->dim, p->nsec, p->geo, p->eps, p->eps_ph, p->omega0,
                        p->amplitude, p->nu, p->pts, p->nmc, p->idx, p->ntup, y, p->fph, dydt, NULL,
                        p->nthreads);
}

/* MaterialStatus::relativeError (MaterialStatus.cpp:189-196): max over the
 * bands, fermion electrons and (when dynamic) boson phonons */
static double mc_dp5_norm(void* u, const double* err, const double* y5) {
    const mc_dp5_user* p = (const mc_dp5_user*)u;
    double e = or_relative_error_dg(p->dim, p->nel, err, y5, -1.0, p->rtol, p->atol, 0.0);
    if (p->dynamic) {
        const int64_t len = p->nel * (p->dim + 1);
        double q = or_relative_error_dg(p->dim, p->nel, err + len, y5 + len, 1.0, p->rtol, p->atol, 0.0);
        if (q >= e) e = q;
    }
    return e;
}

/* MaterialTimeStatus::propagate (MaterialTimeStatus.cpp:102-137) with the
 * tuple operator: DP5(4) over the electron band's DG coefficients, the phonon
 * band a bath of occupation bath_n (its constraint makes its error zero),
 * fermion error norm without penalty (Band.hpp default). */
void or_dp5_mc_eph(int dim, const int32_t* nsec, const double* geo,
                   const double* eps, const double* eps_ph, double omega0, double amplitude,
                   int32_t nu, const double* pts, int32_t nmc, const int32_t* idx, int64_t ntup,
                   int32_t dynamic, double bath_n, double* f, double t0, double t_final,
                   double dense_dt, double rtol, double atol, double* dense_out, double* dense_t,
                   int32_t max_dense, int32_t nthreads, or_dp5_stats* stats) {
    or_dp5_mc_eph_ex(dim, nsec, geo, eps, eps_ph, omega0, amplitude, nu, pts, nmc, idx, ntup,
                     dynamic, bath_n, NULL, 0, 0.0, 1.0, 0.0, 0.0, f, t0, t_final, dense_dt, rtol, atol,
                     dense_out, dense_t, max_dense, nthreads, stats, NULL);
}

void or_dp5_mc_eph_ex(int dim, const int32_t* nsec, const double* geo,
                      const double* eps, const double* eps_ph, double omega0, double amplitude,
                      int32_t nu, const double* pts, int32_t nmc, const int32_t* idx,
                      int64_t ntup, int32_t dynamic, double bath_n, const double* sched,
                      int32_t ns, double t_start, double sdt, double satol, double srtol,
                      double* f, double t0, double t_final, double dense_dt,
                      double rtol, double atol, double* dense_out, double* dense_t,
                      int32_t max_dense, int32_t nthreads, or_dp5_stats* stats, double* ledger) {
    mc_dp5_user u;
    memset(&u, 0, sizeof(u));
    or_bath_sched bs = {sched, ns, t_start, sdt, satol, srtol};
    u.dim = dim; u.nsec = nsec; u.geo = geo; u.eps = eps; u.pts = pts;
    u.eps_ph = eps_ph; u.dynamic = dynamic;
    u.omega0 = omega0; u.amplitude = amplitude; u.rtol = rtol; u.atol = atol;
    u.nu = nu; u.nmc = nmc; u.nthreads = nthreads; u.idx = idx; u.ntup = ntup;
    u.nel = or_num_elements(dim, nsec);
    u.bs = (sched && !dynamic) ? &bs : NULL;
    const int nc = dim + 1;
    u.fph = (double*)calloc((size_t)u.nel * nc, sizeof(double));
    for (int64_t e = 0; e < u.nel; ++e) u.fph[e * nc] = bath_n;
    if (ledger) *ledger = 0.0;
    or_dp5_integrate_ex(u.nel * nc * (dynamic ? 2 : 1), f, t0, t_final, dense_dt, mc_dp5_rhs, mc_dp5_norm, &u,
                        u.bs, (ledger && !dynamic) ? mc_dp5_flux : NULL, ledger, dense_out, dense_t, max_dense,
                        NULL, NULL, 0, stats);
    free(u.fph);
}

void or_dp5_mc_eph_replay(int dim, const int32_t* nsec, const double* geo,
                          const double* eps, const double* eps_ph, double omega0, double amplitude,
                          int32_t nu, const double* pts, int32_t nmc, const int32_t* idx,
                          int64_t ntup, int32_t dynamic, double bath_n, const double* sched,
                          int32_t ns, double t_start, double sdt, double* f, double t0,
                          double t_final, double dense_dt, const double* h_steps, int64_t n_steps,
                          double* dense_out, double* dense_t, int32_t max_dense, int32_t nthreads,
                          double* ledger) {
    mc_dp5_user u;
    memset(&u, 0, sizeof(u));
    or_bath_sched bs = {sched, ns, t_start, sdt, 0.0, 0.0};
    u.dim = dim; u.nsec = nsec; u.geo = geo; u.eps = eps; u.pts = pts;
    u.eps_ph = eps_ph; u.dynamic = dynamic;
    u.omega0 = omega0; u.amplitude = amplitude;
    u.nu = nu; u.nmc = nmc; u.nthreads = nthreads; u.idx = idx; u.ntup = ntup;
    u.nel = or_num_elements(dim, nsec);
    u.bs = (sched && !dynamic) ? &bs : NULL;
    const int nc = dim + 1;
    u.fph = (double*)calloc((size_t)u.nel * nc, sizeof(double));
    for (int64_t e = 0; e < u.nel; ++e) u.fph[e * nc] = bath_n;
    int8_t* acc = (int8_t*)malloc((size_t)(n_steps > 0 ? n_steps : 1));
    memset(acc, 1, (size_t)(n_steps > 0 ? n_steps : 1));
    or_dp5_replay_ex(u.nel * nc * (dynamic ? 2 : 1), f, t0, t_final, dense_dt, mc_dp5_rhs, &u,
                     (ledger && !dynamic) ? mc_dp5_flux : NULL, ledger, h_steps, acc, n_steps, dense_out,
                     dense_t, max_dense);
    free(acc);
    free(u.fph);
}

/* ============================================================ rates */
/* ScatteringChannel::scatteringRates(status, band) (ScatteringChannel.cpp:544-612)
 * for the electron band of the e-ph channel: for each leg the band occupies
 * (0 and 2), singleScatteringRate replaces that leg's forms by
 *   direct:        (leg==0 ? -1 : 1) ((1-dir) stat + (1+dir)) * 1
 *   time-reversed: ((1+dir) stat + (1-dir)) * 1
 * keeps the others as in the propagator, and subtracts dir * result of that
 * leg; the legs' inverse-massed results are summed. */
void or_mc_rates_eph(int dim, const int32_t* nsec, const double* geo,
                     const double* eps, const double* eps_ph, double omega0, double amplitude,
                     int32_t nu, const double* pts, int32_t nmc, const int32_t* idx, int64_t ntup,
                     const double* f_el, const double* f_ph, double* out_el, int32_t nthreads) {
    mc_geo g;
    geo_init(&g, dim, nsec, geo, eps, eps_ph, omega0);
    const int nc = dim + 1;
    const int64_t nel = g.nel, len = nel * nc;
    int nt = nthreads > 0 ? nthreads : 1;
    double* acc = (double*)calloc((size_t)nt * len, sizeof(double));
    double fullamp = amplitude;
    for (int leg = 0; leg < 3; ++leg) fullamp *= g.vol * (double)mc_fact(dim);
    static const double dir[3] = {1.0, 1.0, -1.0};
    static const double stat[3] = {-1.0, 1.0, -1.0};
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int th = 0; th < nt; ++th) {
        double* ae = acc + (size_t)th * len;
        int64_t t0 = ntup * th / nt, t1 = ntup * (th + 1) / nt;
        for (int64_t t = t0; t < t1; ++t) {
            int64_t el[3] = {idx[t * 4], idx[t * 4 + 1], idx[t * 4 + 2]};
            int u[3] = {0, 0, 0};
            umklapp_vec(dim, nu, idx[t * 4 + 3], u);
            double D[36], N[4];
            tuple_DN(&g, el, u, D, N);
            const double* c[3] = {f_el + el[0] * nc, f_ph + el[1] * nc, f_el + el[2] * nc};
            for (int which = 0; which < 3; which += 2) {
                double fa[12], fb[12], res[12];
                for (int leg = 0; leg < 3; ++leg)
                    for (int a = 0; a < nc; ++a) {
                        double one = a == 0 ? 1.0 : 0.0;
                        if (leg == which) {
                            fa[leg * nc + a] = (leg == 0 ? -1.0 : 1.0) *
                                               ((1.0 - dir[leg]) * stat[leg] + (1.0 + dir[leg])) * one;
                            fb[leg * nc + a] = ((1.0 + dir[leg]) * stat[leg] + (1.0 - dir[leg])) * one;
                        } else {
                            fa[leg * nc + a] = (leg == 0 ? -1.0 : 1.0) * 0.5 *
                                               ((1.0 - dir[leg]) * one +
                                                ((1.0 - dir[leg]) * stat[leg] + (1.0 + dir[leg])) * c[leg][a]);
                            fb[leg * nc + a] = 0.5 * ((1.0 + dir[leg]) * one +
                                                      ((1.0 + dir[leg]) * stat[leg] + (1.0 - dir[leg])) * c[leg][a]);
                        }
                    }
                if (or_mc_integrate(dim, 3, D, N, fa, fb, pts, nmc, res) != 0) continue;
                for (int a = 0; a < nc; ++a) ae[el[which] * nc + a] -= dir[which] * fullamp * res[which * nc + a];
            }
        }
    }
    double minv[16];
    or_mc_inverse_mass(dim, g.vol, minv);
    for (int64_t e = 0; e < nel; ++e) {
        double se[4] = {0, 0, 0, 0};
        for (int th = 0; th < nt; ++th)
            for (int a = 0; a < nc; ++a) se[a] += acc[(size_t)th * len + e * nc + a];
        for (int b = 0; b < nc; ++b) {
            double v = 0.0;
            for (int a = 0; a < nc; ++a) v += se[a] * minv[a * nc + b];
            out_el[e * nc + b] = v;
        }
    }
    free(acc);
    geo_free(&g);
}

/* ================================================ 4-leg channel (e-e) */
/* The channel {el in, el in, el out, el out} on the electron band: the
 * reference's ScatteringChannel<NDim,4> (ScatteringChannel.cpp:716-722) with
 * legDirection {+1, +1, -1, -1} and fermion statistics on every leg.
 * functAmplitude (ScatteringChannel.hpp:70, evaluated at the legs'
 * barycentres by constructAmplitude, ScatteringChannel.cpp:399-415) is 1 for
 * kappa <= 0 (the reference default, ScatteringChannel.cpp:144), else the
 * squared screened Coulomb potential 1/(|q|^2 + kappa^2)^2 with q the
 * barycentre of leg 2 minus that of leg 0. */
static const double k_dir4[4] = {1.0, 1.0, -1.0, -1.0};

/* D and N of a 4-leg tuple (ScatteringChannel.cpp:285-352). */
static void tuple4_DN(const mc_geo* g, const int64_t* el, const int* u, double* D, double* N) {
    const int dim = g->dim, ncol = 4 * dim;
    for (int r = 0; r <= dim; ++r) N[r] = 0.0;
    for (int leg = 0; leg < 4; ++leg) {
        double kf[3][4], ef[4];
        elem_forms(g, el[leg], 0, kf, ef);
        for (int d = 0; d < dim; ++d) {
            for (int a = 0; a < dim; ++a) D[d * ncol + leg * dim + a] = k_dir4[leg] * kf[d][1 + a];
            N[d] -= k_dir4[leg] * kf[d][0];
        }
        for (int a = 0; a < dim; ++a) D[dim * ncol + leg * dim + a] = k_dir4[leg] * ef[1 + a];
        N[dim] -= k_dir4[leg] * ef[0];
    }
    for (int d = 0; d < dim; ++d)
        for (int c = 0; c < dim; ++c) N[d] -= (double)u[c] * g->zone[d][c];
}

/* D_C^-1 of a 4-leg tuple; returns det D_C (IntegratorCore.hpp:153-155) */
static double tuple4_dc(int dim, const double* D, double* inv) {
    const int nc = dim + 1;
    double DC[16];
    for (int r = 0; r < nc; ++r)
        for (int q = 0; q < nc; ++q) DC[r * nc + q] = D[r * 4 * dim + 3 * dim - 1 + q];
    return or_mc_invert(nc, DC, inv);
}

/* Screening value of a 4-leg tuple: tuple_screen with c^4 and four volumes;
 * -1 when |det D_C| < min_det. */
static double tuple4_screen(const mc_geo* g, const double* D, const double* N, const double* pts,
                            int32_t nmc, double min_det) {
    const int dim = g->dim;
    double inv[16];
    double det = tuple4_dc(dim, D, inv);
    if (!(fabs(det) >= min_det)) return -1.0;
    int64_t hits = 0;
    for (int32_t m = 0; m < nmc; ++m) {
        double K[4][3];
        hits += shell_point(dim, 4, D, N, inv, pts, nmc, m, K);
    }
    double c = dim == 1 ? 1.0 : sqrt(2.0);
    double cpow = c * c * c * c;
    double amp = 1.0;
    for (int leg = 0; leg < 4; ++leg) amp *= g->vol * (double)mc_fact(dim);
    return (double)hits * cpow * amp / shell_divisor(dim, 4, det, nmc);
}

/* Sections that can hold a point of the box [lo, hi]: the box's own section
 * range on an axis-aligned mesh, else the reduced coordinates of its corners
 * one section wider.  Returns 0 when none. */
static int box_sections(const mc_geo* g, const double* lo, const double* hi, int* slo, int* shi) {
    const int dim = g->dim;
    double smin[3], smax[3];
    for (int d = 0; d < dim; ++d) {
        if (g->axis) {
            double h = g->extent[d] / g->nsec[d];
            smin[d] = floor((lo[d] - g->origin[d]) / h);
            smax[d] = floor((hi[d] - g->origin[d]) / h);
        } else {
            smin[d] = 1e300;
            smax[d] = -1e300;
        }
    }
    if (!g->axis) {
        for (int cm = 0; cm < (1 << dim); ++cm) {
            double x[3];
            for (int d = 0; d < dim; ++d) x[d] = ((cm >> d) & 1 ? hi[d] : lo[d]) - g->origin[d];
            for (int c = 0; c < dim; ++c) {
                double sr = 0.0;
                for (int d = 0; d < dim; ++d) sr = sr + g->dGinv[c][d] * x[d];
                if (sr < smin[c]) smin[c] = sr;
                if (sr > smax[c]) smax[c] = sr;
            }
        }
        for (int c = 0; c < dim; ++c) {
            smin[c] = floor(smin[c]) - 1.0;
            smax[c] = floor(smax[c]) + 1.0;
        }
    }
    for (int d = 0; d < dim; ++d) {
        double a = smin[d], b = smax[d];
        if (b < 0.0 || a > g->nsec[d] - 1) return 0;
        slo[d] = a < 0.0 ? 0 : (int)a;
        shi[d] = b > g->nsec[d] - 1 ? g->nsec[d] - 1 : (int)b;
    }
    return 1;
}

/* Tuples of the 4-leg channel, ordered by (e0, e1, e2, umklapp id, e3):
 * every leg-3 element whose box meets k0 + k1 - k2 + G u and whose energy
 * window admits energy conservation (ScatteringChannel.hpp:212-222), kept when
 * |det D_C| >= min_det and the screening value exceeds `threshold`
 * (ScatteringChannel.cpp:474-503).  brute != 0 tests every leg-3 element
 * without the box and energy pruning (the same list: both prunings keep every
 * tuple whose delta shell meets the simplices).  idx[t*5 + {0..4}] = e0, e1,
 * e2, e3, uid.  Returns the tuple count (idx may be NULL; at most cap written). */
int64_t or_mc_tuples_ee(int dim, const int32_t* nsec, const double* geo, const double* eps, int32_t nu,
                        const double* pts, int32_t nmc, double min_det, double threshold, int32_t brute,
                        int32_t* idx, int64_t cap) {
    mc_geo g;
    geo_init(&g, dim, nsec, geo, eps, NULL, 0.0);
    const int64_t nel = g.nel;
    const int nes = mc_fact(dim);
    int32_t nuv = 1;
    for (int d = 0; d < dim; ++d) nuv *= 2 * nu + 1;
    double* box = (double*)malloc(sizeof(double) * nel * 2 * dim);
    double* er = (double*)malloc(sizeof(double) * nel * 2);
    double emin = 0.0, emax = 0.0;
    for (int64_t e = 0; e < nel; ++e) {
        elem_box(&g, e, box + e * 2 * dim, box + e * 2 * dim + dim);
        elem_erange(&g, eps, e, er + 2 * e, er + 2 * e + 1);
        if (e == 0 || er[2 * e] < emin) emin = er[2 * e];
        if (e == 0 || er[2 * e + 1] > emax) emax = er[2 * e + 1];
    }
    const int64_t npair = nel * nel;
    int64_t* counts = (int64_t*)calloc((size_t)npair + 1, sizeof(int64_t));
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            if (!idx) break;
            for (int64_t p = 0; p < npair; ++p) counts[p + 1] += counts[p];
        }
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t p = 0; p < npair; ++p) {
            const int64_t e0 = p / nel, e1 = p % nel;
            int64_t w = pass == 1 ? counts[p] : 0, cnt = 0;
            const double* b0 = box + e0 * 2 * dim;
            const double* b1 = box + e1 * 2 * dim;
            for (int64_t e2 = 0; e2 < nel; ++e2) {
                /* energy window with leg 3 over the band's whole range */
                if (!brute && !(er[2 * e0] + er[2 * e1] - er[2 * e2 + 1] - emax <= 0.0 &&
                                er[2 * e0 + 1] + er[2 * e1 + 1] - er[2 * e2] - emin >= 0.0))
                    continue;
                const double* b2 = box + e2 * 2 * dim;
                for (int32_t uid = 0; uid < nuv; ++uid) {
                    int u[3] = {0, 0, 0};
                    umklapp_vec(dim, nu, uid, u);
                    int slo[3] = {0, 0, 0}, shi[3] = {0, 0, 0};
                    if (brute) {
                        for (int d = 0; d < dim; ++d) shi[d] = nsec[d] - 1;
                    } else {
                        /* k3 = k0 + k1 - k2 + G u */
                        double lo[3], hi[3];
                        for (int d = 0; d < dim; ++d) {
                            double U = 0.0;
                            for (int c = 0; c < dim; ++c) U = U + (double)u[c] * g.zone[d][c];
                            lo[d] = b0[d] + b1[d] - b2[dim + d] + U;
                            hi[d] = b0[dim + d] + b1[dim + d] - b2[d] + U;
                        }
                        if (!box_sections(&g, lo, hi, slo, shi)) continue;
                        int sc[3] = {slo[0], dim > 1 ? slo[1] : 0, dim > 2 ? slo[2] : 0};
                        for (;;) {
                            int64_t sid = 0, stride = 1;
                            for (int d = 0; d < dim; ++d) { sid += stride * sc[d]; stride *= nsec[d]; }
                            for (int t = 0; t < nes; ++t) {
                                int64_t e3 = sid * nes + t;
                                const double* b3 = box + e3 * 2 * dim;
                                int meet = 1;
                                for (int d = 0; d < dim; ++d) meet &= b3[d] <= hi[d] && b3[dim + d] >= lo[d];
                                if (!meet) continue;
                                if (!(er[2 * e0] + er[2 * e1] - er[2 * e2 + 1] - er[2 * e3 + 1] <= 0.0 &&
                                      er[2 * e0 + 1] + er[2 * e1 + 1] - er[2 * e2] - er[2 * e3] >= 0.0))
                                    continue;
                                int64_t el[4] = {e0, e1, e2, e3};
                                double D[48], N[4];
                                tuple4_DN(&g, el, u, D, N);
                                if (!(tuple4_screen(&g, D, N, pts, nmc, min_det) > threshold)) continue;
                                if (pass == 1 && w < cap) {
                                    int32_t* o = idx + w * 5;
                                    o[0] = (int32_t)e0; o[1] = (int32_t)e1; o[2] = (int32_t)e2;
                                    o[3] = (int32_t)e3; o[4] = uid;
                                }
                                ++w;
                                ++cnt;
                            }
                            int d = 0;
                            while (d < dim && ++sc[d] > shi[d]) { sc[d] = slo[d]; ++d; }
                            if (d == dim) break;
                        }
                        continue;
                    }
                    for (int64_t e3 = 0; e3 < nel; ++e3) {
                        int64_t el[4] = {e0, e1, e2, e3};
                        double D[48], N[4];
                        tuple4_DN(&g, el, u, D, N);
                        if (!(tuple4_screen(&g, D, N, pts, nmc, min_det) > threshold)) continue;
                        if (pass == 1 && w < cap) {
                            int32_t* o = idx + w * 5;
                            o[0] = (int32_t)e0; o[1] = (int32_t)e1; o[2] = (int32_t)e2;
                            o[3] = (int32_t)e3; o[4] = uid;
                        }
                        ++w;
                        ++cnt;
                    }
                }
            }
            if (pass == 0) counts[p + 1] = cnt;
        }
    }
    int64_t total = 0;
    if (idx) total = counts[npair];
    else for (int64_t p = 0; p < npair; ++p) total += counts[p + 1];
    free(counts); free(box); free(er);
    geo_free(&g);
    return total;
}

/* constructAmplitude (ScatteringChannel.cpp:399-415) of a 4-leg tuple:
 * amplitude * prod(vol * dim!) * functAmplitude(barycentres). */
static double tuple4_amplitude(const mc_geo* g, const int64_t* el, double amplitude, double kappa) {
    const int dim = g->dim;
    double a = amplitude;
    for (int leg = 0; leg < 4; ++leg) a *= g->vol * (double)mc_fact(dim);
    if (kappa > 0.0) {
        double q2 = 0.0;
        for (int d = 0; d < dim; ++d) {
            double c0 = 0.0, c2 = 0.0;
            for (int v = 0; v <= dim; ++v) {
                c0 += g->nodes[(int64_t)g->conn[el[0] * (dim + 1) + v] * dim + d];
                c2 += g->nodes[(int64_t)g->conn[el[2] * (dim + 1) + v] * dim + d];
            }
            const double q = (c2 - c0) / (double)(dim + 1);
            q2 += q * q;
        }
        const double v = 1.0 / (q2 + kappa * kappa);
        a *= v * v;
    }
    return a;
}

/* The forms of a 4-leg tuple for DG coefficient blocks c[leg] (the electron
 * band at each leg's element), ScatteringChannel.cpp:628-640 with statistics
 * -1: direct -f0 f1 (1 - f2)(1 - f3), time reversed (1 - f0)(1 - f1) f2 f3. */
static void tuple4_forms(int dim, const double* const* c, double* fa, double* fb) {
    const int nc = dim + 1;
    for (int leg = 0; leg < 4; ++leg)
        for (int a = 0; a < nc; ++a) {
            const double one = a == 0 ? 1.0 : 0.0, x = c[leg][a];
            if (k_dir4[leg] > 0.0) {
                fa[leg * nc + a] = leg == 0 ? -x : x;
                fb[leg * nc + a] = one - x;
            } else {
                fa[leg * nc + a] = one - x;
                fb[leg * nc + a] = x;
            }
        }
}

/* The integral of one 4-leg tuple (singleScatteringPropagator without the
 * amplitude and the leg signs, ScatteringChannel.cpp:618-650): res[4*(dim+1)]
 * for the states of the four legs c (dim+1 coefficients each); D and N
 * returned when not NULL.  Returns -1 when |det D_C| < min_det (res zero). */
int or_mc_tuple_ee(int dim, const int32_t* nsec, const double* geo, const double* eps, int32_t nu,
                   const double* pts, int32_t nmc, double min_det, const int32_t* tup,
                   const double* c, double* res, double* D_out, double* N_out) {
    mc_geo g;
    geo_init(&g, dim, nsec, geo, eps, NULL, 0.0);
    const int nc = dim + 1;
    int64_t el[4] = {tup[0], tup[1], tup[2], tup[3]};
    int u[3] = {0, 0, 0};
    umklapp_vec(dim, nu, tup[4], u);
    double D[48], N[4], inv[16];
    tuple4_DN(&g, el, u, D, N);
    if (D_out) memcpy(D_out, D, sizeof(double) * (dim + 1) * 4 * dim);
    if (N_out) memcpy(N_out, N, sizeof(double) * (dim + 1));
    for (int i = 0; i < 4 * nc; ++i) res[i] = 0.0;
    int rc = -1;
    if (fabs(tuple4_dc(dim, D, inv)) >= min_det) {
        const double* cl[4] = {c, c + nc, c + 2 * nc, c + 3 * nc};
        double fa[16], fb[16];
        tuple4_forms(dim, cl, fa, fb);
        rc = or_mc_integrate(dim, 4, D, N, fa, fb, pts, nmc, res);
    }
    geo_free(&g);
    return rc;
}

/* ScatteringChannel<NDim,4>::propagate (ScatteringChannel.cpp:651-670) over a
 * tuple list [ntup][5] for `batch` members f [batch][nel][dim+1], then
 * applyInverseMass (Function.cpp:801-804): out [batch][nel][dim+1].  Tuples
 * with |det D_C| < min_det are skipped (:621-622).  Per element the
 * contributions are summed in tuple order, leg by leg. */
void or_mc_apply_ee(int dim, const int32_t* nsec, const double* geo, const double* eps, int32_t nu,
                    const double* pts, int32_t nmc, double min_det, double amplitude, double kappa,
                    const int32_t* idx, int64_t ntup, const double* f, int32_t batch, double* out) {
    mc_geo g;
    geo_init(&g, dim, nsec, geo, eps, NULL, 0.0);
    const int nc = dim + 1;
    const int64_t nel = g.nel, len = nel * nc;
    double minv[16];
    or_mc_inverse_mass(dim, g.vol, minv);
#pragma omp parallel for schedule(static)
    for (int32_t b = 0; b < batch; ++b) {
        const double* fb_ = f + (size_t)b * len;
        double* acc = (double*)calloc((size_t)len, sizeof(double));
        for (int64_t t = 0; t < ntup; ++t) {
            const int32_t* tp = idx + t * 5;
            int64_t el[4] = {tp[0], tp[1], tp[2], tp[3]};
            int u[3] = {0, 0, 0};
            umklapp_vec(dim, nu, tp[4], u);
            double D[48], N[4], inv[16];
            tuple4_DN(&g, el, u, D, N);
            if (!(fabs(tuple4_dc(dim, D, inv)) >= min_det)) continue;
            const double* cl[4] = {fb_ + el[0] * nc, fb_ + el[1] * nc, fb_ + el[2] * nc, fb_ + el[3] * nc};
            double fa[16], fbv[16], res[16];
            tuple4_forms(dim, cl, fa, fbv);
            or_mc_integrate(dim, 4, D, N, fa, fbv, pts, nmc, res);
            const double amp = tuple4_amplitude(&g, el, amplitude, kappa);
            for (int leg = 0; leg < 4; ++leg)
                for (int a = 0; a < nc; ++a) acc[el[leg] * nc + a] += k_dir4[leg] * (amp * res[leg * nc + a]);
        }
        for (int64_t e = 0; e < nel; ++e)
            for (int q = 0; q < nc; ++q) {
                double v = 0.0;
                for (int a = 0; a < nc; ++a) v += acc[e * nc + a] * minv[a * nc + q];
                out[(size_t)b * len + e * nc + q] = v;
            }
        free(acc);
    }
    geo_free(&g);
}

/* MaterialTimeStatus::propagate (MaterialTimeStatus.cpp:102-137) with the
 * 4-leg channel alone: DP5(4) over the electron band's DG coefficients with
 * the fermion error norm (Band.cpp:87-101). */
void or_dp5_mc_ee(int dim, const int32_t* nsec, const double* geo, const double* eps, int32_t nu,
                  const double* pts, int32_t nmc, double min_det, double amplitude, double kappa,
                  const int32_t* idx, int64_t ntup, double* f, double t0, double t_final, double dense_dt,
                  double rtol, double atol, double* dense_out, double* dense_t, int32_t max_dense,
                  or_dp5_stats* stats) {
    mc_dp5_user u;
    memset(&u, 0, sizeof(u));
    u.dim = dim; u.nsec = nsec; u.geo = geo; u.eps = eps; u.pts = pts;
    u.amplitude = amplitude; u.rtol = rtol; u.atol = atol;
    u.nu = nu; u.nmc = nmc; u.idx = idx; u.ntup = ntup;
    u.nel = or_num_elements(dim, nsec);
    u.legs4 = 1; u.min_det = min_det; u.kappa = kappa;
    or_dp5_integrate_ex(u.nel * (dim + 1), f, t0, t_final, dense_dt, mc_dp5_rhs, mc_dp5_norm, &u, NULL, NULL,
                        NULL, dense_out, dense_t, max_dense, NULL, NULL, 0, stats);
}
