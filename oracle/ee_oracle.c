/*
 * ee_oracle.c -- CPU FP64 restatement of the electron-electron (4-leg)
 * collision operator on the periodic momentum lattice.
 *
 * TEST INFRASTRUCTURE ONLY (see tortoise_oracle.h).
 *
 * What it computes (no reference function computes it in this form; the
 * reference counterpart is ScatteringChannel<3,4> with functAmplitude =
 * V(q)^2 at the element barycentres, src/Physics/Scattering/
 * ScatteringChannel.cpp:399-415,618-650):
 *
 *   I_1 = sum_{2,3} w^2 A(k1-k3) delta(e1 + e2 - e3 - e4)
 *                   [ f3 f4 (1-f1)(1-f2) - f1 f2 (1-f3)(1-f4) ],   k4 = k1 + k2 - k3
 *
 * on the P periodic points of the mesh box (nsec points per axis; the mesh's
 * duplicate boundary nodes carry the value of their periodic image), w = the
 * volume per point, A(q) = g^2 / (|q|^2 + kappa2)^2 with |q| the minimum
 * image, and the energy delta a Gaussian of width sigma cut at cutoff*sigma.
 *
 * The operator the library evaluates (or_ee_apply) projects the delta on a
 * uniform grid w_m = m dw symmetric about 0 with hat functions h_m:
 *   delta(x - y) -> sum_{m,m'} h_m(x) h_m'(y) G_sigma((m - m') dw),
 * x = e1 - e3, y = e4 - e2.  Grouping by q = k1 - k3 (then k4 = k2 + q):
 *   Ha_q(m) = sum_2 w h_m(e(2+q) - e(2)) f(2+q)(1 - f(2))
 *   Hb_q(m) = sum_2 w h_m(e(2+q) - e(2)) f(2)(1 - f(2+q))
 *   S_q = G * H_q                     (convolution on the grid)
 *   I_1 = sum_3 w A(q) sum_m h_m(e1 - e3) [ (1-f1) f3 Sa_q(m) - f1 (1-f3) Sb_q(m) ]
 * The projected delta is even under (x, y) -> (-x, -y), so sum_1 w I_1 = 0
 * holds exactly (the relabelling (1,2,3,4) -> (3,4,1,2) swaps gain and loss).
 * or_ee_apply_direct evaluates the unprojected Gaussian by brute force,
 * O(P^3), for the convergence pins in tests/test_oracle_ee.py.
 */
#include "tortoise_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int dim;
    int32_t ns[3];
    int64_t P;
    double w;          /* volume per periodic point */
    double dk[3];      /* lattice spacing per axis */
    double box[3];     /* period per axis */
    int64_t* node;     /* periodic point -> canonical mesh node */
} ee_lattice;

static void lattice_make(int dim, const int32_t* nsec, const double* extent, ee_lattice* L) {
    L->dim = dim;
    L->P = 1;
    L->w = 1.0;
    for (int d = 0; d < 3; ++d) {
        L->ns[d] = d < dim ? nsec[d] : 1;
        L->box[d] = d < dim ? extent[d] : 1.0;
        L->dk[d] = L->box[d] / L->ns[d];
    }
    for (int d = 0; d < dim; ++d) {
        L->P *= L->ns[d];
        L->w *= L->dk[d];
    }
    L->node = (int64_t*)malloc(sizeof(int64_t) * L->P);
    for (int64_t p = 0; p < L->P; ++p) {
        int64_t r = p, id = 0, stride = 1;
        for (int d = 0; d < dim; ++d) {
            const int64_t c = r % L->ns[d];
            r /= L->ns[d];
            id += c * stride;
            stride *= nsec[d] + 1;
        }
        L->node[p] = id;
    }
}

/* p2 + q on the periodic lattice */
static int64_t lattice_add(const ee_lattice* L, int64_t a, int64_t b, int sign) {
    int64_t id = 0, stride = 1;
    for (int d = 0; d < L->dim; ++d) {
        const int64_t ca = a % L->ns[d], cb = b % L->ns[d];
        a /= L->ns[d];
        b /= L->ns[d];
        int64_t c = (ca + sign * cb) % L->ns[d];
        if (c < 0) c += L->ns[d];
        id += c * stride;
        stride *= L->ns[d];
    }
    return id;
}

/* A(q) for q = k_a - k_b, minimum image */
static double ee_amp(const ee_lattice* L, int64_t a, int64_t b, double g, double kappa2) {
    double q2 = 0.0;
    for (int d = 0; d < L->dim; ++d) {
        const int64_t ca = a % L->ns[d], cb = b % L->ns[d];
        a /= L->ns[d];
        b /= L->ns[d];
        int64_t c = ca - cb;
        if (c < 0) c += L->ns[d];
        double qd = c * L->dk[d];
        if (qd >= 0.5 * L->box[d]) qd -= L->box[d];
        q2 += qd * qd;
    }
    const double v = 1.0 / (q2 + kappa2);
    return g * g * v * v;
}

void or_ee_grid(int dim, const int32_t* nsec, const double* eps, double dw, int32_t* mh) {
    /* half-width Mh of the grid w_m = m dw, |m| <= Mh, covering every energy
     * difference on the lattice plus one cell */
    int64_t n = or_num_nodes(dim, nsec);
    double lo = eps[0], hi = eps[0];
    for (int64_t i = 1; i < n; ++i) {
        lo = eps[i] < lo ? eps[i] : lo;
        hi = eps[i] > hi ? eps[i] : hi;
    }
    *mh = (int32_t)ceil((hi - lo) / dw) + 2;
}

void or_ee_apply(int dim, const int32_t* nsec, const double* extent, const double* eps,
                 const double* f, double g, double kappa2, double sigma, double dw,
                 double cutoff, double* I) {
    ee_lattice L;
    lattice_make(dim, nsec, extent, &L);
    int32_t Mh;
    or_ee_grid(dim, nsec, eps, dw, &Mh);
    const int64_t M = 2 * (int64_t)Mh + 1;
    const int32_t J = (int32_t)floor(cutoff * sigma / dw);
    double* gt = (double*)malloc(sizeof(double) * (2 * J + 1));
    const double nrm = 1.0 / (sigma * sqrt(2.0 * M_PI));
    for (int32_t j = -J; j <= J; ++j) {
        const double x = j * dw;
        gt[j + J] = nrm * exp(-0.5 * x * x / (sigma * sigma));
    }
    const int64_t P = L.P;
    double* e = (double*)malloc(sizeof(double) * P);
    double* ff = (double*)malloc(sizeof(double) * P);
    for (int64_t p = 0; p < P; ++p) {
        e[p] = eps[L.node[p]];
        ff[p] = f[L.node[p]];
    }
    double* Ip = (double*)calloc((size_t)P, sizeof(double));
    const double w = L.w;
    /* per-thread partial sums over q, combined in thread order (static
     * schedule: the result depends only on the thread count) */
    const int nth = omp_get_max_threads();
    double* accs = (double*)calloc((size_t)nth * P, sizeof(double));
#pragma omp parallel num_threads(nth)
    {
        double* acc = accs + (size_t)omp_get_thread_num() * P;
        double* Ha = (double*)malloc(sizeof(double) * M);
        double* Hb = (double*)malloc(sizeof(double) * M);
        double* Sa = (double*)malloc(sizeof(double) * M);
        double* Sb = (double*)malloc(sizeof(double) * M);
        double* contrib = (double*)malloc(sizeof(double) * P);
#pragma omp for schedule(static)
        for (int64_t q = 0; q < P; ++q) {
            memset(Ha, 0, sizeof(double) * M);
            memset(Hb, 0, sizeof(double) * M);
            for (int64_t p2 = 0; p2 < P; ++p2) {
                const int64_t p4 = lattice_add(&L, p2, q, +1);
                const double y = (e[p4] - e[p2]) / dw;
                const double fl = floor(y);
                const double t = y - fl;
                const int64_t m = (int64_t)fl + Mh;
                const double fa = w * ff[p4] * (1.0 - ff[p2]);
                const double fb = w * ff[p2] * (1.0 - ff[p4]);
                Ha[m] += (1.0 - t) * fa;
                Ha[m + 1] += t * fa;
                Hb[m] += (1.0 - t) * fb;
                Hb[m + 1] += t * fb;
            }
            for (int64_t m = 0; m < M; ++m) {
                double sa = 0.0, sb = 0.0;
                for (int32_t j = -J; j <= J; ++j) {
                    const int64_t mp = m - j;
                    if (mp < 0 || mp >= M) continue;
                    sa += gt[j + J] * Ha[mp];
                    sb += gt[j + J] * Hb[mp];
                }
                Sa[m] = sa;
                Sb[m] = sb;
            }
            /* pairs (1, 3) with k1 - k3 = q */
            for (int64_t p1 = 0; p1 < P; ++p1) {
                const int64_t p3 = lattice_add(&L, p1, q, -1);
                const double x = (e[p1] - e[p3]) / dw;
                const double fl = floor(x);
                const double t = x - fl;
                const int64_t m = (int64_t)fl + Mh;
                const double sa = (1.0 - t) * Sa[m] + t * Sa[m + 1];
                const double sb = (1.0 - t) * Sb[m] + t * Sb[m + 1];
                const double f1 = ff[p1], f3 = ff[p3];
                contrib[p1] = w * ee_amp(&L, p1, p3, g, kappa2) *
                              ((1.0 - f1) * f3 * sa - f1 * (1.0 - f3) * sb);
            }
            for (int64_t p1 = 0; p1 < P; ++p1) acc[p1] += contrib[p1];
        }
        free(Ha); free(Hb); free(Sa); free(Sb); free(contrib);
    }
    for (int t = 0; t < nth; ++t)
        for (int64_t p = 0; p < P; ++p) Ip[p] += accs[(size_t)t * P + p];
    free(accs);
    const int64_t n = or_num_nodes(dim, nsec);
    /* every mesh node takes the value of its periodic image */
    for (int64_t i = 0; i < n; ++i) {
        int64_t r = i, p = 0, stride = 1;
        for (int d = 0; d < dim; ++d) {
            const int64_t c = r % (nsec[d] + 1);
            r /= nsec[d] + 1;
            p += (c % nsec[d]) * stride;
            stride *= nsec[d];
        }
        I[i] = Ip[p];
    }
    free(Ip); free(e); free(ff); free(gt); free(L.node);
}

void or_ee_apply_direct(int dim, const int32_t* nsec, const double* extent, const double* eps,
                        const double* f, double g, double kappa2, double sigma, double cutoff,
                        double* I) {
    ee_lattice L;
    lattice_make(dim, nsec, extent, &L);
    const int64_t P = L.P;
    const double w = L.w;
    const double nrm = 1.0 / (sigma * sqrt(2.0 * M_PI));
    double* Ip = (double*)calloc((size_t)P, sizeof(double));
#pragma omp parallel for schedule(dynamic)
    for (int64_t p1 = 0; p1 < P; ++p1) {
        const double e1 = eps[L.node[p1]], f1 = f[L.node[p1]];
        double s = 0.0;
        for (int64_t p3 = 0; p3 < P; ++p3) {
            const double e3 = eps[L.node[p3]], f3 = f[L.node[p3]];
            const double a = ee_amp(&L, p1, p3, g, kappa2);
            const int64_t q = lattice_add(&L, p1, p3, -1);
            for (int64_t p2 = 0; p2 < P; ++p2) {
                const int64_t p4 = lattice_add(&L, p2, q, +1);
                const double e2 = eps[L.node[p2]], f2 = f[L.node[p2]];
                const double e4 = eps[L.node[p4]], f4 = f[L.node[p4]];
                const double x = e1 + e2 - e3 - e4;
                if (fabs(x) > cutoff * sigma) continue;
                const double d = nrm * exp(-0.5 * x * x / (sigma * sigma));
                s += w * w * a * d *
                     (f3 * f4 * (1.0 - f1) * (1.0 - f2) - f1 * f2 * (1.0 - f3) * (1.0 - f4));
            }
        }
        Ip[p1] = s;
    }
    const int64_t n = or_num_nodes(dim, nsec);
    for (int64_t i = 0; i < n; ++i) {
        int64_t r = i, p = 0, stride = 1;
        for (int d = 0; d < dim; ++d) {
            const int64_t c = r % (nsec[d] + 1);
            r /= nsec[d] + 1;
            p += (c % nsec[d]) * stride;
            stride *= nsec[d];
        }
        I[i] = Ip[p];
    }
    free(Ip); free(L.node);
}
