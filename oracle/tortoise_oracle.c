/*
 * tortoise_oracle.c -- CPU FP64 restatement of the scattering-integral hot
 * path.  TEST INFRASTRUCTURE ONLY (see tortoise_oracle.h for the reference
 * files each part follows and for what is new).  OpenMP over rows where the
 * work is large; every reduction order is fixed so results are repeatable.
 */
#include "tortoise_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ===================================================================== mesh
 * Section s = (sx, sy[, sz]) owns dim! elements.  Element id
 *   2D: 2*(sx + sy*nx) + t            (Mesh.cpp:412-415)
 *   3D: 6*(sx + (sy + sz*ny)*nx) + t  (Mesh.cpp:416-420)
 * Vertex order inside a section: depth-first walk from the origin corner that
 * sets one still-zero axis to 1 at each step, axes tried in increasing order
 * (Mesh.cpp:55-87).  For 2D: t=0 -> (0,0),(1,0),(1,1); t=1 -> (0,0),(0,1),(1,1).
 * Nodes are the (nsec+1)^dim lattice, node id ix + (nx+1)*(iy + (ny+1)*iz).
 */
static int fact(int d) { int r = 1; for (int i = 2; i <= d; ++i) r *= i; return r; }

/* vertex offsets of element t in a section: off[a*3+d] in {0,1} */
static void ref_section_vertices(int dim, int t, int* off) {
    /* permutations of the axes in lexicographic order = DFS order */
    int perm[3] = {0, 1, 2};
    /* enumerate the t-th permutation of {0..dim-1} in lexicographic order */
    int avail[3] = {0, 1, 2};
    int navail = dim;
    int rem = t;
    for (int k = 0; k < dim; ++k) {
        int f = fact(dim - 1 - k);
        int idx = rem / f;
        rem %= f;
        perm[k] = avail[idx];
        for (int q = idx; q < navail - 1; ++q) avail[q] = avail[q + 1];
        --navail;
    }
    int cur[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) off[d] = 0;
    for (int a = 1; a <= dim; ++a) {
        cur[perm[a - 1]] = 1;
        for (int d = 0; d < 3; ++d) off[a * 3 + d] = cur[d];
    }
}

int64_t or_num_nodes(int dim, const int32_t* nsec) {
    int64_t n = 1;
    for (int d = 0; d < dim; ++d) n *= (int64_t)(nsec[d] + 1);
    return n;
}

int64_t or_num_elements(int dim, const int32_t* nsec) {
    int64_t n = fact(dim);
    for (int d = 0; d < dim; ++d) n *= nsec[d];
    return n;
}

double or_element_volume(int dim, const int32_t* nsec, const double* extent) {
    double v = 1.0;
    for (int d = 0; d < dim; ++d) v *= extent[d] / (double)nsec[d];
    return fabs(v / (double)fact(dim));
}

void or_mesh_nodes(int dim, const int32_t* nsec, const double* origin,
                   const double* extent, double* coords) {
    int64_t n = or_num_nodes(dim, nsec);
    for (int64_t i = 0; i < n; ++i) {
        int64_t r = i;
        for (int d = 0; d < dim; ++d) {
            int64_t id = r % (nsec[d] + 1);
            r /= (nsec[d] + 1);
            coords[i * dim + d] = origin[d] + (extent[d] / (double)nsec[d]) * (double)id;
        }
    }
}

static int64_t node_id(int dim, const int32_t* nsec, const int* c) {
    int64_t id = 0, stride = 1;
    for (int d = 0; d < dim; ++d) { id += stride * c[d]; stride *= (nsec[d] + 1); }
    return id;
}

void or_mesh_elements(int dim, const int32_t* nsec, int32_t* conn) {
    int nes = fact(dim);
    int64_t nsecs = 1;
    for (int d = 0; d < dim; ++d) nsecs *= nsec[d];
    for (int64_t s = 0; s < nsecs; ++s) {
        int sc[3] = {0, 0, 0};
        int64_t r = s;
        for (int d = 0; d < dim; ++d) { sc[d] = (int)(r % nsec[d]); r /= nsec[d]; }
        for (int t = 0; t < nes; ++t) {
            int off[12];
            ref_section_vertices(dim, t, off);
            for (int a = 0; a <= dim; ++a) {
                int c[3];
                for (int d = 0; d < dim; ++d) c[d] = sc[d] + off[a * 3 + d];
                conn[(s * nes + t) * (dim + 1) + a] = (int32_t)node_id(dim, nsec, c);
            }
        }
    }
}

void or_node_volumes(int dim, const int32_t* nsec, const double* extent, double* V) {
    int64_t n = or_num_nodes(dim, nsec), nel = or_num_elements(dim, nsec);
    double vol = or_element_volume(dim, nsec, extent);
    int32_t* conn = (int32_t*)malloc(sizeof(int32_t) * nel * (dim + 1));
    or_mesh_elements(dim, nsec, conn);
    for (int64_t i = 0; i < n; ++i) V[i] = 0.0;
    for (int64_t e = 0; e < nel; ++e)
        for (int a = 0; a <= dim; ++a) V[conn[e * (dim + 1) + a]] += vol / (double)(dim + 1);
    free(conn);
}

/* ============================================== hat-averaged occupation
 * fbar_i = int phi_i f / int phi_i for the P1 interpolant f of nodal values:
 * the consistent mass matrix over the lumped one.  On a simplex with d+1
 * vertices int phi_a phi_b = vol (1 + delta_ab) / ((d+1)(d+2)) and
 * int phi_a = vol/(d+1), so row i of the stencil collects, over the elements
 * e holding i, (1 + delta_ib) / ((d+2) |star(i)|) for each vertex b of e. */
or_hat* or_hat_create(int dim, const int32_t* nsec) {
    const int64_t n = or_num_nodes(dim, nsec), nel = or_num_elements(dim, nsec);
    const int nv = dim + 1;
    int32_t* conn = (int32_t*)malloc(sizeof(int32_t) * nel * nv);
    or_mesh_elements(dim, nsec, conn);
    /* dense accumulation per node over a small neighbour list */
    int32_t* star = (int32_t*)calloc(n, sizeof(int32_t));
    for (int64_t e = 0; e < nel; ++e)
        for (int a = 0; a < nv; ++a) ++star[conn[e * nv + a]];
    /* at most (2^dim - 1) * 2 + 1 distinct vertices share an element with a node */
    const int cap = dim == 2 ? 7 : 27;
    int32_t* nb = (int32_t*)malloc(sizeof(int32_t) * n * cap);
    double* wt = (double*)calloc(n * cap, sizeof(double));
    int32_t* cnt = (int32_t*)calloc(n, sizeof(int32_t));
    for (int64_t e = 0; e < nel; ++e)
        for (int a = 0; a < nv; ++a) {
            const int32_t i = conn[e * nv + a];
            const double scale = 1.0 / ((double)(dim + 2) * (double)star[i]);
            for (int b = 0; b < nv; ++b) {
                const int32_t m = conn[e * nv + b];
                const double w = (b == a ? 2.0 : 1.0) * scale;
                int k = 0;
                while (k < cnt[i] && nb[i * cap + k] != m) ++k;
                if (k == cnt[i]) { nb[i * cap + k] = m; ++cnt[i]; }
                wt[i * cap + k] += w;
            }
        }
    or_hat* h = (or_hat*)malloc(sizeof(or_hat));
    h->n = n;
    h->ptr = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
    h->ptr[0] = 0;
    for (int64_t i = 0; i < n; ++i) h->ptr[i + 1] = h->ptr[i] + cnt[i];
    h->col = (int32_t*)malloc(sizeof(int32_t) * h->ptr[n]);
    h->w = (double*)malloc(sizeof(double) * h->ptr[n]);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < cnt[i]; ++k) {
            h->col[h->ptr[i] + k] = nb[i * cap + k];
            h->w[h->ptr[i] + k] = wt[i * cap + k];
        }
    free(conn); free(star); free(nb); free(wt); free(cnt);
    return h;
}

/* Number-defect weight of the hat form: the gross rate node i would have at
 * its own occupation x if every other node were half filled,
 * ((1-x) (W 1)_i + x (W^T 1)_i) / V_i, with W = (n+1) K + n K^T,
 * rk = K 1 and ck = K^T 1 (linear in x). */
double or_hat_defect_weight(double x, double np1, double nb, double rk, double ck, double V) {
    return ((1.0 - x) * (np1 * rk + nb * ck) + x * (np1 * ck + nb * rk)) / V;
}

void or_hat_free(or_hat* h) {
    if (!h) return;
    free(h->ptr); free(h->col); free(h->w); free(h);
}

void or_hat_apply(const or_hat* h, const double* f, double* fbar) {
    for (int64_t i = 0; i < h->n; ++i) {
        double s = 0.0;
        for (int64_t k = h->ptr[i]; k < h->ptr[i + 1]; ++k) s += h->w[k] * f[h->col[k]];
        fbar[i] = s;
    }
}

void or_hat_apply_t(const or_hat* h, const double* z, double* out) {
    for (int64_t i = 0; i < h->n; ++i) out[i] = 0.0;
    for (int64_t i = 0; i < h->n; ++i)
        for (int64_t k = h->ptr[i]; k < h->ptr[i + 1]; ++k) out[h->col[k]] += h->w[k] * z[i];
}

int64_t or_hat_stencil(int dim, const int32_t* nsec, int64_t* ptr, int32_t* col, double* w) {
    or_hat* h = or_hat_create(dim, nsec);
    const int64_t nnz = h->ptr[h->n];
    if (ptr) memcpy(ptr, h->ptr, sizeof(int64_t) * (h->n + 1));
    if (col) memcpy(col, h->col, sizeof(int32_t) * nnz);
    if (w) memcpy(w, h->w, sizeof(double) * nnz);
    or_hat_free(h);
    return nnz;
}

/* ======================================================= energy projection */
void or_egrid_make(const double* eps, int64_t n, double de, or_egrid* g) {
    double lo = eps[0], hi = eps[0];
    for (int64_t i = 1; i < n; ++i) { if (eps[i] < lo) lo = eps[i]; if (eps[i] > hi) hi = eps[i]; }
    g->de = de;
    g->e0 = lo - de;
    g->nbins = (int32_t)floor((hi - g->e0) / de) + 3;
}

static int32_t egrid_cell(const or_egrid* g, double x) {
    int32_t c = (int32_t)floor((x - g->e0) / g->de);
    if (c < 0) c = 0;
    if (c > g->nbins - 2) c = g->nbins - 2;
    return c;
}

/* Cox-de Boor value of the order-k B-spline N_{0,k} on knots t[0..k] at x,
 * where x lies strictly inside the knot span [t[span], t[span+1]). */
static double bspline_value(const double* t, int k, int span, double x) {
    double N[5] = {0, 0, 0, 0, 0};
    N[span] = 1.0;
    for (int q = 2; q <= k; ++q) {
        for (int i = 0; i <= k - q; ++i) {
            double v = 0.0;
            double d1 = t[i + q - 1] - t[i];
            double d2 = t[i + q] - t[i + 1];
            if (d1 > 0.0) v += (x - t[i]) / d1 * N[i];
            if (d2 > 0.0) v += (t[i + q] - x) / d2 * N[i + 1];
            N[i] = v;
        }
    }
    return N[0];
}

static void sort_small(double* v, int n) {
    for (int i = 1; i < n; ++i) {
        double x = v[i];
        int j = i - 1;
        while (j >= 0 && v[j] > x) { v[j + 1] = v[j]; --j; }
        v[j + 1] = x;
    }
}

double or_simplex_density(int dim, const double* evert, int a, double vol, double x) {
    double t[5];
    for (int q = 0; q <= dim; ++q) t[q] = evert[q];
    t[dim + 1] = evert[a];
    sort_small(t, dim + 2);
    double span = t[dim + 1] - t[0];
    if (!(span > 0.0) || x <= t[0] || x >= t[dim + 1]) return 0.0;
    int s = 0;
    while (s < dim && !(x < t[s + 1])) ++s;
    return vol * bspline_value(t, dim + 1, s, x) / span;
}

/* 3-point Gauss-Legendre on [0,1] */
static const double GL3_X[3] = {0.11270166537925831148, 0.5, 0.88729833462074168852};
static const double GL3_W[3] = {0.27777777777777777778, 0.44444444444444444444,
                                0.27777777777777777778};

/* Adds vol/(dim+1) * int M(x) h_b(x) dx for every hat b into acc[b - base]. */
static void project_simplex_basis(int dim, const double* ev, int a, double vol,
                                  const or_egrid* g, double* acc, int32_t base) {
    double t[5];
    for (int q = 0; q <= dim; ++q) t[q] = ev[q];
    t[dim + 1] = ev[a];
    sort_small(t, dim + 2);
    const int nk = dim + 2;
    double span = t[nk - 1] - t[0];
    const double de = g->de;
    if (!(span > 1e-9 * de)) {
        /* flat simplex: the density is vol/(dim+1) * delta(x - t0) */
        double x = t[0];
        int32_t c = egrid_cell(g, x);
        double ec = g->e0 + (double)c * de;
        double w1 = (x - ec) / de;
        double m = vol / (double)(dim + 1);
        acc[c - base] += m * (1.0 - w1);
        acc[c + 1 - base] += m * w1;
        return;
    }
    const double scale = vol / span;
    int32_t c0 = egrid_cell(g, t[0]), c1 = egrid_cell(g, t[nk - 1]);
    for (int32_t c = c0; c <= c1; ++c) {
        double ec = g->e0 + (double)c * de;
        double ec1 = ec + de;
        double lo = t[0] > ec ? t[0] : ec;
        double hi = t[nk - 1] < ec1 ? t[nk - 1] : ec1;
        if (!(hi > lo)) continue;
        /* split [lo,hi] at the knots */
        for (int s = 0; s < nk - 1; ++s) {
            double a0 = t[s] > lo ? t[s] : lo;
            double a1 = t[s + 1] < hi ? t[s + 1] : hi;
            if (!(a1 > a0)) continue;
            double w = a1 - a0;
            double s0 = 0.0, s1 = 0.0;
            for (int q = 0; q < 3; ++q) {
                double x = a0 + GL3_X[q] * w;
                double v = GL3_W[q] * w * scale * bspline_value(t, nk - 1, s, x);
                s0 += v * (ec1 - x);
                s1 += v * (x - ec);
            }
            acc[c - base] += s0 / de;
            acc[c + 1 - base] += s1 / de;
        }
    }
}

/* incident (element, local vertex) pairs of a lattice node; returns count */
static int node_star(int dim, const int32_t* nsec, int64_t node, int64_t* elems, int* locs) {
    int nc[3] = {0, 0, 0};
    int64_t r = node;
    for (int d = 0; d < dim; ++d) { nc[d] = (int)(r % (nsec[d] + 1)); r /= (nsec[d] + 1); }
    int nes = fact(dim);
    int cnt = 0;
    for (int m = 0; m < (1 << dim); ++m) {
        int sc[3], ok = 1;
        int64_t sid = 0, stride = 1;
        for (int d = 0; d < dim; ++d) {
            sc[d] = nc[d] - ((m >> d) & 1);
            if (sc[d] < 0 || sc[d] >= nsec[d]) ok = 0;
            sid += stride * sc[d];
            stride *= nsec[d];
        }
        if (!ok) continue;
        for (int t = 0; t < nes; ++t) {
            int off[12];
            ref_section_vertices(dim, t, off);
            for (int a = 0; a <= dim; ++a) {
                int match = 1;
                for (int d = 0; d < dim; ++d)
                    if (sc[d] + off[a * 3 + d] != nc[d]) match = 0;
                if (match) { elems[cnt] = sid * nes + t; locs[cnt] = a; ++cnt; }
            }
        }
    }
    return cnt;
}

int32_t or_projection(int dim, const int32_t* nsec, const double* extent,
                      const double* eps, const or_egrid* g,
                      int32_t* off, int32_t* len, double* R, int32_t maxw) {
    int64_t n = or_num_nodes(dim, nsec), nel = or_num_elements(dim, nsec);
    double vol = or_element_volume(dim, nsec, extent);
    int32_t* conn = (int32_t*)malloc(sizeof(int32_t) * nel * (dim + 1));
    or_mesh_elements(dim, nsec, conn);
    int32_t maxlen = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : maxlen)
    for (int64_t i = 0; i < n; ++i) {
        int64_t elems[48];
        int locs[48];
        int ns = node_star(dim, nsec, i, elems, locs);
        int32_t lo = 0x7fffffff, hi = -1;
        for (int q = 0; q < ns; ++q) {
            double emin = 1e300, emax = -1e300;
            for (int a = 0; a <= dim; ++a) {
                double e = eps[conn[elems[q] * (dim + 1) + a]];
                if (e < emin) emin = e;
                if (e > emax) emax = e;
            }
            int32_t c0 = egrid_cell(g, emin), c1 = egrid_cell(g, emax) + 1;
            if (c0 < lo) lo = c0;
            if (c1 > hi) hi = c1;
        }
        off[i] = lo;
        len[i] = hi - lo + 1;
        if (len[i] > maxlen) maxlen = len[i];
        if (R && len[i] <= maxw) {
            double* row = R + i * (int64_t)maxw;
            for (int32_t w = 0; w < maxw; ++w) row[w] = 0.0;
            for (int q = 0; q < ns; ++q) {
                double ev[4];
                for (int a = 0; a <= dim; ++a) ev[a] = eps[conn[elems[q] * (dim + 1) + a]];
                project_simplex_basis(dim, ev, locs[q], vol, g, row, lo);
            }
        }
    }
    free(conn);
    return maxlen;
}

/* ============================================================ kernel K */
int32_t or_gauss_table(double de, double shift, double sigma, double cutoff,
                       int32_t* m_lo, double* gm, int32_t cap) {
    int32_t lo = (int32_t)ceil((-cutoff * sigma - shift) / de);
    int32_t hi = (int32_t)floor((cutoff * sigma - shift) / de);
    int32_t cnt = 0;
    const double norm = 1.0 / (sigma * sqrt(2.0 * M_PI));
    for (int32_t m = lo; m <= hi; ++m) {
        double x = (double)m * de + shift;
        if (cnt < cap) gm[cnt] = norm * exp(-0.5 * (x / sigma) * (x / sigma));
        ++cnt;
    }
    *m_lo = lo;
    return cnt;
}

/* Q_j(b) = sum_b' g_{b-b'} R_j(b'), b in [off_j + m_lo, off_j + len_j - 1 + m_hi] */
static void conv_row(const double* Rj, int32_t lenj, const double* gm, int32_t nm, double* Q) {
    int32_t lq = lenj + nm - 1;
    for (int32_t p = 0; p < lq; ++p) {
        double s = 0.0;
        /* Q[p] = sum_w R[w] g[p - w] */
        int32_t w0 = p - (nm - 1) > 0 ? p - (nm - 1) : 0;
        int32_t w1 = p < lenj - 1 ? p : lenj - 1;
        for (int32_t w = w0; w <= w1; ++w) s += Rj[w] * gm[p - w];
        Q[p] = s;
    }
}

static double kernel_entry(const double* Ri, int32_t offi, int32_t leni,
                           const double* Qj, int32_t qoff, int32_t qlen) {
    int32_t b0 = offi > qoff ? offi : qoff;
    int32_t b1 = (offi + leni < qoff + qlen) ? offi + leni : qoff + qlen;
    double s = 0.0;
    for (int32_t b = b0; b < b1; ++b) s += Ri[b - offi] * Qj[b - qoff];
    return s;
}

void or_kernel_row(int64_t n, int64_t i, const int32_t* off, const int32_t* len,
                   const double* R, int32_t maxw, const double* gm,
                   int32_t m_lo, int32_t nm, double amp, double* Krow) {
    double* Q = (double*)malloc(sizeof(double) * (maxw + nm));
    for (int64_t j = 0; j < n; ++j) {
        conv_row(R + j * (int64_t)maxw, len[j], gm, nm, Q);
        Krow[j] = amp * kernel_entry(R + i * (int64_t)maxw, off[i], len[i], Q,
                                     off[j] + m_lo, len[j] + nm - 1);
    }
    free(Q);
}

void or_kernel_full(int64_t n, const int32_t* off, const int32_t* len,
                    const double* R, int32_t maxw, const double* gm,
                    int32_t m_lo, int32_t nm, double amp, double* K) {
    int32_t qw = maxw + nm - 1;
    double* Q = (double*)malloc(sizeof(double) * n * (int64_t)qw);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) conv_row(R + j * (int64_t)maxw, len[j], gm, nm, Q + j * (int64_t)qw);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j)
            K[i * n + j] = amp * kernel_entry(R + i * (int64_t)maxw, off[i], len[i],
                                              Q + j * (int64_t)qw, off[j] + m_lo, len[j] + nm - 1);
    free(Q);
}

/* ========================================================= collision */
void or_collision_apply(int64_t n, const double* K, const double* V,
                        double bath_n, const double* f, double* I, int conserve) {
    or_collision_apply_occ(n, K, V, bath_n, f, NULL, I, conserve);
}

/* The occupation factors of the node the row belongs to (the test function's
 * node: (1 - f_i) of the gain, f_i of the loss) are x = fbar with a hat
 * stencil, else the nodal f; the products with K use the nodal f. */
void or_collision_apply_occ(int64_t n, const double* K, const double* V, double bath_n,
                            const double* f, const or_hat* hat, double* I, int conserve) {
    double* xb = NULL;
    const double* x = f;
    if (hat) {
        xb = (double*)malloc(sizeof(double) * n);
        or_hat_apply(hat, f, xb);
        x = xb;
    }
    double* u = (double*)calloc(n, sizeof(double));
    double* rk = (double*)calloc(n, sizeof(double));
    double* v = (double*)calloc(n, sizeof(double));
    double* ck = (double*)calloc(n, sizeof(double));
    /* One pass over K (row-major): u = K f and rk = K 1 from each row;
     * v = K^T f and ck = K^T 1 accumulated per thread over its row block and
     * summed in thread order afterwards. */
#pragma omp parallel
    {
        double* vp = (double*)calloc(2 * n, sizeof(double));
        double* cp = vp + n;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const double* Ki = K + i * n;
            const double fi = f[i];
            double su = 0.0, sr = 0.0;
            for (int64_t j = 0; j < n; ++j) {
                double k = Ki[j];
                su += k * f[j];
                sr += k;
                vp[j] += k * fi;
                cp[j] += k;
            }
            u[i] = su;
            rk[i] = sr;
        }
#pragma omp for ordered schedule(static, 1)
        for (int t = 0; t < omp_get_num_threads(); ++t) {
#pragma omp ordered
            for (int64_t j = 0; j < n; ++j) { v[j] += vp[j]; ck[j] += cp[j]; }
        }
        free(vp);
    }
    const double np1 = bath_n + 1.0;
    for (int64_t i = 0; i < n; ++i) {
        double wf = np1 * u[i] + bath_n * v[i];        /* (W f)_i   */
        double wt1 = np1 * ck[i] + bath_n * rk[i];     /* (W^T 1)_i */
        double wtf = np1 * v[i] + bath_n * u[i];       /* (W^T f)_i */
        I[i] = ((1.0 - x[i]) * wf - x[i] * (wt1 - wtf)) / V[i];
        if (hat) xb[i] = or_hat_defect_weight(x[i], np1, bath_n, rk[i], ck[i], V[i]);   /* x[i] is not read again */
    }
    if (conserve) {
        /* weights: the hat form's (in xb), else |f(1-f)| */
        double d = 0.0, wsum = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            d += V[i] * I[i];
            wsum += V[i] * (hat ? xb[i] : fabs(x[i] * (1.0 - x[i])));
        }
        if (wsum > 0.0)
            for (int64_t i = 0; i < n; ++i) I[i] -= d * (hat ? xb[i] : fabs(x[i] * (1.0 - x[i]))) / wsum;
    }
    free(u); free(v); free(rk); free(ck); free(xb);
}

/* The same operator from K's band in energy order: row p (node perm[p]) holds
 * the columns [lo[p], lo[p] + rowptr[p+1] - rowptr[p]) contiguously at
 * vals + rowptr[p].  One pass over the band gives u = K f and v = K^T f
 * (rows dealt to threads in blocks, v summed from per-thread partials in
 * thread order); rk = K 1 and ck = K^T 1 (energy order) come precomputed.
 * scratch: nthreads * n doubles. */
void or_collision_apply_band(int64_t n, const int64_t* rowptr, const int32_t* lo, const double* vals,
                             const double* rk, const double* ck, const double* Vs, const int32_t* perm,
                             double bath_n, const double* f, double* I, double* scratch, int nthreads) {
    or_collision_apply_band_occ(n, rowptr, lo, vals, rk, ck, Vs, perm, bath_n, f, NULL, 0, I, scratch, nthreads);
}

/* ... with the occupation factors of or_collision_apply_occ (hat NULL: the
 * nodal f) and the number-defect removal; scratch (nthreads + 3) * n */
void or_collision_apply_band_occ(int64_t n, const int64_t* rowptr, const int32_t* lo, const double* vals,
                                 const double* rk, const double* ck, const double* Vs, const int32_t* perm,
                                 double bath_n, const double* f, const or_hat* hat, int conserve, double* I,
                                 double* scratch, int nthreads) {
    const int nt = nthreads > 0 ? nthreads : 1;
    double* fs = scratch + (size_t)nt * n;       /* f in energy order */
    double* u = fs + n;
    double* xb = u + n;                          /* occupation factors, caller order */
    const double* x = f;
    if (hat) {
        or_hat_apply(hat, f, xb);
        x = xb;
    }
    for (int64_t p = 0; p < n; ++p) fs[p] = f[perm[p]];
#pragma omp parallel num_threads(nt)
    {
        const int th = omp_get_thread_num();
        double* vp = scratch + (size_t)th * n;
        for (int64_t j = 0; j < n; ++j) vp[j] = 0.0;
        const int64_t p0 = n * th / nt, p1 = n * (th + 1) / nt;
        for (int64_t p = p0; p < p1; ++p) {
            const double* r = vals + rowptr[p];
            const int64_t len = rowptr[p + 1] - rowptr[p];
            const int32_t c0 = lo[p];
            const double fp = fs[p];
            double su = 0.0;
            for (int64_t q = 0; q < len; ++q) {
                su += r[q] * fs[c0 + q];
                vp[c0 + q] += r[q] * fp;
            }
            u[p] = su;
        }
    }
    const double np1 = bath_n + 1.0;
    for (int64_t p = 0; p < n; ++p) {
        double v = 0.0;
        for (int t = 0; t < nt; ++t) v += scratch[(size_t)t * n + p];
        const double wf = np1 * u[p] + bath_n * v;
        const double wt1 = np1 * ck[p] + bath_n * rk[p];
        const double wtf = np1 * v + bath_n * u[p];
        const double xp = x[perm[p]];
        I[perm[p]] = ((1.0 - xp) * wf - xp * (wt1 - wtf)) / Vs[p];
        if (hat) xb[perm[p]] = or_hat_defect_weight(xp, np1, bath_n, rk[p], ck[p], Vs[p]);
    }
    if (conserve) {
        /* the same weights as or_collision_apply_occ, sums in energy order */
        double d = 0.0, wsum = 0.0;
        for (int64_t p = 0; p < n; ++p) {
            const double xp = x[perm[p]];
            d += Vs[p] * I[perm[p]];
            wsum += Vs[p] * (hat ? xb[perm[p]] : fabs(xp * (1.0 - xp)));
        }
        if (wsum > 0.0)
            for (int64_t i = 0; i < n; ++i) I[i] -= d * (hat ? xb[i] : fabs(x[i] * (1.0 - x[i]))) / wsum;
    }
}

void or_conservation(int64_t n, const double* V, const double* eps,
                     const double* f, double* number, double* energy) {
    double s = 0.0, e = 0.0;
    for (int64_t i = 0; i < n; ++i) { s += V[i] * f[i]; e += V[i] * eps[i] * f[i]; }
    *number = s;
    *energy = e;
}

/* ======================================================== error norm */
/* Band.cpp:87-101.  Function arithmetic is nodal (Function.cpp:75-109), so
 * every pointwise operation is done on node values; integrate(rhs) of two
 * P1 fields over an element is exact: vol/((d+1)(d+2)) (sum r_a^2 + (sum r_a)^2). */
double or_relative_error(int dim, int64_t nel, const int32_t* conn,
                         const double* err, const double* pop, double stat,
                         double rtol, double atol, double penalty) {
    double s_r = 0.0, s_p = 0.0;
    for (int64_t e = 0; e < nel; ++e) {
        double sr = 0.0, sr2 = 0.0, sp = 0.0, sp2 = 0.0;
        for (int a = 0; a <= dim; ++a) {
            int32_t i = conn[e * (dim + 1) + a];
            double y = pop[i];
            double scale, pen;
            if (stat == -1.0) {
                double m = fabs(y) < fabs(1.0 - y) ? fabs(y) : fabs(1.0 - y);
                scale = rtol * m + atol;
                pen = ((y < 0.0 ? -y : 0.0) + (y - 1.0 > 0.0 ? y - 1.0 : 0.0)) / scale;
            } else {
                scale = rtol * y + atol;
                pen = (y < 0.0 ? -y : 0.0) / scale;
            }
            double r = err[i] / scale;
            sr += r; sr2 += r * r; sp += pen; sp2 += pen * pen;
        }
        s_r += sr2 + sr * sr;
        s_p += sp2 + sp * sp;
    }
    /* vol cancels against Nel*vol in the normalisation */
    double denom = (double)((dim + 1) * (dim + 2)) * (double)nel;
    return sqrt((s_r + penalty * s_p) / denom);
}

/* ======================================================== DP5(4) driver */
/* Butcher tableau, controller and dense output of DormandPrince54.hpp:503-511,
 * :532-543, :588-662, :598-611.  The step vetoer (:562-564, :640-642) is
 * MaterialTimeStatus::timeStepVetoer (MaterialTimeStatus.cpp:60-100); a
 * constrained bath enters through the stage times the rhs receives
 * (enforceValue at t + c_i h, :565-578). */
#define DP_C2 (1. / 5.)
#define DP_C3 (3. / 10.)
#define DP_C4 (4. / 5.)
#define DP_C5 (8. / 9.)
static const double A21 = 1. / 5.;
static const double A31 = 3. / 40., A32 = 9. / 40.;
static const double A41 = 44. / 45., A42 = -56. / 15., A43 = 32. / 9.;
static const double A51 = 19372. / 6561., A52 = -25360. / 2187., A53 = 64448. / 6561., A54 = -212. / 729.;
static const double A61 = 9017. / 3168., A62 = -355. / 33., A63 = 46732. / 5247., A64 = 49. / 176., A65 = -5103. / 18656.;
static const double B1 = 35. / 384., B3 = 500. / 1113., B4 = 125. / 192., B5 = -2187. / 6784., B6 = 11. / 84.;
static const double BH1 = 5179. / 57600., BH3 = 7571. / 16695., BH4 = 393. / 640., BH5 = -92097. / 339200., BH6 = 187. / 2100., BH7 = 1. / 40.;

typedef struct dp5_work {
    int64_t n;
    double* base;
    double *k1, *k2, *k3, *k4, *k5, *k6, *k7, *ys, *y5, *err;
} dp5_work;

static void dp5_alloc(dp5_work* w, int64_t n) {
    w->n = n;
    double* b = (double*)malloc(sizeof(double) * n * 10);
    w->base = b;
    w->k1 = b; w->k2 = b + n; w->k3 = b + 2 * n; w->k4 = b + 3 * n; w->k5 = b + 4 * n;
    w->k6 = b + 5 * n; w->k7 = b + 6 * n; w->ys = b + 7 * n; w->y5 = b + 8 * n; w->err = b + 9 * n;
}
static void dp5_free(dp5_work* w) { free(w->base); }

/* One attempted step: k2..k7, y5 and err (= y5 - y4).  Returns evals (6).
 * The stage sums are written in the reference's order (h*a_ij)*k_j. */
static void dp5_attempt(dp5_work* w, const double* y, double t, double h, or_rhs_fn rhs, void* user) {
    int64_t n = w->n;
    for (int64_t i = 0; i < n; ++i) w->ys[i] = y[i] + (A21 * h) * w->k1[i];
    rhs(user, t + DP_C2 * h, w->ys, w->k2);
    for (int64_t i = 0; i < n; ++i) w->ys[i] = y[i] + (A31 * h) * w->k1[i] + (A32 * h) * w->k2[i];
    rhs(user, t + DP_C3 * h, w->ys, w->k3);
    for (int64_t i = 0; i < n; ++i)
        w->ys[i] = y[i] + (A41 * h) * w->k1[i] + (A42 * h) * w->k2[i] + (A43 * h) * w->k3[i];
    rhs(user, t + DP_C4 * h, w->ys, w->k4);
    for (int64_t i = 0; i < n; ++i)
        w->ys[i] = y[i] + (A51 * h) * w->k1[i] + (A52 * h) * w->k2[i] + (A53 * h) * w->k3[i] + (A54 * h) * w->k4[i];
    rhs(user, t + DP_C5 * h, w->ys, w->k5);
    for (int64_t i = 0; i < n; ++i)
        w->ys[i] = y[i] + (A61 * h) * w->k1[i] + (A62 * h) * w->k2[i] + (A63 * h) * w->k3[i] + (A64 * h) * w->k4[i] + (A65 * h) * w->k5[i];
    rhs(user, t + h, w->ys, w->k6);
    for (int64_t i = 0; i < n; ++i)
        w->y5[i] = y[i] + (h * B1) * w->k1[i] + (h * B3) * w->k3[i] + (h * B4) * w->k4[i] + (h * B5) * w->k5[i] + (h * B6) * w->k6[i];
    rhs(user, t + h, w->y5, w->k7);
    for (int64_t i = 0; i < n; ++i) {
        double y4 = y[i] + (h * BH1) * w->k1[i] + (h * BH3) * w->k3[i] + (h * BH4) * w->k4[i] + (h * BH5) * w->k5[i] + (h * BH6) * w->k6[i] + (h * BH7) * w->k7[i];
        w->err[i] = w->y5[i] - y4;
    }
}

/* dense output at s in (0,1]: DormandPrince54.hpp:601-608 */
static void dp5_dense(const dp5_work* w, const double* y, double h, double s, double* out) {
    double bt1 = -(435. * s * s * s - 1184. * s * s + 1098. * s - 384.) / (384.);
    double bt3 = 500. * s * (6. * s * s - 14. * s + 9.) / (1113.);
    double bt4 = -125. * s * (9. * s * s - 16. * s + 6.) / (192.);
    double bt5 = 729. * s * (35. * s * s - 64. * s + 26.) / (6784.);
    double bt6 = -11. * s * (3. * s - 2.) * (5. * s - 6.) / (84.);
    double bt7 = s * (s - 1.) * (5. * s - 3.) / (2.);
    for (int64_t i = 0; i < w->n; ++i)
        out[i] = y[i] + (s * h * bt1) * w->k1[i] + (s * h * bt3) * w->k3[i] + (s * h * bt4) * w->k4[i] +
                 (s * h * bt5) * w->k5[i] + (s * h * bt6) * w->k6[i] + (s * h * bt7) * w->k7[i];
}

/* piecewise-linear samples, held at the ends (the device's sched_value) */
double or_sched_value(const double* s, int32_t ns, double t_start, double dt, double t) {
    const double u = (t - t_start) / dt;
    if (!(u > 0.0)) return s[0];
    if (u >= (double)(ns - 1)) return s[ns - 1];
    int k = (int)floor(u);
    if (k > ns - 2) k = ns - 2;
    const double w = u - (double)k;
    return fma(w, s[k + 1] - s[k], s[k]);
}

/* MaterialTimeStatus.cpp:60-100 for one uniform constrained band: sample
 * times t0 + dense_dt, t0 + 2 dense_dt, ... run through the five intervals of
 * timesToTest; a sample further from its interval's average than
 * (1 + 0.3)|L - R| + atol + rtol |average| is the critical time. */
double or_step_vetoer(const or_bath_sched* bs, double t, double h, double dense_dt) {
    const double times[6] = {t, t + DP_C2 * h, t + DP_C3 * h, t + DP_C4 * h, t + DP_C5 * h, t + h};
    double crit = times[5];
    if (bs) {
        const double var = 1. + 0.3;
        double cur = times[0] + dense_dt;
        double L = or_sched_value(bs->s, bs->ns, bs->t_start, bs->dt, times[0]);
        double R = or_sched_value(bs->s, bs->ns, bs->t_start, bs->dt, times[1]);
        int found = 0;
        for (int i = 0; i < 5 && !found; ++i) {
            if (i != 0) {
                L = R;
                R = or_sched_value(bs->s, bs->ns, bs->t_start, bs->dt, times[i + 1]);
            }
            const double avg = 0.5 * (L + R);
            const double acc = var * fabs(L - R) + bs->atol + bs->rtol * fabs(avg);
            while (cur < times[i + 1]) {
                const double v = or_sched_value(bs->s, bs->ns, bs->t_start, bs->dt, cur);
                if (fabs(v - avg) / acc > 1.) {
                    crit = cur;
                    found = 1;
                    break;
                }
                const double next = cur + dense_dt;
                if (!(next > cur)) break;
                cur = next;
            }
        }
    }
    return crit - times[0];
}

void or_dp5_integrate(int64_t n, double* y, double t0, double t_final,
                      double dense_dt, or_rhs_fn rhs, or_norm_fn norm, void* user,
                      double* dense_out, double* dense_t, int32_t max_dense,
                      double* h_log, int8_t* acc_log, int64_t log_cap,
                      or_dp5_stats* stats) {
    or_dp5_integrate_ex(n, y, t0, t_final, dense_dt, rhs, norm, user, NULL, NULL, NULL, dense_out,
                        dense_t, max_dense, h_log, acc_log, log_cap, stats);
}

void or_dp5_integrate_ex(int64_t n, double* y, double t0, double t_final,
                         double dense_dt, or_rhs_fn rhs, or_norm_fn norm, void* user,
                         const or_bath_sched* bs, or_flux_fn flux, double* ledger,
                         double* dense_out, double* dense_t, int32_t max_dense,
                         double* h_log, int8_t* acc_log, int64_t log_cap,
                         or_dp5_stats* stats) {
    dp5_work w;
    dp5_alloc(&w, n);
    or_dp5_stats st;
    memset(&st, 0, sizeof(st));
    double t = t0, h = dense_dt, dense_time = t0 + dense_dt;
    const double min_h = dense_dt / 10000.;
    const double minscale = 0.333333, maxscale = 3.0, safety = 0.98;
    const double alpha = 0.7 / 5., beta = 0.4 / 5.;
    double olderror = 1.0;
    int notfailed = 1, rejected = 0;
    double e1 = 0.0;   /* flux of k1 (FSAL with k7) */
    rhs(user, t, y, w.k1);
    if (flux) e1 = flux(user, w.k1);
    st.evals = 1;
    h = fmin(h, or_step_vetoer(bs, t, h, dense_dt));
    while (t < t_final && notfailed) {
        dp5_attempt(&w, y, t, h, rhs, user);
        st.evals += 6;
        if (h_log && st.attempted < log_cap) h_log[st.attempted] = h;
        double en = norm(user, w.err, w.y5);
        int accept = en <= 1.0;
        if (acc_log && st.attempted < log_cap) acc_log[st.attempted] = (int8_t)accept;
        ++st.attempted;
        if (accept) {
            while (dense_time <= t + 1.0001 * h && dense_time <= t_final + .0001 * h) {
                double s = (dense_time - t) / h;
                if (st.n_dense < max_dense && dense_out) {
                    dp5_dense(&w, y, h, s, dense_out + (int64_t)st.n_dense * n);
                    if (dense_t) dense_t[st.n_dense] = dense_time;
                }
                ++st.n_dense;
                dense_time += dense_dt;
            }
            if (flux) {
                const double e3 = flux(user, w.k3), e4 = flux(user, w.k4), e5 = flux(user, w.k5);
                const double e6 = flux(user, w.k6), e7 = flux(user, w.k7);
                const double de = (h * B1) * e1 + (h * B3) * e3 + (h * B4) * e4 + (h * B5) * e5 + (h * B6) * e6;
                *ledger -= de;
                e1 = e7;
            }
            memcpy(y, w.y5, sizeof(double) * n);
            const double hstep = h;
            t += h;
            double* tmp = w.k1; w.k1 = w.k7; w.k7 = tmp;
            double scale;
            if (en == 0.) {
                scale = maxscale;
            } else {
                scale = safety * pow(1. / en, alpha) * pow(olderror, beta);
                scale = fmin(fmax(scale, minscale), maxscale);
            }
            if (rejected) h = hstep * fmin(1.0, scale);
            else h = hstep * scale;
            h = fmin(h, or_step_vetoer(bs, t, h, dense_dt));
            rejected = 0;
            olderror = en;
            ++st.accepted;
        } else {
            double scale = safety * pow(1. / en, 1 / 5.);
            scale = fmin(fmax(scale, minscale), maxscale);
            h *= scale;
            rejected = 1;
            ++st.rejected;
            if (h < min_h) notfailed = 0;
        }
    }
    st.failed = !notfailed;
    if (stats) *stats = st;
    dp5_free(&w);
}

/* ----------------------------------------------- W-form right-hand side */
typedef struct coll_user {
    int dim;
    int64_t n, nel;
    const double* K;
    const double* V;
    double bath_n;
    const or_bath_sched* bs;   /* non-NULL: the bath at the stage time */
    const double* eps;         /* ledger weights V_i eps_i come from V and eps */
    int conserve;
    const or_hat* hat;         /* non-NULL: hat-averaged occupation factors */
    int32_t* conn;
    double rtol, atol;
} coll_user;

static void coll_rhs(void* u, double t, const double* y, double* dydt) {
    coll_user* c = (coll_user*)u;
    const double nb = c->bs ? or_sched_value(c->bs->s, c->bs->ns, c->bs->t_start, c->bs->dt, t) : c->bath_n;
    or_collision_apply_occ(c->n, c->K, c->V, nb, y, c->hat, dydt, c->conserve);
}

/* electron energy flux sum_i V_i eps_i dydt_i (sequential) */
static double coll_flux(void* u, const double* dydt) {
    coll_user* c = (coll_user*)u;
    double s = 0.0;
    for (int64_t i = 0; i < c->n; ++i) s += c->V[i] * c->eps[i] * dydt[i];
    return s;
}

static double coll_norm(void* u, const double* err, const double* y5) {
    coll_user* c = (coll_user*)u;
    return or_relative_error(c->dim, c->nel, c->conn, err, y5, -1.0, c->rtol, c->atol, 0.0);
}

void or_dp5_collision(int dim, const int32_t* nsec, int64_t n, const double* K,
                      const double* V, double bath_n, double* f, double t0,
                      double t_final, double dense_dt, double rtol, double atol,
                      int conserve, double* dense_out, double* dense_t,
                      int32_t max_dense, double* h_log, int8_t* acc_log,
                      int64_t log_cap, or_dp5_stats* stats) {
    or_dp5_collision_ex(dim, nsec, n, K, V, NULL, bath_n, NULL, 0, 0.0, 1.0, 0.0, 0.0, f, t0, t_final,
                        dense_dt, rtol, atol, conserve, dense_out, dense_t, max_dense, h_log, acc_log,
                        log_cap, stats, NULL);
}

void or_dp5_collision_ex(int dim, const int32_t* nsec, int64_t n, const double* K,
                         const double* V, const double* eps, double bath_n,
                         const double* sched, int32_t ns, double t_start, double sdt,
                         double satol, double srtol, double* f, double t0,
                         double t_final, double dense_dt, double rtol, double atol,
                         int conserve, double* dense_out, double* dense_t,
                         int32_t max_dense, double* h_log, int8_t* acc_log,
                         int64_t log_cap, or_dp5_stats* stats, double* ledger) {
    coll_user c;
    memset(&c, 0, sizeof(c));
    or_bath_sched bs = {sched, ns, t_start, sdt, satol, srtol};
    c.dim = dim; c.n = n; c.nel = or_num_elements(dim, nsec);
    c.K = K; c.V = V; c.bath_n = bath_n; c.conserve = conserve & OR_CONSERVE;
    or_hat* hat = (conserve & OR_OCC_HAT) ? or_hat_create(dim, nsec) : NULL;
    c.hat = hat;
    c.bs = sched ? &bs : NULL;
    c.eps = eps;
    c.rtol = rtol; c.atol = atol;
    c.conn = (int32_t*)malloc(sizeof(int32_t) * c.nel * (dim + 1));
    or_mesh_elements(dim, nsec, c.conn);
    if (ledger) *ledger = 0.0;
    or_dp5_integrate_ex(n, f, t0, t_final, dense_dt, coll_rhs, coll_norm, &c, c.bs,
                        (ledger && eps) ? coll_flux : NULL, ledger, dense_out, dense_t, max_dense,
                        h_log, acc_log, log_cap, stats);
    free(c.conn);
    or_hat_free(hat);
}

void or_dp5_collision_replay(int dim, const int32_t* nsec, int64_t n,
                             const double* K, const double* V, double bath_n,
                             double* f, double t0, double t_final,
                             double dense_dt, int conserve, const double* h_log,
                             const int8_t* acc_log, int64_t n_attempts,
                             double* dense_out, double* dense_t,
                             int32_t max_dense) {
    or_dp5_collision_replay_ex(dim, nsec, n, K, V, NULL, bath_n, NULL, 0, 0.0, 1.0, f, t0, t_final,
                               dense_dt, conserve, h_log, acc_log, n_attempts, dense_out, dense_t,
                               max_dense, NULL);
}

/* the fixed-step driver shared by the replays: accepted attempts of
 * (h_log, acc_log) in order, dense outputs and, with flux, the ledger */
void or_dp5_replay_ex(int64_t n, double* f, double t0, double t_final, double dense_dt,
                      or_rhs_fn rhs, void* user, or_flux_fn flux, double* ledger,
                      const double* h_log, const int8_t* acc_log, int64_t n_attempts,
                      double* dense_out, double* dense_t, int32_t max_dense) {
    dp5_work w;
    dp5_alloc(&w, n);
    double t = t0, dense_time = t0 + dense_dt;
    int32_t nd = 0;
    rhs(user, t, f, w.k1);
    double e1 = flux ? flux(user, w.k1) : 0.0;
    if (ledger) *ledger = 0.0;
    for (int64_t a = 0; a < n_attempts && t < t_final; ++a) {
        if (!acc_log[a]) continue;
        double h = h_log[a];
        dp5_attempt(&w, f, t, h, rhs, user);
        if (flux) {
            const double e3 = flux(user, w.k3), e4 = flux(user, w.k4), e5 = flux(user, w.k5);
            const double e6 = flux(user, w.k6), e7 = flux(user, w.k7);
            *ledger -= (h * B1) * e1 + (h * B3) * e3 + (h * B4) * e4 + (h * B5) * e5 + (h * B6) * e6;
            e1 = e7;
        }
        while (dense_time <= t + 1.0001 * h && dense_time <= t_final + .0001 * h) {
            double s = (dense_time - t) / h;
            if (nd < max_dense && dense_out) {
                dp5_dense(&w, f, h, s, dense_out + (int64_t)nd * n);
                if (dense_t) dense_t[nd] = dense_time;
            }
            ++nd;
            dense_time += dense_dt;
        }
        memcpy(f, w.y5, sizeof(double) * n);
        t += h;
        double* tmp = w.k1; w.k1 = w.k7; w.k7 = tmp;
    }
    dp5_free(&w);
}

void or_dp5_collision_replay_ex(int dim, const int32_t* nsec, int64_t n,
                                const double* K, const double* V, const double* eps,
                                double bath_n, const double* sched, int32_t ns,
                                double t_start, double sdt, double* f, double t0,
                                double t_final, double dense_dt, int conserve,
                                const double* h_log, const int8_t* acc_log,
                                int64_t n_attempts, double* dense_out, double* dense_t,
                                int32_t max_dense, double* ledger) {
    coll_user c;
    memset(&c, 0, sizeof(c));
    or_bath_sched bs = {sched, ns, t_start, sdt, 0.0, 0.0};
    c.n = n; c.K = K; c.V = V; c.bath_n = bath_n; c.conserve = conserve & OR_CONSERVE;
    or_hat* hat = (conserve & OR_OCC_HAT) ? or_hat_create(dim, nsec) : NULL;
    c.hat = hat;
    c.bs = sched ? &bs : NULL;
    c.eps = eps;
    or_dp5_replay_ex(n, f, t0, t_final, dense_dt, coll_rhs, &c, (ledger && eps) ? coll_flux : NULL, ledger,
                     h_log, acc_log, n_attempts, dense_out, dense_t, max_dense);
    or_hat_free(hat);
}

/* ============================================================ adjoint */
void or_kernel_domega_full(int64_t n, const int32_t* off, const int32_t* len,
                           const double* R, int32_t maxw, double de, double omega0,
                           double sigma, double cutoff, double amp, double* Kd) {
    /* G'(x) = -x/sigma^2 G(x) sampled on the same offsets as G */
    int32_t m_lo = (int32_t)ceil((-cutoff * sigma - omega0) / de);
    int32_t m_hi = (int32_t)floor((cutoff * sigma - omega0) / de);
    int32_t nm = m_hi - m_lo + 1;
    double* gd = (double*)malloc(sizeof(double) * nm);
    const double norm = 1.0 / (sigma * sqrt(2.0 * M_PI));
    for (int32_t m = m_lo; m <= m_hi; ++m) {
        double x = (double)m * de + omega0;
        gd[m - m_lo] = -x / (sigma * sigma) * norm * exp(-0.5 * (x / sigma) * (x / sigma));
    }
    or_kernel_full(n, off, len, R, maxw, gd, m_lo, nm, amp, Kd);
    free(gd);
}

/* u = K x and v = K^T x for several vectors x at once (one pass over K) */
static void dual_products(int64_t n, const double* K, int nv, const double* const* x,
                          double* const* u, double* const* v) {
    for (int q = 0; q < nv; ++q)
        for (int64_t i = 0; i < n; ++i) { u[q][i] = 0.0; v[q][i] = 0.0; }
    for (int64_t i = 0; i < n; ++i) {
        const double* Ki = K + i * n;
        for (int q = 0; q < nv; ++q) {
            double su = 0.0;
            const double xi = x[q][i];
            double* vq = v[q];
            const double* xq = x[q];
            for (int64_t j = 0; j < n; ++j) {
                su += Ki[j] * xq[j];
                vq[j] += Ki[j] * xi;
            }
            u[q][i] = su;
        }
    }
}

void or_collision_general(int64_t n, const double* K, const double* V, double alpha,
                          double beta, const double* f, double* I) {
    or_collision_general_occ(n, K, V, alpha, beta, f, NULL, I);
}

void or_collision_general_occ(int64_t n, const double* K, const double* V, double alpha,
                              double beta, const double* f, const or_hat* hat, double* I) {
    double* buf = (double*)malloc(sizeof(double) * 6 * n);
    double* x = buf + 5 * n;
    if (hat) or_hat_apply(hat, f, x);
    else memcpy(x, f, sizeof(double) * n);
    double *one = buf, *u = buf + n, *v = buf + 2 * n, *rk = buf + 3 * n, *ck = buf + 4 * n;
    for (int64_t i = 0; i < n; ++i) one[i] = 1.0;
    const double* xs[2] = {f, one};
    double* us[2] = {u, rk};
    double* vs[2] = {v, ck};
    dual_products(n, K, 2, xs, us, vs);
    for (int64_t i = 0; i < n; ++i) {
        double wf = alpha * u[i] + beta * v[i];
        double wt1 = alpha * ck[i] + beta * rk[i];
        double wtf = alpha * v[i] + beta * u[i];
        I[i] = ((1.0 - x[i]) * wf - x[i] * (wt1 - wtf)) / V[i];
    }
    free(buf);
}

/* J^T mu: with m = mu/V, x = m(1-f), y = m f,
 * out_k = -m_k [(W f)_k + (W^T 1)_k - (W^T f)_k] + (W^T x)_k + (W y)_k */
void or_collision_vjp(int64_t n, const double* K, const double* V, double bath_n,
                      const double* f, const double* mu, double* out) {
    or_collision_vjp_occ(n, K, V, bath_n, f, NULL, mu, out);
}

/* With a hat stencil S the occupation factors are fbar = S f, and the term
 * -m (W f + W^T 1 - W^T f) reaches f through S^T. */
void or_collision_vjp_occ(int64_t n, const double* K, const double* V, double bath_n,
                          const double* f, const or_hat* hat, const double* mu, double* out) {
    or_collision_vjp_ex(n, K, V, bath_n, f, hat, 0, mu, out);
}

/* The number-defect removal's pieces at state f: occupation factors x, the
 * products uf = K f, vf = K^T f, rk = K 1, ck = K^T 1, the uncorrected
 * I0, the weights w (0 without conserve), s = sum V I0 / sum V w and, for a
 * cotangent mu, c = sum mu w / sum V w.  buf: 8 n doubles, laid out as
 * x | uf | vf | rk | ck | I0 | w | scratch. */
static void defect_pieces(int64_t n, const double* K, const double* V, double bath_n, const double* f,
                          const or_hat* hat, int conserve, const double* mu, double* buf, double* s,
                          double* c) {
    double *x = buf, *uf = buf + n, *vf = buf + 2 * n, *rk = buf + 3 * n, *ck = buf + 4 * n;
    double *I0 = buf + 5 * n, *w = buf + 6 * n, *one = buf + 7 * n;
    if (hat) or_hat_apply(hat, f, x);
    else memcpy(x, f, sizeof(double) * n);
    for (int64_t i = 0; i < n; ++i) one[i] = 1.0;
    const double* xs[2] = {f, one};
    double* us[2] = {uf, rk};
    double* vs[2] = {vf, ck};
    dual_products(n, K, 2, xs, us, vs);
    const double np1 = bath_n + 1.0;
    double D = 0.0, Z = 0.0, M = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double wf = np1 * uf[i] + bath_n * vf[i];
        const double wt1 = np1 * ck[i] + bath_n * rk[i];
        const double wtf = np1 * vf[i] + bath_n * uf[i];
        I0[i] = ((1.0 - x[i]) * wf - x[i] * (wt1 - wtf)) / V[i];
        w[i] = !conserve ? 0.0
               : hat ? or_hat_defect_weight(x[i], np1, bath_n, rk[i], ck[i], V[i])
                     : fabs(x[i] * (1.0 - x[i]));
        D += V[i] * I0[i];
        Z += V[i] * w[i];
        if (mu) M += mu[i] * w[i];
    }
    *s = (conserve && Z > 0.0) ? D / Z : 0.0;
    *c = (conserve && Z > 0.0) ? M / Z : 0.0;
}

/* d w_i / d x_i of the removal's weights */
static double defect_dweight(const or_hat* hat, double x, double np1, double nb, double rk, double ck, double V) {
    if (hat) return ((np1 * ck + nb * rk) - (np1 * rk + nb * ck)) / V;
    const double p = x * (1.0 - x);
    return p > 0.0 ? 1.0 - 2.0 * x : (p < 0.0 ? 2.0 * x - 1.0 : 0.0);
}

/* With conserve the operator is I = I0 - s w (s = sum V I0 / sum V w), and
 *   mu^T dI = mu'^T dI0 - s mu'^T dw,  mu' = mu - c V,  c = mu^T w / sum V w. */
void or_collision_vjp_ex(int64_t n, const double* K, const double* V, double bath_n, const double* f,
                         const or_hat* hat, int conserve, const double* mu, double* out) {
    double* pc = (double*)malloc(sizeof(double) * 8 * n);
    double sdef = 0.0, cdef = 0.0;
    defect_pieces(n, K, V, bath_n, f, hat, conserve, mu, pc, &sdef, &cdef);
    const double *ob = pc, *uf = pc + n, *vf = pc + 2 * n, *u1 = pc + 3 * n, *v1 = pc + 4 * n;
    double* buf = (double*)malloc(sizeof(double) * 9 * n);
    double *x = buf, *y = buf + n, *mup = buf + 2 * n;
    double *ux = buf + 3 * n, *vx = buf + 4 * n, *uy = buf + 5 * n, *vy = buf + 6 * n;
    double *z = buf + 7 * n, *zt = buf + 8 * n;
    for (int64_t i = 0; i < n; ++i) {
        mup[i] = mu[i] - cdef * V[i];
        const double m = mup[i] / V[i];
        x[i] = m * (1.0 - ob[i]);
        y[i] = m * ob[i];
    }
    const double* xs[2] = {x, y};
    double* us[2] = {ux, uy};
    double* vs[2] = {vx, vy};
    dual_products(n, K, 2, xs, us, vs);
    const double np1 = bath_n + 1.0;
    for (int64_t k = 0; k < n; ++k) {
        const double m = mup[k] / V[k];
        const double wf = np1 * uf[k] + bath_n * vf[k];
        const double wt1 = np1 * v1[k] + bath_n * u1[k];
        const double wtf = np1 * vf[k] + bath_n * uf[k];
        const double wtx = np1 * vx[k] + bath_n * ux[k];
        const double wy = np1 * uy[k] + bath_n * vy[k];
        z[k] = -m * (wf + wt1 - wtf);
        if (conserve) z[k] -= sdef * mup[k] * defect_dweight(hat, ob[k], np1, bath_n, u1[k], v1[k], V[k]);
        out[k] = wtx + wy;
    }
    if (hat) or_hat_apply_t(hat, z, zt);
    else memcpy(zt, z, sizeof(double) * n);
    for (int64_t k = 0; k < n; ++k) out[k] += zt[k];
    free(buf);
    free(pc);
}

typedef struct adj_ctx {
    int64_t n;
    const double *K, *Kd, *V;
    double bath, dbath, g;
    const or_hat* hat;
    int conserve;
} adj_ctx;

static void adj_rhs(const adj_ctx* a, const double* y, double* out) {
    or_collision_apply_occ(a->n, a->K, a->V, a->bath, y, a->hat, out, a->conserve);
}

/* mu^T dF/dtheta at state y: (dL/dg, dL/domega0) contributions.  With the
 * number-defect removal F = I0 - s w:  mu^T dF = mu'^T dI0 - s mu'^T dw
 * (mu' = mu - c V); the hat form's weights scale as g^2 and move with omega0
 * through K 1, K^T 1 and the bath, the nodal weights depend on neither. */
static void adj_param(const adj_ctx* a, const double* y, const double* mu, double* dg, double* dw) {
    int64_t n = a->n;
    double* pc = (double*)malloc(sizeof(double) * 8 * n);
    double sdef = 0.0, cdef = 0.0;
    defect_pieces(n, a->K, a->V, a->bath, y, a->hat, a->conserve, mu, pc, &sdef, &cdef);
    const double *x = pc, *rk = pc + 3 * n, *ck = pc + 4 * n, *I0 = pc + 5 * n, *w = pc + 6 * n;
    double* t = (double*)malloc(sizeof(double) * 5 * n);
    double *mup = t, *t1 = t + n, *t2 = t + 2 * n, *rkd = t + 3 * n, *ckd = t + 4 * n;
    for (int64_t i = 0; i < n; ++i) mup[i] = mu[i] - cdef * a->V[i];
    double sg = 0.0;
    for (int64_t i = 0; i < n; ++i) sg += mup[i] * (I0[i] - (a->hat ? sdef * w[i] : 0.0));
    *dg += 2.0 * sg / a->g;
    /* W' = (n+1) Kd + n Kd^T + n' (K + K^T) */
    or_collision_general_occ(n, a->Kd, a->V, a->bath + 1.0, a->bath, y, a->hat, t1);
    or_collision_general_occ(n, a->K, a->V, a->dbath, a->dbath, y, a->hat, t2);
    double sw = 0.0;
    for (int64_t i = 0; i < n; ++i) sw += mup[i] * (t1[i] + t2[i]);
    if (a->conserve && a->hat) {
        for (int64_t i = 0; i < n; ++i) { rkd[i] = 0.0; ckd[i] = 0.0; }
        for (int64_t i = 0; i < n; ++i)
            for (int64_t j = 0; j < n; ++j) { rkd[i] += a->Kd[i * n + j]; ckd[j] += a->Kd[i * n + j]; }
        const double np1 = a->bath + 1.0, nb = a->bath, dn = a->dbath;
        for (int64_t i = 0; i < n; ++i) {
            const double dA = np1 * rkd[i] + nb * ckd[i] + dn * (rk[i] + ck[i]);
            const double dB = np1 * ckd[i] + nb * rkd[i] + dn * (ck[i] + rk[i]);
            sw -= sdef * mup[i] * ((1.0 - x[i]) * dA + x[i] * dB) / a->V[i];
        }
    }
    *dw += sw;
    free(t);
    free(pc);
}

static const double TA[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {1. / 5., 0, 0, 0, 0, 0},
    {3. / 40., 9. / 40., 0, 0, 0, 0},
    {44. / 45., -56. / 15., 32. / 9., 0, 0, 0},
    {19372. / 6561., -25360. / 2187., 64448. / 6561., -212. / 729., 0, 0},
    {9017. / 3168., -355. / 33., 46732. / 5247., 49. / 176., -5103. / 18656., 0},
    {35. / 384., 0., 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84.}};

/* stage states Y[0..6] (Y[6] = y5) of the step from y with size h */
static void adj_stages(const adj_ctx* a, const double* y, double h, double** Y, double** k) {
    int64_t n = a->n;
    for (int s = 0; s < 7; ++s) {
        for (int64_t i = 0; i < n; ++i) {
            double v = y[i];
            for (int j = 0; j < s; ++j) v += (TA[s][j] * h) * k[j][i];
            Y[s][i] = v;
        }
        adj_rhs(a, Y[s], k[s]);
    }
}

void or_dp5_adjoint(int64_t n, const double* K, const double* Kd, const double* V,
                    double bath_n, double dbath, double g, const double* f0, double t0,
                    double t_final, double dense_dt, const double* h_steps, int64_t n_steps,
                    int32_t ckpt, const double* dL_df, double* dL_df0, double* dL_dg,
                    double* dL_domega0) {
    or_dp5_adjoint_occ(n, K, Kd, V, bath_n, dbath, g, NULL, 0, f0, t0, t_final, dense_dt, h_steps, n_steps,
                       ckpt, dL_df, dL_df0, dL_dg, dL_domega0);
}

void or_dp5_adjoint_occ(int64_t n, const double* K, const double* Kd, const double* V,
                        double bath_n, double dbath, double g, const or_hat* hat, int conserve,
                        const double* f0, double t0, double t_final, double dense_dt,
                        const double* h_steps, int64_t n_steps, int32_t ckpt, const double* dL_df,
                        double* dL_df0, double* dL_dg, double* dL_domega0) {
    adj_ctx a = {n, K, Kd, V, bath_n, dbath, g, hat, conserve};
    /* forward: find the step holding the last dense output and keep checkpoints */
    int64_t nck = n_steps / ckpt + 2;
    double* ck = (double*)malloc(sizeof(double) * n * nck);
    double* y = (double*)malloc(sizeof(double) * n);
    double *Y[7], *k[7];
    double* sb = (double*)malloc(sizeof(double) * 14 * n);
    for (int s = 0; s < 7; ++s) { Y[s] = sb + s * n; k[s] = sb + (7 + s) * n; }
    memcpy(y, f0, sizeof(double) * n);
    double t = t0, dense_time = t0 + dense_dt, s_last = 0.0;
    int64_t m_last = -1;
    for (int64_t m = 0; m < n_steps && t < t_final; ++m) {
        if (m % ckpt == 0) memcpy(ck + (m / ckpt) * n, y, sizeof(double) * n);
        double h = h_steps[m];
        adj_stages(&a, y, h, Y, k);
        while (dense_time <= t + 1.0001 * h && dense_time <= t_final + .0001 * h) {
            m_last = m;
            s_last = (dense_time - t) / h;
            dense_time += dense_dt;
        }
        memcpy(y, Y[6], sizeof(double) * n);
        t += h;
    }
    double* lam = (double*)calloc(n, sizeof(double));
    double* mu[7];
    double* nu = (double*)malloc(sizeof(double) * n);
    double* mb = (double*)malloc(sizeof(double) * 7 * n);
    for (int s = 0; s < 7; ++s) mu[s] = mb + s * n;
    double dg = 0.0, dw = 0.0;
    double* seg = (double*)malloc(sizeof(double) * n * (ckpt + 1));
    for (int64_t c0 = (m_last / ckpt) * ckpt; c0 >= 0 && m_last >= 0; c0 -= ckpt) {
        int64_t c1 = c0 + ckpt - 1 < m_last ? c0 + ckpt - 1 : m_last;
        /* recompute the segment's step-start states from its checkpoint */
        memcpy(seg, ck + (c0 / ckpt) * n, sizeof(double) * n);
        for (int64_t m = c0; m < c1; ++m) {
            adj_stages(&a, seg + (m - c0) * n, h_steps[m], Y, k);
            memcpy(seg + (m - c0 + 1) * n, Y[6], sizeof(double) * n);
        }
        for (int64_t m = c1; m >= c0; --m) {
            const double h = h_steps[m];
            const double* ym = seg + (m - c0) * n;
            adj_stages(&a, ym, h, Y, k);
            for (int s = 0; s < 7; ++s) memset(mu[s], 0, sizeof(double) * n);
            if (m == m_last) {
                /* dense output y + sum_j (s h btilde_j) k_j (DormandPrince54.hpp:601-608) */
                double s = s_last;
                double bt[7];
                bt[0] = -(435. * s * s * s - 1184. * s * s + 1098. * s - 384.) / (384.);
                bt[1] = 0.0;
                bt[2] = 500. * s * (6. * s * s - 14. * s + 9.) / (1113.);
                bt[3] = -125. * s * (9. * s * s - 16. * s + 6.) / (192.);
                bt[4] = 729. * s * (35. * s * s - 64. * s + 26.) / (6784.);
                bt[5] = -11. * s * (3. * s - 2.) * (5. * s - 6.) / (84.);
                bt[6] = s * (s - 1.) * (5. * s - 3.) / (2.);
                for (int64_t i = 0; i < n; ++i) lam[i] = dL_df[i];
                for (int j = 0; j < 7; ++j)
                    for (int64_t i = 0; i < n; ++i) mu[j][i] = (s * h * bt[j]) * dL_df[i];
                /* k7 = F(Y7), Y7 = y + sum_i (h b_i) k_i */
                or_collision_vjp_ex(n, K, V, bath_n, Y[6], hat, conserve, mu[6], nu);
                adj_param(&a, Y[6], mu[6], &dg, &dw);
                for (int64_t i = 0; i < n; ++i) lam[i] += nu[i];
                for (int j = 0; j < 6; ++j)
                    for (int64_t i = 0; i < n; ++i) mu[j][i] += (TA[6][j] * h) * nu[i];
            } else {
                /* y_{m+1} = y_m + sum_i (h b_i) k_i; lam holds dL/dy_{m+1} */
                for (int j = 0; j < 6; ++j)
                    for (int64_t i = 0; i < n; ++i) mu[j][i] = (TA[6][j] * h) * lam[i];
            }
            for (int st = 5; st >= 1; --st) {
                or_collision_vjp_ex(n, K, V, bath_n, Y[st], hat, conserve, mu[st], nu);
                adj_param(&a, Y[st], mu[st], &dg, &dw);
                for (int64_t i = 0; i < n; ++i) lam[i] += nu[i];
                for (int j = 0; j < st; ++j)
                    for (int64_t i = 0; i < n; ++i) mu[j][i] += (TA[st][j] * h) * nu[i];
            }
            or_collision_vjp_ex(n, K, V, bath_n, Y[0], hat, conserve, mu[0], nu);
            adj_param(&a, Y[0], mu[0], &dg, &dw);
            for (int64_t i = 0; i < n; ++i) lam[i] += nu[i];
        }
    }
    memcpy(dL_df0, lam, sizeof(double) * n);
    *dL_dg = dg;
    *dL_domega0 = dw;
    free(ck); free(y); free(sb); free(lam); free(nu); free(mb); free(seg);
}
