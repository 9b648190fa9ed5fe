// tg_adjoint.hip -- discrete adjoint of the collision operator and of DP5.
//
// Gradients of a loss L(f(t1)) with respect to the initial distribution f0,
// the coupling g and the phonon energy omega0, through the fixed-step DP5
// (the accepted step sequence of a forward run, all steps accepted, dense
// output at t1: DormandPrince54.hpp:565-611 with the controller frozen).
//
// Per stage with state Y and stage adjoint mu the reverse sweep needs
//   J(Y)^T mu = -m [(W Y) + (W^T 1) - (W^T Y)] + W^T x + W y,
//     m = mu / V,  x = m (1 - Y),  y = m Y,  W = (n+1) K + n K^T,
//   mu^T dF/dg      = (2/g) mu^T F(Y)            (K scales as g^2)
//   mu^T dF/domega0 = mu^T F[W'](Y),  W' = (n+1) K' + n K'^T + n' (K + K^T),
// (with hat occupation factors x = S Y replaces Y in the factors of the rows'
// nodes, and with the number-defect removal F = I0 - s w adds its own terms:
// vjp below),
// with K' = dK/domega0 (same windows, sampled G') and n' = dn/domega0 given
// by the caller.  Every product with K is one pass over K's band
// (tg_stack.hip), every product with K' one pass of the dual kernel over its
// band (K Y and K^T Y in one read either way).  With the e-e channel on, its operator
// joins the forward stages and its vector-Jacobian product joins J^T mu
// (tg_ee.hip); it depends on neither g nor omega0.  Forward states are
// recomputed from checkpoints every `ckpt` steps; only the step-start states
// of one segment and the seven stage states of one step are held at a time.
#include <cmath>
#include <cstring>
#include <vector>

#include "tg_internal.hpp"

namespace tg {

struct LinArgs {
    const double* p[7];
    double c[7];
    int np;
};

// out = base + sum_j c_j p_j (base may alias out)
__global__ void lincomb_kernel(int64_t n, const double* base, LinArgs a, double* out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v = base ? base[i] : 0.0;
        for (int j = 0; j < a.np; ++j) v += a.c[j] * a.p[j][i];
        out[i] = v;
    }
}

__global__ void vjp_prep_kernel(int64_t n, const double* __restrict__ mu,
                                const double* __restrict__ f, const double* __restrict__ V,
                                double* __restrict__ x, double* __restrict__ y) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = mu[i] / V[i];
    x[i] = m * (1.0 - f[i]);
    y[i] = m * f[i];
}

// with z (hat occupation factors) the first term goes to z (it reaches f
// through S^T) and out gets the rest.  With the number-defect removal (sc:
// s, c on the device) the term -s mu'_k dw_k/dx_k joins the first one (mu
// is then mu' = mu - c V, x the occupation factors).
__global__ void vjp_out_kernel(int64_t n, const double* __restrict__ mu,
                               const double* __restrict__ V, double nb,
                               const double* __restrict__ uf, const double* __restrict__ vf,
                               const double* __restrict__ rk, const double* __restrict__ ck,
                               const double* __restrict__ ux, const double* __restrict__ vx,
                               const double* __restrict__ uy, const double* __restrict__ vy,
                               double* __restrict__ out, double* __restrict__ z,
                               const double* __restrict__ sc, const double* __restrict__ x, bool hat) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double np1 = nb + 1.0;
    const double m = mu[k] / V[k];
    const double wf = np1 * uf[k] + nb * vf[k];
    const double wt1 = np1 * ck[k] + nb * rk[k];
    const double wtf = np1 * vf[k] + nb * uf[k];
    const double wtx = np1 * vx[k] + nb * ux[k];
    const double wy = np1 * uy[k] + nb * vy[k];
    double zk = -m * (wf + wt1 - wtf);
    if (sc) {
        double dw;
        if (hat) {
            dw = ((np1 * ck[k] + nb * rk[k]) - (np1 * rk[k] + nb * ck[k])) / V[k];
        } else {
            const double p = x[k] * (1.0 - x[k]);
            dw = p > 0.0 ? 1.0 - 2.0 * x[k] : (p < 0.0 ? 2.0 * x[k] - 1.0 : 0.0);
        }
        zk -= sc[0] * mu[k] * dw;
    }
    if (z) {
        z[k] = zk;
        out[k] = wtx + wy;
    } else {
        out[k] = zk + wtx + wy;
    }
}

// out_k += (S^T z)_k = sum_i S_ik z_i for the hat stencil S (the stencil is
// symmetric in its pattern and its multiplicities, S_ik = mult_ik / D_i with
// D_i the multiplicity sum of row i)
__global__ void hat_transpose_kernel(int64_t n, int dim, const uint32_t* __restrict__ hs,
                                     const double* __restrict__ z, double* __restrict__ out) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    out[k] += hat_value_dim(dim, hs, k, [&](int64_t i) { return z[i] / (double)hat_den(dim, hs, i); }) *
              (double)hat_den(dim, hs, k);
}

// occupation factors S f of one state (caller order)
__global__ void hat_apply_kernel(int64_t n, int dim, const uint32_t* __restrict__ hs, const double* __restrict__ f,
                                 double* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = hat_value_dim(dim, hs, i, [&](int64_t j) { return f[j]; });
}

constexpr int kAdjT = 256;

// The number-defect removal at a stage state: per-block partials of
// D = sum V I0, Z = sum V w and M = sum mu w (I0 the operator before the
// removal, w its weights: the hat form's, or |x(1-x)|)
__global__ void defect_partial_kernel(int64_t n, const double* __restrict__ mu, const double* __restrict__ x,
                                      const double* __restrict__ V, double nb, const double* __restrict__ uf,
                                      const double* __restrict__ vf, const double* __restrict__ rk,
                                      const double* __restrict__ ck, bool hat, double* __restrict__ part) {
    __shared__ double sh[3][kAdjT / 64];
    const int64_t i = blockIdx.x * (int64_t)kAdjT + threadIdx.x;
    double a = 0.0, b = 0.0, e = 0.0;
    if (i < n) {
        const double np1 = nb + 1.0, xi = x[i];
        const double I0 = ((1.0 - xi) * (np1 * uf[i] + nb * vf[i]) -
                           xi * (np1 * ck[i] + nb * rk[i] - np1 * vf[i] - nb * uf[i])) / V[i];
        const double w = hat ? hat_defect_weight(xi, np1, nb, rk[i], ck[i], V[i]) : fabs(xi * (1.0 - xi));
        a = V[i] * I0;
        b = V[i] * w;
        e = mu[i] * w;
    }
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_down(a, o, 64);
        b += __shfl_down(b, o, 64);
        e += __shfl_down(e, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = a;
        sh[1][threadIdx.x >> 6] = b;
        sh[2][threadIdx.x >> 6] = e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0, se = 0.0;
        for (int w = 0; w < kAdjT / 64; ++w) {
            sa += sh[0][w];
            sb += sh[1][w];
            se += sh[2][w];
        }
        part[3 * blockIdx.x] = sa;
        part[3 * blockIdx.x + 1] = sb;
        part[3 * blockIdx.x + 2] = se;
    }
}

// sc = {s, c}: s = D / Z (the removal's scale), c = M / Z (fixed order, one thread)
__global__ void defect_scale_kernel(const double* __restrict__ part, int nblk, double* __restrict__ sc) {
    double D = 0.0, Z = 0.0, M = 0.0;
    for (int k = 0; k < nblk; ++k) {
        D += part[3 * k];
        Z += part[3 * k + 1];
        M += part[3 * k + 2];
    }
    sc[0] = Z > 0.0 ? D / Z : 0.0;
    sc[1] = Z > 0.0 ? M / Z : 0.0;
}

// mu' = mu - c V
__global__ void mu_shift_kernel(int64_t n, const double* __restrict__ mu, const double* __restrict__ V,
                                const double* __restrict__ sc, double* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = mu[i] - sc[1] * V[i];
}

// per-block partials of mu^T F(Y) and mu^T F[W'](Y)
// With the hat form's number-defect removal (sc non-null) F = I0 - s w and
// mu is mu': the weights scale as g^2 and move with omega0 through K 1,
// K^T 1 and the bath (the nodal weights depend on neither).
__global__ void param_partial_kernel(int64_t n, const double* __restrict__ mu,
                                     const double* __restrict__ f, const double* __restrict__ V,
                                     double nb, double dnb, const double* __restrict__ uf,
                                     const double* __restrict__ vf, const double* __restrict__ rk,
                                     const double* __restrict__ ck, const double* __restrict__ udf,
                                     const double* __restrict__ vdf,
                                     const double* __restrict__ rkd,
                                     const double* __restrict__ ckd, double* __restrict__ part,
                                     const double* __restrict__ sc) {
    __shared__ double sh[2][kAdjT / 64];
    const int64_t i = blockIdx.x * (int64_t)kAdjT + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (i < n) {
        const double np1 = nb + 1.0, fi = f[i];
        const double Fi = ((1.0 - fi) * (np1 * uf[i] + nb * vf[i]) -
                           fi * (np1 * ck[i] + nb * rk[i] - np1 * vf[i] - nb * uf[i])) / V[i];
        const double Fd = ((1.0 - fi) * (np1 * udf[i] + nb * vdf[i]) -
                           fi * (np1 * ckd[i] + nb * rkd[i] - np1 * vdf[i] - nb * udf[i])) / V[i];
        const double Fs = ((1.0 - fi) * (dnb * (uf[i] + vf[i])) -
                           fi * (dnb * (ck[i] + rk[i]) - dnb * (vf[i] + uf[i]))) / V[i];
        a = mu[i] * Fi;
        b = mu[i] * (Fd + Fs);
        if (sc) {
            const double w = hat_defect_weight(fi, np1, nb, rk[i], ck[i], V[i]);
            const double dA = np1 * rkd[i] + nb * ckd[i] + dnb * (rk[i] + ck[i]);
            const double dB = np1 * ckd[i] + nb * rkd[i] + dnb * (ck[i] + rk[i]);
            a -= sc[0] * mu[i] * w;
            b -= sc[0] * mu[i] * ((1.0 - fi) * dA + fi * dB) / V[i];
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_down(a, o, 64);
        b += __shfl_down(b, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = a;
        sh[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
        for (int w = 0; w < kAdjT / 64; ++w) {
            sa += sh[0][w];
            sb += sh[1][w];
        }
        part[2 * blockIdx.x] = sa;
        part[2 * blockIdx.x + 1] = sb;
    }
}

// acc[0] += (2/g) sum_a, acc[1] += sum_b (fixed order, one thread)
__global__ void param_accumulate_kernel(const double* __restrict__ part, int nblk, double two_over_g,
                                        double* __restrict__ acc) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < nblk; ++k) {
        a += part[2 * k];
        b += part[2 * k + 1];
    }
    acc[0] += two_over_g * a;
    acc[1] += b;
}

struct AdjWork {
    int64_t n = 0;
    int32_t ckpt = 0, nck = 0;
    double* Y[7] = {};
    double* k[7] = {};
    double* mu[7] = {};
    double *nu = nullptr, *lam = nullptr, *x = nullptr, *y = nullptr;
    double *uf = nullptr, *vf = nullptr, *ux = nullptr, *vx = nullptr, *uy = nullptr,
           *vy = nullptr, *udf = nullptr, *vdf = nullptr;
    double *xbar = nullptr, *z = nullptr;   // hat occupation factors of a stage state, S^T's input
    double* mup = nullptr;    // mu - c V (number-defect removal)
    double* dpart = nullptr;  // defect partials, 3 per block
    double* sc = nullptr;     // s, c of the removal
    double* seg = nullptr;    // (ckpt+1) x n step-start states of a segment
    double* ckp = nullptr;    // nck x n checkpoints
    double* part = nullptr;
    double* acc = nullptr;
};

void adjoint_destroy(tg_ctx* c) {
    AdjWork* w = (AdjWork*)c->adj;
    if (!w) return;
    for (int s = 0; s < 7; ++s) {
        dev_free(&w->Y[s]);
        dev_free(&w->k[s]);
        dev_free(&w->mu[s]);
    }
    for (double** p : {&w->nu, &w->lam, &w->x, &w->y, &w->uf, &w->vf, &w->ux, &w->vx, &w->uy,
                       &w->vy, &w->udf, &w->vdf, &w->xbar, &w->z, &w->mup, &w->seg, &w->ckp, &w->part,
                       &w->acc, &w->dpart, &w->sc})
        dev_free(p);
    delete w;
    c->adj = nullptr;
}

static tg_status adj_ensure(tg_ctx* c, int32_t ckpt, int32_t nck) {
    AdjWork* w = (AdjWork*)c->adj;
    if (w && w->ckpt >= ckpt && w->nck >= nck) return TG_OK;
    adjoint_destroy(c);
    w = new AdjWork();
    c->adj = w;
    w->n = c->n;
    w->ckpt = ckpt;
    w->nck = nck;
    const size_t n = (size_t)c->n;
    for (int s = 0; s < 7; ++s) {
        TG_TRY(dev_alloc(c, &w->Y[s], n));
        TG_TRY(dev_alloc(c, &w->k[s], n));
        TG_TRY(dev_alloc(c, &w->mu[s], n));
    }
    for (double** p : {&w->nu, &w->lam, &w->x, &w->y, &w->uf, &w->vf, &w->ux, &w->vx, &w->uy,
                       &w->vy, &w->udf, &w->vdf, &w->xbar, &w->z, &w->mup})
        TG_TRY(dev_alloc(c, p, n));
    TG_TRY(dev_alloc(c, &w->dpart, 3 * (size_t)ceil_div(c->n, kAdjT)));
    TG_TRY(dev_alloc(c, &w->sc, 2));
    TG_TRY(dev_alloc(c, &w->seg, n * (ckpt + 1)));
    TG_TRY(dev_alloc(c, &w->ckp, n * nck));
    TG_TRY(dev_alloc(c, &w->part, 2 * (size_t)ceil_div(c->n, kAdjT)));
    TG_TRY(dev_alloc(c, &w->acc, 2));
    return TG_OK;
}

// Butcher rows (DormandPrince54.hpp:503-511); row 6 is b (y5)
static const double kA[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {1. / 5., 0, 0, 0, 0, 0},
    {3. / 40., 9. / 40., 0, 0, 0, 0},
    {44. / 45., -56. / 15., 32. / 9., 0, 0, 0},
    {19372. / 6561., -25360. / 2187., 64448. / 6561., -212. / 729., 0, 0},
    {9017. / 3168., -355. / 33., 46732. / 5247., 49. / 176., -5103. / 18656., 0},
    {35. / 384., 0., 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84.}};

struct AdjRun {
    tg_ctx* c;
    AdjWork* w;
    double nb, dnb, g;
    int conserve;
    unsigned grid() const { return ceil_div(c->n, 256); }

    tg_status lincomb(const double* base, std::initializer_list<std::pair<double, const double*>> t,
                      double* out) {
        LinArgs a{};
        a.np = 0;
        for (auto& q : t) {
            a.c[a.np] = q.first;
            a.p[a.np] = q.second;
            ++a.np;
        }
        hipLaunchKernelGGL(lincomb_kernel, dim3(grid()), dim3(256), 0, c->stream, c->n, base, a, out);
        TG_CHECK_LAUNCH(c);
        return TG_OK;
    }
    // stage states Y[0..last] and k[0..last] of the step from y with size h
    tg_status stages(const double* y, double h, int last) {
        for (int s = 0; s <= last; ++s) {
            LinArgs a{};
            a.np = s;
            for (int j = 0; j < s; ++j) {
                a.c[j] = kA[s][j] * h;
                a.p[j] = w->k[j];
            }
            hipLaunchKernelGGL(lincomb_kernel, dim3(grid()), dim3(256), 0, c->stream, c->n, y, a,
                               w->Y[s]);
            TG_CHECK_LAUNCH(c);
            TG_TRY(plain_apply_device(c, w->Y[s], w->k[s], nb, conserve));
            if (c->ee) TG_TRY(ee_apply_device(c, w->Y[s], w->k[s], 1));
        }
        return TG_OK;
    }
    // nu = J(Y)^T mu, accumulate the parameter products
    // With hat occupation factors xbar = S Y replaces Y in the factors of
    // the rows' nodes, and -m (W Y + W^T 1 - W^T Y) reaches Y through S^T.
    tg_status vjp(const double* Yst, const double* mu) {
        const int64_t n = c->n;
        const bool hat = c->occ == TG_OCC_HAT;
        const double* occ = Yst;
        if (hat) {
            hipLaunchKernelGGL(hat_apply_kernel, dim3(grid()), dim3(256), 0, c->stream, n, c->dim, c->d_hat, Yst,
                               w->xbar);
            TG_CHECK_LAUNCH(c);
            occ = w->xbar;
        }
        TG_TRY(stack_products_device(c, Yst, w->uf, w->vf));
        // with the number-defect removal I = I0 - s w:
        //   mu^T dI = mu'^T dI0 - s mu'^T dw,  mu' = mu - c V,  c = mu^T w / sum V w
        const double* sc = nullptr;
        if (conserve) {
            const int nblk = (int)ceil_div(n, kAdjT);
            hipLaunchKernelGGL(defect_partial_kernel, dim3(nblk), dim3(kAdjT), 0, c->stream, n, mu, occ, c->d_V, nb,
                               w->uf, w->vf, c->d_rk, c->d_ck, hat, w->dpart);
            TG_CHECK_LAUNCH(c);
            hipLaunchKernelGGL(defect_scale_kernel, dim3(1), dim3(1), 0, c->stream, w->dpart, nblk, w->sc);
            TG_CHECK_LAUNCH(c);
            hipLaunchKernelGGL(mu_shift_kernel, dim3(grid()), dim3(256), 0, c->stream, n, mu, c->d_V, w->sc,
                               w->mup);
            TG_CHECK_LAUNCH(c);
            mu = w->mup;
            sc = w->sc;
        }
        hipLaunchKernelGGL(vjp_prep_kernel, dim3(grid()), dim3(256), 0, c->stream, n, mu, occ,
                           c->d_V, w->x, w->y);
        TG_CHECK_LAUNCH(c);
        TG_TRY(stack_products_device(c, w->x, w->ux, w->vx));
        TG_TRY(stack_products_device(c, w->y, w->uy, w->vy));
        hipLaunchKernelGGL(vjp_out_kernel, dim3(grid()), dim3(256), 0, c->stream, n, mu, c->d_V, nb,
                           w->uf, w->vf, c->d_rk, c->d_ck, w->ux, w->vx, w->uy, w->vy, w->nu,
                           hat ? w->z : nullptr, sc, occ, hat);
        TG_CHECK_LAUNCH(c);
        if (hat) {
            hipLaunchKernelGGL(hat_transpose_kernel, dim3(grid()), dim3(256), 0, c->stream, n, c->dim, c->d_hat,
                               w->z, w->nu);
            TG_CHECK_LAUNCH(c);
        }
        // the e-e channel adds its own product (it depends on neither g nor omega0)
        if (c->ee) TG_TRY(ee_vjp_device(c, Yst, mu, w->nu));
        TG_TRY(stack_products_device(c, Yst, w->udf, w->vdf, true));
        const int nblk = (int)ceil_div(n, kAdjT);
        hipLaunchKernelGGL(param_partial_kernel, dim3(nblk), dim3(kAdjT), 0, c->stream, n, mu, occ,
                           c->d_V, nb, dnb, w->uf, w->vf, c->d_rk, c->d_ck, w->udf, w->vdf,
                           c->d_rkd, c->d_ckd, w->part, hat ? sc : nullptr);
        TG_CHECK_LAUNCH(c);
        hipLaunchKernelGGL(param_accumulate_kernel, dim3(1), dim3(1), 0, c->stream, w->part, nblk,
                           2.0 / g, w->acc);
        TG_CHECK_LAUNCH(c);
        return TG_OK;
    }
};

}  // namespace tg

using namespace tg;

extern "C" tg_status tg_dp5_adjoint(tg_ctx* c, const double* f0, int64_t n, double t0, double t1,
                                    double dense_dt, const double* h_steps, int32_t n_steps,
                                    int32_t ckpt, double bath_n, double dbath_domega0,
                                    const double* dL_df, double* dL_df0, double* dL_dparams, int32_t flags) {
    if (!c || !f0 || !h_steps || !dL_df || !dL_df0 || !dL_dparams) return TG_ERR_ARG;
    if (n != c->n) return set_error(c, TG_ERR_ARG, "n mismatch");
    if (n_steps < 1 || ckpt < 1 || !(dense_dt > 0.0)) return set_error(c, TG_ERR_ARG, "bad steps");
    if (!(c->g != 0.0)) return set_error(c, TG_ERR_ARG, "g must be nonzero for dL/dg");
    if (!c->assembled) return set_error(c, TG_ERR_STATE, "tg_kernel_assemble not called");
    TG_TRY(assemble_domega(c));
    const int32_t nck = n_steps / ckpt + 2;
    TG_TRY(adj_ensure(c, ckpt, nck));
    AdjWork* w = (AdjWork*)c->adj;
    if (flags & ~TG_CONSERVE) return set_error(c, TG_ERR_ARG, "tg_dp5_adjoint takes TG_CONSERVE only");
    AdjRun R{c, w, bath_n, dbath_domega0, c->g, (flags & TG_CONSERVE) ? 1 : 0};
    hipStream_t st = c->stream;
    const size_t vb = sizeof(double) * n;

    // forward: checkpoints and the step holding the last dense output
    TG_HIP(c, hipMemcpyAsync(w->seg, f0, vb, hipMemcpyHostToDevice, st));
    double t = t0, dense_time = t0 + dense_dt, s_last = 0.0;
    int64_t m_last = -1;
    for (int64_t m = 0; m < n_steps && t < t1; ++m) {
        if (m % ckpt == 0)
            TG_HIP(c, hipMemcpyAsync(w->ckp + (m / ckpt) * n, w->seg, vb, hipMemcpyDeviceToDevice, st));
        const double h = h_steps[m];
        TG_TRY(R.stages(w->seg, h, 5));
        TG_TRY(R.lincomb(w->seg, {{kA[6][0] * h, w->k[0]}, {kA[6][2] * h, w->k[2]},
                                  {kA[6][3] * h, w->k[3]}, {kA[6][4] * h, w->k[4]},
                                  {kA[6][5] * h, w->k[5]}}, w->seg));
        while (dense_time <= t + 1.0001 * h && dense_time <= t1 + .0001 * h) {
            m_last = m;
            s_last = (dense_time - t) / h;
            dense_time += dense_dt;
        }
        t += h;
    }
    if (m_last < 0) return set_error(c, TG_ERR_ARG, "no dense output at t1 on the given steps");

    TG_HIP(c, hipMemsetAsync(w->acc, 0, 2 * sizeof(double), st));
    for (int64_t c0 = (m_last / ckpt) * ckpt; c0 >= 0; c0 -= ckpt) {
        const int64_t c1 = std::min<int64_t>(c0 + ckpt - 1, m_last);
        TG_HIP(c, hipMemcpyAsync(w->seg, w->ckp + (c0 / ckpt) * n, vb, hipMemcpyDeviceToDevice, st));
        for (int64_t m = c0; m < c1; ++m) {
            const double h = h_steps[m];
            double* ym = w->seg + (m - c0) * n;
            TG_TRY(R.stages(ym, h, 5));
            TG_TRY(R.lincomb(ym, {{kA[6][0] * h, w->k[0]}, {kA[6][2] * h, w->k[2]},
                                  {kA[6][3] * h, w->k[3]}, {kA[6][4] * h, w->k[4]},
                                  {kA[6][5] * h, w->k[5]}}, ym + n));
        }
        for (int64_t m = c1; m >= c0; --m) {
            const double h = h_steps[m];
            const double* ym = w->seg + (m - c0) * n;
            const bool last = m == m_last;
            TG_TRY(R.stages(ym, h, last ? 6 : 5));
            if (last) {
                // dense output y + sum_j (s h btilde_j) k_j (DormandPrince54.hpp:601-608)
                const double s = s_last;
                double bt[7];
                bt[0] = -(435. * s * s * s - 1184. * s * s + 1098. * s - 384.) / (384.);
                bt[1] = 0.0;
                bt[2] = 500. * s * (6. * s * s - 14. * s + 9.) / (1113.);
                bt[3] = -125. * s * (9. * s * s - 16. * s + 6.) / (192.);
                bt[4] = 729. * s * (35. * s * s - 64. * s + 26.) / (6784.);
                bt[5] = -11. * s * (3. * s - 2.) * (5. * s - 6.) / (84.);
                bt[6] = s * (s - 1.) * (5. * s - 3.) / (2.);
                TG_HIP(c, hipMemcpyAsync(w->lam, dL_df, vb, hipMemcpyHostToDevice, st));
                for (int j = 0; j < 7; ++j) TG_TRY(R.lincomb(nullptr, {{s * h * bt[j], w->lam}}, w->mu[j]));
                // k7 = F(Y7), Y7 = y + sum_i (h b_i) k_i
                TG_TRY(R.vjp(w->Y[6], w->mu[6]));
                TG_TRY(R.lincomb(w->lam, {{1.0, w->nu}}, w->lam));
                for (int j = 0; j < 6; ++j)
                    TG_TRY(R.lincomb(w->mu[j], {{kA[6][j] * h, w->nu}}, w->mu[j]));
            } else {
                for (int j = 0; j < 6; ++j) TG_TRY(R.lincomb(nullptr, {{kA[6][j] * h, w->lam}}, w->mu[j]));
            }
            for (int sidx = 5; sidx >= 1; --sidx) {
                TG_TRY(R.vjp(w->Y[sidx], w->mu[sidx]));
                TG_TRY(R.lincomb(w->lam, {{1.0, w->nu}}, w->lam));
                for (int j = 0; j < sidx; ++j)
                    TG_TRY(R.lincomb(w->mu[j], {{kA[sidx][j] * h, w->nu}}, w->mu[j]));
            }
            TG_TRY(R.vjp(w->Y[0], w->mu[0]));
            TG_TRY(R.lincomb(w->lam, {{1.0, w->nu}}, w->lam));
        }
    }
    TG_HIP(c, hipMemcpyAsync(dL_df0, w->lam, vb, hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipMemcpyAsync(dL_dparams, w->acc, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipStreamSynchronize(st));
    return TG_OK;
}
