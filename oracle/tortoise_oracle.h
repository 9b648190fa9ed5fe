/*
 * tortoise_oracle.h -- CPU FP64 restatement of the scattering-integral hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code: it is the checker that tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py compare the HIP path against.
 *
 * What it restates (reference paths relative to the TORTOISE tree):
 *   - structured simplex mesh, section/element numbering and vertex order:
 *       src/Geometry/Structured/DirectSpace/Mesh.cpp:55-87 (vertElemInRefSec),
 *       :222-267 (verticesCoord), :384-433 (section/element ids)
 *   - element volume and lumped P1 masses: Mesh.cpp:268-293 (elemVolume),
 *       Function.cpp:369-372 (integrate() of a P1 field)
 *   - error norm of the adaptive stepper: src/Physics/Material/Band.cpp:87-101
 *       with Function::integrate(rhs) (Function.cpp:373-375) done exactly
 *       (order-2 Gauss is exact for P1*P1)
 *   - max over bands: src/Physics/Status/MaterialStatus.cpp:189-196
 *   - Dormand-Prince 5(4) driver with PI control, reject rule, dense output:
 *       include/Generics/Algorithms/DormandPrince54.hpp:479-665
 *   - channel amplitude convention (amplitude = g^2, full momentum measure,
 *       no 2*pi): src/Physics/Scattering/ScatteringChannel.cpp:399-415
 *
 * What is new (no reference function computes it; pinned by closed forms and
 * invariants in tests/test_oracle_*.py):
 *   - nodal P1 W-form of the e-ph collision operator with a Gaussian-broadened
 *     energy delta and a phonon bath of occupation n:
 *       I_i = [ (1-f_i) (W f)_i - f_i (W^T (1-f))_i ] / V_i,
 *       W   = (n+1) K + n K^T,
 *       K_ij = g^2 sum_{b,b'} R_ib R_jb' G_sigma((b-b') dE + omega0)
 *     where R_ib = int phi_i(k) h_b(eps(k)) dk is the exact projection of the
 *     P1 basis function phi_i (eps linear per element) on the hat functions
 *     h_b of a uniform energy grid, computed with the simplex-spline identity
 *     (basis-weighted energy density of a linear simplex = a B-spline with the
 *     vertex energies plus the basis vertex's energy as knots).
 */
#ifndef TORTOISE_ORACLE_H
#define TORTOISE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- mesh */
int64_t or_num_nodes(int dim, const int32_t* nsec);
int64_t or_num_elements(int dim, const int32_t* nsec);
double  or_element_volume(int dim, const int32_t* nsec, const double* extent);
/* node coordinates, node-major: coords[i*dim + d] */
void    or_mesh_nodes(int dim, const int32_t* nsec, const double* origin,
                      const double* extent, double* coords);
/* element -> node connectivity in the reference's vertex order, conn[e*(dim+1)+a] */
void    or_mesh_elements(int dim, const int32_t* nsec, int32_t* conn);
/* lumped P1 mass: V_i = sum_{e ni i} vol_e/(dim+1) */
void    or_node_volumes(int dim, const int32_t* nsec, const double* extent, double* V);

/* ------------------------------------------------------ energy projection */
typedef struct or_egrid {
    double  e0;      /* energy of grid point 0 */
    double  de;      /* grid spacing */
    int32_t nbins;   /* number of hat functions */
} or_egrid;

void or_egrid_make(const double* eps, int64_t n, double de, or_egrid* g);

/* Projection windows: node i owns hats [off[i], off[i]+len[i]).  Returns the
 * largest len (call with R == NULL to size the buffer).  R is n x maxw, row i
 * holds R_i,off[i]+w at column w. */
int32_t or_projection(int dim, const int32_t* nsec, const double* extent,
                      const double* eps, const or_egrid* g,
                      int32_t* off, int32_t* len, double* R, int32_t maxw);

/* Basis-weighted energy density of one simplex, B-spline form:
 * D_a(x) = vol * N(x | sorted{e_0..e_dim, e_a}) / (t_last - t_first). */
double or_simplex_density(int dim, const double* evert, int a, double vol, double x);

/* Sampled broadened delta: g_m = G_sigma(m*de + shift) for m in [m_lo, m_hi]
 * with |m*de+shift| <= cutoff*sigma.  Returns count, fills m_lo. */
int32_t or_gauss_table(double de, double shift, double sigma, double cutoff,
                       int32_t* m_lo, double* gm, int32_t cap);

/* Emission kernel K (row-major n x n) or a single row of it. */
void or_kernel_full(int64_t n, const int32_t* off, const int32_t* len,
                    const double* R, int32_t maxw, const double* gm,
                    int32_t m_lo, int32_t nm, double amp, double* K);
void or_kernel_row(int64_t n, int64_t i, const int32_t* off, const int32_t* len,
                   const double* R, int32_t maxw, const double* gm,
                   int32_t m_lo, int32_t nm, double amp, double* Krow);

/* ----------------------------------------------- hat-averaged occupation */
/* fbar_i = int phi_i f / int phi_i of the P1 interpolant (consistent mass over
 * lumped mass) as a CSR stencil on the structured mesh: fbar = S f.  New:
 * the reference integrates the P1 products exactly per element tuple
 * (IntegratorCore.hpp:298-316); the W-form takes the test function's
 * occupation factor as this average instead of the nodal value. */
typedef struct or_hat {
    int64_t n;
    int64_t* ptr;
    int32_t* col;
    double* w;
} or_hat;
or_hat* or_hat_create(int dim, const int32_t* nsec);
void    or_hat_free(or_hat* h);
void    or_hat_apply(const or_hat* h, const double* f, double* fbar);   /* S f */
void    or_hat_apply_t(const or_hat* h, const double* z, double* out);  /* S^T z */
/* copies the stencil out (NULL pointers: count only); returns its nonzeros */
int64_t or_hat_stencil(int dim, const int32_t* nsec, int64_t* ptr, int32_t* col, double* w);

/* Number-defect weight of the hat form: the gross rate node i would have at
 * its own occupation x if every other node were half filled,
 * ((1-x) (W 1)_i + x (W^T 1)_i) / V_i with W = (n+1) K + n K^T,
 * rk = K 1, ck = K^T 1 (the nodal form keeps |f(1-f)|). */
double  or_hat_defect_weight(double x, double np1, double nb, double rk, double ck, double V);

/* bits of the `conserve` argument of the DP5 wrappers below */
#define OR_CONSERVE 1   /* number-defect removal */
#define OR_OCC_HAT  2   /* hat-averaged occupation factors */

/* ---------------------------------------------------------- collision */
/* I = [(1-f) (W f) - f (W^T 1 - W^T f)] / V,  W = (n+1)K + n K^T.
 * K row-major n x n.  If conserve != 0 the number defect sum_i V_i I_i is
 * removed with weights |f_i(1-f_i)| (the same rule the GPU paths use). */
void or_collision_apply(int64_t n, const double* K, const double* V,
                        double bath_n, const double* f, double* I, int conserve);
/* The same with the occupation factors of node i (1 - x_i of the gain, x_i of
 * the loss) x = S f when hat != NULL; the number-defect removal then weighs
 * node i with or_hat_defect_weight. */
void or_collision_apply_occ(int64_t n, const double* K, const double* V, double bath_n,
                            const double* f, const or_hat* hat, double* I, int conserve);

/* or_collision_apply from K's band in energy order (the CPU form that reads
 * the same bytes as the GPU's plain evaluation): row p (node perm[p]) holds
 * columns [lo[p], lo[p] + rowptr[p+1] - rowptr[p]) at vals + rowptr[p]; rk,
 * ck, Vs in energy order; scratch (nthreads + 2) * n doubles. */
void or_collision_apply_band(int64_t n, const int64_t* rowptr, const int32_t* lo, const double* vals,
                             const double* rk, const double* ck, const double* Vs, const int32_t* perm,
                             double bath_n, const double* f, double* I, double* scratch, int nthreads);
/* ... with hat occupation factors (hat NULL: nodal) and the number-defect
 * removal; scratch (nthreads + 3) * n doubles */
void or_collision_apply_band_occ(int64_t n, const int64_t* rowptr, const int32_t* lo, const double* vals,
                                 const double* rk, const double* ck, const double* Vs, const int32_t* perm,
                                 double bath_n, const double* f, const or_hat* hat, int conserve, double* I,
                                 double* scratch, int nthreads);

/* sum_i V_i f_i and sum_i V_i eps_i f_i */
void or_conservation(int64_t n, const double* V, const double* eps,
                     const double* f, double* number, double* energy);

/* ------------------------------------------------------- error norm */
/* Band::relativeError for a fermion (stat=-1) or boson (stat=+1) band on
 * nodal data: sqrt((int r^2 + p int pen^2)/(Nel*vol)). */
double or_relative_error(int dim, int64_t nel, const int32_t* conn,
                         const double* err, const double* pop, double stat,
                         double rtol, double atol, double penalty);

/* ------------------------------------------------------- DP5(4) driver */
typedef void   (*or_rhs_fn)(void* user, double t, const double* y, double* dydt);
typedef double (*or_norm_fn)(void* user, const double* err, const double* y5);

typedef struct or_dp5_stats {
    int64_t attempted;
    int64_t accepted;
    int64_t rejected;
    int64_t evals;
    int32_t failed;       /* 1 if the step size underflowed */
    int32_t n_dense;      /* dense outputs written */
} or_dp5_stats;

/* Time-dependent bath of a constrained phonon band (uniform population):
 * samples s[0..ns) at t_start + k*dt, linear between them, held outside
 * (Band::constrainedPopulation, Band.hpp:70-74). */
typedef struct or_bath_sched {
    const double* s;
    int32_t ns;
    double t_start, dt;
    double atol, rtol;      /* the constrained band's error tolerances (vetoer) */
} or_bath_sched;

double or_sched_value(const double* s, int32_t ns, double t_start, double dt, double t);
/* MaterialTimeStatus::timeStepVetoer (MaterialTimeStatus.cpp:60-100) with
 * timesToTest = {t, t+c2 h, t+c3 h, t+c4 h, t+c5 h, t+h}; bs NULL: no
 * constrained band (returns (t+h) - t). */
double or_step_vetoer(const or_bath_sched* bs, double t, double h, double dense_dt);

/* Electron energy flux functional of a derivative (bath energy ledger). */
typedef double (*or_flux_fn)(void* user, const double* dydt);

/* or_dp5_integrate with the step vetoer of a constrained bath (bs, NULL for
 * none: the vetoer then leaves h at (t+h)-t as the reference's does) and,
 * when flux != NULL, the bath energy ledger: *ledger -= sum_j (h b_j) flux(k_j)
 * over the accepted steps. */
void or_dp5_integrate_ex(int64_t n, double* y, double t0, double t_final,
                         double dense_dt, or_rhs_fn rhs, or_norm_fn norm, void* user,
                         const or_bath_sched* bs, or_flux_fn flux, double* ledger,
                         double* dense_out, double* dense_t, int32_t max_dense,
                         double* h_log, int8_t* acc_log, int64_t log_cap,
                         or_dp5_stats* stats);

/* Integrates y' = rhs(t,y) from t0 to t_final.  Dense outputs at t0+k*dense_dt
 * (k>=1) are written to dense_out[(k-1)*n ...] (capacity max_dense) with
 * their times in dense_t.  y is overwritten with the last accepted y5.
 * If h_log != NULL the attempted step sizes are written to h_log[0..cap) and
 * acc_log[k] = 1 for accepted attempts. */
void or_dp5_integrate(int64_t n, double* y, double t0, double t_final,
                      double dense_dt, or_rhs_fn rhs, or_norm_fn norm, void* user,
                      double* dense_out, double* dense_t, int32_t max_dense,
                      double* h_log, int8_t* acc_log, int64_t log_cap,
                      or_dp5_stats* stats);

/* DP5 on the nodal W-form: rhs = or_collision_apply, norm = or_relative_error
 * (fermion).  Convenience wrapper used by the parity tests. */
void or_dp5_collision(int dim, const int32_t* nsec, int64_t n, const double* K,
                      const double* V, double bath_n, double* f, double t0,
                      double t_final, double dense_dt, double rtol, double atol,
                      int conserve, double* dense_out, double* dense_t,
                      int32_t max_dense, double* h_log, int8_t* acc_log,
                      int64_t log_cap, or_dp5_stats* stats);

/* or_dp5_collision with a time-dependent bath (sched NULL: constant bath_n;
 * else ns samples, t_start, sdt, and the band's satol, srtol) and the bath
 * energy ledger (energy received by the bath, with eps the nodal energies:
 * -sum_j (h b_j) sum_i V_i eps_i k_j,i over the accepted steps). */
void or_dp5_collision_ex(int dim, const int32_t* nsec, int64_t n, const double* K,
                         const double* V, const double* eps, double bath_n,
                         const double* sched, int32_t ns, double t_start, double sdt,
                         double satol, double srtol, double* f, double t0,
                         double t_final, double dense_dt, double rtol, double atol,
                         int conserve, double* dense_out, double* dense_t,
                         int32_t max_dense, double* h_log, int8_t* acc_log,
                         int64_t log_cap, or_dp5_stats* stats, double* ledger);

/* Fixed-step replay of a recorded attempt sequence (h_log/acc_log): the same
 * stage arithmetic as or_dp5_integrate with the controller taken out, so a
 * second precision can be compared on identical steps.  Writes the state
 * after the last accepted step to y and dense outputs like or_dp5_integrate. */
void or_dp5_collision_replay(int dim, const int32_t* nsec, int64_t n,
                             const double* K, const double* V, double bath_n,
                             double* f, double t0, double t_final,
                             double dense_dt, int conserve, const double* h_log,
                             const int8_t* acc_log, int64_t n_attempts,
                             double* dense_out, double* dense_t,
                             int32_t max_dense);

/* Fixed-step DP5 over the accepted attempts of (h_log, acc_log) with the
 * same stage arithmetic and dense output as or_dp5_integrate_ex; with flux
 * the bath energy ledger of those steps. */
void or_dp5_replay_ex(int64_t n, double* f, double t0, double t_final, double dense_dt,
                      or_rhs_fn rhs, void* user, or_flux_fn flux, double* ledger,
                      const double* h_log, const int8_t* acc_log, int64_t n_attempts,
                      double* dense_out, double* dense_t, int32_t max_dense);

/* or_dp5_collision_replay with a time-dependent bath (sched NULL: bath_n)
 * and, with eps and ledger, the bath energy ledger of the replayed steps. */
void or_dp5_collision_replay_ex(int dim, const int32_t* nsec, int64_t n,
                                const double* K, const double* V, const double* eps,
                                double bath_n, const double* sched, int32_t ns,
                                double t_start, double sdt, double* f, double t0,
                                double t_final, double dense_dt, int conserve,
                                const double* h_log, const int8_t* acc_log,
                                int64_t n_attempts, double* dense_out, double* dense_t,
                                int32_t max_dense, double* ledger);

/* ------------------------------------------------------------ adjoint */
/* Derivative of K with respect to omega0 (the sampled G' table, same
 * windows): Kd_ij = g^2 sum R_ib R_jb' G'_sigma((b-b') dE + omega0). */
void or_kernel_domega_full(int64_t n, const int32_t* off, const int32_t* len,
                           const double* R, int32_t maxw, double de, double omega0,
                           double sigma, double cutoff, double amp, double* Kd);

/* Collision operator with a general W = alpha K + beta K^T (no correction). */
void or_collision_general(int64_t n, const double* K, const double* V, double alpha,
                          double beta, const double* f, double* I);
void or_collision_general_occ(int64_t n, const double* K, const double* V, double alpha,
                              double beta, const double* f, const or_hat* hat, double* I);

/* Vector-Jacobian product of the operator at f: out = J(f)^T mu. */
void or_collision_vjp(int64_t n, const double* K, const double* V, double bath_n,
                      const double* f, const double* mu, double* out);
void or_collision_vjp_occ(int64_t n, const double* K, const double* V, double bath_n,
                          const double* f, const or_hat* hat, const double* mu, double* out);
/* ... of the operator with the number-defect removal when conserve != 0
 * (I = I0 - s w, s = sum V I0 / sum V w; the weights of
 * or_collision_apply_occ) */
void or_collision_vjp_ex(int64_t n, const double* K, const double* V, double bath_n, const double* f,
                         const or_hat* hat, int conserve, const double* mu, double* out);

/* Discrete adjoint of the fixed-step DP5 (same step sequence as
 * or_dp5_collision_replay, all accepted) with dense output at t_final:
 * given dL/df(t_final) returns dL/df0, dL/dg and dL/domega0.  K scales as g^2;
 * Kd = dK/domega0; dbath = d(bath_n)/domega0.  Forward states are recomputed
 * from checkpoints every `ckpt` steps. */
void or_dp5_adjoint(int64_t n, const double* K, const double* Kd, const double* V,
                    double bath_n, double dbath, double g, const double* f0, double t0,
                    double t_final, double dense_dt, const double* h_steps, int64_t n_steps,
                    int32_t ckpt, const double* dL_df, double* dL_df0, double* dL_dg,
                    double* dL_domega0);
/* ... with hat occupation factors (hat NULL: nodal) and, with conserve, the
 * number-defect removal in every stage evaluation */
void or_dp5_adjoint_occ(int64_t n, const double* K, const double* Kd, const double* V,
                        double bath_n, double dbath, double g, const or_hat* hat, int conserve,
                        const double* f0, double t0, double t_final, double dense_dt,
                        const double* h_steps, int64_t n_steps, int32_t ckpt, const double* dL_df,
                        double* dL_df0, double* dL_dg, double* dL_domega0);

/* ------------------------------------------------------- e-e channel */
/* Electron-electron (4-leg) operator on the periodic lattice of the mesh box
 * (ee_oracle.c): A(q) = g^2/(|q|^2+kappa2)^2, Gaussian energy delta of width
 * sigma cut at cutoff*sigma, projected on the grid m*dw, |m| <= mh
 * (or_ee_grid).  I is written for every mesh node (duplicates of a periodic
 * point share its value).  or_ee_apply_direct: the unprojected delta by
 * brute force (O(P^3)). */
void or_ee_grid(int dim, const int32_t* nsec, const double* eps, double dw, int32_t* mh);
void or_ee_apply(int dim, const int32_t* nsec, const double* extent, const double* eps,
                 const double* f, double g, double kappa2, double sigma, double dw,
                 double cutoff, double* I);
void or_ee_apply_direct(int dim, const int32_t* nsec, const double* extent, const double* eps,
                        const double* f, double g, double kappa2, double sigma, double cutoff,
                        double* I);


/* ------------------------------------- reference tuple operator (mc_oracle.c) */
/* Rows of a channel's MC point matrix: max(dim*(nlegs-1)-1, 2*dim). */
int32_t or_mc_rows(int dim, int nlegs);
/* MC points [rows][nmc] of a channel from a splitmix64 stream. */
void    or_mc_points(int dim, int nlegs, int32_t nmc, uint64_t seed, double* pts);
/* Inverse (row-major, n <= 4) by Gauss-Jordan with partial pivoting; returns det. */
double  or_mc_invert(int n, const double* A, double* inv);
/* scatteringIntegrationTypeB with the standard output forms; out nlegs*(dim+1). */
int     or_mc_integrate(int dim, int nlegs, const double* D, const double* N,
                        const double* fa, const double* fb, const double* pts,
                        int32_t nmc, double* out);
/* The tuple-operator functions below take the zone as a geometry block
 * geo[21]: origin (3), the mesh sides (side c at geo[3 + 3c]; node index i_c
 * along side c at origin + sum_c (i_c / nsec[c]) side c), the zone sides
 * (zone side c at geo[12 + 3c]; umklapp vectors sum_c u_c zone side c). */
/* Tuples (e0, e1, e2, umklapp id) of the e-ph channel {el in, pn in, el out}. */
int64_t or_mc_tuples_eph(int dim, const int32_t* nsec, const double* geo, const double* eps, const double* eps_ph,
                         double omega0, int32_t nu, const double* pts, int32_t nmc,
                         double min_det, double threshold, int32_t* idx, int64_t cap);
/* LF-basis inverse mass matrix of one element. */
void    or_mc_inverse_mass(int dim, double vol, double* minv);
/* Propagator of the e-ph channel over a tuple list, then the inverse mass. */
void    or_mc_apply_eph(int dim, const int32_t* nsec, const double* geo, const double* eps, const double* eps_ph,
                        double omega0, double amplitude, int32_t nu, const double* pts,
                        int32_t nmc, const int32_t* idx, int64_t ntup, const double* f_el,
                        const double* f_ph, double* out_el, double* out_ph,
                        int32_t nthreads);
/* Band::relativeError on DG coefficients [nel][dim+1]. */
double  or_relative_error_dg(int dim, int64_t nel, const double* err, const double* pop,
                             double stat, double rtol, double atol, double penalty);
/* DP5(4) of the electron band's DG coefficients with the tuple operator. */
void    or_dp5_mc_eph(int dim, const int32_t* nsec, const double* geo,
                      const double* eps, const double* eps_ph, double omega0, double amplitude,
                      int32_t nu, const double* pts, int32_t nmc, const int32_t* idx,
                      int64_t ntup, int32_t dynamic, double bath_n, double* f, double t0,
                      double t_final, double dense_dt,
                      double rtol, double atol, double* dense_out, double* dense_t,
                      int32_t max_dense, int32_t nthreads, or_dp5_stats* stats);
/* or_dp5_mc_eph with a time-dependent bath (sched NULL: constant bath_n) and
 * the bath energy ledger (-int eps I dt of the electron band; 0 when dynamic). */
void    or_dp5_mc_eph_ex(int dim, const int32_t* nsec, const double* geo,
                         const double* eps, const double* eps_ph, double omega0, double amplitude,
                         int32_t nu, const double* pts, int32_t nmc, const int32_t* idx,
                         int64_t ntup, int32_t dynamic, double bath_n, const double* sched,
                         int32_t ns, double t_start, double sdt, double satol, double srtol,
                         double* f, double t0, double t_final, double dense_dt,
                         double rtol, double atol, double* dense_out, double* dense_t,
                         int32_t max_dense, int32_t nthreads, or_dp5_stats* stats, double* ledger);
/* or_dp5_mc_eph_ex on a given step sequence (all accepted): the tuple
 * operator's fixed-step replay with the ledger. */
void    or_dp5_mc_eph_replay(int dim, const int32_t* nsec, const double* geo,
                             const double* eps, const double* eps_ph, double omega0, double amplitude,
                             int32_t nu, const double* pts, int32_t nmc, const int32_t* idx,
                             int64_t ntup, int32_t dynamic, double bath_n, const double* sched,
                             int32_t ns, double t_start, double sdt, double* f, double t0,
                             double t_final, double dense_dt, const double* h_steps, int64_t n_steps,
                             double* dense_out, double* dense_t, int32_t max_dense, int32_t nthreads,
                             double* ledger);
/* Electron energy int eps f dk of DG coefficients [nel][dim+1] (exact P1 x P1). */
double  or_energy_dg(int dim, const int32_t* nsec, const double* extent, const double* eps,
                     const double* f);
/* ScatteringChannel::scatteringRates of the electron band (legs 0 and 2). */
void    or_mc_rates_eph(int dim, const int32_t* nsec, const double* geo,
                        const double* eps, const double* eps_ph, double omega0, double amplitude,
                        int32_t nu, const double* pts, int32_t nmc, const int32_t* idx,
                        int64_t ntup, const double* f_el, const double* f_ph, double* out_el,
                        int32_t nthreads);
/* Tuples (e0, e1, e2, e3, umklapp id) of the 4-leg channel {el in, el in,
 * el out, el out}; brute: no box or energy pruning of leg 3. */
int64_t or_mc_tuples_ee(int dim, const int32_t* nsec, const double* geo, const double* eps, int32_t nu,
                        const double* pts, int32_t nmc, double min_det, double threshold,
                        int32_t brute, int32_t* idx, int64_t cap);
/* One 4-leg tuple's integral res[4*(dim+1)] (no amplitude, no leg signs). */
int     or_mc_tuple_ee(int dim, const int32_t* nsec, const double* geo, const double* eps, int32_t nu,
                       const double* pts, int32_t nmc, double min_det, const int32_t* tup,
                       const double* c, double* res, double* D_out, double* N_out);
/* Propagator of the 4-leg channel over a tuple list, then the inverse mass. */
void    or_mc_apply_ee(int dim, const int32_t* nsec, const double* geo, const double* eps, int32_t nu,
                       const double* pts, int32_t nmc, double min_det, double amplitude, double kappa,
                       const int32_t* idx, int64_t ntup, const double* f, int32_t batch, double* out);
/* DP5(4) of the electron band's DG state with the 4-leg channel alone. */
void    or_dp5_mc_ee(int dim, const int32_t* nsec, const double* geo, const double* eps, int32_t nu,
                     const double* pts, int32_t nmc, double min_det, double amplitude, double kappa,
                     const int32_t* idx, int64_t ntup, double* f, double t0, double t_final,
                     double dense_dt, double rtol, double atol, double* dense_out, double* dense_t,
                     int32_t max_dense, or_dp5_stats* stats);
/* Nodal values -> DG linear-form coefficients [nel][dim+1]. */
void    or_mc_nodal_to_dg(int dim, const int32_t* nsec, const double* f, double* dg);

#ifdef __cplusplus
}
#endif
#endif
