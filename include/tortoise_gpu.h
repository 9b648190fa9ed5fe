/*
 * tortoise_gpu.h -- C ABI of the MI355X scattering-integral backend.
 *
 * One context owns one mesh, one electron-phonon channel and its assembled
 * kernel on the current HIP device, plus one HIP stream (the caller may
 * substitute its own).  A context is not thread-safe; calls on one context
 * are serialised by the caller.  Errors are returned as negative status codes
 * and described by tg_last_error(); nothing throws across this boundary.
 *
 * Data conventions
 *   - nodal P1 values on the (nsec+1)^dim vertex lattice of a structured mesh,
 *     node id = ix + (nx+1)*(iy + (ny+1)*iz);
 *   - a batch of B members is column-major n x B: member b at ptr + b*n;
 *   - FP64 at the boundary in every precision mode;
 *   - pointers are host memory unless TG_PTR_DEVICE is set in flags, in which
 *     case they are device memory on the context's device.
 *
 * Reference interfaces replaced (paths relative to the TORTOISE tree):
 *   MaterialStatus<NDim>::propagate()            src/Physics/Status/MaterialStatus.cpp:146-161
 *   ScatteringChannel<NDim,NLegs>::propagate()   src/Physics/Scattering/ScatteringChannel.cpp:651-670
 *   MaterialTimeStatus<NDim>::propagate(tf, dt)  src/Physics/Status/MaterialTimeStatus.cpp:102-137
 *   Algorithms::dormandPrince54 (vetoer form)    include/Generics/Algorithms/DormandPrince54.hpp:479-665
 *   MaterialStatus<NDim>::relativeError()        src/Physics/Status/MaterialStatus.cpp:189-196
 *   Band<NDim>::relativeError()                  src/Physics/Material/Band.cpp:87-101
 *   Function<NDim>::integrate() / integrate(rhs) src/Geometry/Structured/FunctionSpace/Function.cpp:369-375
 */
#ifndef TORTOISE_GPU_H
#define TORTOISE_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tg_ctx tg_ctx;
typedef int32_t tg_status;

#define TG_OK                  0
#define TG_ERR_ARG            -1   /* invalid argument or size */
#define TG_ERR_HIP            -2   /* HIP runtime / kernel launch failure */
#define TG_ERR_NONFINITE      -3   /* NaN/Inf met in a state or error norm */
#define TG_ERR_STEP_UNDERFLOW -4   /* DP5 step fell below dense_dt/1e4 (DormandPrince54.hpp:658-661) */
#define TG_ERR_STATE          -5   /* call out of order (e.g. apply before assemble) */
#define TG_ERR_NOMEM          -6   /* device allocation failed */

/* occupation factors (tg_problem_desc.occupation) */
#define TG_OCC_NODAL 0
#define TG_OCC_HAT   1

/* flags */
#define TG_PTR_DEVICE   0x1   /* f / I / out pointers are device memory */
#define TG_PREC_MIXED   0x2   /* batched fp16 hi+lo split kernel, FP32 MFMA, FP64 epilogue */
#define TG_CONSERVE     0x4   /* remove the particle-number defect of one evaluation */
#define TG_QUIET        0x8   /* (dp5) no per-step console line */
#define TG_TIME_KERNEL  0x10  /* (apply) bracket the dominant kernel with HIP events */
#define TG_BATH_SCHEDULE 0x20 /* (dp5) the bath follows the schedule set by tg_set_bath_schedule /
                                 tg_tuple_set_bath_schedule; bath_n is not read */

typedef struct tg_problem_desc {
    int32_t dim;            /* 2 or 3 */
    int32_t nsec[3];        /* sections per axis (reference Mesh nSecSplits) */
    double  origin[3];      /* physical coordinates of the mesh origin */
    double  extent[3];      /* edge lengths of the (axis-aligned) mesh box */
    const double* eps;      /* nodal band energies, host memory, tg_num_nodes() values */
    double  g;              /* e-ph coupling; channel amplitude = g*g */
    double  omega0;         /* dispersionless optical phonon energy */
    double  sigma;          /* Gaussian broadening of the energy delta */
    double  bath_n;         /* default phonon bath occupation */
    double  de;             /* energy-grid spacing of the overlap quadrature (<= 0: sigma/8) */
    double  cutoff;         /* Gaussian cut-off in units of sigma (<= 0: 8) */
    int32_t max_batch;      /* largest batch the context will be asked for (>= 1) */
    /* Occupation factors of the test function's node i (the (1 - f_i) of the
     * gain and the f_i of the loss):
     *   TG_OCC_NODAL (0): the nodal value f_i (the lumped W-form);
     *   TG_OCC_HAT   (1): the hat average fbar_i = int phi_i f / int phi_i of the
     *                     P1 interpolant (consistent over lumped P1 mass).
     * The reference integrates these P1 products exactly per element tuple
     * (include/Physics/Scattering/IntegratorCore.hpp:298-316); lumping them at
     * the node adds an energy diffusion of the size of the hat's energy spread,
     * which the hat average removes.  The hat form does not conserve the
     * particle number exactly by itself: pair it with TG_CONSERVE. */
    int32_t occupation;
    /* electron-electron (4-leg) channel on the periodic lattice of the box
     * (off when ee_g == 0): amplitude ee_g^2 / (|q|^2 + ee_kappa2)^2, energy
     * delta a Gaussian of width ee_sigma (cut at `cutoff` widths) projected on
     * a grid of spacing ee_dw (<= 0: ee_sigma/4).  Replaces a 4-leg
     * ScatteringChannel<NDim,4> with functAmplitude = V(q)^2
     * (src/Physics/Scattering/ScatteringChannel.cpp:399-415,618-650). */
    double  ee_g;
    double  ee_kappa2;
    double  ee_sigma;
    double  ee_dw;
} tg_problem_desc;

typedef struct tg_step_stats {
    int64_t attempted;      /* attempted DP5 steps */
    int64_t accepted;
    int64_t rejected;
    int64_t evals;          /* collision-operator evaluations of this member */
    int32_t failed;         /* 1: step underflow */
    int32_t n_dense;        /* dense outputs produced */
    double  bath_energy;    /* energy ledger of the phonon bath: the energy it received over the
                               accepted steps, i.e. minus the e-ph channel's electron energy flux
                               (sum_i V_i eps_i I_i, FP64 per stage) integrated with the step's
                               weights, so E_electrons + bath_energy stays at its initial value */
} tg_step_stats;

/* Time-dependent bath: the phonon band constrained to a uniform population
 * n_b(t) per member (Band::constrained / constrainedPopulation,
 * include/Physics/Material/Band.hpp:70-74, enforced at every stage time by
 * MaterialStatus::applyConstraints, src/Physics/Status/MaterialStatus.cpp:198-206).
 * Member b's occupation is sampled at t_start + k*dt, k = 0..n_samples-1, at
 * samples + b*n_samples; between samples it is linear, outside it holds the
 * end values.  The step vetoer (MaterialTimeStatus::timeStepVetoer,
 * src/Physics/Status/MaterialTimeStatus.cpp:60-100) reads the band's error
 * tolerances atol and rtol (Band::absoluteErrorTolerance /
 * relativeErrorTolerance, Band.hpp:77-78). */
typedef struct tg_bath_schedule {
    const double* samples;  /* host memory, batch x n_samples */
    int32_t n_samples;      /* >= 2 */
    int32_t batch;          /* members covered (a DP5 call with TG_BATH_SCHEDULE takes at most this many) */
    double  t_start;
    double  dt;             /* > 0 */
    double  atol;
    double  rtol;
} tg_bath_schedule;

/* ---------------------------------------------------------------- setup */
/* Number of lattice nodes of a structured mesh (no context needed). */
int64_t     tg_num_nodes(int32_t dim, const int32_t* nsec);

/* Builds the mesh, uploads nodal data and sizes the energy projection.
 * Replaces the construction of Mesh/Band/Material + ScatteringChannel
 * (src/Physics/Material/Material.cpp:244-301). */
tg_status   tg_ctx_create(const tg_problem_desc* desc, tg_ctx** out);
void        tg_ctx_destroy(tg_ctx* ctx);
const char* tg_last_error(const tg_ctx* ctx);
/* Launch all work of this context on `hip_stream` (a hipStream_t); NULL
 * restores the context's own stream. */
tg_status   tg_ctx_set_stream(tg_ctx* ctx, void* hip_stream);
void*       tg_ctx_stream(const tg_ctx* ctx);
tg_status   tg_mesh_info(const tg_ctx* ctx, int64_t* n_nodes, int64_t* n_elements,
                         double* elem_volume);
/* Lumped P1 masses V_i (host out, n values). */
tg_status   tg_node_volumes(tg_ctx* ctx, double* V);

/* Assembles the emission kernel K (FP64, once) and, when max_batch > 1 or
 * TG_PREC_MIXED is in flags, the stacked [K; K^T] fp16 hi+lo split used by the
 * batched kernel.  Replaces the tuple screening of
 * ScatteringChannel::precalculateIndices (ScatteringChannel.cpp:421-538). */
tg_status   tg_kernel_assemble(tg_ctx* ctx, int32_t flags);
typedef struct tg_kernel_stats {
    int64_t n;              /* nodes */
    int64_t nnz;            /* structurally nonzero entries of K */
    int64_t plain_band;     /* entries of K's band (128-row panels) the plain FP64 evaluation streams */
    int64_t mixed_band;     /* entries of [K; K^T] the batched kernel streams per evaluation */
    int32_t mixed_splits;   /* reduction splits of the batched kernel */
    int32_t mixed_width;    /* members per batched tile (64, 128 or 256) */
    int32_t mixed_scale;    /* power-of-two exponent applied to K before the fp16 split */
    int32_t window;         /* widest energy-projection window (hats per node) */
    double  kmax;           /* largest row sum of K */
    int32_t mixed_tile;     /* members per workgroup of the batched kernel */
    int32_t mixed_fold;     /* iterations per FP64 fold of its FP32 sums (0: FP32 partials) */
    int32_t mixed_waves;    /* waves per workgroup of the batched kernel */
    int32_t mixed_depth;    /* iterations of register prefetch of the batched kernel */
    float   mixed_tuned_ms; /* autotuner's time of the chosen configuration (0: set explicitly) */
    int32_t mixed_persist;  /* persistent kernel with a balanced panel schedule, A fragments straight
                               into registers (1) or A tile through LDS (2); 0: tiled grid */
    int32_t mixed_skip;     /* 1: panels whose fp16 operands are zero for every member are left out of
                               the persistent kernel (their products are exactly 0) */
    int64_t mixed_last_band;/* entries of [K; K^T] the last batched evaluation streamed (its panels
                               with work; mixed_band when nothing was left out) */
    int32_t mixed_cuts;     /* panels whose band the persistent kernel computes as two work units
                               (summed into two partial-sum slots the epilogue adds) */
    int64_t mixed_nnz;      /* nonzero entries of [K; K^T] in the batched kernel's band (2 nnz less
                               the tail entries beyond the extents of the fp16 split) */
    int64_t mixed_last_nnz; /* of those, the entries in the iterations the last batched launch
                               computed (mixed_nnz when nothing was left out) */
} tg_kernel_stats;

/* Sizes of the assembled kernel (for throughput and roofline accounting). */
tg_status   tg_kernel_stats_get(tg_ctx* ctx, tg_kernel_stats* out);
/* Copies rows [row0, row0+nrows) of K (FP64) to host `out` (nrows x n). */
tg_status   tg_kernel_copy_rows(tg_ctx* ctx, int64_t row0, int64_t nrows, double* out);

/* -------------------------------------------------------- the hot path */
/* One collision-operator evaluation per member:
 *   I_b = [(1-x_b)(W_b f_b) - x_b (W_b^T (1-f_b))] / V,  W_b = (n_b+1) K + n_b K^T,
 * x_b = f_b (TG_OCC_NODAL) or its hat average (TG_OCC_HAT).  With TG_CONSERVE
 * the number defect sum_i V_i I_i is removed as I -= s |x (1 - x)|.
 * bath_n: per-member bath occupations (batch values) or NULL for the default.
 * Default path: FP64 streaming pass over K's band giving K f and K^T f
 * in one read (any batch, members in turn).  TG_PREC_MIXED: batched MFMA kernel
 * (batch <= max_batch). */
tg_status   tg_collision_apply(tg_ctx* ctx, const double* f, double* I, int64_t n,
                               int32_t batch, const double* bath_n, int32_t flags);

/* tg_collision_apply (plain FP64 path) that allocates nothing: no
 * hipMalloc, hipFree or allocating library call and no stream
 * synchronisation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops).  No counterpart in the reference.  f and I
 * are device memory (TG_PTR_DEVICE is required); bath_n is host memory (batch
 * values) or NULL; TG_CONSERVE is honoured; TG_PREC_MIXED, TG_TIME_KERNEL and
 * contexts with the e-e channel are refused.  Its scratch (the gathered state
 * and the partial sums) is a 16 MiB device array per GPU, held until the
 * process ends: a context whose band needs more (beyond about 128 x 128
 * sections) is refused with TG_ERR_ARG.  Every call on a GPU shares that
 * GPU's array, so calls that overlap on one GPU (other streams or contexts)
 * race; calls on one stream are ordered.  The work is enqueued on the
 * context's stream and the call returns without waiting. */
tg_status   tg_collision_apply_noalloc(tg_ctx* ctx, const double* f, double* I, int64_t n,
                                       int32_t batch, const double* bath_n, int32_t flags);

/* Duration in ms of the dominant kernel of the last tg_collision_apply made
 * with TG_TIME_KERNEL (plain: the K-band K f / K^T f pass; mixed: the
 * MFMA GEMM), from HIP events on the context's stream.  Waits for it. */
tg_status   tg_last_kernel_ms(tg_ctx* ctx, double* ms);

/* Adaptive DP5(4) from t0 to t1 for every member (per-member step size and
 * accept/reject, error norm of Band::relativeError for a fermion band).
 * f (n x batch) holds the initial state and receives the state after the
 * last accepted step.  dense_out (optional, max_dense x n x batch, member-
 * major: member b's k-th output at dense_out + (b*max_dense + k)*n) receives
 * the dense outputs at t0 + k*dense_dt.  stats: batch entries (optional).
 * h_accepted (optional, batch x h_cap): the accepted step sizes of member b
 * in order at h_accepted + b*h_cap (stats[b].accepted of them, up to h_cap).
 * A step whose error norm is not finite (its stages overflowed) is rejected
 * with the smallest step scale; a member whose step then falls below
 * dense_dt/1e4 fails with TG_ERR_NONFINITE (TG_ERR_STEP_UNDERFLOW when the
 * last norm was finite). */
tg_status   tg_dp5_integrate(tg_ctx* ctx, double* f, int64_t n, int32_t batch,
                             double t0, double t1, double dense_dt, double rtol,
                             double atol, const double* bath_n, int32_t flags,
                             double* dense_out, int32_t max_dense,
                             tg_step_stats* stats, double* h_accepted, int32_t h_cap);

/* Sets (or, with NULL, clears) the time-dependent bath the DP5 entry points
 * use when TG_BATH_SCHEDULE is in their flags: every stage of member b
 * evaluates the operator with n_b(t) at the stage time, and the step vetoer
 * caps the steps (the vetoer runs in every adaptive DP5; without a schedule
 * it returns the step unchanged up to t + h - t, as the reference's does
 * with no constrained band).  Replaces MaterialStatus::applyConstraints and
 * MaterialTimeStatus::timeStepVetoer (MaterialStatus.cpp:198-206,
 * MaterialTimeStatus.cpp:60-100) for a constrained phonon bath. */
tg_status   tg_set_bath_schedule(tg_ctx* ctx, const tg_bath_schedule* sched);

/* The same stage arithmetic on a given step sequence: member b takes the
 * n_steps[b] steps h_steps[b*h_cap ...] (all accepted, no error control),
 * with dense output as tg_dp5_integrate.  Used to compare two precisions on
 * identical steps.  No reference counterpart (test and measurement aid). */
tg_status   tg_dp5_replay(tg_ctx* ctx, double* f, int64_t n, int32_t batch, double t0,
                          double t1, double dense_dt, const double* h_steps,
                          const int32_t* n_steps, int32_t h_cap, const double* bath_n,
                          int32_t flags, double* dense_out, int32_t max_dense);

/* Discrete adjoint of the fixed-step DP5 (one member, FP64).  The forward
 * map is tg_dp5_replay on the step sequence h_steps[0..n_steps) with dense
 * output at t1; given dL/df(t1) (host, n values) it returns dL/df0 (host, n
 * values) and dL_dparams = {dL/dg, dL/domega0}.  dbath_domega0 is the
 * derivative of the bath occupation with respect to omega0 (e.g. of
 * n_BE(omega0, T) at fixed T).  Forward states are recomputed from
 * checkpoints every ckpt steps.  flags: TG_CONSERVE (every stage evaluation
 * removes the number defect, as tg_dp5_replay with TG_CONSERVE does, and the
 * gradients include that removal).  No reference counterpart (the reference
 * has no differentiation); the forward map it differentiates replaces
 * MaterialTimeStatus::propagate (MaterialTimeStatus.cpp:102-137). */
tg_status   tg_dp5_adjoint(tg_ctx* ctx, const double* f0, int64_t n, double t0, double t1,
                           double dense_dt, const double* h_steps, int32_t n_steps,
                           int32_t ckpt, double bath_n, double dbath_domega0,
                           const double* dL_df, double* dL_df0, double* dL_dparams, int32_t flags);

/* FP64 sums per member: number = sum_i V_i f_i, energy = sum_i V_i eps_i f_i. */
tg_status   tg_conservation(tg_ctx* ctx, const double* f, int64_t n, int32_t batch,
                            double* number, double* energy, int32_t flags);

/* Band::relativeError (fermion) of err against pop, FP64 (host results). */
tg_status   tg_relative_error(tg_ctx* ctx, const double* err, const double* pop,
                              int64_t n, int32_t batch, double rtol, double atol,
                              double* out, int32_t flags);


/* ==================================================================
 * Tuple operator: the reference's own e-ph scattering integral.
 *
 * The channel {electron in, phonon in, electron out} of
 * ScatteringChannel<NDim,3> with a dispersionless phonon band held as a bath,
 * evaluated the reference's way: a list of element tuples (e0, e1, e2,
 * umklapp) and, per tuple, the Monte-Carlo integral over the energy-momentum
 * delta shell with a fixed matrix of MC points, on DG (linear-form) element
 * coefficients.  Replaces (paths relative to the TORTOISE tree):
 *   ScatteringChannel::precalculateIndices     src/Physics/Scattering/ScatteringChannel.cpp:421-538
 *   ScatteringChannel::constructMCPoints       src/Physics/Scattering/ScatteringChannel.cpp:677-688
 *   ScatteringChannel::propagate               src/Physics/Scattering/ScatteringChannel.cpp:618-670
 *   scatteringIntegrationTypeB                 include/Physics/Scattering/IntegratorCore.hpp:260-357
 *   Function::applyInverseMass                 src/Geometry/Structured/FunctionSpace/Function.cpp:801-804
 *   .scattIndices file                         src/Physics/Scattering/ScatteringChannel.cpp:422-459,520-537
 *
 * State layout: member b, element e, linear-form coefficient a at
 * ptr[(b*nel + e)*(dim+1) + a] -- the reference Function's element view
 * [f(v0), f(v1)-f(v0), ...] (Function.hpp:326-344).  With dynamic_phonons the
 * phonon band's block follows the electron band's for every member and
 * bath_n may be NULL.
 * ================================================================== */
typedef struct tg_tuple_ctx tg_tuple_ctx;

typedef struct tg_tuple_desc {
    int32_t dim;            /* 2 or 3 */
    int32_t nsec[3];        /* sections per axis */
    double  origin[3];
    double  extent[3];
    const double* eps;      /* nodal electron energies (the P1 dispersion), host memory */
    double  omega0;         /* phonon energy (dispersionless band) */
    double  amplitude;      /* ScatteringChannel::amplitude (g^2); functAmplitude = 1 */
    int32_t n_umklapp;      /* umklapp vectors {-n..n}^dim (findUmklappVectors(n)), 0..3 */
    int32_t num_mc;         /* numMCPoints (Material default 50) */
    double  min_det;        /* minimumDeterminant (Material default 1e-9) */
    double  threshold;      /* scattElemeThreshold (Material default 1e-9) */
    int32_t max_batch;      /* members per propagate call */
    int32_t dynamic_phonons;/* 0: the phonon band is a bath of occupation bath_n per member;
                               1: it evolves -- states and outputs hold both bands,
                               [member][band: electron, phonon][element][dim+1] */
    const double* eps_ph;   /* phonon nodal energies (a P1 dispersion), or NULL: omega0 */
    /* A general zone (replaces Region<NDim> / Mesh<NDim>'s parallelepipeds,
     * include/Geometry/Structured/DirectSpace/Region.hpp:29-34,
     * src/Geometry/Structured/DirectSpace/Mesh.cpp:268-281): the mesh spans
     * origin + sum_c t_c sides[c] (t in [0,1]^dim; sides[c][d] is component
     * d of side c, Mesh::gVec = Region::gVec * diag(relative extent)), node
     * index i_c along side c at origin + sum_c (i_c / nsec[c]) sides[c]; the
     * umklapp vectors are sum_c u_c zone[c] (zone[c] = Region::gVec column
     * c, ScatteringChannel.cpp:330-334).  All-zero sides: the axis-aligned
     * box origin + extent, which is also the zone. */
    double  sides[3][3];
    double  zone[3][3];
} tg_tuple_desc;

tg_status   tg_tuple_create(const tg_tuple_desc* desc, tg_tuple_ctx** out);
/* The 4-leg channel {el in, el in, el out, el out} of the electron band:
 * ScatteringChannel<NDim,4> (src/Physics/Scattering/ScatteringChannel.cpp:716-722,
 * created by Material::addScatteringChannelsWildcards<4>,
 * src/Physics/Material/Material.cpp:304-347) with fermion statistics on every
 * leg.  desc as for tg_tuple_create (omega0 and eps_ph are not read;
 * dynamic_phonons must be 0).  functAmplitude (ScatteringChannel.hpp:70, at
 * the legs' barycentres): 1 for kappa <= 0, the reference default
 * (ScatteringChannel.cpp:144); else the squared screened Coulomb potential
 * 1/(|q|^2 + kappa^2)^2, q = barycentre of leg 2 - barycentre of leg 0.
 * On such a context the point matrix has tg_tuple_point_rows(dim, 4) rows,
 * tuples are [n][5] = e0, e1, e2, e3, uid (any order; the built list is in
 * (e0, e1, e2, uid, e3) order), tg_tuple_propagate and tg_tuple_dp5 ignore
 * bath_n (the DP5 bath ledger stays zero), and tg_tuple_rates and
 * tg_tuple_set_bath_schedule return TG_ERR_STATE. */
tg_status   tg_tuple4_create(const tg_tuple_desc* desc, double kappa, tg_tuple_ctx** out);
void        tg_tuple_destroy(tg_tuple_ctx* ctx);
const char* tg_tuple_last_error(const tg_tuple_ctx* ctx);
/* rows of the MC point matrix of a channel: max(dim*(nlegs-1)-1, 2*dim) */
int32_t     tg_tuple_point_rows(int32_t dim, int32_t nlegs);
/* Sets the MC point matrix [rows][num_mc] (host, row-major): leg 0's points
 * in rows 0..dim-1, the partially free leg's in the bottom dim-1 rows.  The
 * first call also builds the tuple list (unless one was loaded); every call
 * re-integrates the per-tuple moments (MaterialTimeStatus::propagate redraws
 * the points, MaterialTimeStatus.cpp:102-105). */
tg_status   tg_tuple_set_points(tg_tuple_ctx* ctx, const double* pts, int32_t flags);
/* Tuple list (scatteringMatrixIndices + umklapp ids): [n][4] = e0, e1, e2, uid,
 * ordered by e0.  count is -1 before a list exists. */
tg_status   tg_tuple_count(const tg_tuple_ctx* ctx, int64_t* count);
tg_status   tg_tuple_copy_indices(tg_tuple_ctx* ctx, int32_t* idx, int64_t cap);
tg_status   tg_tuple_set_indices(tg_tuple_ctx* ctx, const int32_t* idx, int64_t count);
tg_status   tg_tuple_save_indices(tg_tuple_ctx* ctx, const char* path);
tg_status   tg_tuple_load_indices(tg_tuple_ctx* ctx, const char* path);
tg_status   tg_tuple_set_amplitude(tg_tuple_ctx* ctx, double amplitude);
/* dN/dt of the electron band (after the inverse mass) for `batch` members
 * with phonon occupations bath_n[b].  flags: TG_PTR_DEVICE, TG_TIME_KERNEL. */
tg_status   tg_tuple_propagate(tg_tuple_ctx* ctx, const double* f, const double* bath_n, double* out,
                               int64_t nel, int32_t batch, int32_t flags);
/* ScatteringChannel::scatteringRates(status, band) (ScatteringChannel.cpp:544-612)
 * of the electron band: the sum over its legs (0 and 2) of the rate with that
 * leg's forms replaced by constants, after the inverse mass.  Same input and
 * flags as tg_tuple_propagate; out holds the electron band only,
 * [batch][nel][dim+1]. */
tg_status   tg_tuple_rates(tg_tuple_ctx* ctx, const double* f, const double* bath_n, double* out,
                           int64_t nel, int32_t batch, int32_t flags);
tg_status   tg_tuple_last_kernel_ms(tg_tuple_ctx* ctx, double* ms);
/* MaterialTimeStatus::propagate (MaterialTimeStatus.cpp:102-137) with the tuple
 * operator: DP5(4) (DormandPrince54.hpp:479-665) of the electron band's DG
 * coefficients f [batch][nel][dim+1] from t0 to t1, per-member steps, the
 * fermion error norm of Band::relativeError (Band.cpp:87-101) on DG data, the
 * phonon band a bath of occupation bath_n[b].  Same outputs as
 * tg_dp5_integrate (dense_out [batch][max_dense][nel][dim+1]).  The MC points
 * are the ones last set (the caller redraws them per propagate, as
 * MaterialTimeStatus.cpp:104 does).  flags: TG_PTR_DEVICE, TG_QUIET,
 * TG_BATH_SCHEDULE.  h_accepted (optional, batch x h_cap): the accepted step
 * sizes of member b at h_accepted + b*h_cap, as in tg_dp5_integrate. */
tg_status   tg_tuple_dp5(tg_tuple_ctx* ctx, double* f, int64_t nel, int32_t batch, double t0,
                         double t1, double dense_dt, double rtol, double atol,
                         const double* bath_n, int32_t flags, double* dense_out,
                         int32_t max_dense, tg_step_stats* stats, double* h_accepted,
                         int32_t h_cap);
/* tg_set_bath_schedule for the tuple operator's bath (channels with a phonon
 * bath; not with dynamic_phonons).  tg_tuple_dp5 with TG_BATH_SCHEDULE in
 * flags evaluates every stage with the member's n_b(t) at the stage time and
 * its step vetoer reads the schedule.  The bath ledger of tg_tuple_dp5
 * (stats[b].bath_energy) integrates minus the electron energy flux
 * int eps(k) I(k) dk of the DG result (zero with dynamic_phonons, where the
 * phonon band is part of the state). */
tg_status   tg_tuple_set_bath_schedule(tg_tuple_ctx* ctx, const tg_bath_schedule* sched);
void*       tg_tuple_stream(const tg_tuple_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
