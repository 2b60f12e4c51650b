/*
 * spinach_powder.h - C-ABI of the MI355X powder-averaged time-domain
 * propagation library (libspinach_powder.so).
 *
 * The library replaces the orientation loop and the weighted sum of
 * Spinach's static powder context for the case
 *   powder(spin_system, @acquire, parameters, 'nmr')
 * in the sphten-liouv formalism with no kinetics and no rotating frames
 * (orientation-dependent relaxation: sp_plan_powder_relax; the states handed
 * between the stages of a pulse sequence: sp_plan_set_initial_states and
 * sp_plan_propagate). It receives the isotropic generator I and the 34
 * anisotropic components Q{rank}{k,m} that hamiltonian() returns, the
 * orientation grid, the initial state and the coil, and returns the
 * powder-averaged free induction decay.
 *
 * Conventions
 *  - Sparse matrices use MATLAB's CSC layout (mxGetJc / mxGetIr /
 *    mxGetComplexDoubles with -R2018a interleaved complex), so a MEX gateway
 *    passes MATLAB's own arrays without conversion.
 *  - Complex vectors are interleaved (re, im) doubles.
 *  - Every entry point returns 0 on success or an SP_E* code; the message of
 *    the last failure on a context is available from sp_last_error(). No C++
 *    exception crosses this boundary.
 *  - The library owns nothing the caller passes; it copies what it needs.
 *  - A context is bound to one device and one HIP stream; calls on one
 *    context must not overlap (MATLAB calls MEX files from one thread).
 */
#ifndef SPINACH_POWDER_H
#define SPINACH_POWDER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SP_OK 0
#define SP_EINVAL 1   /* invalid argument (shape, NULL, non-finite) */
#define SP_EDEVICE 2  /* HIP runtime failure */
#define SP_ENOMEM 3   /* device or host allocation failed */
#define SP_ERANGE 4   /* |L*dt| or step count outside the supported range */

#define SP_MODE_FP64 0 /* complex FP64 throughout */
#define SP_MODE_FP32 1 /* complex FP32 bulk; FP64 norms, step counts and orientation sum */
/* Flag OR-ed into the mode: keep the general complex-generator arithmetic
 * even where the generator is real. Without it, a system whose generator is
 * real at every orientation (I real and Q{r}{k,m} = (-1)^(k-m) conj(Q{r}{-k,-m}),
 * which secular 'nmr' Hamiltonians in the sphten-liouv basis satisfy) is
 * propagated with real matrix products, P = cos(L dt) - i sin(L dt); the
 * result agrees with the complex path to rounding. */
#define SP_MODE_COMPLEX 16
/* Flag OR-ed into the mode: use the propagator power scheme for every system.
 * Without it, FP64 systems with a real generator and a padded coil-connected
 * dimension of at most 448 take the spectral path: one Householder
 * tridiagonalisation and one divide-and-conquer tridiagonal eigensolution per
 * orientation give f(t) = sum_k g_k exp(-i lambda_k t) directly (agreement with the
 * propagator path: ~1e-12 relative over thousands of steps). */
#define SP_MODE_PROPAGATOR 32
/* Flag OR-ed into the mode of a plan (sp_plan_powder, sp_plan_powder_batch,
 * sp_plan_mas): the first sp_execute records the plan's whole pass (queue reset,
 * propagation kernels, orientation sum) as a HIP graph on the context's stream,
 * and every sp_execute launches that graph (one launch per pass instead of one
 * per kernel). Results are identical to the plan without the flag. The stats of
 * a graph execute carry the whole pass in kernel_ms and zeros in stage_ms.
 * Replaces: nothing in the reference (an execution option for callers that
 * replay a plan). */
#define SP_MODE_GRAPH 64
/* Flag OR-ed into the mode: propagate every system by time stepping, one
 * step(L,rho,dt) per time step with detection after each (kernel/krylov.m:
 * 139-163 'observable', kernel/step.m:204-375), as evolution() does when
 * spin_system.sys.enable contains 'krylov' (kernel/evolution.m:531-533,574).
 * Without the flag, systems whose coil-connected dimension is at least
 * tols->krylov_tol, or above the dense limit of 1024, are stepped; the rest
 * take the spectral or propagator path. The stepping path runs in FP64 in
 * both precision modes. */
#define SP_MODE_KRYLOV 128

#define SP_NQ 34 /* Q slots: rank 1 (k,m) row-major 3x3, then rank 2 (k,m) row-major 5x5 */

typedef struct sp_ctx sp_ctx;
typedef struct sp_plan sp_plan;

/* Square sparse matrix in MATLAB CSC layout. n == 0 or nnz == 0 means an
 * empty matrix (an empty Q{r}{k,m} slot). */
typedef struct sp_csc {
  int64_t n;           /* dimension (rows == cols) */
  const uint64_t* jc;  /* n+1 column pointers (mwIndex) */
  const uint64_t* ir;  /* row indices, jc[n] entries */
  const double* pr;    /* values: 2*jc[n] interleaved doubles if is_complex, else jc[n] reals */
  int32_t is_complex;
} sp_csc;

/* Tolerances from spin_system.tols (kernel/utilities/tolerances.m). NULL
 * means the reference defaults below. The library uses krylov_tol for the path
 * choice (see SP_MODE_KRYLOV); the coil-connected reduction is exact on the
 * union pattern and the dense paths compute the exact exponential, so
 * liouv_zero, zte_tol, prop_chop and small_matrix are validated (finite,
 * non-negative) but change nothing. */
typedef struct sp_tols {
  double liouv_zero;   /* 1e-10 */
  double zte_tol;      /* 1e-24 */
  double prop_chop;    /* 1e-10 */
  double small_matrix; /* 200   */
  double krylov_tol;   /* 10000 */
} sp_tols;

typedef struct sp_stats {
  double elapsed_s;          /* wall time of the GPU pipeline (HIP events) */
  double orient_steps_per_s; /* n_orientations * nsteps / elapsed_s */
  double kernel_ms;          /* duration of the propagation kernel */
  int64_t full_dim;          /* Liouville dimension of I and Q */
  int64_t reduced_dim;       /* dimension of the coil-connected subspace */
  int64_t reduced_nnz;       /* structural nonzeros of the reduced union pattern */
  int32_t padded_dim;        /* reduced dimension rounded up to the tile size (sp_mas_fid: 16 when
                                the invariant subspaces are propagated as 16 x 16 tiles) */
  int32_t block;             /* time-block length B of the power scheme */
  int32_t taylor_terms;      /* Taylor degree of the scaled exponential */
  int32_t max_squarings;     /* largest scaling exponent over orientations */
  int32_t ps_block;          /* Paterson-Stockmeyer block length of the Taylor evaluation */
  int32_t sum_squarings_lo;  /* sum of scaling exponents over orientations (low 32 bits) */
  int32_t real_systems;      /* systems whose generator is real at every orientation */
  int32_t spectral_systems;  /* systems evaluated on the spectral path */
  int32_t eig_max_sweeps;    /* most secular-equation iterations any root of the spectral path needed */
  int32_t krylov_systems;    /* systems propagated by time stepping (SP_MODE_KRYLOV) */
  double stage_ms[4];        /* propagator and stepping kernels, tridiagonalisation, tridiagonal
                                eigensolver, synthesis */
  int64_t step_products;     /* sparse matrix-vector products the stepping path made in the pass
                                (Taylor terms over all substeps, steps and orientations);
                                sp_mas_fid on tiles: the 16 x 16 real tile products (a
                                complex product counts 4) */
  double plan_ms;            /* host planning (reduction, layouts) and upload of the plan */
  int32_t eig_fallbacks;     /* spectral orientations whose tridiagonal eigensolution did not
                                converge, re-run on the propagator path (real products) */
  int32_t mas_tiles;         /* sp_mas_fid: 16 x 16 tiles the invariant subspaces are packed into
                                (0: the dense kernel ran) */
  int32_t mas_pairs;         /* sp_mas_fid: tile pairs where both coil and rho0 have entries */
  int32_t mas_detected;      /* sp_mas_fid: elements of those pairs the detection sums over (the
                                component blocks where both have entries), padded to 16 */
} sp_stats;

/* ---- context --------------------------------------------------------- */

/* Replaces: the worker pool set up by kernel/create.m:243-371 and used by the
 * parfor of kernel/contexts/powder.m:226. One context per process and
 * device; NULL on failure (no device, HIP error). */
sp_ctx* sp_create(int device);

/* Releases all device memory and the stream (a MEX gateway calls this from
 * mexAtExit). NULL is accepted. */
void sp_destroy(sp_ctx* ctx);

/* Message of the last failure on this context (never NULL). With ctx==NULL
 * returns the message of the last failed context-free call (sp_create,
 * sp_generator_is_real) in this thread. */
const char* sp_last_error(const sp_ctx* ctx);

/* Host-only query (no device needed): sets *is_real to 1 when
 * H(Omega) = I + sum D^r(k,m)(Omega) Q{r}{k,m} is real at every orientation,
 * i.e. I is real and Q{r}{k,m} = (-1)^(k-m) conj(Q{r}{-k,-m}) for every slot,
 * to 8 ulp of the largest entry; those systems take the real-product path
 * (see SP_MODE_COMPLEX). Replaces: nothing in the reference, which always
 * propagates in complex arithmetic (kernel/propagator.m:40-42); it reports
 * the path choice the planner makes. */
int sp_generator_is_real(const sp_csc* I, const sp_csc* Q, int32_t* is_real);

/* ---- native host setup (no device needed) ------------------------------ */

/* One term of the spin Hamiltonian (kernel/hamiltonian.m:673-690's descriptor): the product of
 * one or two single-spin ISTs (ist: 1 = T(1,+1), 2 = T(1,0), 3 = T(1,-1); coef: the product of
 * the operator names' coefficients, 'L+' = -sqrt(2) T(1,+1), 'L-' = sqrt(2) T(1,-1),
 * 'Lz' = T(1,0)), its isotropic coefficient and its rank-1 (0..2) and rank-2 (3..7) IST
 * coefficients and irreducible components, complex interleaved. */
typedef struct sp_ham_term {
  int32_t nspin;         /* 1 or 2 */
  int32_t spin[2];       /* 0-based spin numbers */
  int32_t ist[2];
  int32_t pad;
  double coef;
  double iso[2];
  double ist_coef[16];
  double irr[16];
} sp_ham_term;

typedef struct sp_ham sp_ham; /* opaque: the built I and Q */

/* I and the 34 Q slots of the Liouvillian commutation superoperator in the spin-1/2
 * sphten-liouv basis (basis: dim rows of nspins IST indices 0..3, rows ascending as
 * kernel/basis.m sorts them), each the sum over the terms of the coefficient times the term's
 * commutation superoperator, quantised to tol (clean_up.m). NULL on failure (message from
 * sp_last_error(NULL)). Replaces: kernel/hamiltonian.m:673-792 with the product-table
 * construction of kernel/operators/superop.m:62-170. */
sp_ham* sp_ham_build(int32_t nspins, const int8_t* basis, int64_t dim, const sp_ham_term* terms, int32_t nterms,
                     double tol);

/* Matrix `which` (0 = I, 1..34 = the Q slots in the order of sp_powder_fid's Q array) as a
 * complex CSC view into h (valid until sp_ham_free). */
int sp_ham_matrix(const sp_ham* h, int32_t which, sp_csc* out);

/* Releases h (NULL is accepted). */
void sp_ham_free(sp_ham* h);

/* ---- one-shot call ----------------------------------------------------- */

/* Replaces: kernel/contexts/powder.m:226-286 (parfor body with
 * H=I+orientation(Q,[a b g]), H=(H+H')/2, R=relaxation()=0, acquire() ->
 * evolution(L,coil,rho0,dt,nsteps,'observable'), then the weighted sum
 * answer=sum_n w_n*ans{n}).
 *   I          isotropic part of the commutation superoperator (hamiltonian.m:718-738,
 *              after frqoffset.m)
 *   Q          SP_NQ matrices Q{r}{k,m} (hamiltonian.m:741-792); empty slots allowed
 *   alphas..weights  grid, n_or entries each (powder.m:173-179)
 *   rho0, coil Liouville state vectors, dimension I->n, interleaved complex
 *   dt, nsteps time step 1/sweep and npoints-1 (acquire.m:70-71)
 *   mode       SP_MODE_FP64 or SP_MODE_FP32
 *   fid_out    nsteps+1 interleaved complex: sum_n w_n * coil'*P_n^k*rho0
 *   per_orient_out  optional (NULL allowed): n_or*(nsteps+1) interleaved complex,
 *              orientation-major, unweighted (parameters.sum_up=false)
 *   stats      optional (NULL allowed)
 */
int sp_powder_fid(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* alphas,
                  const double* betas, const double* gammas, const double* weights, int64_t n_or,
                  const double* rho0, const double* coil, double dt, int64_t nsteps,
                  const sp_tols* tols, int mode, double* fid_out, double* per_orient_out,
                  sp_stats* stats);

/* Variant of sp_powder_fid that calls no allocator (no hipMalloc, hipFree or
 * allocating library routine): every device buffer of the call - the staged
 * inputs, per-orientation spectra and FIDs, orientation-sum partials and
 * work-queue heads - is carved from a static device array of 64 MiB that each
 * device holds from its first sp_create until the process ends. No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (serving loops, harnesses that forbid it). Same arguments and results as
 * sp_powder_fid (results identical bit for bit), for the FP64 spectral path
 * only: mode must be SP_MODE_FP64 and the system real with a coil-connected
 * dimension of at most 192 (SP_EINVAL otherwise). Orientations run in blocks
 * sized to the array; a block of 64 needs about 1 KiB per time point, so
 * nsteps beyond ~60000 returns SP_ERANGE; an orientation whose tridiagonal
 * eigensolution does not converge returns SP_ERANGE too (the plan path re-runs it on
 * the propagator path, which needs workspace). Calls on one device share the array:
 * calls that overlap on one device (through different contexts) race. */
int sp_powder_fid_noalloc(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* alphas,
                          const double* betas, const double* gammas, const double* weights, int64_t n_or,
                          const double* rho0, const double* coil, double dt, int64_t nsteps,
                          const sp_tols* tols, int mode, double* fid_out, double* per_orient_out,
                          sp_stats* stats);

/* ---- plan / execute split (inputs resident on the device) -------------- */

/* Replaces: the host-side part of powder.m:155-225 plus the per-orientation
 * reduce(L',coil) of evolution.m:141-145, done once on the union sparsity
 * pattern of I and Q. Uploads the reduced generator components, grid and
 * states to the device. Returns NULL on failure (see sp_last_error(ctx)). */
sp_plan* sp_plan_powder(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* alphas,
                        const double* betas, const double* gammas, const double* weights,
                        int64_t n_or, const double* rho0, const double* coil, double dt,
                        int64_t nsteps, const sp_tols* tols, int mode);

/* Relaxation superoperator R(Omega) of one spin system, so that the generator
 * of each orientation is L = H + i*R (kernel/evolution.m: L = H + 1i*R + 1i*K;
 * R from relaxation(spin_system,[alpha beta gamma]) at kernel/contexts/
 * powder.m:263). Orientation-dependent terms are those of kernel/relaxation.m
 * whose rates are quadratic forms ort*T*ort' of the orientation vector
 * ort = [0 0 1]*euler2dcm(alpha,beta,gamma) = (-sin b cos g, sin b sin g, cos b):
 * the extended T1/T2 model with 3x3 rate tensors (kernel/utilities/
 * rlx_t1_t2.m) and anisotropic damping (relaxation.m, 'damp'); both are
 * diagonal. Per orientation
 *   R(Omega) = R0 + diag(g0 + sum_p q_p(Omega) g_p),
 *   q = (ox^2, oy^2, oz^2, 2 ox oy, 2 ox oz, 2 oy oz),
 * with R0 (any sparse matrix, e.g. orientation-independent theories) and the
 * seven vectors g0..g6 of the full dimension in `diag` (7*dim doubles, g_p at
 * p*dim). Either member may be NULL. */
typedef struct sp_relax {
  const sp_csc* R0;
  const double* diag;
} sp_relax;

/* sp_plan_powder with a relaxation superoperator (rlx may be NULL: then the
 * same as sp_plan_powder). The generator is not Hermitian, so such systems take
 * the propagator path in complex arithmetic (reduced dimension <= 1024), or the
 * time-stepping path (SP_MODE_KRYLOV, krylov_tol, or larger reduced dimensions);
 * a dead time (sp_plan_set_dead_time) steps under the same L = H + iR. R0's
 * pattern, and the diagonal when rlx->diag is given, join the union pattern of
 * the coil-connected reduction. Replaces: the relaxation part of the parfor
 * body of powder.m:262-269. */
sp_plan* sp_plan_powder_relax(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* alphas,
                              const double* betas, const double* gammas, const double* weights,
                              int64_t n_or, const double* rho0, const double* coil, double dt,
                              int64_t nsteps, const sp_tols* tols, int mode, const sp_relax* rlx);

/* Runs the propagation kernels and the orientation sum on the device and
 * waits for them. Replaces: the parfor of powder.m:226-274 and the sum of
 * powder.m:280-286. */
int sp_execute(sp_plan* plan, sp_stats* stats);

/* Copies the results of the last sp_execute to the host (arguments as in
 * sp_powder_fid). */
int sp_plan_result(sp_plan* plan, double* fid_out, double* per_orient_out);

void sp_plan_destroy(sp_plan* plan);

/* Sets the dead time of a plan (seconds, >= 0; 0 removes it): every
 * orientation's initial state is first propagated by step(L,rho0,dead_time)
 * and the acquisition starts from the result. Replaces: experiments/acquire.m:
 * 65-67 (parameters.dead_time). Takes effect from the next sp_execute; applies
 * to every system of a batch plan. Not available on MAS plans. */
int sp_plan_set_dead_time(sp_plan* plan, double dead_time);

/* Per-orientation initial states of a single-system plan: rho_or holds n_or
 * complex state vectors of the full dimension (interleaved, orientation-major);
 * from the next sp_execute the acquisition of orientation o starts from its
 * state instead of rho0 (NULL restores rho0). The plan's coil-connected
 * reduction keeps what the coil can see: the parts of a state outside it never
 * reach the coil under L. Replaces: the state handed from one stage of a pulse
 * sequence to the next acquisition, as in experiments/sp_acquire.m (after the
 * shaped pulse) and experiments/cpmg.m (after every refocusing pulse). Not
 * with a dead time; not on MAS or batch plans. */
int sp_plan_set_initial_states(sp_plan* plan, const double* rho_or);

/* Propagates every orientation's state by nsteps steps of dt under its own
 * generator, rho(Omega) <- step(L(Omega), rho, dt) nsteps times (kernel/step.m:
 * 204-375, the reordered Taylor series with the substeps of its 1-norm; SP_ERANGE
 * when one step needs more than 10000 substeps), from the plan's rho0 or the
 * states of sp_plan_set_initial_states, and writes the results (n_or complex
 * vectors of the full dimension, zero outside the reduced space) to rho_out.
 * Reduction follows the plan's coil: a plan built to propagate states passes
 * as its coil a vector whose support seeds every state the sequence can
 * reach. Replaces: the per-orientation evolution(...,'final') and step() of
 * experiments/cpmg.m:41-55 and kernel/pulses/shaped_pulse_af.m (its 'expv'
 * branch); with a relaxation superoperator under L = H + iR. Not with a dead
 * time. */
int sp_plan_propagate(sp_plan* plan, double dt, int64_t nsteps, double* rho_out);

/* ---- launch configuration and autotuning ------------------------------ */

/* Launch configuration of a plan's kernels. Replaces: nothing in the
 * reference (an execution option). Fields that a plan's paths do not use are
 * ignored by sp_plan_set_tuning and returned as the planner's defaults. */
typedef struct sp_tuning {
  int32_t log2_block;      /* propagator path: time block B = 2^log2_block of the power scheme (0..10) */
  int32_t prop_wg_per_cu;  /* propagator path: resident workgroups per CU (1..8) */
  int32_t tri_wg_per_cu;   /* spectral path: tridiagonalisation workgroups per CU (1..4) */
  int32_t syn_wg_per_cu;   /* spectral path: synthesis workgroups per CU (1..8) */
  int32_t step_wg_per_cu;  /* stepping path: workgroups per CU (1..4, capped by its LDS) */
  int32_t candidates;      /* sp_plan_autotune: configurations timed (0: the cached choice was applied) */
  double default_ms;       /* sp_plan_autotune: pass time of the configuration the plan had before */
  double best_ms;          /* sp_plan_autotune: pass time of the chosen configuration */
} sp_tuning;

/* Times the plan's pass over candidate launch configurations (coordinate
 * descent: time block, then workgroups per CU of each path, passes >= 1
 * timed passes per candidate after one warm-up), keeps the fastest and
 * records it in a per-process cache keyed by the plan's shape (dimensions,
 * step count, item counts per path, precision); plans built later with the
 * same shape start from the cached choice, and a call on such a plan applies
 * it without timing (candidates = 0) unless passes < 0 forces a new search.
 * The per-orientation arithmetic does not depend on the workgroup counts; the
 * time block changes the propagator path's rounding only (~1e-15). */
int sp_plan_autotune(sp_plan* plan, int32_t passes, sp_tuning* chosen);

/* The plan's current launch configuration / apply one (validated, clamped). */
int sp_plan_get_tuning(const sp_plan* plan, sp_tuning* out);
int sp_plan_set_tuning(sp_plan* plan, const sp_tuning* in);

/* ---- batched uneven workload ------------------------------------------ */

/* One spin system of a batch: the arguments of sp_powder_fid for that system. */
typedef struct sp_system {
  const sp_csc* I;
  const sp_csc* Q; /* SP_NQ entries */
  const double* alphas;
  const double* betas;
  const double* gammas;
  const double* weights;
  int64_t n_or;
  const double* rho0;
  const double* coil;
  double* fid_out; /* nsteps+1 interleaved complex */
} sp_system;

/* Many spin systems of uneven dimension, all with the same dt and nsteps,
 * propagated by one persistent launch over a cost-ordered work queue of
 * (system, orientation) items. Replaces: a MATLAB loop of powder() calls,
 * one per system. */
sp_plan* sp_plan_powder_batch(sp_ctx* ctx, const sp_system* systems, int64_t n_systems, double dt,
                              int64_t nsteps, const sp_tols* tols, int mode);

/* Copies each system's FID to its sp_system.fid_out (the pointers given at
 * planning time). */
int sp_plan_result_batch(sp_plan* plan);

/* ---- magic-angle spinning, Hilbert space ------------------------------ */

#define SP_MAS_FRAME_MAGNET 0 /* D_rot2lab*D_rotor*D_lab2rot*D_initial */
#define SP_MAS_FRAME_ROTOR 1  /* D_rot2lab*D_rotor*D_initial */

typedef struct sp_rotor {
  double axis[3];   /* spinning axis in the lab frame, e.g. {1,1,1} for the magic angle */
  double rate_hz;   /* spinning rate */
  int32_t ticks;    /* rotor phases per period; dt = 1/(rate_hz*ticks) */
  int32_t frame;    /* SP_MAS_FRAME_MAGNET or SP_MAS_FRAME_ROTOR */
} sp_rotor;

/* Replaces: the rotor-phase Hamiltonian stack of kernel/utilities/rotor_stack.m:63-151
 * (H_j = I + sum D(phi_j)(k,m) Q{r}{k,m}, phi_j = 2*pi*j/ticks) propagated in
 * Hilbert space tick by tick with P_j = expm(-1i*H_j*dt) (kernel/propagator.m:40-42),
 * rho = P*rho*P' (kernel/step.m:115-137) and detection hdot(coil,rho)
 * (kernel/utilities/hdot.m:23-29), summed over the powder grid with its weights
 * (kernel/contexts/powder.m:280-286). I and Q are Hilbert-space operators
 * (formalism 'zeeman-hilb', dimension N); rho0 and coil are dense N x N
 * column-major interleaved complex (MATLAB layout). The FID has nsteps+1 points
 * f(n*dt), n = 0..nsteps. */
sp_plan* sp_plan_mas(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* alphas,
                     const double* betas, const double* gammas, const double* weights, int64_t n_or,
                     const double* rho0, const double* coil, const sp_rotor* rotor, int64_t nsteps,
                     int mode);

/* One-shot MAS call: plan, execute, copy the result, destroy. */
int sp_mas_fid(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* alphas,
               const double* betas, const double* gammas, const double* weights, int64_t n_or,
               const double* rho0, const double* coil, const sp_rotor* rotor, int64_t nsteps,
               int mode, double* fid_out, double* per_orient_out, sp_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* SPINACH_POWDER_H */
