/* frsdg.h — C-ABI of the MI355X FR residual library (libfrsdg.so).
 *
 * Drop-in replacement for the residual callbacks that the reference
 * (mvtteoleo/fr_sdg) registers with PETSc TS/SNES.  All state arrays are
 * DEVICE pointers (HBM-resident), in the reference's natural DMDA layout:
 *   2D: element (i,j) is the contiguous block starting at (j*nx + i)*4*n*n,
 *       node (k,l) and equation eq at offset ((l*n + k)*4 + eq)
 *       (include/physics_impl_2D.hpp:27-35 Indexer);
 *   1D: element e at e*3*n, equation/node at (eq*n + node)
 *       (include/physics_impl_1D.hpp:19-24).
 * n = p+1.  Arrays must be 16-byte aligned.  Owned elements only (no ghosts):
 * periodic wrap and boundary states are resolved inside.
 *
 * Every entry point returns 0 (FRSDG_OK) on success and a nonzero code
 * otherwise (mirrors PetscErrorCode; nothing throws across the ABI).  One
 * context is not reentrant: do not call it concurrently from two threads.
 * `stream` is a hipStream_t (NULL = default stream); calls are asynchronous
 * on that stream unless stated otherwise.
 */
#ifndef FRSDG_H
#define FRSDG_H
#ifdef __cplusplus
extern "C" {
#endif

#define FRSDG_OK 0
#define FRSDG_ERR_INVALID 1     /* bad descriptor / argument */
#define FRSDG_ERR_HIP 2         /* HIP runtime error */
#define FRSDG_ERR_UNSUPPORTED 3 /* valid but not implemented on this path */
#define FRSDG_ERR_ALIGN 4       /* pointer not 16-byte aligned */
#define FRSDG_ERR_NUMERIC 5     /* zero pivot in a factorisation; failed explicit step */

/* boundary types: values of numPDE::BCType (boundary_conditions.hpp:8-16) plus
 * FRSDG_BC_SLIPWALL, the curvilinear slip wall of tests/bump.cpp:250-272 */
#define FRSDG_BC_WALL 0
#define FRSDG_BC_OUTFLOW 1
#define FRSDG_BC_DIRICHLET 2
#define FRSDG_BC_PERIODIC 5
#define FRSDG_BC_SLIPWALL 6

/* geometry models (geometry_models.hpp:72-195, tests/bump.cpp:13-171) */
#define FRSDG_GEOM_CARTESIAN 0
#define FRSDG_GEOM_DELAVAL 1 /* GeometryDeLavalNozzle2D with StraightLineProfile */
#define FRSDG_GEOM_WEDGE 2   /* GeometryCurvedWedge2D (tests/bump.cpp) */
#define FRSDG_GEOM_BUMP 3    /* channel bump: param {A, x0, xc, w} */
#define FRSDG_GEOM_WARPED 4  /* sine warp: param {A, x0, y0, -} */

typedef struct frsdg_ctx frsdg_ctx;

/* Problem description: the ConfigOptions / ProblemTraits / BoundaryConditions
 * content of the reference (parser.hpp:45-191, physics_evaluator.hpp:37-52,
 * boundary_conditions.hpp:214-324). */
typedef struct frsdg_desc
{
    int    dims;       /* 1 or 2 */
    int    nx, ny;     /* -n_elems per axis */
    int    p;          /* -poly_deg, 1..8 (reference_element.hpp:24, MaxNumElem = 9) */
    double lx, ly;     /* -box_l per axis */
    int    gas_kind;   /* -gas_mod_law: 0 ideal, 1 van der Waals */
    double gamma, R_spec, vdw_a, vdw_b;
    int    riemann;    /* 0 Rusanov, 1 HLLC (riemann_solver.hpp:13-17) */
    int    geometry;   /* FRSDG_GEOM_* */
    double geom_param[4];
    int    bc_type[4]; /* Left, Right, Top, Bottom (boundary_conditions.hpp:17-23) */
    double bc_state[4][4]; /* Dirichlet states (rho, rhoU, rhoV, rhoE) */
    double c_vcjh;     /* -radau_corr */
    int    apply_filter, filt_cutoff, filt_strength; /* -apply_filter -filt_cut_off_deg -filt_strength */
    double filt_alpha;
    int    jac_type;   /* MatShell J*v: 0 forward tangent, 1 complex step (-jac_type complexfd) */
    int    tangent_bc_mode; /* 0 exact linearisation of lambda BCs; 1 reference-compatible:
                               lambda-bound BCs frozen as complex step sees them
                               (boundary_conditions.hpp:270-294) */
    /* WENO limiter with shock sensor in limit_solution (filter.hpp:486-755;
     * parser.hpp:53-62, 277-297): -apply_weno, -weno_gamma (0 < gamma < 1),
     * -shock_sens_theta (> 0), -shock_sens_activation_value.  With the
     * spectral filter also on, the filter's output replaces the limiter's, as
     * in the reference. */
    int    apply_weno;
    double weno_gamma, shock_theta, shock_threshold;
    /* -cond_sys: the implicit path's state holds (log rho, u, [v,] log e) per
     * node (TSSetTransientVariable, petsc_solver_impl.hpp:572-575, 626-655):
     * frsdg_ifunction, the J*v products, frsdg_vjp, frsdg_beuler_step,
     * frsdg_impose_ic and frsdg_l2_error then work through the change of
     * variables C(W) (problem_def.hpp:36-70).  Explicit stepping and
     * frsdg_residual/frsdg_rhs see the stored vector as conservative, as the
     * reference's rhs_func does. */
    int    cond_sys;
    /* coordinate system (Traits::coord_sys, macro_and_libc.hpp:154): 0
     * Cartesian, 1 radial — the residual gains the source terms of
     * physics_impl_1D.hpp:70-90 (r = x; interior elements only, as there) and
     * physics_impl_2D.hpp:306-319 (axisymmetric, r = y/2 at the mapped node) */
    int    coord_sys;
} frsdg_desc;

/* PhysicsEvaluator ctor (physics_evaluator.hpp:188-201) + the once-only
 * precompute of GLL/D/VCJH/filter tables and curved-mesh metrics in HBM. */
int  frsdg_create(const frsdg_desc* d, frsdg_ctx** out);
int  frsdg_destroy(frsdg_ctx* ctx);
long long frsdg_ndofs(const frsdg_ctx* ctx);
/* host copies of the setup tables (GLL csi/w, D (n*n row-major), dL, dR, filter (n*n)) */
int  frsdg_get_tables(const frsdg_ctx* ctx, double* csi, double* w, double* D, double* dL,
                      double* dR, double* filt);

/* residual_kernel<double> (physics_impl_2D.hpp:644-902, physics_impl_1D.hpp:8-94):
 * R = +div F (reference sign). */
int frsdg_residual(frsdg_ctx* ctx, double t, const double* u, double* R, void* stream);
/* rhs_func (petsc_solver_impl.hpp:657-670): rhs = -R(u). */
int frsdg_rhs(frsdg_ctx* ctx, double t, const double* u, double* rhs, void* stream);
/* fun_implicit (petsc_solver_impl.hpp:626-655): F = udot + R(u). */
int frsdg_ifunction(frsdg_ctx* ctx, double t, const double* u, const double* udot, double* F,
                    void* stream);

/* Forward tangent: out = sigma*v + (dR/du)(u) v.  Replaces the complex-step
 * MatShell product (petsc_solver_impl.hpp:270-359) with an exact J*v. */
int frsdg_jvp(frsdg_ctx* ctx, double t, const double* u, const double* v, double sigma, double* out,
              void* stream);
/* Complex-step product of petsc_solver_impl.hpp:270-359 on the GPU:
 * out = sigma*Im(u+ihv)/h + Im R(u+ihv)/h (h = 1e-7 in the reference, :179). */
int frsdg_jvp_cstep(frsdg_ctx* ctx, double t, const double* u, const double* v, double sigma,
                    double h, double* out, void* stream);
/* Adjoint: out = sigma*w + (dR/du)(u)^T w. No reference counterpart (new). */
int frsdg_vjp(frsdg_ctx* ctx, double t, const double* u, const double* w, double sigma, double* out,
              void* stream);

/* pre_jv_eval (petsc_solver_impl.hpp:223-248): snapshot of the linearisation
 * point (u stays owned by the caller and must stay valid), sigma and time. */
int frsdg_pre_jv_eval(frsdg_ctx* ctx, double t, const double* u, double sigma);
/* matmult (petsc_solver_impl.hpp:250-267): out = sigma*v + J v at the snapshot,
 * by forward tangent (jac_type 0) or complex step with h = 1e-7 (jac_type 1). */
int frsdg_matmult(frsdg_ctx* ctx, const double* v, double* out, void* stream);

/* limit_solution with the spectral filter (physics_impl_2D.hpp:76-124,
 * physics_impl_1D.hpp:96-121, filter.hpp:425-445,714-755).  out may equal u. */
int frsdg_filter(frsdg_ctx* ctx, const double* u, double* out, void* stream);
/* apply_stabilization (petsc_solver_impl.hpp:683-702): filter u in place if
 * the filter is active, no-op otherwise. */
int frsdg_apply_stabilization(frsdg_ctx* ctx, double* u, void* stream);

/* Transpose of the (linear) filter map, for adjoints: out = F^T w. */
int frsdg_filter_transpose(frsdg_ctx* ctx, const double* w, double* out, void* stream);
/* Quadrature weight of every DOF (device array of ndofs): w_k w_l detJ in 2D
 * (compute_error, physics_impl_2D.hpp:253), (h/2) w_k in 1D. */
int frsdg_mass(frsdg_ctx* ctx, double* m, void* stream);
/* Boundary quadrature of face direction dir (0 Left, 1 Right, 2 Top, 3 Bottom):
 * per face node along that boundary (element-major, then node), the pair
 * (w ds n_x, w ds n_y) with n the face normal of the geometry model. */
int frsdg_boundary_weights(frsdg_ctx* ctx, int dir, double* out, void* stream);

/* One explicit step u <- S(u) with rhs = -R: stages 3 = Shu-Osher SSPRK(3,3),
 * stages 4 = SSPRK(4,3) (PETSc TSSSP rks3 default), each stage fused with its
 * residual; post-step filter (TSSetPostStep, petsc_solver_impl.hpp:604-605). */
int frsdg_ssprk_step(frsdg_ctx* ctx, double t, double dt, double* u, int stages, void* stream);
/* nsteps steps of frsdg_ssprk_step, captured once into a hipGraph (16 steps) and
 * replayed.  Every stage output is checked where it is stored (health flags:
 * bit 1 a non-finite value, bit 2 rho <= 0, bit 4 internal energy <= 0 at a
 * node); the run stops after the first 16-step graph (or the remaining plain
 * steps) that raised a flag of the stop mask and returns FRSDG_ERR_NUMERIC
 * (-ts_error_if_step_fails, tools/check_radial.py:208-210), u holding the last
 * state computed.  Synchronises the stream once per graph. */
int frsdg_ssprk_run(frsdg_ctx* ctx, double t, double dt, double* u, int stages, int nsteps,
                    void* stream);
/* Health flags that stop frsdg_ssprk_run (a subset of 1 | 2 | 4); the default 1
 * stops on non-finite states only: the reference's explicit runs carry on
 * through transient non-positive nodal values (the floors of the flux keep
 * them finite), which frsdg_step_health still reports.  No counterpart in the
 * reference (its step-failure test is PETSc's). */
int frsdg_set_stop_mask(frsdg_ctx* ctx, int mask);
/* The stage-output health flags that frsdg_ssprk_step calls accumulated (bit 1
 * non-finite, bit 2 rho <= 0, bit 4 internal energy <= 0), cleared when clear
 * is set; frsdg_ssprk_run clears them when it starts.  Synchronises the stream. */
int frsdg_step_health(frsdg_ctx* ctx, int* flags, int clear, void* stream);
/* The same flags for a conservative state u.  Synchronises the stream. */
int frsdg_check_state(frsdg_ctx* ctx, const double* u, int* flags, void* stream);

/* impose_ic (physics_impl_2D.hpp:156-201, physics_impl_1D.hpp:207-244) for the
 * presets of factory.hpp: kind 0 uniform (rho,u,v,p), 1 isentropic vortex
 * (u0,v0,beta,r_v,l), 2 circular explosion (rhoL,pL,rhoR,pR,radius,xc,yc,width: a
 * width > 0 smooths the jump over a tanh layer),
 * 3 Sod 1D (rhoL,uL,pL,rhoR,uR,pR), and 4 a smooth pressure pulse at rest
 * (A,w,xc,yc,rho0,p0: rho = rho0(1+A g), p = p0(1+A g)^gamma, g = exp(-r^2/w^2)).
 * prm is a host array of 8 doubles. */
int frsdg_impose_ic(frsdg_ctx* ctx, int kind, const double* prm, double* u, void* stream);
/* compute_error (physics_impl_2D.hpp:203-265) against the same presets at time
 * t; out is a HOST array of 4 doubles.  Synchronises the stream. */
int frsdg_l2_error(frsdg_ctx* ctx, int kind, const double* prm, double t, const double* u,
                   double* out, void* stream);

/* Allocation-free frsdg_l2_error (same kernels, same result, bitwise): no
 * counterpart in the reference; for callers that cannot allocate device memory
 * (serving loops, harnesses that forbid hipMalloc, code that must not grow the
 * device heap).  Its reduction buffer is a file-scope device array of
 * 4 * (FRSDG_L2_SCRATCH_BLOCKS + 1) doubles (2 MiB) that every GPU the library
 * runs on holds until the process ends; a context with more than
 * 256 * FRSDG_L2_SCRATCH_BLOCKS = 2^24 nodes (E n^2) is rejected with
 * FRSDG_ERR_INVALID.  All calls on one GPU share that array: calls that overlap
 * on one GPU (two streams, two threads) race.  Synchronises the stream. */
#define FRSDG_L2_SCRATCH_BLOCKS 65536
int frsdg_l2_error_noalloc(frsdg_ctx* ctx, int kind, const double* prm, double t, const double* u, double* out,
                           void* stream);

/* ---------------- implicit steps: Jacobian-free Newton-Krylov ----------------
 * The TSSetIFunction / TSSetIJacobian path of the reference
 * (petsc_solver_impl.hpp:572-579, 626-655, 168-267) with PETSc's SNES newtonls
 * + KSP GMRES replaced by a device solver: Newton with PETSc's default
 * cubic backtracking line search (SNESLINESEARCHBT: Armijo alpha 1e-4,
 * steps clamped to [0.1, 0.5] of the previous, 40 tries), restarted GMRES (classical Gram-Schmidt, right preconditioning) whose
 * operator is the MatShell product sigma*v + J v from the forward-tangent or
 * the complex-step kernel. */
typedef struct frsdg_newton_opts
{
    int    jac_type;    /* 0 forward tangent, 1 complex step (h = 1e-7) */
    int    precond;     /* 0 none (-pc_type none), 1 element block-Jacobi, 2 red-black
                           element block Gauss-Seidel (one forward sweep, matrix-free
                           off-diagonal coupling: one extra J*v per application),
                           3 ILU(0) of the coloured finite-difference matrix (the
                           reference's preconditioning matrix with PETSc's default
                           one-process PC), 4 ILU(0) of the same pattern with exact
                           tangent columns (frsdg_color_jacobian methods 0 and 2) */
    double snes_rtol;   /* stop when ||F|| <= snes_rtol * ||F_0|| (PETSc default 1e-8) */
    double snes_atol;   /* or ||F|| <= snes_atol (PETSc default 1e-50) */
    int    snes_max_it; /* PETSc default 50 */
    int    ksp_restart; /* PETSc default 30 */
    int    ksp_max_it;  /* PETSc default 10000 */
    double ksp_rtol;    /* PETSc default 1e-5 */
    int    ksp_cgs_refine; /* 0 one classical Gram-Schmidt pass (PETSc default), 1 always two */
    double snes_stol;   /* also stop when ||lambda du|| <= snes_stol * ||u|| after a step
                           (SNES_CONVERGED_SNORM_RELATIVE; PETSc default 1e-8, 0 off) */
    int    ksp_pc_side; /* 0 right preconditioning, test on the true residual norm;
                           1 left, test on the preconditioned residual norm (PETSc's
                           GMRES default, -ksp_pc_side left); 2 flexible right
                           preconditioning (-ksp_type fgmres): the solution is formed
                           from the stored preconditioned basis, so its true residual
                           is the estimate even when the preconditioner's apply carries
                           large rounding (ILU(0) of a badly conditioned matrix) */
    int    pc_lag;      /* rebuild the preconditioner every pc_lag-th Newton iteration,
                           counted across consecutive frsdg_beuler_step calls and the
                           steps of one frsdg_pseudo_run (-snes_lag_preconditioner k
                           -snes_lag_preconditioner_persists); <= 1 every iteration
                           (PETSc default) */
    int    pc_fp32;     /* preconditioners 3/4: apply single-precision copies of the ILU(0)
                           factors (accumulation in double); halves the bytes each
                           application reads, but where the factors are badly
                           conditioned (the 256x64 p=5 bump matrix at free stream)
                           their rounding leaves no useful preconditioner.  The
                           transposed apply stays double. */
} frsdg_newton_opts;

typedef struct frsdg_newton_stats
{
    /* converged: 1 by the residual tests (rtol / atol), 2 by the step-norm test
       (snes_stol), 0 not converged */
    int    newton_its, ksp_its, jv_evals, res_evals, converged;
    double fnorm0, fnorm;
    double seconds; /* wall time of the solve, stream-synchronised */
    /* where the time went (GPU seconds between events on the stream): building
       the preconditioner (pc_builds times: the coloured matrix or the probed
       blocks, then their ILU(0) / inverses), applying it, the J*v products of
       GMRES and its Gram-Schmidt orthogonalisation; the rest of `seconds` is
       residual evaluations, the line search and host work */
    int    pc_builds;
    double t_pc_assemble, t_pc_factor, t_pc_apply, t_jv, t_orth;
} frsdg_newton_stats;

/* One backward-Euler step: solves F(u) = (u - u^n)/dt + R(u) = 0 for u^{n+1}.
 * u holds u^n on entry and u^{n+1} on return.  With the context's filter or
 * WENO limiter on, frsdg_apply_stabilization runs on u^{n+1} after the solve,
 * as the reference's TSSetPostStep / TSSetPostEvaluate callback does once per
 * implicit step (petsc_solver_impl.hpp:582-596).  Synchronises the stream. */
int frsdg_beuler_step(frsdg_ctx* ctx, double t, double dt, double* u, const frsdg_newton_opts* opts,
                      frsdg_newton_stats* stats, void* stream);

/* Pseudo-transient continuation, PETSc TSPSEUDO as the reference's steady
 * harness drives it (-ts_type pseudo -ts_pseudo_increment inc
 * [-ts_pseudo_increment_dt_from_initial_dt] -ts_pseudo_fatol 0
 * -ts_pseudo_frtol 0, tools/boundary_conforming_claude.py:113-127):
 * backward-Euler steps (frsdg_beuler_step) whose size follows
 * TSPseudoTimeStepDefault, evaluated before each step as TSStep_Pseudo does:
 * with f_k = ||F(u_k, U_t = 0)||_2 (the IFunction at the current state, i.e.
 * the steady residual) and f_0 that of the initial state, step k takes
 * inc * dt0 * f_0 / f_k (from_initial_dt) or inc * dt_{k-1} * f_{k-1} / f_k
 * (dt_{-1} = dt0, f_{-1} = f_0), capped at dt_max when dt_max > 0 (so the
 * first step takes inc * dt0).  Stops after max_steps steps or when t reaches
 * t_max (t_max > 0; the last step is shortened to land on it,
 * -ts_exact_final_time matchstep).  dt_hist, fnorm_hist and time_hist
 * (max_steps entries, may be NULL) get each step's dt, the f after it and the
 * wall seconds since the start (stream-synchronised); *steps the number of
 * steps taken; stats (may be NULL) the per-step Newton statistics summed over
 * the run, with fnorm0 = f_0 and fnorm the last f.  u is updated in place.
 * Synchronises the stream. */
int frsdg_pseudo_run(frsdg_ctx* ctx, double t, double dt0, double increment, int from_initial_dt, double dt_max,
                     double t_max, int max_steps, double* u, const frsdg_newton_opts* opts, double* dt_hist,
                     double* fnorm_hist, double* time_hist, int* steps, frsdg_newton_stats* stats, void* stream);

/* The preconditioning matrix of the reference's implicit path
 * (IJacobianStrategy::setup_colored_preconditioner and MatFDColoringApply
 * through ts_coloring_wrapper, petsc_solver_impl.hpp:55-95, 126-160, 221-237):
 * P = dF/du at u of F(x) = IFunction(t, x, sigma (T(x) - T(u_n))), T the
 * transient variable (identity, or prim -> cons under cond_sys); u_n NULL
 * means u_n = u.  The DMDA box stencil of width 1 makes P block sparse:
 * P holds E * S blocks of nd x nd doubles (S = 9 in 2D, 3 in 1D; nd = DOFs per
 * element), block (e, slot) = dF_e/du_m for the neighbour m at slot
 * (dj+1)*3 + (di+1) (1D: di+1), column-major, zero where there is no
 * neighbour.  method 0: coloured finite differences as MatFDColoringApply
 * with PETSc's defaults (colouring of DMCreateColoring, step
 * h = sqrt(eps) sqrt(1 + ||u||_2), "wp"); 1: per-column step
 * h = sqrt(eps) u_i with |u_i| raised to 100 sqrt(eps) ("ds"); 2: exact
 * columns from the tangent product, same colouring.  Periodic directions
 * need a multiple of 3 elements (PETSc's colouring requires it), else
 * FRSDG_ERR_UNSUPPORTED.  Synchronises the stream. */
int frsdg_color_jacobian(frsdg_ctx* ctx, double t, const double* u, const double* u_n, double sigma, int method,
                         double* P, void* stream);

/* ILU(0) of a matrix in the frsdg_color_jacobian layout, in place, natural
 * element order, no pivoting: PETSc's default PC of a one-process run
 * (-pc_type ilu, levels 0) on the reference's preconditioning matrix.  On
 * return each diagonal block holds its LU factors (unit-lower L_II below the
 * diagonal, U_II on and above), each block below the diagonal L_II^-1 L_IK
 * and each block above U_II^-1 U_IJ (scaled so that the apply's rows are
 * independent); Dinv (E * 2 * nd * nd doubles) receives L_II^-1 and U_II^-1.
 * FRSDG_ERR_NUMERIC on a pivot of magnitude <= 1e-12 (PETSc's default
 * zero-pivot tolerance, MAT_FACTOR_NUMERIC_ZEROPIVOT).  Synchronises the
 * stream. */
int frsdg_ilu0_factor(frsdg_ctx* ctx, double* P, double* Dinv, void* stream);

/* z = (L U)^-1 x with the factors of frsdg_ilu0_factor (PCApply of ILU);
 * z may alias x.  Synchronises the stream (it checks that the
 * dependence-ordered sweeps of 2D blocks completed every wait). */
int frsdg_ilu0_apply(frsdg_ctx* ctx, const double* P, const double* Dinv, const double* x, double* z,
                     void* stream);

/* z = (L U)^-T x with the same factors (PCApplyTranspose of ILU; the
 * preconditioner of frsdg_adjoint_solve with precond 3/4); z may alias x. */
int frsdg_ilu0_apply_transpose(frsdg_ctx* ctx, const double* P, const double* Dinv, const double* x, double* z,
                               void* stream);

/* The linear solve of one Newton iteration on its own: (sigma I + dR/du(u)) x = b
 * by the device GMRES (PETSc KSPSolve with the options' KSP settings and
 * preconditioner; precond 3/4 build the coloured matrix of
 * F(x) = IFunction(t, x, sigma (T(x) - T(u))) at u).  hist (host, hist_cap
 * entries, may be NULL with hist_cap 0) receives the residual history as
 * -ksp_monitor / KSPGetResidualHistory report it for right-preconditioned
 * GMRES: ||b||, then the GMRES estimate of ||b - A x_j|| after every
 * iteration j; *hist_len its length.  stats: ksp_its, fnorm0 = ||b||,
 * fnorm = the true final residual.  Synchronises the stream. */
int frsdg_linear_solve(frsdg_ctx* ctx, double t, double sigma, const double* u, const double* b, double* x,
                       const frsdg_newton_opts* opts, frsdg_newton_stats* stats, double* hist, int hist_cap,
                       int* hist_len, void* stream);

/* Adjoint linear solve (sigma I + dR/du(u))^T lambda = rhs by GMRES on the
 * adjoint kernel (opts: ksp_* fields; precond 0..2 on probed blocks, 3/4 the
 * transposed ILU(0) of the coloured matrix of F at u; jac_type for the block
 * probes).  No reference counterpart.  Synchronises the stream. */
int frsdg_adjoint_solve(frsdg_ctx* ctx, double t, double sigma, const double* u, const double* rhs,
                        double* lambda, const frsdg_newton_opts* opts, frsdg_newton_stats* stats,
                        void* stream);

/* Unsteady discrete adjoint of explicit SSPRK(3,3) stepping (frsdg_ssprk_step
 * with stages 3: rhs = -R, post-step spectral filter when active): lam_0 =
 * (d u_N / d u_0)^T lam_T for u_N = S^N(u_0) with steps of size dt.  No
 * reference counterpart.  The states the reverse sweep needs are recomputed
 * from at most max_snapshots stored states (binomial / revolve schedule of
 * Griewank: a segment of T steps with s free slots and r repetitions,
 * C(s + r, s) >= T, reverses its last min(T - 1, C(s - 1 + r, s - 1)) steps
 * first from a new snapshot).  Each step back recomputes the step's first two
 * stage states and applies the transposed stage combinations through three
 * adjoint-kernel launches (the combinations fused into their outputs) and the
 * transposed filter.  Device memory: (min(max_snapshots, N - 1) + 6) states,
 * kept by the context for later calls.  lam_0 may alias lam_T.  The WENO
 * limiter has no transpose: FRSDG_ERR_UNSUPPORTED when it is the post-step
 * stabiliser.  Synchronises the stream. */
typedef struct frsdg_adjoint_stats
{
    int    forward_steps;  /* forward steps recomputed from snapshots */
    int    adjoint_steps;  /* steps reversed (= N) */
    int    snapshots_used; /* largest number of snapshots held at once */
    double seconds;        /* wall time, stream-synchronised */
} frsdg_adjoint_stats;
int frsdg_unsteady_adjoint(frsdg_ctx* ctx, double t, double dt, int nsteps, int max_snapshots, const double* u0,
                           const double* lam_T, double* lam_0, frsdg_adjoint_stats* stats, void* stream);

/* ---------------- batched sweep of independent problems ----------------
 * No reference counterpart (the reference runs one problem per process).
 * Instances may differ in mesh size, order, gas model and BCs (2D only).  Their
 * states are concatenated in one device array; frsdg_batch_offsets gives the
 * first DOF of each (count+1 entries).  One SSP-RK3 step = 3 persistent-kernel
 * launches over all (instance, element-tile) work items, the post-step filter
 * fused into the last one. */
typedef struct frsdg_batch frsdg_batch;
int       frsdg_batch_create(const frsdg_desc* descs, int count, frsdg_batch** out);
int       frsdg_batch_destroy(frsdg_batch* b);
long long frsdg_batch_ndofs(const frsdg_batch* b);
int       frsdg_batch_offsets(const frsdg_batch* b, long long* off);
/* the per-instance context (owned by the batch; do not destroy) */
int       frsdg_batch_instance(frsdg_batch* b, int k, frsdg_ctx** out);
/* per-instance time steps (host array of count doubles), copied to the device; synchronous */
int       frsdg_batch_set_dt(frsdg_batch* b, const double* dt);
/* launch shape: persistent workgroups (0: one workgroup per element tile), element
 * tile = default tile of the order (epb_div 1) or that tile with its rows (columns if it
 * has one row) halved (epb_div 2), and mode 0 (one kernel per stage over all orders),
 * 1 (one kernel per stage, order and mesh kind) or 2 (as 1, cut into runs of instances
 * whose stage states fit in about half of the last-level cache, each run taking its
 * three stages back to back) */
int       frsdg_batch_set_tuning(frsdg_batch* b, int grid, int epb_div, int mode);
/* element tile of the instances of order n = p+1 (after frsdg_batch_set_tuning,
 * which resets every order to its epb_div choice): shape 0 the default tile of
 * the order, 1 that tile halved, 2 the other tile with the fewest rim faces per
 * element (n = 7: 2 x 2 elements instead of 5 x 1) */
int       frsdg_batch_set_tile_shape(frsdg_batch* b, int n, int shape);
/* Every stage output is checked in the stage kernel's store, as in
 * frsdg_ssprk_step, into a flag word per instance (bit 1 non-finite, bit 2
 * rho <= 0, bit 4 internal energy <= 0). */
int       frsdg_batch_ssprk_step(frsdg_batch* b, double* u, void* stream);
/* nsteps steps replayed from a captured hipGraph; clears the flags when it
 * starts and stops with FRSDG_ERR_NUMERIC after the first graph of steps in
 * which an instance raised a flag of its own stop mask (frsdg_set_stop_mask on
 * frsdg_batch_instance; default: non-finite states).  Synchronises the stream
 * once per graph. */
int       frsdg_batch_ssprk_run(frsdg_batch* b, double* u, int nsteps, void* stream);
/* the flag words of the instances (host array of count ints), cleared when clear
 * is set; synchronises the stream.  No counterpart in the reference. */
int       frsdg_batch_health(frsdg_batch* b, int* flags, int clear, void* stream);
/* the same step with one set of launches per instance (comparison point) */
int       frsdg_batch_ssprk_step_each(frsdg_batch* b, const double* dt, double* u, void* stream);

const char* frsdg_strerror(int code);

#ifdef __cplusplus
}
#endif
#endif
