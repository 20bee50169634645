/* fr_oracle.h — C entry points of the CPU ORACLE (test infrastructure only).
 *
 * The oracle restates the reference FR residual on the CPU so that tests,
 * the smoke check and the bench's CPU-baseline leg can check / time the HIP
 * product against it.  Nothing in the product links or loads this library.
 *
 * Layout of every state array is the reference's natural DMDA order
 * (physics_impl_2D.hpp:27-35, physics_impl_1D.hpp:19-24):
 *   2D: element (i,j) at ((j*nx + i) * 4*n*n), node (k,l), eq at ((l*n+k)*4+eq)
 *   1D: element e at (e * 3*n), eq/node at (eq*n + node)
 */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

typedef struct fro_desc
{
    int    dims;        /* 1 or 2 */
    int    nx, ny;      /* elements per axis (ny ignored in 1D) */
    int    p;           /* polynomial degree, 1..8 */
    double lx, ly;      /* box lengths (parser.hpp -box_l) */
    int    gas_kind;    /* 0 ideal, 1 van der Waals */
    double gamma, R_spec, vdw_a, vdw_b;
    int    riemann;     /* 0 Rusanov, 1 HLLC */
    int    geometry;    /* 0 Cartesian, 1 de Laval (straight), 2 wedge, 3 bump, 4 warped */
    double geom_param[4];
    int    bc_type[4];  /* Left, Right, Top, Bottom: 0 Wall 1 Outflow 2 Dirichlet 5 Periodic 6 SlipWall */
    double bc_state[4][4];
    double c_vcjh;      /* -radau_corr */
    int    apply_filter;
    int    filt_cutoff, filt_strength;
    double filt_alpha;
    int    apply_weno;       /* -apply_weno */
    double weno_gamma;       /* -weno_gamma (linear weight of the own element) */
    double shock_theta;      /* -shock_sens_theta */
    double shock_threshold;  /* -shock_sens_activation_value */
    int    cond_sys;         /* -cond_sys: implicit path on (log rho, u, [v,] log e) */
    int    coord_sys;        /* 0 Cartesian, 1 radial (source terms) */
} fro_desc;

void* fro_create(const fro_desc* d);
/* threads of the OpenMP element loops; returns the previous setting */
int   fro_set_threads(int nthreads);
void  fro_destroy(void* h);
long  fro_ndofs(void* h);
void  fro_tables(void* h, double* csi, double* w, double* D, double* dL, double* dR, double* filt);

/* residual_kernel<double>: R = +div F */
int fro_residual(void* h, double t, const double* u, double* R);
/* residual_kernel<complex<double>> on (ure + i uim) */
int fro_residual_cplx(void* h, double t, const double* ure, const double* uim, double* Rre,
                      double* Rim);
/* ComplexStepJacobianStrategy::matmult: out = sigma*Im(q)/h + Im R(q)/h with
 * q = u + i h v, or q = C(u + i h v) when cond_sys */
int fro_jvp_cstep(void* h, double t, const double* u, const double* v, double sigma,
                  double hstep, double* out);
/* fun_implicit (petsc_solver_impl.hpp:626-655): F = U_t + R(C(U)) with C the
 * prim->cons map when cond_sys, identity otherwise */
int fro_ifunction(void* h, double t, const double* u, const double* udot, double* F);
/* prim_to_cons_kernel / cons_to_prim_elem node by node (problem_def.hpp:36-70) */
int fro_prim_to_cons(void* h, const double* w, double* c);
int fro_cons_to_prim(void* h, const double* c, double* w);
/* limit_solution: spectral filter and/or WENO limiter with shock sensor */
int fro_filter(void* h, const double* u, double* out);
/* one explicit step (stages 3: Shu-Osher SSPRK(3,3); 4: SSPRK(4,3)) + post-step filter */
int fro_ssprk_step(void* h, double t, double dt, double* u, int stages);
/* impose_ic: kind 0 uniform(rho,u,v,p) 1 vortex(u0,v0,beta,rv,l) 2 explosion(rhoL,pL,rhoR,pR,r,xc,yc)
 * 3 sod1d(rhoL,uL,pL,rhoR,uR,pR) 4 smooth pulse(A,w,xc,yc,rho0,p0) */
int fro_impose_ic(void* h, int kind, const double* prm, double* u);
/* compute_error against the same kinds as exact solution at time t */
int fro_l2_error(void* h, int kind, const double* prm, double t, const double* u, double* out);
/* pointwise pieces */
int fro_flux2d(void* h, const double* U, double* F, double* G);
int fro_riemann2d(void* h, const double* UL, const double* UR, double nx, double ny, double* out);
int fro_riemann1d(void* h, const double* UL, const double* UR, double* out);
/* composite metrics at computational coords: xi_x, xi_y, eta_x, eta_y, detJ, nx_x, ny_x, nx_y, ny_y */
int fro_metrics(void* h, double cx, double cy, double* out);

#ifdef __cplusplus
}
#endif
