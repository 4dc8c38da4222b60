/*
 * CPU oracle for the SIMPLE / pressure-correction path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the timed CPU baseline.  The
 * product (cfd_solver_amd, libcfd_gpu.so) never links or calls it.
 *
 * Two groups of functions:
 *
 *  or_*_ref  -- restatements of the reference solver's own stages on its own
 *               periodic O-grid (solver.cpp line numbers cited per function),
 *               iterate-for-iterate where the reference iterates.  Pinned by
 *               golden vectors produced by the compiled reference (oracle/_ref)
 *               and committed under tests/golden/.
 *  or_cav_*  -- the steady cavity SIMPLE algorithm the GPU path implements,
 *               built from the same operators.  The reference has no cavity;
 *               these are pinned by closed forms and invariants (uniform-grid
 *               stencils, row sums, pure conduction Nu = 1, Nu balance between
 *               hot and cold walls, brute-force dense solves on small grids).
 *
 * Layout: every 2-D array is n1 rows of n0 doubles, element (i, j) at j*n0 + i
 * (xi fastest).  A 25-point operator is 25 planes of N = n0*n1 doubles, plane
 * k = (dj+2)*5 + (di+2) holding the coefficient of phi(i+di, j+dj); a 9-point
 * operator is 9 planes, plane k = (dj+1)*3 + (di+1).
 */
#ifndef CFD_ORACLE_H
#define CFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_K25(di, dj) (((dj) + 2) * 5 + ((di) + 2))
#define OR_K9(di, dj) (((dj) + 1) * 3 + ((di) + 1))

/* ---------------- reference mode (periodic xi, O-grid) ---------------- */

/* Momentum / temperature operators I - dt/Re L (solver.cpp:751-939).
 * AU, AT: 25 planes each.  BU, BT: 6 arrays of n0 -- wall/far-field couplings
 * removed from rows j=1 and j=n1-2, order {s, se, sw, n, ne, nw}. */
void or_assemble_mom_ref(int n0, int n1, double h1, double h2,
                         const double *alph, const double *beta, const double *gamma,
                         const double *p1, const double *q1,
                         double dt, double Re, double Pr,
                         double *AU, double *AT, double *BU, double *BT);

/* Pressure-correction 9-point operator (solver.cpp:943-1039), 9 planes. */
void or_assemble_p9_ref(int n0, int n1, double h1, double h2,
                        const double *xix, const double *xiy,
                        const double *etx, const double *ety, double *A9);

/* Contravariant velocities + hybrid convection + RHS (solver.cpp:1062-1236). */
void or_rhs_ref(int n0, int n1, double h1, double h2,
                const double *u, const double *v, const double *T,
                const double *up, const double *vp, const double *p,
                const double *xix, const double *xiy, const double *etx, const double *ety,
                const double *alph, const double *gamma,
                const double *BU, const double *BT,
                double Re, double Pr, double Ri, double dt,
                double *uxi, double *uet,
                double *qu, double *qv, double *qt, double *qup, double *qvp);

/* Lexicographic Gauss-Seidel on a 25/9-point operator (solver.cpp:2279-2445).
 * nw_for_nnw != 0 reproduces the temperature call, which passes the (-1,+1)
 * plane in the (-1,+2) slot (solver.cpp:1289).  Returns iterations used. */
int or_gauss_ref(int n0, int n1, const double *A, int nw_for_nnw,
                 double *phi, const double *q, double tol, int maxit,
                 double *last_ratio);

/* Face-velocity divergence q = div(u_f)/dt (solver.cpp:1389-1448). */
void or_face_div_ref(int n0, int n1, double h1, double h2,
                     const double *up, const double *vp, const double *p,
                     const double *xix, const double *xiy, const double *etx, const double *ety,
                     double dt, double *q);

/* Stone's SIP for the 9-point system (solver.cpp:2066-2277). */
int or_sip9p_ref(int n0, int n1, const double *A9, double *phi, const double *q,
                 double tol, int maxit, double alp, double *last_ratio);

/* Predictor velocities on the outer row: current velocity where the free
 * stream enters, quadratic extrapolation where it leaves (solver.cpp:1351-1381). */
void or_outflow_up_ref(int n0, int n1, double *up, double *vp, const double *u, const double *v,
                       const double *xnox, const double *xnoy, double uinf, double vinf);

/* p' boundary rows after the solve (solver.cpp:1474-1497). */
void or_pcor_bc_ref(int n0, int n1, double *pcor, const double *xnox, const double *xnoy,
                    double uinf, double vinf);

/* u = us - dt grad p', p += p' (solver.cpp:1504-1541). */
void or_correct_ref(int n0, int n1, double h1, double h2,
                    const double *us, const double *vs, const double *pcor,
                    const double *xix, const double *xiy, const double *etx, const double *ety,
                    double dt, double *u, double *v, double *p);

/* Boundary velocities after the correction: far-field prediction from the
 * radial/tangential velocity of row n1-2 and the circulation switch, the
 * oscillating-cylinder wall (cos_ft = cos(2 pi F t)), convected outflow T
 * (solver.cpp:1543-1655).  Returns the circulation along row n1-2. */
double or_farfield_ref(int n0, int n1, double h2, const double *x, const double *y,
                       double *u, double *v, double *T, double *up, double *vp,
                       const double *Told, const double *uet, const double *etx, const double *ety,
                       const double *ajac, const double *xnox, const double *xnoy,
                       double uinf, double vinf, double dt, double speed_amp, double cos_ft);

/* Wall pressure from the normal momentum equation at j = 0 (solver.cpp:1667-1728).
 * alc_u/alc_v: the local-acceleration terms per wall node (zero when F = 0). */
void or_wall_p_ref(int n0, int n1, double h1, double h2,
                   const double *u, const double *v, const double *T, double *p,
                   const double *xix, const double *xiy, const double *etx, const double *ety,
                   const double *alph, const double *beta, const double *gamma,
                   const double *q1, const double *ajac,
                   const double *alc_u, const double *alc_v, double Re, double Ri);

/* Outer-row pressure: normal momentum equation where the free stream enters,
 * Gresho's condition where it leaves (solver.cpp:1731-1782). */
void or_exit_p_ref(int n0, int n1, double h1, double h2,
                   const double *u, const double *v, const double *T, const double *uold,
                   const double *vold, double *p,
                   const double *xix, const double *xiy, const double *etx, const double *ety,
                   const double *alph, const double *beta, const double *gamma,
                   const double *q1, const double *ajac, const double *xnox, const double *xnoy,
                   double uinf, double vinf, double dt, double Re, double Ri);

/* Stream function (solver.cpp:1784-1800). */
void or_si_ref(int n0, int n1, double h2, const double *u, const double *v, const double *xix,
               const double *xiy, const double *ajac, double *si);

/* Vorticity, interior and wall/outer rows (solver.cpp:1802-1887). */
void or_vort_ref(int n0, int n1, double h1, double h2, const double *u, const double *v,
                 const double *xix, const double *xiy, const double *etx, const double *ety, double *vort);

/* cl, cd, cm, CL_pr, CD_pr, CL_vor, CD_vor, Nu into out[8] (solver.cpp:1891-1982). */
void or_forces_ref(int n0, int n1, double h1, double h2, const double *x, const double *y,
                   const double *p, const double *vort, const double *T, const double *ajac,
                   const double *etx, const double *ety, double Re, double alpha, double ar, double *out);

/* Dilation and its signed maximum floored at 0 (solver.cpp:1806-1852). */
double or_dil_ref(int n0, int n1, double h1, double h2,
                  const double *u, const double *v,
                  const double *xix, const double *xiy, const double *etx, const double *ety,
                  double *dil);

/* Wall Nusselt number on the j = 0 wall (solver.cpp:1951-1982) and the local
 * dT/dn profile written to SURF_DIST.dat (solver.cpp:2031-2038). */
double or_nusselt_ref(int n0, int n1, double h1, double h2, const double *T,
                      const double *ajac, const double *etx, const double *ety,
                      double ar, double *dthdn);

/* ---------------- cavity SIMPLE (the GPU path's algorithm) ---------------- */

typedef struct {
    double Re, Ri, Pr;
    double dt;          /* pseudo-time step of the momentum/energy predictor      */
    double alpha_u;     /* implicit (Patankar) relaxation of u, v                 */
    double alpha_p;     /* pressure relaxation p += alpha_p p'                    */
    double alpha_t;     /* implicit relaxation of T                               */
    double lid_u;       /* lid velocity at j = n1-1                               */
    double t_hot, t_cold;
    int mom_sweeps;     /* Jacobi sweeps of the low-order predictor operator      */
    double p_rtol;      /* p' BiCGStab relative tolerance (2-norm)                */
    int p_maxit;        /* p' BiCGStab iteration cap                              */
    int mg_levels;      /* multigrid levels (1 = Jacobi preconditioner only)      */
    int mg_pre, mg_post, mg_coarse; /* smoothing sweeps                            */
    double mg_omega;    /* Jacobi damping in the multigrid smoother               */
    int wall_upwind;    /* 1: rows next to a wall use the hybrid switch too (KK where
                           the five-point stencil fits, first-order upwind where it
                           does not, when |Pe| > 2); 0: central there at any Pe
                           (the reference's rows j = 1, n1-2)                     */
    int fas_levels;     /* grid levels of the nonlinear (FAS) multigrid outer cycle;
                           1 = plain SIMPLE iterations                            */
    int fas_pre, fas_post, fas_coarse; /* SIMPLE iterations per level and cycle   */
    int fas_min_n;      /* no FAS level coarser than this many nodes per side; <= 0:
                           the smallest 2^k + 1 >= 17 (>= 33 when Pr > 1 or
                           |Ri| Re^2 Pr > 2e5) with Re / 2^k <= 32                   */
    int fas_coarse_scheme; /* convection scheme of the coarse FAS levels (wall_upwind
                           values; the finer of it and wall_upwind is used)         */
    double hyb_blend;   /* 0: the reference's switch at |Pe| = 2; > 0: blend from the
                           central to the upwind form over 2 < |Pe| < 2 + hyb_blend  */
    int fas_coarse_pmaxit; /* p' BiCGStab iteration cap on the coarse FAS levels (0:
                           p_maxit; < 0: p' is one V-cycle, vcycle_solve())         */
    int plant;          /* tests only, a deliberate mistake (0: none): 1-4 flips the
                           sign of the normal-momentum term of the wall pressure on
                           wall 0-3 (j=0, j=n1-1, i=0, i=n0-1), 5 the sign of the
                           tangential term of the adiabatic side walls, 6 the sign
                           of the velocity correction                              */
} or_cav_params;

typedef struct {
    double mass_l1;     /* sum |q| of this iteration                              */
    double energy_l1;   /* sum |R_T| of this iteration                            */
    double mom_l1;      /* sum |R_u| + |R_v|                                      */
    double dmax;        /* signed max of the nodal dilation, floored at 0         */
    double nu_hot, nu_cold;
    int p_iters;
    double p_relres;
    double mass_res;    /* mass_l1 / mass_l1 of the level's first SIMPLE iteration */
    double energy_res;  /* energy_l1 likewise                                      */
} or_cav_stats;

typedef struct or_cav_case or_cav_case;

or_cav_case *or_cav_create(int n0, int n1, double h1, double h2,
                           const double *x, const double *y,
                           const double *xix, const double *xiy,
                           const double *etx, const double *ety,
                           const double *alph, const double *beta, const double *gamma,
                           const double *ajac, const double *p1, const double *q1,
                           const or_cav_params *prm);
void or_cav_destroy(or_cav_case *c);
void or_cav_set_fields(or_cav_case *c, const double *u, const double *v,
                       const double *T, const double *p);
void or_cav_get_fields(const or_cav_case *c, double *u, double *v, double *T, double *p);
void or_cav_apply_bcs(or_cav_case *c);
int or_cav_outer(or_cav_case *c, or_cav_stats *st);
/* One outer iteration: a FAS V-cycle (fas_pre SIMPLE iterations, the
 * coarse-grid correction, fas_post SIMPLE iterations; fas_coarse on the
 * coarsest level) when the case has coarse levels, else one SIMPLE iteration;
 * *st gets the last fine SIMPLE iteration's statistics. */
int or_cav_fas_cycle(or_cav_case *c, or_cav_stats *st);
int or_cav_fas_depth(const or_cav_case *c);
/* Residuals of the steady equations at the current state, sources included:
 * R/dt of u, v, T and the mass residual q - s_m (interior; 0 on walls). */
void or_cav_state_residual(or_cav_case *c, double *ru, double *rv, double *rT, double *rm);

/* stage-level entry points of the cavity path (tests compare them one by one) */
void or_cav_lo_diag(const or_cav_case *c, double *d_out);
void or_cav_p_apply(const or_cav_case *c, const double *d, const double *x, double *y);
/* maxit < 0: x = one V-cycle on the mean-free q (vcycle_solve), x overwritten */
int or_cav_p_solve(or_cav_case *c, const double *d, const double *q, double *x,
                   double rtol, int maxit, double *relres);
void or_cav_residuals(or_cav_case *c, double *Ru, double *Rv, double *RT, double *d);
void or_cav_last_q(const or_cav_case *c, double *q, double *pc);
/* q = J-weighted face divergence of the face velocities built from the nodal
 * velocities us, vs, pressure p and factors d (minus the level's mass source). */
void or_cav_face_div(or_cav_case *c, const double *us, const double *vs, const double *p,
                     const double *d, double *q);
/* The correction of an outer iteration: p' copied to the walls (pc_out), then
 * u = us - d grad p', v likewise, p += alpha_p p' on the interior (p in/out). */
void or_cav_correct(or_cav_case *c, const double *us, const double *vs, const double *pc,
                    const double *d, double *u, double *v, double *p, double *pc_out);

#ifdef __cplusplus
}
#endif
#endif
