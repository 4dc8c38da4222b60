/*
 * cfd_gpu.h -- C ABI of the MI355X (gfx950) SIMPLE / pressure-correction path.
 *
 * All arrays are FP64, contiguous, n1 rows of n0 values: node (i, j) at
 * j*n0 + i (xi fastest).  Host pointers unless stated otherwise.  Every
 * function returns CFD_OK (0) or a negative CFD_E_* code; nothing throws.
 *
 * Two groups of entry points:
 *
 *  1. Reference stages (cfd_ref_*, cfd_sip9p, cfd_gauss) -- one call per
 *     stage of the reference solver's time step on its periodic O-grid
 *     (column n0-1 duplicates column 0).  Each replaces the cited region of
 *     solver.cpp (PSolver.cpp is the same code one line lower) and takes the
 *     same arrays the reference's Solver members hold, converted from the
 *     reference's double** rows (a[i][j]) to the contiguous layout above.
 *     25-point operators are 25 planes of n0*n1 values, plane (dj+2)*5+(di+2)
 *     holding the coefficient of phi(i+di, j+dj); 9-point ones 9 planes,
 *     plane (dj+1)*3+(di+1).
 *
 *  2. Cavity cases (cfd_case_*) -- a device-resident steady mixed-convection
 *     cavity solved by the SIMPLE-type outer iteration built from the same
 *     operators; the case owns its HBM buffers and its HIP stream.
 *
 * Threading: one case per host thread at a time; stage entry points are
 * re-entrant (they allocate their own device scratch per call).
 */
#ifndef CFD_GPU_H
#define CFD_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define CFD_OK 0
#define CFD_E_ARG (-1)      /* bad size / null pointer / unsupported shape        */
#define CFD_E_OOM (-2)      /* device allocation failed                           */
#define CFD_E_NAN (-3)      /* a residual or field became NaN/Inf                 */
#define CFD_E_NOTCONV (-4)  /* iteration cap reached before the tolerance         */
#define CFD_E_HIP (-5)      /* HIP runtime error (message on stderr)              */
#define CFD_E_STATE (-6)    /* call out of order (e.g. fields never set)          */
#define CFD_E_FULL (-7)     /* no free slot for the case (cfd_pack_add_case)      */

/* Library / device information. */
const char *cfd_version(void);
int cfd_device_count(int *count);

/* ------------------------------------------------------------------ */
/* 1. Reference stages                                                 */
/* ------------------------------------------------------------------ */

/* Momentum / temperature operators I - dt/Re L, second order at j = 1 and
 * j = n1-2, fourth order inside; wall/far-field couplings of those rows moved
 * into BU/BT (6 arrays of n0: s, se, sw, n, ne, nw).
 * Replaces solver.cpp:751-939. */
int cfd_ref_assemble_mom(int n0, int n1, double dxi, double deta,
                         const double *alph, const double *beta, const double *gamma,
                         const double *p1, const double *q1,
                         double dt, double Re, double Pr,
                         double *AU, double *AT, double *BU, double *BT);

/* Pressure-correction 9-point operator.  Replaces solver.cpp:943-1039. */
int cfd_ref_assemble_p9(int n0, int n1, double dxi, double deta,
                        const double *xix, const double *xiy,
                        const double *etx, const double *ety, double *A9);

/* Contravariant velocities and the u/v/T right-hand sides (hybrid convection,
 * pressure gradient, buoyancy, boundary-row sums).
 * Replaces solver.cpp:1062-1236. */
int cfd_ref_rhs(int n0, int n1, double dxi, double deta,
                const double *u, const double *v, const double *T,
                const double *up, const double *vp, const double *p,
                const double *xix, const double *xiy, const double *etx, const double *ety,
                const double *alph, const double *gamma,
                const double *BU, const double *BT,
                double Re, double Pr, double Ri, double dt,
                double *uxi, double *uet,
                double *qu, double *qv, double *qt, double *qup, double *qvp);

/* Solve A phi = q for the interior rows 1..n1-2 of the periodic O-grid with a
 * 25-point operator given as the reference's 25 arrays, in the argument order
 * of Solver::gauss (solver.cpp:2279-2285).  phi: in = initial guess and the
 * fixed boundary rows, out = solution.  Solved on the GPU by BiCGStab to
 * ||q - A phi||_2 <= rtol ||q - A phi0||_2 (replaces the Gauss-Seidel sweeps
 * of solver.cpp:2279-2445, whose fixed point it shares). */
int cfd_gauss(int n0, int n1,
              const double *ap, const double *ae, const double *as, const double *an,
              const double *aw, const double *ase, const double *asw, const double *ane,
              const double *anw, const double *ass, const double *assee, const double *assww,
              const double *asse, const double *assw, const double *asee, const double *asww,
              const double *ann, const double *annee, const double *annww, const double *anne,
              const double *annw, const double *anee, const double *anww, const double *aee,
              const double *aww, double *phi, const double *q,
              double rtol, int maxit, int *iters, double *relres);

/* Face-velocity divergence q = div(u_f)/dt.  Replaces solver.cpp:1389-1448. */
int cfd_ref_face_div(int n0, int n1, double dxi, double deta,
                     const double *up, const double *vp, const double *p,
                     const double *xix, const double *xiy, const double *etx, const double *ety,
                     double dt, double *q);

/* Solve the 9-point pressure-correction system, argument order of
 * Solver::sip9p (solver.cpp:2066-2068).  Same conventions as cfd_gauss;
 * replaces Stone's SIP of solver.cpp:2066-2277.  BiCGStab preconditioned by exact
 * tridiagonal solves along the eta lines (columns) when n1 - 2 <= 1024, else by the
 * diagonal; the iteration is controlled on the device (the host waits on a flag in
 * mapped memory, one iteration behind the launches). */
int cfd_sip9p(int n0, int n1,
              const double *ap, const double *ae, const double *as, const double *an,
              const double *aw, const double *ase, const double *asw, const double *ane,
              const double *anw, double *phi, const double *q,
              double rtol, int maxit, int *iters, double *relres);

/* cfd_sip9p that allocates no memory (no hipMalloc, no library allocation): the
 * device buffers of the solve live in one static array of the library, sized for
 * n0 * n1 <= CFD_SIP9P_NOALLOC_MAX_NODES (the reference's array bounds, 350 x 570;
 * about 33 MB, held on every device the library runs on until the process ends).
 * Larger systems return CFD_E_ARG.  Same arguments, results and iteration counts
 * as cfd_sip9p.  Calls that overlap on one device share the array and race.
 * No counterpart in the reference: this entry point is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation). */
#define CFD_SIP9P_NOALLOC_MAX_NODES (350 * 570)
int cfd_sip9p_noalloc(int n0, int n1,
                      const double *ap, const double *ae, const double *as, const double *an,
                      const double *aw, const double *ase, const double *asw, const double *ane,
                      const double *anw, double *phi, const double *q,
                      double rtol, int maxit, int *iters, double *relres);
/* Predictor velocities up/vp on the outer row j = n1-1: u/v where the free
 * stream enters (uinf*xnox + vinf*xnoy >= 0), else quadratic extrapolation from
 * the three rows inside.  Replaces solver.cpp:1351-1381. */
int cfd_ref_outflow_up(int n0, int n1, double *up, double *vp, const double *u, const double *v,
                       const double *xnox, const double *xnoy, double uinf, double vinf);

/* p' boundary rows after the solve: wall copy and far-field 0/copy by the
 * sign of the free-stream normal velocity.  Replaces solver.cpp:1474-1497. */
int cfd_ref_pcor_bc(int n0, int n1, double *pcor, const double *xnox, const double *xnoy,
                    double uinf, double vinf);

/* u = us - dt grad p', v likewise, p += p'.  Replaces solver.cpp:1504-1541. */
int cfd_ref_correct(int n0, int n1, double dxi, double deta,
                    const double *us, const double *vs, const double *pcor,
                    const double *xix, const double *xiy, const double *etx, const double *ety,
                    double dt, double *u, double *v, double *p);

/* Wall (j = 0) pressure from the normal momentum equation; alc_u/alc_v are the
 * per-node local accelerations (may be NULL for zero).
 * Replaces solver.cpp:1667-1728. */
int cfd_ref_wall_pressure(int n0, int n1, double dxi, double deta,
                          const double *u, const double *v, const double *T, double *p,
                          const double *xix, const double *xiy, const double *etx, const double *ety,
                          const double *alph, const double *beta, const double *gamma,
                          const double *q1, const double *ajac,
                          const double *alc_u, const double *alc_v, double Re, double Ri);

/* Boundary velocities after the correction: on the outer row the free stream
 * (T = 0) where it enters, where it leaves the far-field prediction from row
 * n1-2's radial/tangential velocity (tangential decay exponent 1 when the
 * circulation along row n1-2 exceeds 1e-2 in magnitude, else 2) with T convected
 * from Told by uet (both the values at the start of the step); on the wall row
 * the oscillating cylinder's velocity (-a y, a x), a = speed_amp cos(2 pi F time),
 * copied into up/vp.  circ (may be NULL) receives the circulation.
 * Replaces solver.cpp:1543-1655. */
int cfd_ref_farfield(int n0, int n1, double deta, const double *x, const double *y,
                     double *u, double *v, double *T, double *up, double *vp,
                     const double *Told, const double *uet, const double *etx, const double *ety,
                     const double *ajac, const double *xnox, const double *xnoy,
                     double uinf, double vinf, double dt, double speed_amp, double F, double time,
                     double *circ);

/* Outer-row (j = n1-1) pressure: normal momentum equation with local
 * acceleration (u - uold)/dt where the free stream enters, Gresho's condition
 * where it leaves.  q1 may be NULL (zero).  Replaces solver.cpp:1731-1782. */
int cfd_ref_exit_pressure(int n0, int n1, double dxi, double deta,
                          const double *u, const double *v, const double *T,
                          const double *uold, const double *vold, double *p,
                          const double *xix, const double *xiy, const double *etx, const double *ety,
                          const double *alph, const double *beta, const double *gamma,
                          const double *q1, const double *ajac, const double *xnox, const double *xnoy,
                          double uinf, double vinf, double dt, double Re, double Ri);

/* Stream function si, integrated along each eta line from 0 on the wall.
 * Replaces solver.cpp:1784-1800. */
int cfd_ref_stream_function(int n0, int n1, double deta, const double *u, const double *v,
                            const double *xix, const double *xiy, const double *ajac, double *si);

/* Vorticity: central differences inside, one-sided in eta on the wall and outer
 * rows.  Replaces solver.cpp:1802-1887 (its vorticity part). */
int cfd_ref_vorticity(int n0, int n1, double dxi, double deta, const double *u, const double *v,
                      const double *xix, const double *xiy, const double *etx, const double *ety,
                      double *vort);

/* coef[8] = cl, cd, cm, CL_pr, CD_pr, CL_vor, CD_vor, Nu from the wall row (free
 * stream at alpha_deg degrees, ellipse aspect ratio ar).  Replaces
 * solver.cpp:1891-1982. */
int cfd_ref_forces(int n0, int n1, double dxi, double deta, const double *x, const double *y,
                   const double *p, const double *vort, const double *T, const double *ajac,
                   const double *etx, const double *ety, double Re, double alpha_deg, double ar,
                   double *coef);

/* Dilation field and its signed maximum floored at 0 (the mass-imbalance
 * monitor).  Replaces solver.cpp:1806-1852. */
int cfd_ref_dilation(int n0, int n1, double dxi, double deta,
                     const double *u, const double *v,
                     const double *xix, const double *xiy, const double *etx, const double *ety,
                     double *dil, double *dmax);

/* Mean wall Nusselt number of the j = 0 wall, normalised by the ellipse
 * perimeter of aspect ratio ar, and the local dT/dn profile (may be NULL).
 * Replaces solver.cpp:1951-1982 and 2031-2038. */
int cfd_ref_nusselt(int n0, int n1, double dxi, double deta, const double *T,
                    const double *ajac, const double *etx, const double *ety,
                    double ar, double *nu, double *dthdn);

/* Reference-mode time steps resident on the device: the reference's O-grid
 * case (its constructor's operators, solver.cpp:751-1039, assembled once) and
 * its whole time step, solver.cpp:1057-1852, stage for stage in the reference's
 * order -- right-hand sides, the u/v/T and predictor solves, outer-row
 * predictors, face divergence, p' solve and boundary rows, correction, far-field
 * and wall velocities, wall and outer-row pressure, dilation.  The reference's
 * Gauss-Seidel (solver.cpp:2279) and SIP (solver.cpp:2066) solves, which stop at
 * 0.75e-2 of the first L1 residual, are replaced by BiCGStab to the relative
 * 2-norm tolerances below: same systems, same fixed points.  time is held fixed
 * as in the reference (it declares time int, so time += dt leaves it at 0) unless
 * time_accurate is set: then each step first advances time by dt
 * (solver.cpp:1058 with a double time), and the oscillating cylinder's wall
 * velocity and acceleration (solver.cpp:1616-1629, 1715-1718) follow it. */
typedef struct {
    double Re, Pr, Ri, dt;           /* the reference's defaults: 1000, 0.71, 0, 1e-4      */
    double uinf, vinf;               /* free stream (sin 82 deg, cos 82 deg)              */
    double speed_amp, accn_amp, F, time;  /* cylinder oscillation (0 when F = 0)          */
    double mom_rtol, p_rtol;         /* BiCGStab tolerances of the u/v/T/predictor and p' solves */
    int mom_maxit, p_maxit;
    int time_accurate;               /* 1: time += dt at the start of every step               */
    double alpha_deg;                /* free-stream angle for the force coefficients (82)     */
    double ar;                       /* ellipse aspect ratio (Nusselt normalisation)          */
} cfd_ref_params;

typedef struct {
    double dmax;       /* the dilation maximum the reference prints per step   */
    double circ;       /* circulation along row n1-2 (far-field switch)        */
    int iters[6];      /* BiCGStab iterations: u, v, T, up, vp, p'             */
    double relres[6];  /* their final relative residuals                      */
    double cl, cd, cm; /* lift, drag, moment coefficients (solver.cpp:1891-1949) */
    double nu;         /* mean wall Nusselt number (solver.cpp:1951-1982)       */
    double cl_pr, cd_pr, cl_vor, cd_vor;  /* pressure and vorticity parts of cl, cd */
} cfd_ref_stats;

typedef struct cfd_ref_case_s *cfd_ref_case;

/* The O-grid of INP.DAT (solver.cpp:516-553): n0 x n1 nodes, x/y and metric
 * arrays, p1/q1 (NULL: zero), outer-boundary normals xnox/xnoy (n0). */
int cfd_ref_case_create(int n0, int n1, double dxi, double deta,
                        const double *x, const double *y,
                        const double *xix, const double *xiy, const double *etx, const double *ety,
                        const double *alph, const double *beta, const double *gamma, const double *ajac,
                        const double *p1, const double *q1, const double *xnox, const double *xnoy,
                        const cfd_ref_params *prm, cfd_ref_case *out);
/* State at the end of a step: u, v, T, p and the predictor velocities up, vp. */
int cfd_ref_case_set_state(cfd_ref_case h, const double *u, const double *v, const double *T,
                           const double *p, const double *up, const double *vp);
int cfd_ref_case_get_state(cfd_ref_case h, double *u, double *v, double *T, double *p,
                           double *up, double *vp);
/* nsteps time steps; last (may be NULL) receives the last step's statistics. */
int cfd_ref_case_step(cfd_ref_case h, int nsteps, cfd_ref_stats *last);
/* Stream function and vorticity of the last step (solver.cpp:1784-1887); either
 * pointer may be NULL. */
int cfd_ref_case_get_post(cfd_ref_case h, double *si, double *vort);
/* The case's current time (advanced by the steps when time_accurate is set). */
int cfd_ref_case_time(cfd_ref_case h, double *time);
void cfd_ref_case_destroy(cfd_ref_case h);

/* ------------------------------------------------------------------ */
/* 2. Cavity cases                                                     */
/* ------------------------------------------------------------------ */

#define CFD_FP64 0   /* every stage in FP64                                       */
#define CFD_MIXED 1  /* FP32 p' Krylov/multigrid and predictor sweeps, FP64
                        residuals, refinement, accumulation and reductions     */

typedef struct {
    double Re, Ri, Pr;
    double dt;          /* pseudo-time step of the predictor (large = steady SIMPLE) */
    double alpha_u;     /* implicit relaxation of u, v                           */
    double alpha_p;     /* p += alpha_p p'                                       */
    double alpha_t;     /* implicit relaxation of T                              */
    double lid_u;       /* lid (j = n1-1) velocity                               */
    double t_hot, t_cold;
    int mom_sweeps;     /* Jacobi sweeps of the low-order predictor              */
    double p_rtol;      /* p' relative tolerance (2-norm)                        */
    int p_maxit;        /* p' BiCGStab iteration cap (default 2)                 */
    int mg_levels;      /* multigrid levels of the p' preconditioner             */
    int mg_pre, mg_post, mg_coarse; /* Jacobi sweeps before / after the coarse-grid
                           correction on each level, and on the coarsest (defaults
                           2, 2, 4)                                                 */
    double mg_omega;    /* damping of the Jacobi smoother                        */
    int precision;      /* CFD_FP64 or CFD_MIXED                                 */
    int graph;          /* 0 (default): launch eagerly, reading the p' solver's device
                           flags once per BiCGStab iteration to stop launching;
                           1: capture one outer iteration as a hipGraph and replay it;
                           the p' BiCGStab then runs p_maxit unrolled iterations with
                           device-side convergence tests (no host synchronisation,
                           so batches on several streams overlap)                   */
    int mg_tune;        /* 1 (default): each multigrid level's block size (256, 512
                           or 1024 threads per tile) is timed once, before the first
                           solve, and the fastest kept (results do not depend on
                           it); 0: the launch-size rule                             */
    int wall_upwind;    /* convection in the rows next to a wall: 0 the reference's
                           central differences at any Peclet number (solver.cpp:
                           1120-1175 rows j = 1, n1-2); 1 (default) the hybrid switch
                           there too (Kawamura-Kuwahara where its five points fit,
                           first-order upwind where they do not); 2 the switch to
                           first-order upwind everywhere                            */
    double hyb_blend;   /* 0: the reference's switch at |Pe| = 2; > 0 (default 1):
                           a linear blend from central to upwind over
                           2 < |Pe| < 2 + hyb_blend, so the discrete equations are
                           continuous in the state and have an exact fixed point   */
    int fas_levels;     /* grid levels of the nonlinear multigrid (FAS) outer cycle
                           with SIMPLE iterations as its smoother (default 8, as
                           many as the grid allows); 1: plain SIMPLE iterations.
                           An outer iteration (cfd_case_outer) is one cycle        */
    int fas_pre, fas_post; /* SIMPLE iterations per level before / after the
                           coarse-grid correction (default 2, 2)                  */
    int fas_coarse;     /* SIMPLE iterations on the coarsest level (default 40)   */
    int fas_min_n;      /* no level coarser than this many nodes per side; <= 0
                           (default): the smallest 2^k + 1 >= 17 whose cell Reynolds
                           number Re / 2^k is at most 32, and >= 33 when Pr > 1 or
                           |Ri| Re^2 Pr > 2e5, from params (a batch then takes only
                           slots whose physics fits its coarsest level)            */
    int fas_coarse_scheme; /* wall_upwind value of the coarse levels (default 2;
                           the more upwinded of it and wall_upwind is used)        */
    int fas_small;      /* 1 (default): a coarse FAS level of at most 33 x 33 nodes
                           (and at most 4 p' multigrid levels) runs its SIMPLE
                           iterations in one workgroup per case, in FP64; 0: in the
                           stage kernels of the fine levels                        */
    int fas_coarse_pmaxit; /* p' BiCGStab iteration cap on the coarse FAS levels
                           (their SIMPLE iterations only smooth; 0: p_maxit; < 0:
                           no Krylov iteration, p' is one multigrid V-cycle on the
                           mean-free q)                                            */
} cfd_params;

typedef struct {
    double mass_l1;     /* sum |q| (J-weighted face divergence) this iteration  */
    double energy_l1;   /* sum |R_T|                                             */
    double mom_l1;      /* sum |R_u| + |R_v|                                     */
    double dmax;        /* max nodal dilation, floored at 0                      */
    double nu_hot, nu_cold;
    int p_iters;
    double p_relres;
    double mass_res;    /* mass_l1 / mass_l1 of the first iteration              */
    double energy_res;  /* energy_l1 / energy_l1 of the first iteration          */
    int iter;           /* outer iterations done so far                          */
} cfd_stats;

typedef struct cfd_case_s *cfd_case;
typedef struct cfd_batch_s *cfd_batch;

/* Fill *p with the defaults used by the cavity configurations. */
void cfd_params_default(cfd_params *p);
/* Nodes per side (the smaller) of the coarsest level of the FAS hierarchy a case or
 * batch of n0 x n1 nodes with these parameters gets (n0 or n1 when it has none). */
int cfd_fas_coarsest(int n0, int n1, const cfd_params *p);

/* Create a cavity case on the current device from INP.DAT-style grid arrays
 * (x, y, metrics, alph/beta/gamma, Jacobian) plus the Laplacian control
 * functions p1 = lap(xi), q1 = lap(eta) (NULL = 0, as the reference reads them).
 * Fields start at rest with the wall conditions applied.  n0, n1 >= 5 and
 * n0 * n1 <= 2^28 (cases, batches and packs alike; CFD_E_ARG otherwise). */
int cfd_case_create(int n0, int n1, double dxi, double deta,
                    const double *x, const double *y,
                    const double *xix, const double *xiy, const double *etx, const double *ety,
                    const double *alph, const double *beta, const double *gamma,
                    const double *ajac, const double *p1, const double *q1,
                    const cfd_params *params, cfd_case *out);
int cfd_case_destroy(cfd_case c);
int cfd_case_set_fields(cfd_case c, const double *u, const double *v, const double *T,
                        const double *p);
int cfd_case_get_fields(cfd_case c, double *u, double *v, double *T, double *p);
int cfd_case_apply_bcs(cfd_case c);
/* Run n outer iterations; *st (may be NULL) receives the last iteration's stats. */
int cfd_case_outer(cfd_case c, int n, cfd_stats *st);
/* Iterate until mass_res <= tol (and energy_res <= tol) or max_outer. */
int cfd_case_solve(cfd_case c, double tol, int max_outer, cfd_stats *st);
/* Bytes of HBM the case holds. */
int cfd_case_bytes(cfd_case c, size_t *bytes);

/* Stage-level access for tests (host arrays). */
int cfd_case_residuals(cfd_case c, double *Ru, double *Rv, double *RT, double *d);
int cfd_case_p_apply(cfd_case c, const double *d, const double *x, double *y);
/* p' solve L x = q (x from zero) with the case's solver to ||r|| <= rtol ||q||
 * (FP64 residual); maxit caps BiCGStab iterations (per refinement step in the
 * mixed path).  d, q: host arrays, or NULL to keep the case's current device
 * values (d from the last residual evaluation, q from the last call); x: host
 * output, or NULL to leave the solution on the device. */
int cfd_case_p_solve(cfd_case c, const double *d, const double *q, double *x,
                     double rtol, int maxit, int *iters, double *relres);

/* Live kernel timing.  While enabled (enable = 1), outer iterations run
 * eagerly (not from the captured graph) and every fine-level launch of the
 * hot kernels is bracketed by HIP events on the case's stream;
 * cfd_case_profile_read returns, per kind, the summed device milliseconds and
 * the launch count since the last cfd_case_profile call (enable = 2 only
 * resets the counters).
 * Kinds: 0 p' operator apply (with its dot product), 1 predictor Jacobi sweep,
 * 2 residual/RHS, 3 face divergence, 4 multigrid down (pre-smoothing, residual,
 * restriction) and 5 multigrid up (prolongation, post-smoothing) on the fine level. */
int cfd_case_profile(cfd_case c, int enable);
int cfd_case_profile_read(cfd_case c, double *ms_total, long *launches, int nkinds);

/* Average device time of one launch of a hot kernel on the case's stream,
 * measured with HIP events around `reps` back-to-back launches on current
 * data.  kernel: 0 = p' operator apply, 1 = predictor (all sweeps),
 * 2 = residual/RHS, 3 = face divergence, 4 = multigrid down, 5 = multigrid up
 * (fine level). */
int cfd_case_time_kernel(cfd_case c, int kernel, int reps, double *ms_per_launch);

/* ------------------------------------------------------------------ */
/* 3. Batches: slots of equally sized cavity cases advanced together   */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference (one case per process, solver.cpp:2041-2062
 * main).  A batch holds `slots` cases of n0 x n1 nodes with spacing (dxi, deta);
 * every kernel of an outer iteration serves all slots in one launch.  A single
 * cfd_case is a batch of one slot, and the cfd_case_* calls below apply to it.
 * Batch-wide settings come from `params` at creation (predictor sweeps, p'
 * solver, multigrid, precision, graph); each slot's physics (Re, Ri, Pr, dt,
 * alpha_u, alpha_p, alpha_t, lid_u, t_hot, t_cold) from the params given to
 * cfd_batch_set_case. */
int cfd_batch_create(int n0, int n1, double dxi, double deta, int slots,
                     const cfd_params *params, cfd_batch *out);
int cfd_batch_destroy(cfd_batch b);
/* Load a case into a slot (grid arrays as for cfd_case_create); its fields start
 * at rest with the wall conditions and the slot becomes active.  Other slots are
 * not touched. */
int cfd_batch_set_case(cfd_batch b, int slot, const double *x, const double *y,
                       const double *xix, const double *xiy, const double *etx, const double *ety,
                       const double *alph, const double *beta, const double *gamma,
                       const double *ajac, const double *p1, const double *q1,
                       const cfd_params *params);
/* The slot stops taking part in outer iterations (its arrays stay readable). */
int cfd_batch_release(cfd_batch b, int slot);
/* n outer iterations of every active slot; st (may be NULL) receives `slots`
 * entries: the last iteration's statistics per slot (zeros for inactive slots).
 * A slot whose residuals turned non-finite shows it in its statistics (it keeps
 * iterating until released); the call itself succeeds. */
int cfd_batch_outer(cfd_batch b, int n, cfd_stats *st);
/* The two halves of one cfd_batch_outer iteration: enqueue it on the batch's
 * stream without waiting, then wait and read the statistics (several batches
 * can be launched before any is finished). */
int cfd_batch_launch(cfd_batch b);
int cfd_batch_finish(cfd_batch b, cfd_stats *st);
int cfd_batch_get_fields(cfd_batch b, int slot, double *u, double *v, double *T, double *p);
int cfd_batch_set_fields(cfd_batch b, int slot, const double *u, const double *v, const double *T,
                         const double *p);
int cfd_batch_bytes(cfd_batch b, size_t *bytes);
/* HBM bytes a batch of this shape would hold (no device work). */
int cfd_batch_bytes_needed(int n0, int n1, double dxi, double deta, int slots,
                           const cfd_params *params, size_t *bytes);

/* ------------------------------------------------------------------ */
/* 3b. Packs: cavity cases of uneven size sharing every launch          */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference (one case per process, solver.cpp:2041-2062
 * main).  A pack is a chain of `nslabs` slabs: slab k holds slots[k] grids of
 * ((n0-1) >> k) + 1 x ((n1-1) >> k) + 1 nodes with spacing 2^k (dxi, deta) (n0 and
 * n1 odd down the chain).  A case whose grid is slab k's size takes one slot of
 * slab k for its own grid and one slot of each slab below it down to its coarsest
 * FAS level (the level a batch of its own would have, from its physics), so the
 * slots of one slab hold the finest grids of the cases of that size beside the
 * coarse levels of every larger case.  One outer iteration (one FAS cycle of every
 * case) is one launch sequence down and up the chain: every SIMPLE stage of a slab
 * is one launch over all its slots, and each case computes what it computes in a
 * batch of its own (the same bits).  A case's own grid must not be a small level
 * (more than 33 x 33 nodes) and its coarsest level must be one (at most 33 x 33;
 * Re <= 1024 with the automatic depth).  Pack-wide settings come from `params` at
 * creation (graph: 1 captures the cycle once and replays it); each case's physics
 * from the params given to cfd_pack_add_case. */
typedef struct cfd_pack_s *cfd_pack;

int cfd_pack_create(int n0, int n1, double dxi, double deta, int nslabs, const int *slots,
                    const cfd_params *params, cfd_pack *out);
int cfd_pack_destroy(cfd_pack p);
/* HBM bytes a pack of this shape would hold (no device work). */
int cfd_pack_bytes_needed(int n0, int n1, double dxi, double deta, int nslabs, const int *slots,
                          const cfd_params *params, size_t *bytes);
int cfd_pack_bytes(cfd_pack p, size_t *bytes);
/* Number of case ids (the sum of slots): statistics arrays hold this many entries. */
int cfd_pack_capacity(cfd_pack p, int *ids);
/* Load a case of n0 x n1 nodes (grid arrays as for cfd_case_create) into free
 * slots of its slabs; its fields start at rest with the wall conditions and *id
 * receives its case id.  CFD_E_FULL: a slab of its chain has no free slot;
 * CFD_E_ARG: no slab of its size, or a hierarchy the pack cannot hold. */
int cfd_pack_add_case(cfd_pack p, int n0, int n1, const double *x, const double *y,
                      const double *xix, const double *xiy, const double *etx, const double *ety,
                      const double *alph, const double *beta, const double *gamma,
                      const double *ajac, const double *p1, const double *q1,
                      const cfd_params *params, int *id);
/* The case stops taking part and its slots become free (its fields stay readable
 * until the slots are reused). */
int cfd_pack_release(cfd_pack p, int id);
/* n outer iterations of every case; st (may be NULL) receives one entry per case
 * id (cfd_pack_capacity), zeros where there is no case. */
int cfd_pack_outer(cfd_pack p, int n, cfd_stats *st);
int cfd_pack_launch(cfd_pack p);
int cfd_pack_finish(cfd_pack p, cfd_stats *st);
int cfd_pack_get_fields(cfd_pack p, int id, double *u, double *v, double *T, double *p_);

/* ------------------------------------------------------------------ */
/* 4. Serving: continuous batching of cavity solves on one device      */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference.  Solves are queued by cfd_server_submit and
 * advanced by cfd_server_step, one outer iteration of every running solve per
 * step: queued requests take free slots of the batch of their grid shape
 * (`slots_per_batch` slots; a new shape gets a batch when it fits
 * `hbm_budget` bytes, empty batches being released to make room), all batches
 * are launched before any is awaited, and a solve leaves its slot when its mass
 * and energy residuals reach `tol`, at `max_outer` iterations, or on a
 * non-finite residual.  A request's result does not depend on what else is
 * served.  Batch-wide settings come from the server's params; each request's
 * physics from its own. */
typedef struct cfd_server_s *cfd_server;

#define CFD_REQ_QUEUED 0
#define CFD_REQ_RUNNING 1
#define CFD_REQ_DONE 2      /* converged to tol                         */
#define CFD_REQ_NOTCONV 3   /* stopped at max_outer                      */
#define CFD_REQ_FAILED 4    /* non-finite residual or not servable       */

int cfd_server_create(const cfd_params *params, size_t hbm_budget, int slots_per_batch,
                      cfd_server *out);
/* Packed mode: the server holds one pack (cfd_pack_create's arguments; its bytes
 * are the server's HBM use) and every request takes free slots of it (a request
 * waits while its slabs are full; one the pack cannot hold fails); each
 * cfd_server_step is one outer iteration of the whole pack. */
int cfd_server_create_packed(const cfd_params *params, int n0, int n1, double dxi, double deta, int nslabs,
                             const int *slots, cfd_server *out);
int cfd_server_destroy(cfd_server s);
/* Queue a solve (the arrays are copied); *ticket identifies it. */
int cfd_server_submit(cfd_server s, int n0, int n1, double dxi, double deta,
                      const double *x, const double *y,
                      const double *xix, const double *xiy, const double *etx, const double *ety,
                      const double *alph, const double *beta, const double *gamma,
                      const double *ajac, const double *p1, const double *q1,
                      const cfd_params *params, double tol, int max_outer, long *ticket);
/* One scheduling round; *running / *queued (may be NULL) receive the counts after it. */
int cfd_server_step(cfd_server s, int *running, int *queued);
/* State of a request; once finished (DONE/NOTCONV/FAILED) the statistics, the
 * submit-to-finish latency in seconds and the fields (any pointer may be NULL)
 * are returned once and the request is forgotten. */
int cfd_server_poll(cfd_server s, long ticket, int *state, cfd_stats *st, double *latency_s,
                    double *u, double *v, double *T, double *p);
/* Batches held, HBM bytes held, requests finished so far (any pointer may be NULL). */
int cfd_server_info(cfd_server s, int *batches, size_t *bytes, long *finished);

#ifdef __cplusplus
}
#endif
#endif
