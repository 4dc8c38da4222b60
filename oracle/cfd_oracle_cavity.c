/*
 * CPU oracle, cavity SIMPLE mode -- TEST INFRASTRUCTURE ONLY (see cfd_oracle.h).
 *
 * The steady mixed-convection cavity iteration the GPU path implements, written
 * as plain serial FP64 loops.  It reuses the reference's operators and changes
 * only what a steady, walled, under-relaxed iteration needs:
 *
 *   residual   R = q_ref(phi) - A_ref phi        q_ref: solver.cpp:1074-1192 (hybrid
 *              convection, central grad p, Ri*T); A_ref: I - dt/Re L with the
 *              solver.cpp:751-848 coefficients (9-point next to a wall, 25-point
 *              inside), wall couplings kept in the stencil.
 *   predictor  (A_lo / alpha) delta = R by Jacobi sweeps, A_lo = I + dt C_upwind
 *              - dt/Re L2 (the solver.cpp:764-786 row everywhere + first-order upwind
 *              convection); phi* = phi + delta.  d = dt*alpha_u/diag(A_lo).
 *   face div   solver.cpp:1389-1448 with the pressure-gradient-free velocity taken as
 *              u* + d grad p and dt replaced by the face value d_f = (d_P + d_nb)/2;
 *              q = div(u_f) (no 1/dt).
 *   p' solve   L_d p' = q, the solver.cpp:943-1025 9-point operator written in flux
 *              form with each face scaled by d_f (equal to dt * the reference
 *              operator when d == dt); p' = 0 on walls during the solve as the
 *              reference holds it (solver.cpp:1452-1465).  BiCGStab, right-
 *              preconditioned by a geometric multigrid V-cycle.
 *   correction p' copied to the walls (solver.cpp:1474-1482 on all four walls),
 *              u = u* - d grad p' (solver.cpp:1504-1531), p += alpha_p p'.
 *   BCs        walls no-slip, lid u = lid_u, T hot/cold, adiabatic sides; wall
 *              pressure from the normal momentum equation (solver.cpp:1676-1728 on
 *              the hot wall, its mirror solver.cpp:1750-1774 on the lid, and the same
 *              construction with xi and eta swapped on the sides).
 *   reductions sum|q|, sum|R_T|, dmax (solver.cpp:1806-1852), wall Nusselt
 *              (solver.cpp:1951-1982 integrand, normalised by the wall length).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "cfd_oracle.h"

#define ID(i, j) ((size_t)(j) * (size_t)n0 + (size_t)(i))
#define MAXLEV 16

typedef struct {
    int n0, n1;
    double h1, h2;
    double *xix, *xiy, *etx, *ety, *jac, *d;   /* level operator data */
    double *x, *b, *r;                    /* level vectors */
} mg_level;

struct or_cav_case {
    int n0, n1;
    double h1, h2;
    or_cav_params prm;
    double *xg, *yg, *xix, *xiy, *etx, *ety, *alph, *beta, *gamma, *ajac, *p1, *q1;
    double *u, *v, *T, *p;
    double *U, *V, *Ru, *Rv, *RT, *du, *dv, *dT, *tu, *tv, *tT, *d, *us, *vs, *q, *pc;
    double *bi[8]; /* BiCGStab work */
    mg_level lev[MAXLEV];
    int nlev;
    double mass0, energy0;
    int iter;
    double wall_len_hot, wall_len_cold;
    /* FAS: source terms of this level's equations (NULL on a level without them),
     * the next coarser level, and (on a coarse level) the state restricted to it */
    double *su, *sv, *sT, *sm;
    or_cav_case *coarse;
    double *u0, *v0, *T0, *p0;
};

static double *dup_arr(const double *a, size_t n) {
    double *r = (double *)malloc(n * sizeof(double));
    if (a) memcpy(r, a, n * sizeof(double));
    else memset(r, 0, n * sizeof(double));
    return r;
}

static double *zalloc(size_t n) { return (double *)calloc(n, sizeof(double)); }

/* ------------------------------------------------------------------------ */
/* Conservative flux-form pressure-correction operator.                       */
/*                                                                            */
/* The face velocity of solver.cpp:1402-1436, with dt replaced by the face    */
/* d_f = (d_P + d_nb)/2, has the pressure part                                 */
/*   u_e^p = d_e/2 [(xix_P + xix_E) p_xi,e + (etx_P + etx_E) p_eta,e]          */
/* (v with xiy/ety), p_xi,e = (pE - pP)/h1, p_eta,e = (pNE + pN - pS - pSE)/4h2 */
/* (north/south faces swap the roles).  The face flux is the contravariant    */
/*   F_e = (J xix)_e u_e + (J xiy)_e v_e,  (J m)_e = ((J m)_P + (J m)_E)/2,      */
/* and (L p)_P = (F_e - F_w)/h1 + (G_n - G_s)/h2, the J-weighted divergence.  */
/* A face touching a wall node carries no flux (no penetration), so every     */
/* row sums to zero and so does every column: constants span both null       */
/* spaces.  p at wall nodes is read as the adjacent interior value (the       */
/* Neumann ghost), which only enters through tangential face gradients.       */
/* Coefficients c[k], k = (dj+1)*3 + (di+1), act on the clamped neighbours.   */
static void p9_coefs(int n0, int n1, const double *xix, const double *xiy, const double *etx,
                     const double *ety, const double *jac, const double *d, double h1, double h2,
                     int i, int j, double c[9])
{
    const size_t P = ID(i, j), E = ID(i + 1, j), W = ID(i - 1, j), Nn = ID(i, j + 1), S = ID(i, j - 1);
    const int we = (i + 1 == n0 - 1), ww = (i - 1 == 0), wn = (j + 1 == n1 - 1), ws = (j - 1 == 0);
    for (int k = 0; k < 9; k++) c[k] = 0.0;
#define C(di, dj) c[((dj) + 1) * 3 + ((di) + 1)]
    const double jxP = jac[P] * xix[P], jyP = jac[P] * xiy[P];
    const double kxP = jac[P] * etx[P], kyP = jac[P] * ety[P];
    if (!we) { /* face e: +F_e/h1 */
        const double mx = 0.5 * (jxP + jac[E] * xix[E]), my = 0.5 * (jyP + jac[E] * xiy[E]);
        const double a = mx * (xix[P] + xix[E]) + my * (xiy[P] + xiy[E]);
        const double b = mx * (etx[P] + etx[E]) + my * (ety[P] + ety[E]);
        const double de = 0.5 * (d[P] + d[E]);
        const double kn = 0.5 * de * a / (h1 * h1), kt = 0.5 * de * b / (4.0 * h1 * h2);
        C(1, 0) += kn; C(0, 0) -= kn;
        C(1, 1) += kt; C(0, 1) += kt; C(0, -1) -= kt; C(1, -1) -= kt;
    }
    if (!ww) { /* face w: -F_w/h1 */
        const double mx = 0.5 * (jxP + jac[W] * xix[W]), my = 0.5 * (jyP + jac[W] * xiy[W]);
        const double a = mx * (xix[P] + xix[W]) + my * (xiy[P] + xiy[W]);
        const double b = mx * (etx[P] + etx[W]) + my * (ety[P] + ety[W]);
        const double dw = 0.5 * (d[P] + d[W]);
        const double kn = 0.5 * dw * a / (h1 * h1), kt = 0.5 * dw * b / (4.0 * h1 * h2);
        C(0, 0) -= kn; C(-1, 0) += kn;
        C(0, 1) -= kt; C(-1, 1) -= kt; C(0, -1) += kt; C(-1, -1) += kt;
    }
    if (!wn) { /* face n: +G_n/h2 */
        const double mx = 0.5 * (kxP + jac[Nn] * etx[Nn]), my = 0.5 * (kyP + jac[Nn] * ety[Nn]);
        const double a = mx * (etx[P] + etx[Nn]) + my * (ety[P] + ety[Nn]);
        const double b = mx * (xix[P] + xix[Nn]) + my * (xiy[P] + xiy[Nn]);
        const double dn = 0.5 * (d[P] + d[Nn]);
        const double kn = 0.5 * dn * a / (h2 * h2), kt = 0.5 * dn * b / (4.0 * h1 * h2);
        C(0, 1) += kn; C(0, 0) -= kn;
        C(1, 1) += kt; C(-1, 1) -= kt; C(1, 0) += kt; C(-1, 0) -= kt;
    }
    if (!ws) { /* face s: -G_s/h2 */
        const double mx = 0.5 * (kxP + jac[S] * etx[S]), my = 0.5 * (kyP + jac[S] * ety[S]);
        const double a = mx * (etx[P] + etx[S]) + my * (ety[P] + ety[S]);
        const double b = mx * (xix[P] + xix[S]) + my * (xiy[P] + xiy[S]);
        const double ds = 0.5 * (d[P] + d[S]);
        const double kn = 0.5 * ds * a / (h2 * h2), kt = 0.5 * ds * b / (4.0 * h1 * h2);
        C(0, 0) -= kn; C(0, -1) += kn;
        C(1, 0) -= kt; C(-1, 0) += kt; C(1, -1) -= kt; C(-1, -1) += kt;
    }
#undef C
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* y = L x on interior nodes (neighbours clamped into the interior); y = 0 on the border */
static void p9_apply_lvl(const mg_level *L, const double *x, double *y)
{
    const int n0 = L->n0, n1 = L->n1;
    for (int j = 0; j < n1; j++)
        for (int i = 0; i < n0; i++) {
            const size_t P = ID(i, j);
            if (i == 0 || j == 0 || i == n0 - 1 || j == n1 - 1) { y[P] = 0.0; continue; }
            double c[9];
            p9_coefs(n0, n1, L->xix, L->xiy, L->etx, L->ety, L->jac, L->d, L->h1, L->h2, i, j, c);
            double s = 0.0;
            for (int dj = -1; dj <= 1; dj++)
                for (int di = -1; di <= 1; di++) {
                    const int ii = clampi(i + di, 1, n0 - 2), jj = clampi(j + dj, 1, n1 - 2);
                    s += c[(dj + 1) * 3 + (di + 1)] * x[ID(ii, jj)];
                }
            y[P] = s;
        }
}

/* diagonal of the clamped operator: every coefficient whose neighbour clamps onto P */
static double p9_diag(const mg_level *L, int i, int j)
{
    const int n0 = L->n0, n1 = L->n1;
    double c[9];
    p9_coefs(n0, n1, L->xix, L->xiy, L->etx, L->ety, L->jac, L->d, L->h1, L->h2, i, j, c);
    double s = 0.0;
    for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++)
            if (clampi(i + di, 1, n0 - 2) == i && clampi(j + dj, 1, n1 - 2) == j)
                s += c[(dj + 1) * 3 + (di + 1)];
    return s;
}

/* remove the interior mean (the constant null space of the Neumann operator) */
static void project_mean(int n0, int n1, double *v)
{
    double s = 0.0;
    for (int j = 1; j < n1 - 1; j++)
        for (int i = 1; i < n0 - 1; i++) s += v[ID(i, j)];
    s /= (double)(n0 - 2) * (double)(n1 - 2);
    for (int j = 1; j < n1 - 1; j++)
        for (int i = 1; i < n0 - 1; i++) v[ID(i, j)] -= s;
}

/* damped Jacobi sweeps on level L: x <- x + w D^-1 (b - A x) */
static void mg_smooth(mg_level *L, int sweeps, double w, int x_is_zero)
{
    const int n0 = L->n0, n1 = L->n1;
    for (int s = 0; s < sweeps; s++) {
        if (x_is_zero && s == 0) {
            for (int j = 1; j < n1 - 1; j++)
                for (int i = 1; i < n0 - 1; i++)
                    L->x[ID(i, j)] = w * L->b[ID(i, j)] / p9_diag(L, i, j);
            continue;
        }
        p9_apply_lvl(L, L->x, L->r);
        for (int j = 1; j < n1 - 1; j++)
            for (int i = 1; i < n0 - 1; i++) {
                const size_t P = ID(i, j);
                L->x[P] += w * (L->b[P] - L->r[P]) / p9_diag(L, i, j);
            }
    }
}

/* full-weighting restriction of the residual r of level f into b of level c */
static void mg_restrict(const mg_level *f, mg_level *c)
{
    const int n0 = f->n0;
    const int m0 = c->n0, m1 = c->n1;
    for (int J = 0; J < m1; J++)
        for (int I = 0; I < m0; I++) {
            double *bc = &c->b[(size_t)J * m0 + I];
            if (I == 0 || J == 0 || I == m0 - 1 || J == m1 - 1) { *bc = 0.0; continue; }
            const int i = 2 * I, j = 2 * J;
            const double *r = f->r;
            *bc = (4.0 * r[ID(i, j)] +
                   2.0 * (r[ID(i - 1, j)] + r[ID(i + 1, j)] + r[ID(i, j - 1)] + r[ID(i, j + 1)]) +
                   (r[ID(i - 1, j - 1)] + r[ID(i + 1, j - 1)] + r[ID(i - 1, j + 1)] + r[ID(i + 1, j + 1)])) /
                  16.0;
        }
    /* the coarse Neumann system is solvable only for a zero-sum right-hand side */
    project_mean(m0, m1, c->b);
}

/* bilinear prolongation of level c's x, added to level f's x (interior only);
 * coarse border nodes read as their interior neighbour (Neumann ghost) */
static void mg_prolong_add(const mg_level *c, mg_level *f)
{
    const int n0 = f->n0, n1 = f->n1, m0 = c->n0, m1 = c->n1;
    const double *xc = c->x;
#define XC(I, J) xc[(size_t)clampi((J), 1, m1 - 2) * m0 + clampi((I), 1, m0 - 2)]
    for (int j = 1; j < n1 - 1; j++)
        for (int i = 1; i < n0 - 1; i++) {
            const int I = i >> 1, J = j >> 1;
            double e;
            if ((i & 1) == 0 && (j & 1) == 0) e = XC(I, J);
            else if ((j & 1) == 0) e = 0.5 * (XC(I, J) + XC(I + 1, J));
            else if ((i & 1) == 0) e = 0.5 * (XC(I, J) + XC(I, J + 1));
            else e = 0.25 * (XC(I, J) + XC(I + 1, J) + XC(I, J + 1) + XC(I + 1, J + 1));
            f->x[ID(i, j)] += e;
        }
#undef XC
}

static void mg_vcycle(or_cav_case *c, int l)
{
    mg_level *L = &c->lev[l];
    const or_cav_params *pr = &c->prm;
    const size_t n = (size_t)L->n0 * L->n1;
    memset(L->x, 0, n * sizeof(double));
    if (l == c->nlev - 1) {
        mg_smooth(L, pr->mg_coarse, pr->mg_omega, 1);
        return;
    }
    mg_smooth(L, pr->mg_pre, pr->mg_omega, 1);
    p9_apply_lvl(L, L->x, L->r);
    for (int j = 1; j < L->n1 - 1; j++)
        for (int i = 1; i < L->n0 - 1; i++) {
            const size_t P = (size_t)j * L->n0 + i;
            L->r[P] = L->b[P] - L->r[P];
        }
    mg_restrict(L, &c->lev[l + 1]);
    mg_vcycle(c, l + 1);
    mg_prolong_add(&c->lev[l + 1], L);
    mg_smooth(L, pr->mg_post, pr->mg_omega, 0);
}

/* z = M^-1 y on the fine level (multigrid V-cycle, or one Jacobi step if nlev==1) */
static void precond(or_cav_case *c, const double *y, double *z)
{
    mg_level *L = &c->lev[0];
    const size_t n = (size_t)L->n0 * L->n1;
    memcpy(L->b, y, n * sizeof(double));
    mg_vcycle(c, 0);
    memcpy(z, L->x, n * sizeof(double));
}

static double dot_int(int n0, int n1, const double *a, const double *b)
{
    double s = 0.0;
    for (int j = 1; j < n1 - 1; j++)
        for (int i = 1; i < n0 - 1; i++) s += a[ID(i, j)] * b[ID(i, j)];
    return s;
}

/* BiCGStab, right preconditioned; x starts at its given value (0 on the border). */
static int bicgstab(or_cav_case *c, const double *bvec, double *x, double rtol, int maxit,
                    double *relres)
{
    const int n0 = c->n0, n1 = c->n1;
    const size_t N = (size_t)n0 * n1;
    mg_level *L0 = &c->lev[0];
    double *r = c->bi[0], *r0 = c->bi[1], *pv = c->bi[2], *vv = c->bi[3];
    double *s = c->bi[4], *t = c->bi[5], *ph = c->bi[6], *sh = c->bi[7];
    p9_apply_lvl(L0, x, r);
    for (size_t k = 0; k < N; k++) r[k] = bvec[k] - r[k];
    for (int j = 0; j < n1; j++)
        for (int i = 0; i < n0; i++)
            if (i == 0 || j == 0 || i == n0 - 1 || j == n1 - 1) r[ID(i, j)] = 0.0;
    memcpy(r0, r, N * sizeof(double));
    memset(pv, 0, N * sizeof(double));
    memset(vv, 0, N * sizeof(double));
    project_mean(n0, n1, r);
    memcpy(r0, r, N * sizeof(double));
    const double bnorm = sqrt(dot_int(n0, n1, bvec, bvec));
    double rnorm = sqrt(dot_int(n0, n1, r, r));
    if (bnorm == 0.0) { if (relres) *relres = 0.0; return 0; }
    const double tol = rtol * bnorm;
    if (rnorm <= tol) { if (relres) *relres = rnorm / bnorm; return 0; }
    double rho_old = 1.0, alpha = 1.0, omega = 1.0;
    int it;
    for (it = 1; it <= maxit; it++) {
        const double rho = dot_int(n0, n1, r0, r);
        if (rho == 0.0) break;
        if (it == 1) {
            memcpy(pv, r, N * sizeof(double));
        } else {
            const double beta = (rho / rho_old) * (alpha / omega);
            for (size_t k = 0; k < N; k++) pv[k] = r[k] + beta * (pv[k] - omega * vv[k]);
        }
        precond(c, pv, ph);
        p9_apply_lvl(L0, ph, vv);
        alpha = rho / dot_int(n0, n1, r0, vv);
        for (size_t k = 0; k < N; k++) s[k] = r[k] - alpha * vv[k];
        const double snorm = sqrt(dot_int(n0, n1, s, s));
        if (snorm <= tol) {
            for (size_t k = 0; k < N; k++) x[k] += alpha * ph[k];
            rnorm = snorm;
            break;
        }
        precond(c, s, sh);
        p9_apply_lvl(L0, sh, t);
        const double tt = dot_int(n0, n1, t, t);
        omega = (tt > 0.0) ? dot_int(n0, n1, t, s) / tt : 0.0;
        for (size_t k = 0; k < N; k++) {
            x[k] += alpha * ph[k] + omega * sh[k];
            r[k] = s[k] - omega * t[k];
        }
        rnorm = sqrt(dot_int(n0, n1, r, r));
        if (rnorm <= tol || omega == 0.0) break;
        rho_old = rho;
    }
    if (relres) *relres = rnorm / bnorm;
    return it > maxit ? maxit : it;
}

/* p' by one multigrid V-cycle from zero on the mean-free right-hand side (no
   Krylov iteration; the coarse FAS levels' SIMPLE iterations only smooth). */
static int vcycle_solve(or_cav_case *c, const double *bvec, double *x, double *relres)
{
    const int n0 = c->n0, n1 = c->n1;
    double *r = c->bi[0];
    for (int j = 0; j < n1; j++)
        for (int i = 0; i < n0; i++)
            r[ID(i, j)] = (i == 0 || j == 0 || i == n0 - 1 || j == n1 - 1) ? 0.0 : bvec[ID(i, j)];
    project_mean(n0, n1, r);
    precond(c, r, x);
    if (relres) *relres = 0.0;
    return 1;
}

/* ------------------------------------------------------------------------ */
static void build_levels(or_cav_case *c)
{
    const int want = c->prm.mg_levels < 1 ? 1 : c->prm.mg_levels;
    c->lev[0].n0 = c->n0; c->lev[0].n1 = c->n1;
    c->lev[0].h1 = c->h1; c->lev[0].h2 = c->h2;
    c->lev[0].xix = c->xix; c->lev[0].xiy = c->xiy; c->lev[0].etx = c->etx; c->lev[0].ety = c->ety;
    c->lev[0].jac = c->ajac;
    c->lev[0].d = c->d;
    size_t N = (size_t)c->n0 * c->n1;
    c->lev[0].x = zalloc(N); c->lev[0].b = zalloc(N); c->lev[0].r = zalloc(N);
    int l = 1;
    while (l < want && l < MAXLEV) {
        const mg_level *f = &c->lev[l - 1];
        if ((f->n0 & 1) == 0 || (f->n1 & 1) == 0 || f->n0 < 9 || f->n1 < 9) break;
        mg_level *L = &c->lev[l];
        L->n0 = (f->n0 + 1) / 2; L->n1 = (f->n1 + 1) / 2;
        L->h1 = 2.0 * f->h1; L->h2 = 2.0 * f->h2;
        const size_t M = (size_t)L->n0 * L->n1;
        L->xix = zalloc(M); L->xiy = zalloc(M); L->etx = zalloc(M); L->ety = zalloc(M);
        L->jac = zalloc(M); L->d = zalloc(M);
        L->x = zalloc(M); L->b = zalloc(M); L->r = zalloc(M);
        for (int J = 0; J < L->n1; J++)
            for (int I = 0; I < L->n0; I++) {
                const size_t src = (size_t)(2 * J) * f->n0 + 2 * I, dst = (size_t)J * L->n0 + I;
                L->xix[dst] = f->xix[src]; L->xiy[dst] = f->xiy[src];
                L->etx[dst] = f->etx[src]; L->ety[dst] = f->ety[src];
                L->jac[dst] = f->jac[src];
            }
        l++;
    }
    c->nlev = l;
}

/* inject the fine d into the coarse levels (every outer iteration) */
static void inject_d(or_cav_case *c)
{
    for (int l = 1; l < c->nlev; l++) {
        const mg_level *f = &c->lev[l - 1];
        mg_level *L = &c->lev[l];
        for (int J = 0; J < L->n1; J++)
            for (int I = 0; I < L->n0; I++)
                L->d[(size_t)J * L->n0 + I] = f->d[(size_t)(2 * J) * f->n0 + 2 * I];
    }
}

/* ------------------------------------------------------------------------ */
/* FAS hierarchy: level l+1 takes every other node of level l (grids of      */
/* 2^k + 1 nodes coarsen vertex-wise), its grid and metric arrays injected,  */
/* h doubled; it carries source arrays for the FAS right-hand sides and a    */
/* copy of the state restricted to it.                                       */
static double *inject_new(const double *f, int n0, int m0, int m1)
{
    double *r = (double *)malloc((size_t)m0 * m1 * sizeof(double));
    for (int J = 0; J < m1; J++)
        for (int I = 0; I < m0; I++) r[(size_t)J * m0 + I] = f[(size_t)(2 * J) * n0 + 2 * I];
    return r;
}

static void fas_build(or_cav_case *c, const double *x, const double *y, const double *xix,
                      const double *xiy, const double *etx, const double *ety, const double *alph,
                      const double *beta, const double *gamma, const double *ajac, const double *p1,
                      const double *q1)
{
    const int n0 = c->n0, n1 = c->n1;
    const int m0 = (n0 + 1) / 2, m1 = (n1 + 1) / 2;
    int lo = c->prm.fas_min_n;
    if (lo <= 0) { /* the smallest 2^k + 1 >= 17 (>= 33 when Pr > 1 or Ra = |Ri| Re^2 Pr > 2e5)
                      with cell Reynolds number Re / 2^k <= 32 */
        const double ra = fabs(c->prm.Ri) * c->prm.Re * c->prm.Re * c->prm.Pr;
        int m = (c->prm.Pr <= 1.0 && ra <= 2.0e5) ? 16 : 32;
        while (32.0 * m < c->prm.Re && m < (1 << 20)) m *= 2;
        lo = m + 1;
    }
    if (lo < 5) lo = 5;
    if (c->prm.fas_levels <= 1 || (n0 & 1) == 0 || (n1 & 1) == 0 || m0 < lo || m1 < lo) return;
    or_cav_params cp = c->prm;
    cp.fas_levels = c->prm.fas_levels - 1;
    if (c->prm.fas_coarse_scheme > cp.wall_upwind) cp.wall_upwind = c->prm.fas_coarse_scheme;
    if (c->prm.fas_coarse_pmaxit > 0) cp.p_maxit = c->prm.fas_coarse_pmaxit;
    if (c->prm.fas_coarse_pmaxit < 0) cp.p_maxit = -1; /* one V-cycle */
    const double *src[12] = {x, y, xix, xiy, etx, ety, alph, beta, gamma, ajac, p1, q1};
    double *inj[12];
    for (int k = 0; k < 12; k++) inj[k] = inject_new(src[k], n0, m0, m1);
    or_cav_case *cc = or_cav_create(m0, m1, 2.0 * c->h1, 2.0 * c->h2, inj[0], inj[1], inj[2], inj[3],
                                    inj[4], inj[5], inj[6], inj[7], inj[8], inj[9], inj[10], inj[11], &cp);
    for (int k = 0; k < 12; k++) free(inj[k]);
    const size_t M = (size_t)m0 * m1;
    cc->su = zalloc(M); cc->sv = zalloc(M); cc->sT = zalloc(M); cc->sm = zalloc(M);
    cc->u0 = zalloc(M); cc->v0 = zalloc(M); cc->T0 = zalloc(M); cc->p0 = zalloc(M);
    c->coarse = cc;
}

or_cav_case *or_cav_create(int n0, int n1, double h1, double h2,
                           const double *x, const double *y,
                           const double *xix, const double *xiy,
                           const double *etx, const double *ety,
                           const double *alph, const double *beta, const double *gamma,
                           const double *ajac, const double *p1, const double *q1,
                           const or_cav_params *prm)
{
    or_cav_case *c = (or_cav_case *)calloc(1, sizeof(*c));
    const size_t N = (size_t)n0 * n1;
    c->n0 = n0; c->n1 = n1; c->h1 = h1; c->h2 = h2; c->prm = *prm;
    c->xg = dup_arr(x, N); c->yg = dup_arr(y, N);
    c->xix = dup_arr(xix, N); c->xiy = dup_arr(xiy, N); c->etx = dup_arr(etx, N); c->ety = dup_arr(ety, N);
    c->alph = dup_arr(alph, N); c->beta = dup_arr(beta, N); c->gamma = dup_arr(gamma, N);
    c->ajac = dup_arr(ajac, N); c->p1 = dup_arr(p1, N); c->q1 = dup_arr(q1, N);
    double **all[] = {&c->u, &c->v, &c->T, &c->p, &c->U, &c->V, &c->Ru, &c->Rv, &c->RT,
                      &c->du, &c->dv, &c->dT, &c->tu, &c->tv, &c->tT, &c->d, &c->us, &c->vs,
                      &c->q, &c->pc};
    for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); k++) *all[k] = zalloc(N);
    for (int k = 0; k < 8; k++) c->bi[k] = zalloc(N);
    build_levels(c);
    c->wall_len_hot = c->wall_len_cold = 0.0;
    for (int i = 0; i < n0 - 1; i++) {
        const size_t a = ID(i, 0), b = ID(i + 1, 0), a1 = ID(i, n1 - 1), b1 = ID(i + 1, n1 - 1);
        c->wall_len_hot += hypot(x[b] - x[a], y[b] - y[a]);
        c->wall_len_cold += hypot(x[b1] - x[a1], y[b1] - y[a1]);
    }
    fas_build(c, x, y, xix, xiy, etx, ety, alph, beta, gamma, ajac, p1, q1);
    return c;
}

void or_cav_destroy(or_cav_case *c)
{
    if (!c) return;
    double *all[] = {c->xg, c->yg, c->xix, c->xiy, c->etx, c->ety, c->alph, c->beta, c->gamma,
                     c->ajac, c->p1, c->q1, c->u, c->v, c->T, c->p, c->U, c->V, c->Ru, c->Rv,
                     c->RT, c->du, c->dv, c->dT, c->tu, c->tv, c->tT, c->d, c->us, c->vs, c->q, c->pc};
    for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); k++) free(all[k]);
    for (int k = 0; k < 8; k++) free(c->bi[k]);
    or_cav_destroy(c->coarse);
    {
        double *fas[] = {c->su, c->sv, c->sT, c->sm, c->u0, c->v0, c->T0, c->p0};
        for (size_t k = 0; k < sizeof(fas) / sizeof(fas[0]); k++) free(fas[k]);
    }
    for (int l = 0; l < c->nlev; l++) {
        free(c->lev[l].x); free(c->lev[l].b); free(c->lev[l].r);
        if (l > 0) {
            free(c->lev[l].xix); free(c->lev[l].xiy); free(c->lev[l].etx); free(c->lev[l].ety);
            free(c->lev[l].jac); free(c->lev[l].d);
        }
    }
    free(c);
}

void or_cav_set_fields(or_cav_case *c, const double *u, const double *v, const double *T,
                       const double *p)
{
    const size_t N = (size_t)c->n0 * c->n1;
    memcpy(c->u, u, N * sizeof(double));
    memcpy(c->v, v, N * sizeof(double));
    memcpy(c->T, T, N * sizeof(double));
    memcpy(c->p, p, N * sizeof(double));
    c->iter = 0;
    c->mass0 = c->energy0 = 0.0;
}

void or_cav_get_fields(const or_cav_case *c, double *u, double *v, double *T, double *p)
{
    const size_t N = (size_t)c->n0 * c->n1;
    if (u) memcpy(u, c->u, N * sizeof(double));
    if (v) memcpy(v, c->v, N * sizeof(double));
    if (T) memcpy(T, c->T, N * sizeof(double));
    if (p) memcpy(p, c->p, N * sizeof(double));
}

/* ------------------------------------------------------------------------ */
/* boundary conditions                                                        */

static void velocity_temperature_bcs(or_cav_case *c)
{
    const int n0 = c->n0, n1 = c->n1;
    const or_cav_params *pr = &c->prm;
    for (int i = 0; i < n0; i++) {
        c->u[ID(i, 0)] = 0.0; c->v[ID(i, 0)] = 0.0;
        c->u[ID(i, n1 - 1)] = (i == 0 || i == n0 - 1) ? 0.0 : pr->lid_u;
        c->v[ID(i, n1 - 1)] = 0.0;
        c->T[ID(i, 0)] = pr->t_hot;
        c->T[ID(i, n1 - 1)] = pr->t_cold;
    }
    for (int j = 0; j < n1; j++) {
        c->u[ID(0, j)] = c->v[ID(0, j)] = 0.0;
        c->u[ID(n0 - 1, j)] = c->v[ID(n0 - 1, j)] = 0.0;
    }
}

/* adiabatic side walls: grad(xi).grad(T) = alph T_xi - beta T_eta = 0 with a
 * one-sided second-order T_xi and a central T_eta along the wall (old values). */
static void temperature_sides(or_cav_case *c)
{
    const int n0 = c->n0, n1 = c->n1;
    const double h1 = c->h1, h2 = c->h2;
    double *T = c->T;
    double *tl = c->tu, *tr = c->tv; /* scratch: new wall values */
    for (int j = 1; j < n1 - 1; j++) {
        size_t P = ID(0, j);
        double te = (T[ID(0, j + 1)] - T[ID(0, j - 1)]) / (2.0 * h2);
        if (c->prm.plant == 5) te = -te;
        double txi = c->beta[P] / c->alph[P] * te;
        tl[j] = (4.0 * T[ID(1, j)] - T[ID(2, j)] - 2.0 * h1 * txi) / 3.0;
        P = ID(n0 - 1, j);
        te = (T[ID(n0 - 1, j + 1)] - T[ID(n0 - 1, j - 1)]) / (2.0 * h2);
        if (c->prm.plant == 5) te = -te;
        txi = c->beta[P] / c->alph[P] * te;
        tr[j] = (4.0 * T[ID(n0 - 2, j)] - T[ID(n0 - 3, j)] + 2.0 * h1 * txi) / 3.0;
    }
    for (int j = 1; j < n1 - 1; j++) {
        T[ID(0, j)] = tl[j];
        T[ID(n0 - 1, j)] = tr[j];
        if (c->sT) { /* FAS: boundary sources of a coarse level */
            T[ID(0, j)] += c->sT[ID(0, j)];
            T[ID(n0 - 1, j)] += c->sT[ID(n0 - 1, j)];
        }
    }
}

/* Wall pressure from the normal momentum equation.  side: 0 = j=0 (eta wall,
 * inward +eta), 1 = j=n1-1, 2 = i=0 (xi wall), 3 = i=n0-1.  For an eta wall
 * this is solver.cpp:1697-1725 (1751-1774 mirrored); for a xi wall the roles of
 * (xi, alph, P, U) and (eta, gamma, Q, V) swap and the tangent becomes
 * (x_xi, y_xi) = (eta_y J, -eta_x J). */
static double wall_p_value(const or_cav_case *c, int side, int i, int j)
{
    const int n0 = c->n0;
    const double h1 = c->h1, h2 = c->h2;
    const double Re = c->prm.Re, Ri = c->prm.Ri;
    const size_t P = ID(i, j);
    const double *fl[2] = {c->u, c->v};
    const double U = c->xix[P] * c->u[P] + c->xiy[P] * c->v[P];
    const double V = c->etx[P] * c->u[P] + c->ety[P] * c->v[P];
    double dpx = 0.0, dpy = 0.0;
    for (int k = 0; k < 2; k++) {
        const double *f = fl[k];
        double d2u, conv;
        if (side <= 1) {
            const int s = (side == 0) ? 1 : -1; /* inward step in j */
            const double aa = c->alph[P] * (f[ID(i + 1, j)] + f[ID(i - 1, j)] - 2 * f[P]) / (h1 * h1);
            const double gg = c->gamma[P] * (f[ID(i, j + 2 * s)] + f[P] - 2 * f[ID(i, j + s)]) / (h2 * h2);
            double bb;
            if (side == 0)
                bb = c->beta[P] * (f[ID(i + 1, j + 1)] + f[ID(i - 1, j)] - f[ID(i - 1, j + 1)] - f[ID(i + 1, j)]) / (2 * h1 * h2);
            else
                bb = c->beta[P] * (f[ID(i + 1, j)] + f[ID(i - 1, j - 1)] - f[ID(i + 1, j - 1)] - f[ID(i - 1, j)]) / (2 * h1 * h2);
            const double qqq = c->q1[P] * s * (-3 * f[P] + 4 * f[ID(i, j + s)] - f[ID(i, j + 2 * s)]) / (2 * h2);
            d2u = aa + gg - 2 * bb + qqq;
            conv = U * 0.5 * (f[ID(i + 1, j)] - f[ID(i - 1, j)]) / h1;
            conv = conv + V * s * (-3 * f[P] + 4 * f[ID(i, j + s)] - f[ID(i, j + 2 * s)]) / (2 * h2);
        } else {
            const int s = (side == 2) ? 1 : -1; /* inward step in i */
            const double aa = c->gamma[P] * (f[ID(i, j + 1)] + f[ID(i, j - 1)] - 2 * f[P]) / (h2 * h2);
            const double gg = c->alph[P] * (f[ID(i + 2 * s, j)] + f[P] - 2 * f[ID(i + s, j)]) / (h1 * h1);
            double bb;
            if (side == 2)
                bb = c->beta[P] * (f[ID(i + 1, j + 1)] + f[ID(i, j - 1)] - f[ID(i + 1, j - 1)] - f[ID(i, j + 1)]) / (2 * h1 * h2);
            else
                bb = c->beta[P] * (f[ID(i, j + 1)] + f[ID(i - 1, j - 1)] - f[ID(i, j - 1)] - f[ID(i - 1, j + 1)]) / (2 * h1 * h2);
            const double ppp = c->p1[P] * s * (-3 * f[P] + 4 * f[ID(i + s, j)] - f[ID(i + 2 * s, j)]) / (2 * h1);
            d2u = aa + gg - 2 * bb + ppp;
            conv = V * 0.5 * (f[ID(i, j + 1)] - f[ID(i, j - 1)]) / h2;
            conv = conv + U * s * (-3 * f[P] + 4 * f[ID(i + s, j)] - f[ID(i + 2 * s, j)]) / (2 * h1);
        }
        if (k == 0) dpx = d2u / Re - conv;
        else dpy = d2u / Re - conv + Ri * c->T[P];
    }
    const double J = c->ajac[P];
    if (c->prm.plant == side + 1) { dpx = -dpx; dpy = -dpy; }
    if (side <= 1) {
        const double tx = -c->xiy[P] * J, ty = c->xix[P] * J; /* (x_eta, y_eta) */
        const double dpn = (dpx * tx + dpy * ty) * h2;
        return side == 0 ? c->p[ID(i, j + 1)] - dpn : c->p[ID(i, j - 1)] + dpn;
    }
    const double tx = c->ety[P] * J, ty = -c->etx[P] * J; /* (x_xi, y_xi) */
    const double dpn = (dpx * tx + dpy * ty) * h1;
    return side == 2 ? c->p[ID(i + 1, j)] - dpn : c->p[ID(i - 1, j)] + dpn;
}

static void pressure_walls(or_cav_case *c)
{
    const int n0 = c->n0, n1 = c->n1;
    double *pb = c->tT; /* scratch for new wall values: [bottom | top | left | right] */
    for (int i = 1; i < n0 - 1; i++) {
        pb[i] = wall_p_value(c, 0, i, 0);
        pb[n0 + i] = wall_p_value(c, 1, i, n1 - 1);
    }
    for (int j = 1; j < n1 - 1; j++) {
        pb[2 * n0 + j] = wall_p_value(c, 2, 0, j);
        pb[2 * n0 + n1 + j] = wall_p_value(c, 3, n0 - 1, j);
    }
    for (int i = 1; i < n0 - 1; i++) {
        c->p[ID(i, 0)] = pb[i];
        c->p[ID(i, n1 - 1)] = pb[n0 + i];
    }
    for (int j = 1; j < n1 - 1; j++) {
        c->p[ID(0, j)] = pb[2 * n0 + j];
        c->p[ID(n0 - 1, j)] = pb[2 * n0 + n1 + j];
    }
    c->p[ID(0, 0)] = 0.5 * (c->p[ID(1, 0)] + c->p[ID(0, 1)]);
    c->p[ID(n0 - 1, 0)] = 0.5 * (c->p[ID(n0 - 2, 0)] + c->p[ID(n0 - 1, 1)]);
    c->p[ID(0, n1 - 1)] = 0.5 * (c->p[ID(1, n1 - 1)] + c->p[ID(0, n1 - 2)]);
    c->p[ID(n0 - 1, n1 - 1)] = 0.5 * (c->p[ID(n0 - 2, n1 - 1)] + c->p[ID(n0 - 1, n1 - 2)]);
    if (c->sm) { /* FAS: boundary sources of a coarse level (border entries of sm) */
        for (int i = 0; i < n0; i++) {
            c->p[ID(i, 0)] += c->sm[ID(i, 0)];
            c->p[ID(i, n1 - 1)] += c->sm[ID(i, n1 - 1)];
        }
        for (int j = 1; j < n1 - 1; j++) {
            c->p[ID(0, j)] += c->sm[ID(0, j)];
            c->p[ID(n0 - 1, j)] += c->sm[ID(n0 - 1, j)];
        }
    }
}

void or_cav_apply_bcs(or_cav_case *c)
{
    velocity_temperature_bcs(c);
    temperature_sides(c);
    pressure_walls(c);
}

/* ------------------------------------------------------------------------ */
/* Derivative along one grid direction at node P of the hybrid scheme.
 * mode 2: hybrid (KK upwind when |Pe| > 2), 1: 4th-order central,
 * 0: 2nd-order central.  Same arithmetic as solver.cpp:1124-1172.
 * mode 3 (cavity rows next to a wall with wall_upwind): the switch of
 * solver.cpp:1124 with first-order upwinding in place of KK where the
 * five-point stencil does not fit. */
static inline double conv_deriv(int mode, double fm2, double fm1, double f0, double fp1, double fp2,
                                 double vel, double pec, double h, double blend)
{
    const int central = (pec <= 2 && pec > -2);
    const double w = blend > 0.0 ? fmin(1.0, (fabs(pec) - 2.0) / blend) : 1.0;
    const double sg = vel > 0.0 ? 1.0 : -1.0;
    if (mode == 3) { /* next to a wall: 2nd-order central, first-order upwind when |Pe| > 2 */
        const double c2 = 0.5 * (fp1 - fm1) / h;
        if (central) return c2;
        if (blend > 0.0) return c2 + w * sg * (2.0 * f0 - fm1 - fp1) / (2.0 * h);
        return vel > 0.0 ? (f0 - fm1) / h : (fp1 - f0) / h;
    }
    if (mode == 0) return 0.5 * (fp1 - fm1) / h;
    const double xpp = 8.0 * (fp1 - fm1);
    const double xnn = fp2 - fm2;
    const double c4 = (1.0 / 12.0) * (xpp - xnn) / h;
    if (mode == 1 || central) return c4;
    if (blend > 0.0) return c4 + w * sg * (fp2 - 4 * fp1 + 6 * f0 - 4 * fm1 + fm2) / (4.0 * h);
    double ak1 = vel * (-fp2 + 8 * fp1 - 8 * fm1 + fm2) / (12.0 * h);
    const double ak2 = fabs(vel) * (fp2 - 4 * fp1 + 6 * f0 - 4 * fm1 + fm2) / (4.0 * h);
    ak1 = ak1 + ak2;
    return ak1 / vel;
}

/* A_ref phi - phi at an interior node, the solver.cpp:764-847 coefficients
 * (`nd` = Re for u/v, Re*Pr for T); `wide` selects the 25-point row. */
static inline double aref_offdiag(int n0, const double *f, int i, int j, size_t P,
                                  double al, double be, double ga, double pp, double qq,
                                  double h1, double h2, double dt, double nd, int wide)
{
#define F(di, dj) f[ID(i + (di), j + (dj))]
    if (!wide) {
        const double ce = -dt * (al / (h1 * h1) + pp / (2.0 * h1)) / nd;
        const double cw = -dt * (al / (h1 * h1) - pp / (2.0 * h1)) / nd;
        const double cn = -dt * (ga / (h2 * h2) + qq / (2.0 * h2)) / nd;
        const double cs = -dt * (ga / (h2 * h2) - qq / (2.0 * h2)) / nd;
        const double cx = dt * be / (2.0 * h1 * h2 * nd);
        const double cp = dt * 2.0 * (al / (h1 * h1) + ga / (h2 * h2)) / nd;
        return cp * f[P] + ce * F(1, 0) + cw * F(-1, 0) + cn * F(0, 1) + cs * F(0, -1) +
               cx * (F(1, 1) + F(-1, -1) - F(-1, 1) - F(1, -1));
    }
    const double ce = (-dt) * ((4.0 * al) / (3.0 * h1 * h1) + (2.0 * pp) / (3.0 * h1)) / nd;
    const double cw = (-dt) * ((4.0 * al) / (3.0 * h1 * h1) - (2.0 * pp) / (3.0 * h1)) / nd;
    const double cn = (-dt) * ((4.0 * ga) / (3.0 * h2 * h2) + (2.0 * qq) / (3.0 * h2)) / nd;
    const double cs = (-dt) * ((4.0 * ga) / (3.0 * h2 * h2) - (2.0 * qq) / (3.0 * h2)) / nd;
    const double cne = (-dt) * (-8.0 * be / (9.0 * h1 * h2)) / nd;   /* ne, sw */
    const double cnw = (-dt) * (8.0 * be / (9.0 * h1 * h2)) / nd;    /* nw, se */
    const double cnn = (-dt) * (-ga / (12.0 * h2 * h2) - qq / (12.0 * h2)) / nd;
    const double css = (-dt) * (-ga / (12.0 * h2 * h2) + qq / (12.0 * h2)) / nd;
    const double cee = (-dt) * (-al / (12.0 * h1 * h1) - pp / (12.0 * h1)) / nd;
    const double cww = (-dt) * (-al / (12.0 * h1 * h1) + pp / (12.0 * h1)) / nd;
    const double c22 = (-dt) * (-be / (72.0 * h1 * h2)) / nd;        /* nnee, ssww */
    const double c2m = (-dt) * (be / (72.0 * h1 * h2)) / nd;         /* nnww, ssee */
    const double c12 = (-dt) * (be / (9.0 * h1 * h2)) / nd;          /* nne, ssw, nee, sww */
    const double c1m = (-dt) * (-be / (9.0 * h1 * h2)) / nd;         /* nnw, sse, nww, see */
    const double cp = dt * (5.0 * al / (2.0 * h1 * h1) + 5.0 * ga / (2.0 * h2 * h2)) / nd;
    return cp * f[P] + ce * F(1, 0) + cw * F(-1, 0) + cn * F(0, 1) + cs * F(0, -1) +
           cne * (F(1, 1) + F(-1, -1)) + cnw * (F(-1, 1) + F(1, -1)) +
           cnn * F(0, 2) + css * F(0, -2) + cee * F(2, 0) + cww * F(-2, 0) +
           c22 * (F(2, 2) + F(-2, -2)) + c2m * (F(-2, 2) + F(2, -2)) +
           c12 * (F(1, 2) + F(-1, -2) + F(2, 1) + F(-2, -1)) +
           c1m * (F(-1, 2) + F(1, -2) + F(-2, 1) + F(2, -1));
#undef F
}

/* Low-order predictor operator at node P: 9 coefficients (diagonal in c[4]). */
static inline void alo_coefs(double al, double be, double ga, double pp, double qq,
                             double U, double V, double h1, double h2, double dt, double nd,
                             double c[9])
{
    c[OR_K9(1, 0)] = -dt * (al / (h1 * h1) + pp / (2.0 * h1)) / nd + dt * fmin(U, 0.0) / h1;
    c[OR_K9(-1, 0)] = -dt * (al / (h1 * h1) - pp / (2.0 * h1)) / nd - dt * fmax(U, 0.0) / h1;
    c[OR_K9(0, 1)] = -dt * (ga / (h2 * h2) + qq / (2.0 * h2)) / nd + dt * fmin(V, 0.0) / h2;
    c[OR_K9(0, -1)] = -dt * (ga / (h2 * h2) - qq / (2.0 * h2)) / nd - dt * fmax(V, 0.0) / h2;
    const double cx = dt * be / (2.0 * h1 * h2 * nd);
    c[OR_K9(1, 1)] = cx; c[OR_K9(-1, -1)] = cx;
    c[OR_K9(-1, 1)] = -cx; c[OR_K9(1, -1)] = -cx;
    c[OR_K9(0, 0)] = 1.0 + dt * 2.0 * (al / (h1 * h1) + ga / (h2 * h2)) / nd +
                     dt * (fabs(U) / h1 + fabs(V) / h2);
}

void or_cav_residuals(or_cav_case *c, double *Ru, double *Rv, double *RT, double *dout)
{
    const int n0 = c->n0, n1 = c->n1;
    const double h1 = c->h1, h2 = c->h2;
    const or_cav_params *pr = &c->prm;
    const double Re = pr->Re, Pr = pr->Pr, Ri = pr->Ri, dt = pr->dt;
    double *U = c->U, *V = c->V;
    for (size_t k = 0; k < (size_t)n0 * n1; k++) {
        U[k] = c->xix[k] * c->u[k] + c->xiy[k] * c->v[k];
        V[k] = c->etx[k] * c->u[k] + c->ety[k] * c->v[k];
    }
    const double *fl[3] = {c->u, c->v, c->T};
    for (int j = 1; j < n1 - 1; j++)
        for (int i = 1; i < n0 - 1; i++) {
            const size_t P = ID(i, j);
            const int fit_i = (i >= 2 && i <= n0 - 3), fit_j = (j >= 2 && j <= n1 - 3);
            const int hyb = fit_i && fit_j;
            int mode_i = hyb ? 2 : (fit_i ? 1 : 0);
            int mode_j = hyb ? 2 : (fit_j ? 1 : 0);
            if (pr->wall_upwind == 1) {
                mode_i = fit_i ? 2 : 3;
                mode_j = fit_j ? 2 : 3;
            } else if (pr->wall_upwind == 2) {
                mode_i = mode_j = 3;
            }
            const double al = c->alph[P], be = c->beta[P], ga = c->gamma[P];
            const double pp = c->p1[P], qq = c->q1[P];
            double conv[3];
            for (int k = 0; k < 3; k++) {
                const double *f = fl[k];
                const double sc = (k <= 1) ? Re : Re * Pr;
                const double pec1 = U[P] * sc * h1 / al;
                const double pec2 = V[P] * sc * h2 / ga;
                const double fm2i = fit_i ? f[ID(i - 2, j)] : 0.0, fp2i = fit_i ? f[ID(i + 2, j)] : 0.0;
                const double fm2j = fit_j ? f[ID(i, j - 2)] : 0.0, fp2j = fit_j ? f[ID(i, j + 2)] : 0.0;
                const double dxi = conv_deriv(mode_i, fm2i, f[ID(i - 1, j)], f[P], f[ID(i + 1, j)], fp2i, U[P], pec1, h1, pr->hyb_blend);
                const double det = conv_deriv(mode_j, fm2j, f[ID(i, j - 1)], f[P], f[ID(i, j + 1)], fp2j, V[P], pec2, h2, pr->hyb_blend);
                conv[k] = U[P] * dxi + V[P] * det;
            }
            const double *p = c->p;
            const double dp_dxi = (p[ID(i + 1, j)] - p[ID(i - 1, j)]) / (2.0 * h1);
            const double dp_de = (p[ID(i, j + 1)] - p[ID(i, j - 1)]) / (2.0 * h2);
            const double dp_dx = c->xix[P] * dp_dxi + c->etx[P] * dp_de;
            const double dp_dy = c->xiy[P] * dp_dxi + c->ety[P] * dp_de;
            const int wide = hyb;
            Ru[P] = dt * (-conv[0] - dp_dx) - aref_offdiag(n0, c->u, i, j, P, al, be, ga, pp, qq, h1, h2, dt, Re, wide);
            Rv[P] = dt * (-conv[1] - dp_dy + Ri * c->T[P]) - aref_offdiag(n0, c->v, i, j, P, al, be, ga, pp, qq, h1, h2, dt, Re, wide);
            RT[P] = -dt * conv[2] - aref_offdiag(n0, c->T, i, j, P, al, be, ga, pp, qq, h1, h2, dt, Re * Pr, wide);
            if (c->su) { /* FAS source terms of a coarse level */
                Ru[P] += dt * c->su[P];
                Rv[P] += dt * c->sv[P];
                RT[P] += dt * c->sT[P];
            }
            double cl[9];
            alo_coefs(al, be, ga, pp, qq, U[P], V[P], h1, h2, dt, Re, cl);
            dout[P] = dt * pr->alpha_u / cl[4];
        }
    /* d on the border: copy of the nearest interior node */
    for (int i = 1; i < n0 - 1; i++) {
        dout[ID(i, 0)] = dout[ID(i, 1)];
        dout[ID(i, n1 - 1)] = dout[ID(i, n1 - 2)];
    }
    for (int j = 0; j < n1; j++) {
        const int jj = j == 0 ? 1 : (j == n1 - 1 ? n1 - 2 : j);
        dout[ID(0, j)] = dout[ID(1, jj)];
        dout[ID(n0 - 1, j)] = dout[ID(n0 - 2, jj)];
    }
}

void or_cav_lo_diag(const or_cav_case *c, double *d_out)
{
    const int n0 = c->n0, n1 = c->n1;
    for (int j = 1; j < n1 - 1; j++)
        for (int i = 1; i < n0 - 1; i++) {
            const size_t P = ID(i, j);
            double cl[9];
            const double U = c->xix[P] * c->u[P] + c->xiy[P] * c->v[P];
            const double V = c->etx[P] * c->u[P] + c->ety[P] * c->v[P];
            alo_coefs(c->alph[P], c->beta[P], c->gamma[P], c->p1[P], c->q1[P], U, V, c->h1, c->h2,
                      c->prm.dt, c->prm.Re, cl);
            d_out[P] = cl[4];
        }
}

/* Jacobi sweeps on (A_lo / alpha) delta = R for u, v (shared operator) and T. */
static void predictor(or_cav_case *c)
{
    const int n0 = c->n0, n1 = c->n1;
    const size_t N = (size_t)n0 * n1;
    const or_cav_params *pr = &c->prm;
    double *du = c->du, *dv = c->dv, *dT = c->dT, *nu = c->tu, *nv = c->tv, *nT = c->tT;
    memset(du, 0, N * sizeof(double)); memset(dv, 0, N * sizeof(double)); memset(dT, 0, N * sizeof(double));
    memset(nu, 0, N * sizeof(double)); memset(nv, 0, N * sizeof(double)); memset(nT, 0, N * sizeof(double));
    for (int s = 0; s < pr->mom_sweeps; s++) {
        for (int j = 1; j < n1 - 1; j++)
            for (int i = 1; i < n0 - 1; i++) {
                const size_t P = ID(i, j);
                double cu[9], ct[9];
                alo_coefs(c->alph[P], c->beta[P], c->gamma[P], c->p1[P], c->q1[P], c->U[P], c->V[P],
                          c->h1, c->h2, pr->dt, pr->Re, cu);
                alo_coefs(c->alph[P], c->beta[P], c->gamma[P], c->p1[P], c->q1[P], c->U[P], c->V[P],
                          c->h1, c->h2, pr->dt, pr->Re * pr->Pr, ct);
                double su = c->Ru[P], sv = c->Rv[P], st = c->RT[P];
                for (int dj = -1; dj <= 1; dj++)
                    for (int di = -1; di <= 1; di++) {
                        if (di == 0 && dj == 0) continue;
                        const int k = (dj + 1) * 3 + (di + 1);
                        const size_t Q = ID(i + di, j + dj);
                        su -= cu[k] * du[Q];
                        sv -= cu[k] * dv[Q];
                        st -= ct[k] * dT[Q];
                    }
                nu[P] = su * pr->alpha_u / cu[4];
                nv[P] = sv * pr->alpha_u / cu[4];
                nT[P] = st * pr->alpha_t / ct[4];
            }
        double *t;
        t = du; du = nu; nu = t;
        t = dv; dv = nv; nv = t;
        t = dT; dT = nT; nT = t;
    }
    /* results end up in c->du/dv/dT regardless of the swap parity */
    if (du != c->du) { memcpy(c->du, du, N * sizeof(double)); memcpy(c->dv, dv, N * sizeof(double)); memcpy(c->dT, dT, N * sizeof(double)); }
}

/* q = J-weighted divergence of the face velocities of solver.cpp:1389-1448,
 * with the pressure-gradient-free velocity taken as u* + d grad p (nodal,
 * central) and dt replaced by the face d_f; faces touching a wall carry no
 * flux.  Returns sum |q| over the interior. */
static double face_div(or_cav_case *c)
{
    const int n0 = c->n0, n1 = c->n1;
    const double h1 = c->h1, h2 = c->h2;
    const double *p = c->p, *d = c->d, *xix = c->xix, *xiy = c->xiy, *etx = c->etx, *ety = c->ety;
    const double *jac = c->ajac;
    double *uh = c->tu, *vh = c->tv;
    for (int j = 1; j < n1 - 1; j++)
        for (int i = 1; i < n0 - 1; i++) {
            const size_t P = ID(i, j);
            const double dp_dxi = (p[ID(i + 1, j)] - p[ID(i - 1, j)]) / (2.0 * h1);
            const double dp_de = (p[ID(i, j + 1)] - p[ID(i, j - 1)]) / (2.0 * h2);
            uh[P] = c->us[P] + d[P] * (xix[P] * dp_dxi + etx[P] * dp_de);
            vh[P] = c->vs[P] + d[P] * (xiy[P] * dp_dxi + ety[P] * dp_de);
        }
    double sum = 0.0;
    for (int j = 1; j < n1 - 1; j++)
        for (int i = 1; i < n0 - 1; i++) {
            const int ipp = i + 1, inn = i - 1, jpp = j + 1, jnn = j - 1;
            const size_t P = ID(i, j), E = ID(ipp, j), W = ID(inn, j), Nn = ID(i, jpp), S = ID(i, jnn);
            double fe = 0.0, fw = 0.0, gn = 0.0, gs = 0.0;
            if (ipp != n0 - 1) {
                const double de = 0.5 * (d[P] + d[E]);
                const double dpdxi = (p[E] - p[P]) / h1;
                const double dpde = (p[ID(ipp, jpp)] + p[Nn] - p[S] - p[ID(ipp, jnn)]) / (4.0 * h2);
                const double uf = 0.5 * (uh[P] + uh[E]) - 0.5 * de * ((xix[P] + xix[E]) * dpdxi + (etx[P] + etx[E]) * dpde);
                const double vf = 0.5 * (vh[P] + vh[E]) - 0.5 * de * ((xiy[P] + xiy[E]) * dpdxi + (ety[P] + ety[E]) * dpde);
                fe = 0.5 * (jac[P] * xix[P] + jac[E] * xix[E]) * uf + 0.5 * (jac[P] * xiy[P] + jac[E] * xiy[E]) * vf;
            }
            if (inn != 0) {
                const double dw = 0.5 * (d[P] + d[W]);
                const double dpdxi = (p[P] - p[W]) / h1;
                const double dpde = (p[Nn] + p[ID(inn, jpp)] - p[S] - p[ID(inn, jnn)]) / (4.0 * h2);
                const double uf = 0.5 * (uh[P] + uh[W]) - 0.5 * dw * ((xix[P] + xix[W]) * dpdxi + (etx[P] + etx[W]) * dpde);
                const double vf = 0.5 * (vh[P] + vh[W]) - 0.5 * dw * ((xiy[P] + xiy[W]) * dpdxi + (ety[P] + ety[W]) * dpde);
                fw = 0.5 * (jac[P] * xix[P] + jac[W] * xix[W]) * uf + 0.5 * (jac[P] * xiy[P] + jac[W] * xiy[W]) * vf;
            }
            if (jpp != n1 - 1) {
                const double dn = 0.5 * (d[P] + d[Nn]);
                const double dpdxi = (p[ID(ipp, jpp)] - p[ID(inn, jpp)] + p[E] - p[W]) / (4.0 * h1);
                const double dpde = (p[Nn] - p[P]) / h2;
                const double uf = 0.5 * (uh[P] + uh[Nn]) - 0.5 * dn * ((xix[P] + xix[Nn]) * dpdxi + (etx[P] + etx[Nn]) * dpde);
                const double vf = 0.5 * (vh[P] + vh[Nn]) - 0.5 * dn * ((xiy[P] + xiy[Nn]) * dpdxi + (ety[P] + ety[Nn]) * dpde);
                gn = 0.5 * (jac[P] * etx[P] + jac[Nn] * etx[Nn]) * uf + 0.5 * (jac[P] * ety[P] + jac[Nn] * ety[Nn]) * vf;
            }
            if (jnn != 0) {
                const double ds = 0.5 * (d[P] + d[S]);
                const double dpdxi = (p[E] - p[W] + p[ID(ipp, jnn)] - p[ID(inn, jnn)]) / (4.0 * h1);
                const double dpde = (p[P] - p[S]) / h2;
                const double uf = 0.5 * (uh[P] + uh[S]) - 0.5 * ds * ((xix[P] + xix[S]) * dpdxi + (etx[P] + etx[S]) * dpde);
                const double vf = 0.5 * (vh[P] + vh[S]) - 0.5 * ds * ((xiy[P] + xiy[S]) * dpdxi + (ety[P] + ety[S]) * dpde);
                gs = 0.5 * (jac[P] * etx[P] + jac[S] * etx[S]) * uf + 0.5 * (jac[P] * ety[P] + jac[S] * ety[S]) * vf;
            }
            double qq = (fe - fw) / h1 + (gn - gs) / h2;
            if (c->sm) qq -= c->sm[P];
            c->q[P] = qq;
            sum += fabs(qq);
        }
    return sum;
}

static void correct(or_cav_case *c)
{
    const int n0 = c->n0, n1 = c->n1;
    const double h1 = c->h1, h2 = c->h2;
    double *pc = c->pc;
    /* p' wall copies (solver.cpp:1474-1482 on each wall), corners from the diagonal */
    for (int i = 1; i < n0 - 1; i++) {
        pc[ID(i, 0)] = pc[ID(i, 1)];
        pc[ID(i, n1 - 1)] = pc[ID(i, n1 - 2)];
    }
    for (int j = 1; j < n1 - 1; j++) {
        pc[ID(0, j)] = pc[ID(1, j)];
        pc[ID(n0 - 1, j)] = pc[ID(n0 - 2, j)];
    }
    pc[ID(0, 0)] = pc[ID(1, 1)];
    pc[ID(n0 - 1, 0)] = pc[ID(n0 - 2, 1)];
    pc[ID(0, n1 - 1)] = pc[ID(1, n1 - 2)];
    pc[ID(n0 - 1, n1 - 1)] = pc[ID(n0 - 2, n1 - 2)];
    for (int j = 1; j < n1 - 1; j++)
        for (int i = 1; i < n0 - 1; i++) {
            const size_t P = ID(i, j);
            double dpx = 0.5 * (pc[ID(i + 1, j)] - pc[ID(i - 1, j)]) / h1;
            double dpe = 0.5 * (pc[ID(i, j + 1)] - pc[ID(i, j - 1)]) / h2;
            if (c->prm.plant == 6) { dpx = -dpx; dpe = -dpe; }
            c->u[P] = c->us[P] - c->d[P] * (c->xix[P] * dpx + c->etx[P] * dpe);
            c->v[P] = c->vs[P] - c->d[P] * (c->xiy[P] * dpx + c->ety[P] * dpe);
            c->p[P] = c->p[P] + c->prm.alpha_p * pc[P];
        }
}

static double dmax_int(or_cav_case *c)
{
    const int n0 = c->n0, n1 = c->n1;
    const double h1 = c->h1, h2 = c->h2;
    const double *u = c->u, *v = c->v;
    double dmax = 0.0;
    for (int j = 1; j < n1 - 1; j++)
        for (int i = 1; i < n0 - 1; i++) {
            const size_t P = ID(i, j);
            const double d = c->xix[P] * (u[ID(i + 1, j)] - u[ID(i - 1, j)]) / (2 * h1) +
                             c->etx[P] * (u[ID(i, j + 1)] - u[ID(i, j - 1)]) / (2 * h2) +
                             c->ety[P] * (v[ID(i, j + 1)] - v[ID(i, j - 1)]) / (2 * h2) +
                             c->xiy[P] * (v[ID(i + 1, j)] - v[ID(i - 1, j)]) / (2 * h1);
            if (d > dmax) dmax = d;
        }
    return dmax;
}

/* Nusselt number of the j=0 (hot) or j=n1-1 (cold) wall: the solver.cpp:1964-1978
 * trapezoid of J*gamma*dT/deta, signed so that heat entering the fluid from the
 * hot wall is positive, divided by the wall length. */
static double nusselt(const or_cav_case *c, int top)
{
    const int n0 = c->n0, n1 = c->n1;
    const double h1 = c->h1, h2 = c->h2;
    const double *T = c->T;
    const int j = top ? n1 - 1 : 0, s = top ? -1 : 1;
    double acc = 0.0;
    for (int i = 0; i < n0 - 1; i++) {
        const int ip = i + 1;
        const size_t A = ID(i, j), B = ID(ip, j);
        const double tj1 = c->ajac[A] * (c->etx[A] * c->etx[A] + c->ety[A] * c->ety[A]);
        const double tj2 = c->ajac[B] * (c->etx[B] * c->etx[B] + c->ety[B] * c->ety[B]);
        const double fh1 = tj1 * (4 * T[ID(i, j + s)] - 3 * T[A] - T[ID(i, j + 2 * s)]) / (2 * h2);
        const double fh2 = tj2 * (4 * T[ID(ip, j + s)] - 3 * T[B] - T[ID(ip, j + 2 * s)]) / (2 * h2);
        acc = acc + 0.5 * (fh1 + fh2) * h1;
    }
    const double dT = c->prm.t_hot - c->prm.t_cold;
    const double len = top ? c->wall_len_cold : c->wall_len_hot;
    return -acc / (len * (dT != 0.0 ? dT : 1.0));
}

int or_cav_outer(or_cav_case *c, or_cav_stats *st)
{
    const int n0 = c->n0, n1 = c->n1;
    const size_t N = (size_t)n0 * n1;
    const or_cav_params *pr = &c->prm;
    /* 1-2: residuals of the reference operators and the d factors */
    or_cav_residuals(c, c->Ru, c->Rv, c->RT, c->d);
    double el = 0.0, ml = 0.0;
    for (int j = 1; j < n1 - 1; j++)
        for (int i = 1; i < n0 - 1; i++) {
            const size_t P = ID(i, j);
            el += fabs(c->RT[P]);
            ml += fabs(c->Ru[P]) + fabs(c->Rv[P]);
        }
    /* 3: predictor */
    predictor(c);
    memcpy(c->us, c->u, N * sizeof(double));
    memcpy(c->vs, c->v, N * sizeof(double));
    for (int j = 1; j < n1 - 1; j++)
        for (int i = 1; i < n0 - 1; i++) {
            const size_t P = ID(i, j);
            c->us[P] += c->du[P];
            c->vs[P] += c->dv[P];
            c->T[P] += c->dT[P];
        }
    /* 4: face divergence */
    const double mass = face_div(c);
    /* 5: p' solve */
    inject_d(c);
    memset(c->pc, 0, N * sizeof(double));
    double relres = 0.0;
    const int its = pr->p_maxit < 0 ? vcycle_solve(c, c->q, c->pc, &relres)
                                     : bicgstab(c, c->q, c->pc, pr->p_rtol, pr->p_maxit, &relres);
    project_mean(n0, n1, c->pc);
    /* 6-7: correction and boundary conditions */
    correct(c);
    temperature_sides(c);
    pressure_walls(c);
    c->iter++;
    if (c->iter == 1) { c->mass0 = mass; c->energy0 = el; }
    if (st) {
        st->mass_l1 = mass;
        st->energy_l1 = el;
        st->mom_l1 = ml;
        st->dmax = dmax_int(c);
        st->nu_hot = nusselt(c, 0);
        st->nu_cold = -nusselt(c, 1);
        st->p_iters = its;
        st->p_relres = relres;
        st->mass_res = c->mass0 > 0.0 ? mass / c->mass0 : 0.0;
        st->energy_res = c->energy0 > 0.0 ? el / c->energy0 : 0.0;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FAS multigrid outer cycle with SIMPLE iterations as the smoother.          */
/*                                                                            */
/* The steady equations of a level are rho(x) = 0 with                        */
/*   rho_u = R_u/dt + s_u, rho_v = R_v/dt + s_v, rho_T = R_T/dt + s_T          */
/*   (R: the residual of or_cav_residuals, dt the pseudo-time step) and         */
/*   rho_m = q(x) - s_m (q: the J-weighted face divergence of the state's face */
/*   velocities with its own d factors).  A coarse level's sources are set so  */
/* that its rho at the restricted state equals the full-weighting restriction */
/* of the fine rho (the FAS tau correction); a SIMPLE iteration of a level    */
/* drives its rho to zero, and the coarse change of the state, interpolated   */
/* bilinearly, corrects the fine interior.                                    */

/* rho of the current state on the interior (0 on walls) */
static void state_residual(or_cav_case *c, double *ru, double *rv, double *rT, double *rm)
{
    const int n0 = c->n0, n1 = c->n1;
    const size_t N = (size_t)n0 * n1;
    const double idt = 1.0 / c->prm.dt;
    or_cav_residuals(c, c->Ru, c->Rv, c->RT, c->d);
    memcpy(c->us, c->u, N * sizeof(double));
    memcpy(c->vs, c->v, N * sizeof(double));
    face_div(c);
    memset(ru, 0, N * sizeof(double)); memset(rv, 0, N * sizeof(double));
    memset(rT, 0, N * sizeof(double)); memset(rm, 0, N * sizeof(double));
    for (int j = 1; j < n1 - 1; j++)
        for (int i = 1; i < n0 - 1; i++) {
            const size_t P = ID(i, j);
            ru[P] = c->Ru[P] * idt;
            rv[P] = c->Rv[P] * idt;
            rT[P] = c->RT[P] * idt;
            rm[P] = c->q[P];
        }
}

/* Residuals of the wall relations at the current state: the change that the
 * adiabatic side-wall update of a SIMPLE iteration would make (wT, side walls),
 * and the change that its wall-pressure update would make after that side-wall
 * update (wp, the whole border), each with the level's sources as they are set;
 * the state is left as it was. */
static void wall_residual(or_cav_case *c, double *wT, double *wp)
{
    const int n0 = c->n0, n1 = c->n1;
    const size_t N = (size_t)n0 * n1;
    double *Ts = c->bi[6], *ps = c->bi[7];
    memcpy(Ts, c->T, N * sizeof(double));
    memcpy(ps, c->p, N * sizeof(double));
    temperature_sides(c);
    if (wT)
        for (size_t k = 0; k < N; k++) wT[k] = c->T[k] - Ts[k];
    if (wp) {
        pressure_walls(c);
        for (size_t k = 0; k < N; k++) wp[k] = c->p[k] - ps[k];
    }
    memcpy(c->T, Ts, N * sizeof(double));
    memcpy(c->p, ps, N * sizeof(double));
}

/* full weighting of fine interior values onto coarse interior node (I, J) */
static double fw(const double *r, int n0, int I, int J)
{
    const int i = 2 * I, j = 2 * J;
    return (4.0 * r[ID(i, j)] + 2.0 * (r[ID(i - 1, j)] + r[ID(i + 1, j)] + r[ID(i, j - 1)] + r[ID(i, j + 1)]) +
            (r[ID(i - 1, j - 1)] + r[ID(i + 1, j - 1)] + r[ID(i - 1, j + 1)] + r[ID(i + 1, j + 1)])) / 16.0;
}

/* bilinear interpolation of the coarse change e = xc - x0 at fine node (i, j) */
static double prolong_at(const double *xc, const double *x0, int m0, int i, int j)
{
#define E(I, J) (xc[(size_t)(J) * m0 + (I)] - x0[(size_t)(J) * m0 + (I)])
    const int I = i >> 1, J = j >> 1;
    if ((i & 1) == 0 && (j & 1) == 0) return E(I, J);
    if ((j & 1) == 0) return 0.5 * (E(I, J) + E(I + 1, J));
    if ((i & 1) == 0) return 0.5 * (E(I, J) + E(I, J + 1));
    return 0.25 * (E(I, J) + E(I + 1, J) + E(I, J + 1) + E(I + 1, J + 1));
#undef E
}

static void fas_cycle(or_cav_case *c, or_cav_stats *st)
{
    const or_cav_params *pr = &c->prm;
    or_cav_case *cc = c->coarse;
    if (!cc) {
        for (int k = 0; k < pr->fas_coarse; k++) or_cav_outer(c, st);
        return;
    }
    for (int k = 0; k < pr->fas_pre; k++) or_cav_outer(c, st);
    const int n0 = c->n0, n1 = c->n1, m0 = cc->n0, m1 = cc->n1;
    const size_t M = (size_t)m0 * m1;
    double *ru = c->bi[0], *rv = c->bi[1], *rT = c->bi[2], *rm = c->bi[3];
    state_residual(c, ru, rv, rT, rm);
    double *wT = c->bi[4], *wp = c->bi[5];
    wall_residual(c, wT, wp);
    /* restricted state (injection) */
    for (int J = 0; J < m1; J++)
        for (int I = 0; I < m0; I++) {
            const size_t Q = (size_t)J * m0 + I, P = ID(2 * I, 2 * J);
            cc->u[Q] = c->u[P]; cc->v[Q] = c->v[P]; cc->T[Q] = c->T[P]; cc->p[Q] = c->p[P];
        }
    memcpy(cc->u0, cc->u, M * sizeof(double)); memcpy(cc->v0, cc->v, M * sizeof(double));
    memcpy(cc->T0, cc->T, M * sizeof(double)); memcpy(cc->p0, cc->p, M * sizeof(double));
    /* coarse residuals of the restricted state without sources */
    double *cu = cc->bi[0], *cv = cc->bi[1], *cT = cc->bi[2], *cm = cc->bi[3];
    double *cwT = cc->bi[4], *cwp = cc->bi[5];
    double *keep[4] = {cc->su, cc->sv, cc->sT, cc->sm};
    cc->su = cc->sv = cc->sT = cc->sm = NULL;
    state_residual(cc, cu, cv, cT, cm);
    wall_residual(cc, cwT, NULL);
    cc->su = keep[0]; cc->sv = keep[1]; cc->sT = keep[2]; cc->sm = keep[3];
    memset(cc->su, 0, M * sizeof(double)); memset(cc->sv, 0, M * sizeof(double));
    memset(cc->sT, 0, M * sizeof(double)); memset(cc->sm, 0, M * sizeof(double));
    for (int J = 1; J < m1 - 1; J++)
        for (int I = 1; I < m0 - 1; I++) {
            const size_t Q = (size_t)J * m0 + I;
            cc->su[Q] = fw(ru, n0, I, J) - cu[Q];
            cc->sv[Q] = fw(rv, n0, I, J) - cv[Q];
            cc->sT[Q] = fw(rT, n0, I, J) - cT[Q];
            cc->sm[Q] = cm[Q] - fw(rm, n0, I, J);
        }
    /* boundary sources: the injected fine wall residual minus the coarse one; the
     * pressure relation is evaluated after the side-wall update with its source */
    for (int J = 1; J < m1 - 1; J++) {
        const size_t P0 = ID(0, 2 * J), P1 = ID(n0 - 1, 2 * J);
        cc->sT[(size_t)J * m0] = wT[P0] - cwT[(size_t)J * m0];
        cc->sT[(size_t)J * m0 + m0 - 1] = wT[P1] - cwT[(size_t)J * m0 + m0 - 1];
    }
    cc->sm = NULL;
    wall_residual(cc, NULL, cwp);
    cc->sm = keep[3];
    for (int J = 0; J < m1; J++)
        for (int I = 0; I < m0; I++) {
            if (I != 0 && J != 0 && I != m0 - 1 && J != m1 - 1) continue;
            const size_t Q = (size_t)J * m0 + I;
            cc->sm[Q] = wp[ID(2 * I, 2 * J)] - cwp[Q];
        }
    fas_cycle(cc, NULL);
    /* coarse-grid correction of the fine interior, then the fine wall conditions */
    for (int j = 1; j < n1 - 1; j++)
        for (int i = 1; i < n0 - 1; i++) {
            const size_t P = ID(i, j);
            c->u[P] += prolong_at(cc->u, cc->u0, m0, i, j);
            c->v[P] += prolong_at(cc->v, cc->v0, m0, i, j);
            c->T[P] += prolong_at(cc->T, cc->T0, m0, i, j);
            c->p[P] += prolong_at(cc->p, cc->p0, m0, i, j);
        }
    or_cav_apply_bcs(c);
    for (int k = 0; k < pr->fas_post; k++) or_cav_outer(c, st);
}

int or_cav_fas_cycle(or_cav_case *c, or_cav_stats *st)
{
    or_cav_stats tmp;
    if (!c->coarse) return or_cav_outer(c, st ? st : &tmp);
    fas_cycle(c, st ? st : &tmp);
    return 0;
}

int or_cav_fas_depth(const or_cav_case *c)
{
    int d = 0;
    for (; c; c = c->coarse) d++;
    return d;
}

void or_cav_state_residual(or_cav_case *c, double *ru, double *rv, double *rT, double *rm)
{
    state_residual(c, ru, rv, rT, rm);
}

void or_cav_p_apply(const or_cav_case *c, const double *d, const double *x, double *y)
{
    mg_level L = c->lev[0];
    L.d = (double *)d;
    p9_apply_lvl(&L, x, y);
}

int or_cav_p_solve(or_cav_case *c, const double *d, const double *q, double *x,
                   double rtol, int maxit, double *relres)
{
    const size_t N = (size_t)c->n0 * c->n1;
    memcpy(c->d, d, N * sizeof(double));
    inject_d(c);
    if (maxit < 0) return vcycle_solve(c, q, x, relres);
    return bicgstab(c, q, x, rtol, maxit, relres);
}

void or_cav_face_div(or_cav_case *c, const double *us, const double *vs, const double *p,
                     const double *d, double *q)
{
    const size_t N = (size_t)c->n0 * c->n1;
    memcpy(c->us, us, N * sizeof(double));
    memcpy(c->vs, vs, N * sizeof(double));
    memcpy(c->p, p, N * sizeof(double));
    memcpy(c->d, d, N * sizeof(double));
    memset(c->q, 0, N * sizeof(double));
    face_div(c);
    memcpy(q, c->q, N * sizeof(double));
}

void or_cav_correct(or_cav_case *c, const double *us, const double *vs, const double *pc,
                    const double *d, double *u, double *v, double *p, double *pc_out)
{
    const size_t N = (size_t)c->n0 * c->n1;
    memcpy(c->us, us, N * sizeof(double));
    memcpy(c->vs, vs, N * sizeof(double));
    memcpy(c->pc, pc, N * sizeof(double));
    memcpy(c->d, d, N * sizeof(double));
    memcpy(c->u, u, N * sizeof(double));
    memcpy(c->v, v, N * sizeof(double));
    memcpy(c->p, p, N * sizeof(double));
    correct(c);
    memcpy(u, c->u, N * sizeof(double));
    memcpy(v, c->v, N * sizeof(double));
    memcpy(p, c->p, N * sizeof(double));
    if (pc_out) memcpy(pc_out, c->pc, N * sizeof(double));
}

/* last p' right-hand side and solution (diagnostics for tests) */
void or_cav_last_q(const or_cav_case *c, double *q, double *pc)
{
    const size_t N = (size_t)c->n0 * c->n1;
    if (q) memcpy(q, c->q, N * sizeof(double));
    if (pc) memcpy(pc, c->pc, N * sizeof(double));
}
