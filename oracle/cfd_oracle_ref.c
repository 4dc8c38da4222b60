/*
 * CPU oracle, reference mode -- TEST INFRASTRUCTURE ONLY (see cfd_oracle.h).
 *
 * Restates the reference solver's hot-path stages on its periodic O-grid in a
 * contiguous (j*n0 + i) layout.  Loop orders, index wraps and floating-point
 * expression order follow the cited solver.cpp lines so that results agree
 * with the compiled reference to rounding (usually bit for bit).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "cfd_oracle.h"

#define ID(i, j) ((size_t)(j) * (size_t)n0 + (size_t)(i))

/* Neighbour columns of column i (0 <= i <= n0-2) on the periodic O-grid.
 * Column n0-1 duplicates column 0; the east neighbours are the literal columns
 * i+1 / i+2 (so n0-1 is addressed directly), except i+2 at i = n0-2 which the
 * reference maps to 1 (solver.cpp:1076-1102, 2316-2328). */
typedef struct { int inn, ipp, inn2, ipp2; } cols_t;

static inline cols_t pcols(int i, int n0) {
    cols_t c;
    c.inn = (i == 0) ? n0 - 2 : i - 1;
    c.ipp = i + 1;
    c.inn2 = (i == 0) ? n0 - 3 : (i == 1) ? n0 - 2 : i - 2;
    c.ipp2 = (i == n0 - 2) ? 1 : i + 2;
    return c;
}

/* ------------------------------------------------------------------------ */
/* solver.cpp:751-939                                                        */
void or_assemble_mom_ref(int n0, int n1, double h1, double h2,
                         const double *alph, const double *beta, const double *gamma,
                         const double *p1, const double *q1,
                         double dt, double Re, double Pr,
                         double *AU, double *AT, double *BU, double *BT)
{
    const size_t N = (size_t)n0 * n1;
    memset(AU, 0, 25 * N * sizeof(double));
    memset(AT, 0, 25 * N * sizeof(double));
    memset(BU, 0, 6 * (size_t)n0 * sizeof(double));
    memset(BT, 0, 6 * (size_t)n0 * sizeof(double));
#define U(di, dj) AU[(size_t)OR_K25(di, dj) * N + id]
#define T(di, dj) AT[(size_t)OR_K25(di, dj) * N + id]
    for (int j = 1; j < n1 - 1; j++) {
        for (int i = 0; i < n0 - 1; i++) {
            const size_t id = ID(i, j);
            const double al = alph[id], be = beta[id], ga = gamma[id];
            const double pp = p1[id], qq = q1[id];
            if (j == 1 || j == n1 - 2) {
                /* second-order 9-point rows, solver.cpp:764-786 */
                U(1, 0) = -dt * (al / (h1 * h1) + pp / (2.0 * h1)) / Re;
                U(-1, 0) = -dt * (al / (h1 * h1) - pp / (2.0 * h1)) / Re;
                U(0, 1) = -dt * (ga / (h2 * h2) + qq / (2.0 * h2)) / Re;
                U(0, -1) = -dt * (ga / (h2 * h2) - qq / (2.0 * h2)) / Re;
                U(1, 1) = dt * be / (2.0 * h1 * h2 * Re);
                U(-1, -1) = U(1, 1);
                U(-1, 1) = -dt * be / (2.0 * h1 * h2 * Re);
                U(1, -1) = U(-1, 1);
                U(0, 0) = 1 + dt * 2.0 * (al / (h1 * h1) + ga / (h2 * h2)) / Re;

                T(1, 0) = -dt * (al / (h1 * h1) + pp / (2.0 * h1)) / (Re * Pr);
                T(-1, 0) = -dt * (al / (h1 * h1) - pp / (2.0 * h1)) / (Re * Pr);
                T(0, 1) = -dt * (ga / (h2 * h2) + qq / (2.0 * h2)) / (Re * Pr);
                T(0, -1) = -dt * (ga / (h2 * h2) - qq / (2.0 * h2)) / (Re * Pr);
                T(1, 1) = dt * (be / (2.0 * h1 * h2)) / (Re * Pr);
                T(-1, -1) = T(1, 1);
                T(-1, 1) = -dt * (be / (2.0 * h1 * h2)) / (Re * Pr);
                T(1, -1) = T(-1, 1);
                T(0, 0) = 1 + dt * 2.0 * (al / (h1 * h1) + ga / (h2 * h2)) / (Re * Pr);
            } else {
                /* fourth-order 25-point rows, solver.cpp:790-847 */
                U(1, 0) = (-dt) * ((4.0 * al) / (3.0 * h1 * h1) + (2.0 * pp) / (3.0 * h1)) / Re;
                U(-1, 0) = (-dt) * ((4.0 * al) / (3.0 * h1 * h1) - (2.0 * pp) / (3.0 * h1)) / Re;
                U(0, 1) = (-dt) * ((4.0 * ga) / (3.0 * h2 * h2) + (2.0 * qq) / (3.0 * h2)) / Re;
                U(0, -1) = (-dt) * ((4.0 * ga) / (3.0 * h2 * h2) - (2.0 * qq) / (3.0 * h2)) / Re;

                U(1, 1) = (-dt) * (-8.0 * be / (9.0 * h1 * h2)) / Re;
                U(-1, 1) = (-dt) * (8.0 * be / (9.0 * h1 * h2)) / Re;
                U(1, -1) = U(-1, 1);
                U(-1, -1) = U(1, 1);

                U(0, 2) = (-dt) * (-ga / (12.0 * h2 * h2) - qq / (12.0 * h2)) / Re;
                U(0, -2) = (-dt) * (-ga / (12.0 * h2 * h2) + qq / (12.0 * h2)) / Re;
                U(2, 0) = (-dt) * (-al / (12.0 * h1 * h1) - pp / (12.0 * h1)) / Re;
                U(-2, 0) = (-dt) * (-al / (12.0 * h1 * h1) + pp / (12.0 * h1)) / Re;

                U(2, 2) = (-dt) * (-be / (72.0 * h1 * h2)) / Re;
                U(-2, 2) = (-dt) * (be / (72.0 * h1 * h2)) / Re;
                U(2, -2) = U(-2, 2);
                U(-2, -2) = U(2, 2);

                U(1, 2) = (-dt) * (be / (9.0 * h1 * h2)) / Re;
                U(-1, 2) = (-dt) * (-be / (9.0 * h1 * h2)) / Re;
                U(1, -2) = U(-1, 2);
                U(-1, -2) = U(1, 2);

                U(2, 1) = U(1, 2);
                U(-2, 1) = U(-1, 2);
                U(2, -1) = U(-1, 2);
                U(-2, -1) = U(1, 2);

                U(0, 0) = 1 + dt * (5.0 * al / (2.0 * h1 * h1) + 5.0 * ga / (2.0 * h2 * h2)) / Re;

                for (int k = 0; k < 25; k++)
                    if (k != OR_K25(0, 0))
                        AT[(size_t)k * N + id] = AU[(size_t)k * N + id] / Pr;
                T(0, 0) = 1 + dt * (5.0 * al / (2.0 * h1 * h1) + 5.0 * ga / (2.0 * h2 * h2)) / (Re * Pr);
            }
            if (j == 1) { /* solver.cpp:850-865: move wall couplings to BU/BT */
                BU[0 * n0 + i] = U(0, -1); BU[1 * n0 + i] = U(1, -1); BU[2 * n0 + i] = U(-1, -1);
                BT[0 * n0 + i] = T(0, -1); BT[1 * n0 + i] = T(1, -1); BT[2 * n0 + i] = T(-1, -1);
                U(0, -1) = U(1, -1) = U(-1, -1) = 0;
                T(0, -1) = T(1, -1) = T(-1, -1) = 0;
            }
            if (j == n1 - 2) { /* solver.cpp:867-881 */
                BU[3 * n0 + i] = U(0, 1); BU[4 * n0 + i] = U(1, 1); BU[5 * n0 + i] = U(-1, 1);
                BT[3 * n0 + i] = T(0, 1); BT[4 * n0 + i] = T(1, 1); BT[5 * n0 + i] = T(-1, 1);
                U(0, 1) = U(1, 1) = U(-1, 1) = 0;
                T(0, 1) = T(1, 1) = T(-1, 1) = 0;
            }
            if (i == 0) { /* periodic duplicate column, solver.cpp:883-937 */
                const size_t dup = ID(n0 - 1, j);
                for (int k = 0; k < 25; k++) {
                    AU[(size_t)k * N + dup] = AU[(size_t)k * N + id];
                    AT[(size_t)k * N + dup] = AT[(size_t)k * N + id];
                }
            }
        }
    }
#undef U
#undef T
}

/* ------------------------------------------------------------------------ */
/* solver.cpp:943-1039                                                       */
void or_assemble_p9_ref(int n0, int n1, double h1, double h2,
                        const double *xix, const double *xiy,
                        const double *etx, const double *ety, double *A9)
{
    const size_t N = (size_t)n0 * n1;
    memset(A9, 0, 9 * N * sizeof(double));
    for (int j = 1; j < n1 - 1; j++) {
        for (int i = 0; i < n0 - 1; i++) {
            const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1;
            const size_t P = ID(i, j), E = ID(ipp, j), W = ID(inn, j);
            const size_t Nn = ID(i, j + 1), S = ID(i, j - 1);
            const double h11 = 2.0 * h1 * h1, h22 = 2.0 * h2 * h2, h12 = 8.0 * h1 * h2;
            double a, b, c, d;
            /* east */
            a = (xix[P] / h11) * (xix[P] + xix[E]);
            b = (etx[P] / h12) * (xix[Nn] - xix[S]);
            c = (xiy[P] / h11) * (xiy[P] + xiy[E]);
            d = (ety[P] / h12) * (xiy[Nn] - xiy[S]);
            A9[OR_K9(1, 0) * N + P] = a + b + c + d;
            /* west */
            a = (xix[P] / h11) * (xix[P] + xix[W]);
            b = (etx[P] / h12) * (xix[S] - xix[Nn]);
            c = (xiy[P] / h11) * (xiy[P] + xiy[W]);
            d = (ety[P] / h12) * (xiy[S] - xiy[Nn]);
            A9[OR_K9(-1, 0) * N + P] = a + b + c + d;
            /* north */
            a = (xix[P] / h12) * (etx[E] - etx[W]);
            b = (etx[P] / h22) * (etx[P] + etx[Nn]);
            c = (xiy[P] / h12) * (ety[E] - ety[W]);
            d = (ety[P] / h22) * (ety[P] + ety[Nn]);
            A9[OR_K9(0, 1) * N + P] = a + b + c + d;
            /* south */
            a = (xix[P] / h12) * (etx[W] - etx[E]);
            b = (etx[P] / h22) * (etx[P] + etx[S]);
            c = (xiy[P] / h12) * (ety[W] - ety[E]);
            d = (ety[P] / h22) * (ety[P] + ety[S]);
            A9[OR_K9(0, -1) * N + P] = a + b + c + d;
            /* north-east */
            a = (xix[P] / h12) * (etx[P] + etx[E]);
            b = (etx[P] / h12) * (xix[P] + xix[Nn]);
            c = (xiy[P] / h12) * (ety[P] + ety[E]);
            d = (ety[P] / h12) * (xiy[P] + xiy[Nn]);
            A9[OR_K9(1, 1) * N + P] = a + b + c + d;
            /* south-west */
            a = (xix[P] / h12) * (etx[P] + etx[W]);
            b = (etx[P] / h12) * (xix[P] + xix[S]);
            c = (xiy[P] / h12) * (ety[P] + ety[W]);
            d = (ety[P] / h12) * (xiy[P] + xiy[S]);
            A9[OR_K9(-1, -1) * N + P] = a + b + c + d;
            /* north-west */
            a = -(xix[P] / h12) * (etx[P] + etx[W]);
            b = -(etx[P] / h12) * (xix[P] + xix[Nn]);
            c = -(xiy[P] / h12) * (ety[P] + ety[W]);
            d = -(ety[P] / h12) * (xiy[P] + xiy[Nn]);
            A9[OR_K9(-1, 1) * N + P] = a + b + c + d;
            /* south-east */
            a = -(xix[P] / h12) * (etx[P] + etx[E]);
            b = -(etx[P] / h12) * (xix[P] + xix[S]);
            c = -(xiy[P] / h12) * (ety[P] + ety[E]);
            d = -(ety[P] / h12) * (xiy[P] + xiy[S]);
            A9[OR_K9(1, -1) * N + P] = a + b + c + d;
            /* centre */
            {
                const double pxi = 1.0 / (2. * h1 * h1), pet = 1.0 / (2. * h2 * h2);
                const double ap = -xix[P] * (2. * xix[P] + xix[W] + xix[E]);
                const double bp = -etx[P] * (2. * etx[P] + etx[S] + etx[Nn]);
                const double cp = -xiy[P] * (2. * xiy[P] + xiy[W] + xiy[E]);
                const double dp = -ety[P] * (2. * ety[P] + ety[S] + ety[Nn]);
                A9[OR_K9(0, 0) * N + P] = ap * pxi + bp * pet + cp * pxi + dp * pet;
            }
            if (i == 0) {
                const size_t dup = ID(n0 - 1, j);
                for (int k = 0; k < 9; k++) A9[(size_t)k * N + dup] = A9[(size_t)k * N + P];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Hybrid convective derivative along one direction, solver.cpp:1120-1173.
 * f2m..f2p: phi at offsets -2..+2; vel: contravariant velocity; pec: cell
 * Peclet number; hybrid: 1 where the 5-point upwind form is allowed. */
static inline double hyb_deriv(double fm2, double fm1, double f0, double fp1, double fp2,
                               double vel, double pec, double h)
{
    if (pec <= 2 && pec > -2) {
        const double xpp = 8.0 * (fp1 - fm1);
        const double xnn = fp2 - fm2;
        return (1.0 / 12.0) * (xpp - xnn) / h;
    }
    double ak1 = vel * (-fp2 + 8 * fp1 - 8 * fm1 + fm2) / (12.0 * h);
    const double ak2 = fabs(vel) * (fp2 - 4 * fp1 + 6 * f0 - 4 * fm1 + fm2) / (4.0 * h);
    ak1 = ak1 + ak2;
    return ak1 / vel;
}

/* solver.cpp:1062-1236 */
void or_rhs_ref(int n0, int n1, double h1, double h2,
                const double *u, const double *v, const double *T,
                const double *up, const double *vp, const double *p,
                const double *xix, const double *xiy, const double *etx, const double *ety,
                const double *alph, const double *gamma,
                const double *BU, const double *BT,
                double Re, double Pr, double Ri, double dt,
                double *uxi, double *uet,
                double *qu, double *qv, double *qt, double *qup, double *qvp)
{
    for (int j = 0; j < n1; j++)
        for (int i = 0; i < n0; i++) {
            const size_t id = ID(i, j);
            uxi[id] = xix[id] * u[id] + xiy[id] * v[id];
            uet[id] = etx[id] * u[id] + ety[id] * v[id];
        }
    const double *fld[3] = {u, v, T};
    for (int j = 1; j < n1 - 1; j++) {
        for (int i = 0; i < n0 - 1; i++) {
            const cols_t c = pcols(i, n0);
            const size_t P = ID(i, j);
            const int interior = (j >= 2 && j <= n1 - 3);
            double conv[3];
            for (int k = 0; k < 3; k++) {
                const double *f = fld[k];
                double pec1, pec2;
                if (k <= 1) {
                    pec1 = uxi[P] * Re * h1 / alph[P];
                    pec2 = uet[P] * Re * h2 / gamma[P];
                } else {
                    pec1 = uxi[P] * Re * Pr * h1 / alph[P];
                    pec2 = uet[P] * Re * Pr * h2 / gamma[P];
                }
                double du_xi, du_et;
                if (interior) {
                    du_xi = hyb_deriv(f[ID(c.inn2, j)], f[ID(c.inn, j)], f[P], f[ID(c.ipp, j)],
                                      f[ID(c.ipp2, j)], uxi[P], pec1, h1);
                    du_et = hyb_deriv(f[ID(i, j - 2)], f[ID(i, j - 1)], f[P], f[ID(i, j + 1)],
                                      f[ID(i, j + 2)], uet[P], pec2, h2);
                } else {
                    const double xpp = 8.0 * (f[ID(c.ipp, j)] - f[ID(c.inn, j)]);
                    const double xnn = f[ID(c.ipp2, j)] - f[ID(c.inn2, j)];
                    du_xi = (1.0 / 12.0) * (xpp - xnn) / h1;
                    du_et = 0.5 * (f[ID(i, j + 1)] - f[ID(i, j - 1)]) / h2;
                }
                conv[k] = uxi[P] * du_xi + uet[P] * du_et;
            }
            const double dp_dxi = (p[ID(c.ipp, j)] - p[ID(c.inn, j)]) / (2.0 * h1);
            const double dp_de = (p[ID(i, j + 1)] - p[ID(i, j - 1)]) / (2.0 * h2);
            const double dp_dx = xix[P] * dp_dxi + etx[P] * dp_de;
            const double dp_dy = xiy[P] * dp_dxi + ety[P] * dp_de;

            qu[P] = dt * (-conv[0] - dp_dx) + u[P];
            qv[P] = dt * (-conv[1] - dp_dy + Ri * T[P]) + v[P];
            qt[P] = -dt * conv[2] + T[P];
            qup[P] = qu[P] + dt * dp_dx;
            qvp[P] = qv[P] + dt * dp_dy;

            if (j == 1) {
                const size_t s = ID(i, j - 1), se = ID(c.ipp, j - 1), sw = ID(c.inn, j - 1);
                const double bs = BU[0 * n0 + i], bse = BU[1 * n0 + i], bsw = BU[2 * n0 + i];
                const double ts = BT[0 * n0 + i], tse = BT[1 * n0 + i], tsw = BT[2 * n0 + i];
                qu[P] = qu[P] - (bs * u[s] + bse * u[se] + bsw * u[sw]);
                qv[P] = qv[P] - (bs * v[s] + bse * v[se] + bsw * v[sw]);
                qt[P] = qt[P] - (ts * T[s] + tse * T[se] + tsw * T[sw]);
                qup[P] = qup[P] - (bs * up[s] + bse * up[se] + bsw * up[sw]);
                qvp[P] = qvp[P] - (bs * vp[s] + bse * vp[se] + bsw * vp[sw]);
            }
            if (j == n1 - 2) {
                const size_t nn = ID(i, j + 1), ne = ID(c.ipp, j + 1), nw = ID(c.inn, j + 1);
                const double bn = BU[3 * n0 + i], bne = BU[4 * n0 + i], bnw = BU[5 * n0 + i];
                const double tn = BT[3 * n0 + i], tne = BT[4 * n0 + i], tnw = BT[5 * n0 + i];
                qu[P] = qu[P] - (bn * u[nn] + bne * u[ne] + bnw * u[nw]);
                qv[P] = qv[P] - (bn * v[nn] + bne * v[ne] + bnw * v[nw]);
                qt[P] = qt[P] - (tn * T[nn] + tne * T[ne] + tnw * T[nw]);
                qup[P] = qup[P] - (bn * up[nn] + bne * up[ne] + bnw * up[nw]);
                qvp[P] = qvp[P] - (bn * vp[nn] + bne * vp[ne] + bnw * vp[nw]);
            }
            if (i == 0) {
                const size_t dup = ID(n0 - 1, j);
                qu[dup] = qu[P]; qv[dup] = qv[P]; qt[dup] = qt[P];
                qup[dup] = qup[P]; qvp[dup] = qvp[P];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Residual q - (A phi) of one 9- or 25-point row in the reference's term
 * order (solver.cpp:2333-2352); nnw_plane selects the plane used at (-1,+2). */
static inline double row_res25(double qP, const double *A, size_t N, size_t P, const double *phi,
                               int n0, int i, int j, const cols_t *c, int second_order,
                               int nnw_plane)
{
#define A_(di, dj) A[(size_t)OR_K25(di, dj) * N + P]
#define F_(ii, jj) phi[ID(ii, jj)]
    double r = qP - A_(0, 0) * F_(i, j) - A_(1, 0) * F_(c->ipp, j) -
               A_(0, 1) * F_(i, j + 1) - A_(0, -1) * F_(i, j - 1) -
               A_(-1, 0) * F_(c->inn, j) - A_(-1, 1) * F_(c->inn, j + 1) -
               A_(1, 1) * F_(c->ipp, j + 1) - A_(-1, -1) * F_(c->inn, j - 1) -
               A_(1, -1) * F_(c->ipp, j - 1);
    if (!second_order) {
        r = r - A_(2, 0) * F_(c->ipp2, j) -
            A_(-2, 0) * F_(c->inn2, j) - A_(2, 2) * F_(c->ipp2, j + 2) -
            A_(2, 1) * F_(c->ipp2, j + 1) - A_(2, -1) * F_(c->ipp2, j - 1) -
            A_(2, -2) * F_(c->ipp2, j - 2) - A_(1, 2) * F_(c->ipp, j + 2) -
            A_(1, -2) * F_(c->ipp, j - 2) - A[(size_t)nnw_plane * N + P] * F_(c->inn, j + 2) -
            A_(-1, -2) * F_(c->inn, j - 2) - A_(-2, 2) * F_(c->inn2, j + 2) -
            A_(-2, 1) * F_(c->inn2, j + 1) - A_(-2, -1) * F_(c->inn2, j - 1) -
            A_(-2, -2) * F_(c->inn2, j - 2) - A_(0, 2) * F_(i, j + 2) -
            A_(0, -2) * F_(i, j - 2);
    }
    return r;
#undef A_
#undef F_
}

/* solver.cpp:2279-2445 */
int or_gauss_ref(int n0, int n1, const double *A, int nw_for_nnw,
                 double *phi, const double *q, double tol, int maxit, double *last_ratio)
{
    const size_t N = (size_t)n0 * n1;
    const int nnw_plane = nw_for_nnw ? OR_K25(-1, 1) : OR_K25(-1, 2);
    double sumnor = 1.0, sumav = 0.0;
    int iter;
    for (iter = 0; iter < maxit; iter++) {
        double ssum = 0.0;
        /* residual-only sweep, solver.cpp:2304-2363 */
        for (int j = 1; j < n1 - 1; j++) {
            const int so = (j == 1 || j == n1 - 2);
            for (int i = 0; i < n0 - 1; i++) {
                const cols_t c = pcols(i, n0);
                const size_t P = ID(i, j);
                const double res = row_res25(q[P], A, N, P, phi, n0, i, j, &c, so, nnw_plane);
                ssum += fabs(res);
            }
        }
        if (iter == 0) sumnor = (ssum != 0.0) ? ssum : 1.0;
        sumav = ssum / sumnor;
        /* in-place Gauss-Seidel, solver.cpp:2378-2432 */
        for (int j = 1; j < n1 - 1; j++) {
            const int so = (j == 1 || j == n1 - 2);
            for (int i = 0; i < n0 - 1; i++) {
                const cols_t c = pcols(i, n0);
                const size_t P = ID(i, j);
#define A_(di, dj) A[(size_t)OR_K25(di, dj) * N + P]
#define F_(ii, jj) phi[ID(ii, jj)]
                double s = q[P] - A_(1, 0) * F_(c.ipp, j) - A_(0, 1) * F_(i, j + 1) -
                           A_(0, -1) * F_(i, j - 1) - A_(-1, 0) * F_(c.inn, j) -
                           A_(-1, 1) * F_(c.inn, j + 1) - A_(1, 1) * F_(c.ipp, j + 1) -
                           A_(-1, -1) * F_(c.inn, j - 1) - A_(1, -1) * F_(c.ipp, j - 1);
                if (!so) {
                    s = s - A_(2, 0) * F_(c.ipp2, j) - A_(-2, 0) * F_(c.inn2, j) -
                        A_(2, 2) * F_(c.ipp2, j + 2) - A_(2, 1) * F_(c.ipp2, j + 1) -
                        A_(2, -1) * F_(c.ipp2, j - 1) - A_(2, -2) * F_(c.ipp2, j - 2) -
                        A_(1, 2) * F_(c.ipp, j + 2) - A_(1, -2) * F_(c.ipp, j - 2) -
                        A[(size_t)nnw_plane * N + P] * F_(c.inn, j + 2) -
                        A_(-1, -2) * F_(c.inn, j - 2) - A_(-2, 2) * F_(c.inn2, j + 2) -
                        A_(-2, 1) * F_(c.inn2, j + 1) - A_(-2, -1) * F_(c.inn2, j - 1) -
                        A_(-2, -2) * F_(c.inn2, j - 2) - A_(0, 2) * F_(i, j + 2) -
                        A_(0, -2) * F_(i, j - 2);
                }
                phi[P] = s / A_(0, 0);
#undef A_
#undef F_
                if (i == 0) phi[ID(n0 - 1, j)] = phi[P];
            }
        }
        if (sumav < tol) { iter++; break; }
    }
    if (last_ratio) *last_ratio = sumav;
    return iter;
}

/* ------------------------------------------------------------------------ */
/* solver.cpp:1389-1448 */
void or_face_div_ref(int n0, int n1, double h1, double h2,
                     const double *up, const double *vp, const double *p,
                     const double *xix, const double *xiy, const double *etx, const double *ety,
                     double dt, double *q)
{
    for (int j = 1; j < n1 - 1; j++) {
        for (int i = 0; i < n0 - 1; i++) {
            const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1;
            const int jpp = j + 1, jnn = j - 1;
            const size_t P = ID(i, j), E = ID(ipp, j), W = ID(inn, j), Nn = ID(i, jpp), S = ID(i, jnn);
            const double dpdxi_ip = (p[E] - p[P]) / h1;
            const double dpde_ip = (p[ID(ipp, jpp)] + p[Nn] - p[S] - p[ID(ipp, jnn)]) / (4.0 * h2);
            const double dpdxi_in = (p[P] - p[W]) / h1;
            const double dpde_in = (p[Nn] + p[ID(inn, jpp)] - p[S] - p[ID(inn, jnn)]) / (4.0 * h2);
            const double dpdxi_jp = (p[ID(ipp, jpp)] - p[ID(inn, jpp)] + p[E] - p[W]) / (4.0 * h1);
            const double dpde_jp = (p[Nn] - p[P]) / h2;
            const double dpdxi_jn = (p[E] - p[W] + p[ID(ipp, jnn)] - p[ID(inn, jnn)]) / (4.0 * h1);
            const double dpde_jn = (p[P] - p[S]) / h2;

            const double us_ip = 0.5 * (up[P] + up[E]) - 0.5 * dt * ((xix[P] + xix[E]) * dpdxi_ip + (etx[P] + etx[E]) * dpde_ip);
            const double us_in = 0.5 * (up[P] + up[W]) - 0.5 * dt * ((xix[P] + xix[W]) * dpdxi_in + (etx[P] + etx[W]) * dpde_in);
            const double us_jp = 0.5 * (up[P] + up[Nn]) - 0.5 * dt * ((xix[P] + xix[Nn]) * dpdxi_jp + (etx[P] + etx[Nn]) * dpde_jp);
            const double us_jn = 0.5 * (up[P] + up[S]) - 0.5 * dt * ((xix[P] + xix[S]) * dpdxi_jn + (etx[P] + etx[S]) * dpde_jn);
            const double vs_ip = 0.5 * (vp[P] + vp[E]) - 0.5 * dt * ((xiy[P] + xiy[E]) * dpdxi_ip + (ety[P] + ety[E]) * dpde_ip);
            const double vs_in = 0.5 * (vp[P] + vp[W]) - 0.5 * dt * ((xiy[P] + xiy[W]) * dpdxi_in + (ety[P] + ety[W]) * dpde_in);
            const double vs_jp = 0.5 * (vp[P] + vp[Nn]) - 0.5 * dt * ((xiy[P] + xiy[Nn]) * dpdxi_jp + (ety[P] + ety[Nn]) * dpde_jp);
            const double vs_jn = 0.5 * (vp[P] + vp[S]) - 0.5 * dt * ((xiy[P] + xiy[S]) * dpdxi_jn + (ety[P] + ety[S]) * dpde_jn);

            const double dusdxi = (us_ip - us_in) / h1;
            const double dusde = (us_jp - us_jn) / h2;
            const double dvsdxi = (vs_ip - vs_in) / h1;
            const double dvsde = (vs_jp - vs_jn) / h2;
            double qq = (xix[P] * dusdxi) + (etx[P] * dusde) + (xiy[P] * dvsdxi) + (ety[P] * dvsde);
            q[P] = qq / dt;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* solver.cpp:2066-2277 */
int or_sip9p_ref(int n0, int n1, const double *A9, double *phi, const double *q,
                 double tol, int maxit, double alp, double *last_ratio)
{
    const size_t N = (size_t)n0 * n1;
    double *buf = (double *)calloc(10 * N, sizeof(double));
    double *bw = buf, *bs = buf + N, *bn = buf + 2 * N, *be = buf + 3 * N, *bne = buf + 4 * N;
    double *bsw = buf + 5 * N, *bp = buf + 6 * N, *qp = buf + 7 * N, *del = buf + 8 * N, *res = buf + 9 * N;
#define A_(di, dj, id) A9[(size_t)OR_K9(di, dj) * N + (id)]
    /* incomplete factorisation, solver.cpp:2104-2153 */
    for (int j = 1; j < n1 - 1; j++) {
        for (int i = 0; i < n0 - 1; i++) {
            const int inn = (i == 0) ? n0 - 2 : i - 1;
            const size_t P = ID(i, j), Wc = ID(inn, j), SWc = ID(inn, j - 1), Sc = ID(i, j - 1);
            bsw[P] = A_(-1, -1, P);
            bw[P] = (A_(-1, 0, P) + alp * A_(-1, 1, P) - bsw[P] * bn[SWc]) / (1.0 + alp * bn[Wc]);
            bs[P] = (A_(0, -1, P) + alp * A_(1, -1, P) - bsw[P] * be[SWc]) / (1.0 + alp * be[Sc]);
            const double ad = A_(-1, 1, P) + A_(1, -1, P) - bs[P] * be[Sc] - bw[P] * bn[Wc];
            bp[P] = A_(0, 0, P) - alp * ad - bs[P] * bn[Sc] - bw[P] * be[Wc] - bsw[P] * bne[SWc];
            bn[P] = (A_(0, 1, P) + alp * A_(-1, 1, P) - alp * bw[P] * bn[Wc] - bw[P] * bne[Wc]) / bp[P];
            be[P] = (A_(1, 0, P) + alp * A_(1, -1, P) - alp * bs[P] * be[Sc] - bs[P] * bne[Sc]) / bp[P];
            bne[P] = A_(1, 1, P) / bp[P];
            if (i == 0) {
                const size_t dup = ID(n0 - 1, j);
                bsw[dup] = bsw[P]; bn[dup] = bn[P]; bs[dup] = bs[P]; bne[dup] = bne[P];
                be[dup] = be[P]; bw[dup] = bw[P]; bp[dup] = bp[P];
            }
        }
    }
    double sumnor = 1.0, sumav = 0.0;
    int iter;
    for (iter = 0; iter < maxit; iter++) {
        double ssum = 0.0;
        /* residual + forward substitution, solver.cpp:2175-2211 */
        for (int j = 1; j < n1 - 1; j++) {
            for (int i = 0; i < n0 - 1; i++) {
                const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1;
                const size_t P = ID(i, j);
                const double r = q[P] - A_(0, 0, P) * phi[P] - A_(1, 0, P) * phi[ID(ipp, j)] -
                                 A_(0, 1, P) * phi[ID(i, j + 1)] - A_(0, -1, P) * phi[ID(i, j - 1)] -
                                 A_(-1, 0, P) * phi[ID(inn, j)] - A_(-1, 1, P) * phi[ID(inn, j + 1)] -
                                 A_(1, 1, P) * phi[ID(ipp, j + 1)] - A_(-1, -1, P) * phi[ID(inn, j - 1)] -
                                 A_(1, -1, P) * phi[ID(ipp, j - 1)];
                res[P] = r;
                ssum += fabs(r);
                qp[P] = (r - bs[P] * qp[ID(i, j - 1)] - bw[P] * qp[ID(inn, j)] -
                         bsw[P] * qp[ID(inn, j - 1)]) / bp[P];
                if (i == 0) {
                    res[ID(n0 - 1, j)] = r;
                    qp[ID(n0 - 1, j)] = qp[P];
                }
            }
        }
        if (iter == 0) sumnor = (ssum != 0.0) ? ssum : 1.0;
        sumav = ssum / sumnor;
        /* backward substitution, solver.cpp:2227-2253 */
        for (int j = n1 - 2; j >= 1; j--) {
            for (int i = n0 - 2; i >= 0; i--) {
                const int ipp = i + 1;
                const size_t P = ID(i, j);
                del[P] = qp[P] - bn[P] * del[ID(i, j + 1)] - be[P] * del[ID(ipp, j)] -
                         bne[P] * del[ID(ipp, j + 1)];
                phi[P] = phi[P] + del[P];
                if (i == 0) phi[ID(n0 - 1, j)] = phi[P];
            }
        }
        if (sumav < tol) { iter++; break; }
    }
#undef A_
    free(buf);
    if (last_ratio) *last_ratio = sumav;
    return iter;
}

/* ------------------------------------------------------------------------ */
/* solver.cpp:1351-1381: predictor velocities on the outer boundary j = n1-1 --
 * the current velocity where the free stream enters (vnn >= 0), else a
 * quadratic extrapolation from the three rows inside. */
void or_outflow_up_ref(int n0, int n1, double *up, double *vp, const double *u, const double *v,
                       const double *xnox, const double *xnoy, double uinf, double vinf)
{
    const int j = n1 - 1;
    for (int i = 0; i < n0 - 1; i++) {
        const double vnn = uinf * xnox[i] + vinf * xnoy[i];
        const size_t P = ID(i, j);
        if (vnn >= 0) {
            up[P] = u[P];
            vp[P] = v[P];
        } else {
            up[P] = (5.0 * up[ID(i, j - 1)] - 4.0 * up[ID(i, j - 2)] + up[ID(i, j - 3)]) / 2.0;
            vp[P] = (5.0 * vp[ID(i, j - 1)] - 4.0 * vp[ID(i, j - 2)] + vp[ID(i, j - 3)]) / 2.0;
        }
        if (i == 0) {
            up[ID(n0 - 1, j)] = up[P];
            vp[ID(n0 - 1, j)] = vp[P];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* solver.cpp:1474-1497 */
void or_pcor_bc_ref(int n0, int n1, double *pcor, const double *xnox, const double *xnoy,
                    double uinf, double vinf)
{
    for (int i = 0; i < n0 - 1; i++) {
        pcor[ID(i, 0)] = pcor[ID(i, 1)];
        if (i == 0) pcor[ID(n0 - 1, 0)] = pcor[ID(i, 0)];
    }
    const int j = n1 - 1;
    for (int i = 0; i < n0 - 1; i++) {
        const double vnn = uinf * xnox[i] + vinf * xnoy[i];
        pcor[ID(i, j)] = 0;
        if (vnn >= 0) pcor[ID(i, j)] = pcor[ID(i, j - 1)];
        if (i == 0) pcor[ID(n0 - 1, j)] = pcor[ID(i, j)];
    }
}

/* solver.cpp:1504-1541 */
void or_correct_ref(int n0, int n1, double h1, double h2,
                    const double *us, const double *vs, const double *pcor,
                    const double *xix, const double *xiy, const double *etx, const double *ety,
                    double dt, double *u, double *v, double *p)
{
    for (int i = 0; i < n0 - 1; i++) {
        for (int j = 1; j < n1 - 1; j++) {
            const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1;
            const size_t P = ID(i, j);
            const double dpx = 0.5 * (pcor[ID(ipp, j)] - pcor[ID(inn, j)]) / h1;
            const double dpe = 0.5 * (pcor[ID(i, j + 1)] - pcor[ID(i, j - 1)]) / h2;
            u[P] = us[P] - dt * (xix[P] * dpx + etx[P] * dpe);
            v[P] = vs[P] - dt * (xiy[P] * dpx + ety[P] * dpe);
            if (i == 0) {
                u[ID(n0 - 1, j)] = u[P];
                v[ID(n0 - 1, j)] = v[P];
            }
        }
    }
    for (int i = 0; i < n0 - 1; i++) {
        for (int j = 1; j < n1 - 1; j++) {
            const size_t P = ID(i, j);
            p[P] = p[P] + pcor[P];
            if (i == 0) p[ID(n0 - 1, j)] = p[P];
        }
    }
}

/* solver.cpp:1543-1655: boundary velocities after the correction.  Radial and
 * tangential velocity on row n1-2, the circulation along that row, the far-field
 * prediction vr = vr0 cr^2 + vr_inf (1 - cr^2), vth = vth0 cr^k + vth_inf (1 - cr^k)
 * with k = 1 when |circulation| > 1e-2 else 2 (cr = radius ratio of rows n1-2 and
 * n1-1); the wall row j = 0 gets the oscillating-cylinder velocity
 * (-a y, a x) with a = speed_amp cos_ft (cos_ft = cos(2 pi F t), evaluated by the
 * caller), copied into up/vp; the outer row gets the free stream (T = 0) where it
 * enters and the predicted velocity with convected T (Told, uet: the values at the
 * start of the step) where it leaves.  Returns the circulation. */
double or_farfield_ref(int n0, int n1, double h2, const double *x, const double *y,
                       double *u, double *v, double *T, double *up, double *vp,
                       const double *Told, const double *uet, const double *etx, const double *ety,
                       const double *ajac, const double *xnox, const double *xnoy,
                       double uinf, double vinf, double dt, double speed_amp, double cos_ft)
{
    double *vr0 = (double *)calloc((size_t)n0 * 4, sizeof(double));
    double *vt0 = vr0 + n0, *vr1 = vr0 + 2 * n0, *vt1 = vr0 + 3 * n0;
    int j = n1 - 2;
    for (int i = 0; i < n0 - 1; i++) {
        const size_t P = ID(i, j);
        const double r = sqrt(x[P] * x[P] + y[P] * y[P]);
        const double c = x[P] / r, s = y[P] / r;
        vr0[i] = u[P] * c + v[P] * s;
        vt0[i] = -u[P] * s + v[P] * c;
    }
    vr0[n0 - 1] = vr0[0];
    vt0[n0 - 1] = vt0[0];
    double circ = 0.0;
    for (int i = 0; i < n0 - 1; i++) {
        const double de = 1.0 / (n0 - 2);
        const size_t A = ID(i, j), B = ID(i + 1, j);
        const double f1 = (u[A] * ety[A] - v[A] * etx[A]) * fabs(ajac[A]);
        const double f2 = (u[B] * ety[B] - v[B] * etx[B]) * fabs(ajac[B]);
        circ = circ + de * 0.5 * (f1 + f2);
    }
    j = n1 - 1;
    const int kk = fabs(circ) > 1e-2 ? 1 : 2;
    for (int i = 0; i < n0 - 1; i++) {
        const size_t P = ID(i, j), Q = ID(i, j - 1);
        const double r = sqrt(x[P] * x[P] + y[P] * y[P]);
        const double cr = sqrt(x[Q] * x[Q] + y[Q] * y[Q]) / r;
        const double c = x[P] / r, s = y[P] / r;
        const double vrinf = uinf * c + vinf * s, vtinf = -uinf * s + vinf * c;
        vr1[i] = vr0[i] * pow(cr, 2.0) + vrinf * (1 - pow(cr, 2.0));
        vt1[i] = vt0[i] * pow(cr, (double)kk) + vtinf * (1 - pow(cr, (double)kk));
    }
    vr1[n0 - 1] = vr1[0];
    vt1[n0 - 1] = vt1[0];
    for (int i = 0; i < n0; i++) {
        const size_t P = ID(i, 0);
        u[P] = -speed_amp * cos_ft * y[P];
        up[P] = u[P];
    }
    for (int i = 0; i < n0; i++) {
        const size_t P = ID(i, 0);
        v[P] = speed_amp * cos_ft * x[P];
        vp[P] = v[P];
    }
    for (int i = 0; i < n0 - 1; i++) {
        const size_t P = ID(i, j);
        const double vnn = uinf * xnox[i] + vinf * xnoy[i];
        if (vnn >= 0) {
            u[P] = uinf;
            v[P] = vinf;
            T[P] = 0.0;
        } else {
            const double r = sqrt(x[P] * x[P] + y[P] * y[P]);
            const double c = x[P] / r, s = y[P] / r;
            u[P] = c * vr1[i] - s * vt1[i];
            v[P] = s * vr1[i] + c * vt1[i];
            T[P] = Told[P] - (uet[P] * dt / h2) * (Told[P] - Told[ID(i, j - 1)]);
        }
        if (i == 0) {
            u[ID(n0 - 1, j)] = u[ID(0, j)];
            v[ID(n0 - 1, j)] = v[ID(0, j)];
            T[ID(n0 - 1, j)] = T[ID(0, j)];
        }
    }
    free(vr0);
    return circ;
}

/* solver.cpp:1667-1728 */
void or_wall_p_ref(int n0, int n1, double h1, double h2,
                   const double *u, const double *v, const double *T, double *p,
                   const double *xix, const double *xiy, const double *etx, const double *ety,
                   const double *alph, const double *beta, const double *gamma,
                   const double *q1, const double *ajac,
                   const double *alc_u, const double *alc_v, double Re, double Ri)
{
    (void)n1;
    const double *fld[2] = {u, v};
    const int j = 0;
    for (int i = 0; i < n0 - 1; i++) {
        const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1;
        const size_t P = ID(i, j);
        const double uxi = xix[P] * u[P] + xiy[P] * v[P];
        const double uet = etx[P] * u[P] + ety[P] * v[P];
        double dp_dx = 0.0, dp_dy = 0.0;
        for (int k = 0; k < 2; k++) {
            const double *f = fld[k];
            const double aa = alph[P] * (f[ID(ipp, j)] + f[ID(inn, j)] - 2 * f[P]) / (h1 * h1);
            const double gg = gamma[P] * (f[ID(i, j + 2)] + f[P] - 2 * f[ID(i, j + 1)]) / (h2 * h2);
            const double bb = beta[P] * (f[ID(ipp, j + 1)] + f[ID(inn, j)] - f[ID(inn, j + 1)] - f[ID(ipp, j)]) /
                              (2 * h1 * h2);
            const double qqq = q1[P] * (-3 * f[P] + 4 * f[ID(i, j + 1)] - f[ID(i, j + 2)]) / (2 * h2);
            const double d2u = aa + gg - 2 * bb + qqq;
            double conv = uxi * 0.5 * (f[ID(ipp, j)] - f[ID(inn, j)]) / h1;
            conv = conv + uet * (f[ID(i, j + 1)] - f[P]) / h2;
            if (k == 0) dp_dx = 1.0 * d2u / Re - conv - alc_u[i];
            else dp_dy = 1.0 * d2u / Re - conv - alc_v[i] + Ri * T[P];
        }
        p[P] = p[ID(i, j + 1)] - (dp_dx * (-xiy[P] * ajac[P]) + dp_dy * (xix[P] * ajac[P])) * h2;
        if (i == 0) p[ID(n0 - 1, j)] = p[P];
    }
}

/* solver.cpp:1731-1782: pressure on the outer row j = n1-1 -- the normal
 * momentum equation with one-sided eta differences where the free stream enters
 * (local acceleration (u - uold)/dt), Gresho's condition
 * p = (3 uet_j - 4 uet_{j-1} + uet_{j-2}) / (2 Re h2) where it leaves; uxi and uet
 * are those of the current (boundary-updated) velocity. */
void or_exit_p_ref(int n0, int n1, double h1, double h2,
                   const double *u, const double *v, const double *T, const double *uold,
                   const double *vold, double *p,
                   const double *xix, const double *xiy, const double *etx, const double *ety,
                   const double *alph, const double *beta, const double *gamma,
                   const double *q1, const double *ajac, const double *xnox, const double *xnoy,
                   double uinf, double vinf, double dt, double Re, double Ri)
{
    const double *fld[2] = {u, v}, *old[2] = {uold, vold};
    const int j = n1 - 1;
    for (int i = 0; i < n0 - 1; i++) {
        const size_t P = ID(i, j);
        const double vnn = uinf * xnox[i] + vinf * xnoy[i];
        if (vnn >= 0) {
            const int ipp = i + 1, inn = (i == 0) ? n0 - 2 : i - 1;
            const double uxi = xix[P] * u[P] + xiy[P] * v[P];
            const double uet = etx[P] * u[P] + ety[P] * v[P];
            double dp_dx = 0.0, dp_dy = 0.0;
            for (int k = 0; k < 2; k++) {
                const double *f = fld[k];
                const double aa = alph[P] * (f[ID(ipp, j)] + f[ID(inn, j)] - 2 * f[P]) / (h1 * h1);
                const double gg = gamma[P] * (f[P] + f[ID(i, j - 2)] - 2 * f[ID(i, j - 1)]) / (h2 * h2);
                const double bb = beta[P] * (f[ID(ipp, j)] + f[ID(inn, j - 1)] - f[ID(ipp, j - 1)] - f[ID(inn, j)]) /
                                  (2 * h1 * h2);
                const double qqq = q1[P] * (3 * f[P] - 4 * f[ID(i, j - 1)] + f[ID(i, j - 2)]) / (2 * h2);
                const double d2u = aa + gg - 2 * bb + qqq;
                double conv = uxi * 0.5 * (f[ID(ipp, j)] - f[ID(inn, j)]) / h1;
                conv = conv + uet * (3.0 * f[P] - 4 * f[ID(i, j - 1)] + f[ID(i, j - 2)]) / (2 * h2);
                const double alc = (f[P] - old[k][P]) / dt;
                if (k == 0) dp_dx = 1.0 * d2u / Re - conv - alc;
                else dp_dy = 1.0 * d2u / Re - conv - alc + Ri * T[P];
            }
            p[P] = p[ID(i, j - 1)] + (dp_dx * (-xiy[P] * ajac[P]) + dp_dy * (xix[P] * ajac[P])) * h2;
        } else {
            const size_t Q = ID(i, j - 1), R = ID(i, j - 2);
            const double e0 = etx[P] * u[P] + ety[P] * v[P];
            const double e1 = etx[Q] * u[Q] + ety[Q] * v[Q];
            const double e2 = etx[R] * u[R] + ety[R] * v[R];
            p[P] = 0.5 * (1.0 / Re) * ((3 * e0 - 4 * e1 + e2) / h2);
        }
        if (i == 0) p[ID(n0 - 1, j)] = p[P];
    }
}

/* solver.cpp:1784-1800: stream function, the trapezoid integral of the
 * J-weighted contravariant flux along each eta line from si = 0 on the wall. */
void or_si_ref(int n0, int n1, double h2, const double *u, const double *v, const double *xix,
               const double *xiy, const double *ajac, double *si)
{
    for (int i = 0; i < n0; i++) si[ID(i, 0)] = 0;
    for (int i = 0; i < n0; i++) {
        for (int j = 1; j < n1; j++) {
            const size_t P = ID(i, j), Q = ID(i, j - 1);
            const double ca = (xix[P] * u[P] * fabs(ajac[P]) + xix[Q] * u[Q] * fabs(ajac[Q]));
            const double cb = (xiy[P] * v[P] * fabs(ajac[P]) + xiy[Q] * v[Q] * fabs(ajac[Q]));
            si[P] = si[Q] + (ca + cb) * 0.5 * h2;
        }
    }
}

/* solver.cpp:1802-1887 (vorticity part): dv/dx - du/dy by the chain rule, central
 * differences inside, one-sided eta differences on the wall and outer rows. */
void or_vort_ref(int n0, int n1, double h1, double h2, const double *u, const double *v,
                 const double *xix, const double *xiy, const double *etx, const double *ety, double *vort)
{
    for (int i = 0; i < n0 - 1; i++) {
        for (int j = 1; j < n1 - 1; j++) {
            const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1;
            const size_t P = ID(i, j);
            const double dv_dxi = 0.5 / h1 * (v[ID(ipp, j)] - v[ID(inn, j)]);
            const double dv_det = 0.5 / h2 * (v[ID(i, j + 1)] - v[ID(i, j - 1)]);
            const double dv_dx = xix[P] * dv_dxi + etx[P] * dv_det;
            const double du_dxi = 0.5 / h1 * (u[ID(ipp, j)] - u[ID(inn, j)]);
            const double du_det = 0.5 / h2 * (u[ID(i, j + 1)] - u[ID(i, j - 1)]);
            const double du_dy = xiy[P] * du_dxi + ety[P] * du_det;
            vort[P] = dv_dx - du_dy;
            if (i == 0) vort[ID(n0 - 1, j)] = vort[P];
        }
    }
    for (int j = 0; j < n1; j += n1 - 1) {
        for (int i = 0; i < n0 - 1; i++) {
            const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1;
            const size_t P = ID(i, j);
            const double dv_dxi = 0.5 / h1 * (v[ID(ipp, j)] - v[ID(inn, j)]);
            const double dv_det = (j == 0) ? 1.0 / h2 * (v[ID(i, j + 1)] - v[P]) : 1.0 / h2 * (v[P] - v[ID(i, j - 1)]);
            const double dv_dx = xix[P] * dv_dxi + etx[P] * dv_det;
            const double du_dxi = 0.5 / h1 * (u[ID(ipp, j)] - u[ID(inn, j)]);
            const double du_det = (j == 0) ? 1.0 / h2 * (u[ID(i, j + 1)] - u[P]) : 1.0 / h2 * (u[P] - u[ID(i, j - 1)]);
            const double du_dy = xiy[P] * du_dxi + ety[P] * du_det;
            vort[P] = dv_dx - du_dy;
            if (i == 0) vort[ID(n0 - 1, j)] = vort[P];
        }
    }
}

/* solver.cpp:1891-1982: force and moment coefficients and the mean wall Nusselt
 * number from the wall row's pressure, vorticity and temperature gradient.
 * out: cl, cd, cm, CL_pr, CD_pr, CL_vor, CD_vor, Nu (alpha in degrees). */
void or_forces_ref(int n0, int n1, double h1, double h2, const double *x, const double *y,
                   const double *p, const double *vort, const double *T, const double *ajac,
                   const double *etx, const double *ety, double Re, double alpha, double ar, double *out)
{
    (void)n1;
    const double Pi = acos(-1.0);
    const int j = 0;
    double pr_x = 0.0, pr_y = 0.0, vor_x = 0.0, vor_y = 0.0;
    for (int i = 0; i < n0 - 1; i++) {
        const size_t A = ID(i, j), B = ID(i + 1, j);
        const double PJ1 = p[A] * ajac[A], pj2 = p[B] * ajac[B];
        const double VJ1 = vort[A] * ajac[A], VJ2 = vort[B] * ajac[B];
        pr_x = pr_x + 0.5 * h1 * (PJ1 * etx[A] + pj2 * etx[B]);
        pr_y = pr_y + 0.5 * h1 * (PJ1 * ety[A] + pj2 * ety[B]);
        vor_x = vor_x + 0.5 * h1 * (VJ1 * ety[A] + VJ2 * ety[B]);
        vor_y = vor_y + 0.5 * h1 * (VJ1 * etx[A] + VJ2 * etx[B]);
    }
    const double sa = sin(alpha * Pi / 180.0), ca = cos(alpha * Pi / 180.0);
    const double cx = 2 * pr_x + (2.0 / Re) * vor_x;
    const double cy = 2 * pr_y - (2.0 / Re) * vor_y;
    out[3] = 2 * pr_y * sa - 2 * pr_x * ca;
    out[4] = 2 * pr_y * ca + 2 * pr_x * sa;
    out[5] = -(2.0 / Re) * vor_y * sa - (2.0 / Re) * vor_x * ca;
    out[6] = -(2.0 / Re) * vor_y * ca + (2.0 / Re) * vor_x * sa;
    out[0] = cy * sa - cx * ca;
    out[1] = cy * ca + cx * sa;
    double press_i = 0.0, vor_i = 0.0, temp_i = 0.0;
    for (int i = 0; i < n0 - 1; i++) {
        const int ip = i + 1;
        const size_t A = ID(i, j), B = ID(ip, j);
        const double PJ1 = p[A] * ajac[A], pj2 = p[B] * ajac[B];
        const double VJ1 = vort[A] * ajac[A], VJ2 = vort[B] * ajac[B];
        const double TJ1 = ajac[A] * (etx[A] * etx[A] + ety[A] * ety[A]);
        const double TJ2 = ajac[B] * (etx[B] * etx[B] + ety[B] * ety[B]);
        const double fp1 = PJ1 * (x[A] * ety[A] - y[A] * etx[A]);
        const double fp2 = pj2 * (x[B] * ety[B] - y[B] * etx[B]);
        const double fv1 = VJ1 * (x[A] * etx[A] + y[A] * ety[A]);
        const double fv2 = VJ2 * (x[B] * etx[B] + y[B] * ety[B]);
        const double fh1 = TJ1 * (4 * T[ID(i, j + 1)] - 3 * T[A] - T[ID(i, j + 2)]) / (2 * h2);
        const double fh2 = TJ2 * (4 * T[ID(ip, j + 1)] - 3 * T[B] - T[ID(ip, j + 2)]) / (2 * h2);
        press_i = press_i + 0.5 * h1 * (fp1 + fp2);
        vor_i = vor_i + 0.5 * h1 * (fv1 + fv2);
        temp_i = temp_i + 0.5 * (fh1 + fh2) * h1;
    }
    out[2] = 2 * press_i - (2.0 / Re) * vor_i;
    out[7] = (2 * temp_i) / (Pi * (3 * (1 + (1.0 / ar)) - sqrt((3 + (1.0 / ar)) * ((3.0 / ar) + 1))));
}

/* solver.cpp:1806-1852 */
double or_dil_ref(int n0, int n1, double h1, double h2,
                  const double *u, const double *v,
                  const double *xix, const double *xiy, const double *etx, const double *ety,
                  double *dil)
{
    double dmax = 0.0;
    for (int i = 0; i < n0 - 1; i++) {
        for (int j = 1; j < n1 - 1; j++) {
            const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1;
            const size_t P = ID(i, j);
            const double d = xix[P] * (u[ID(ipp, j)] - u[ID(inn, j)]) / (2 * h1) +
                             etx[P] * (u[ID(i, j + 1)] - u[ID(i, j - 1)]) / (2 * h2) +
                             ety[P] * (v[ID(i, j + 1)] - v[ID(i, j - 1)]) / (2 * h2) +
                             xiy[P] * (v[ID(ipp, j)] - v[ID(inn, j)]) / (2 * h1);
            dil[P] = d;
            if (i == 0) dil[ID(n0 - 1, j)] = d;
            if (d > dmax) dmax = d;
        }
    }
    return dmax;
}

/* solver.cpp:1951-1982, 2031-2038 */
double or_nusselt_ref(int n0, int n1, double h1, double h2, const double *T,
                      const double *ajac, const double *etx, const double *ety,
                      double ar, double *dthdn)
{
    (void)n1;
    const double Pi = acos(-1.0);
    double temp_i = 0.0;
    const int j = 0;
    for (int i = 0; i < n0 - 1; i++) {
        const int ip = i + 1;
        const size_t A = ID(i, j), B = ID(ip, j);
        const double tj1 = ajac[A] * (etx[A] * etx[A] + ety[A] * ety[A]);
        const double tj2 = ajac[B] * (etx[B] * etx[B] + ety[B] * ety[B]);
        const double fh1 = tj1 * (4 * T[ID(i, j + 1)] - 3 * T[A] - T[ID(i, j + 2)]) / (2 * h2);
        const double fh2 = tj2 * (4 * T[ID(ip, j + 1)] - 3 * T[B] - T[ID(ip, j + 2)]) / (2 * h2);
        temp_i = temp_i + 0.5 * (fh1 + fh2) * h1;
    }
    if (dthdn) {
        for (int i = 0; i < n0; i++) {
            double d = -(4 * T[ID(i, 1)] - 3 * T[ID(i, 0)] - T[ID(i, 2)]) / (2 * h2);
            dthdn[i] = d * sqrt(etx[ID(i, 0)] * etx[ID(i, 0)] + ety[ID(i, 0)] * ety[ID(i, 0)]);
        }
    }
    return (2 * temp_i) / (Pi * (3 * (1 + (1.0 / ar)) - sqrt((3 + (1.0 / ar)) * ((3.0 / ar) + 1))));
}
