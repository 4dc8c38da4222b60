// Device-side stencil formulas of the cavity SIMPLE path (gfx950).
//
// Every formula here is the one the CPU oracle uses, node for node, with one
// change of form: a division by a grid or physics constant (h1, h2, Re, Re Pr,
// their products) is a multiplication by its reciprocal, computed once on the
// host (Recip, and dt/nd per slot).  The results differ from the oracle's in the
// last bits only; a division by a per-node value stays a division.
//  * conv_deriv   -- hybrid 4th-order central / Kawamura-Kuwahara upwind switch
//                    on the cell Peclet number (solver.cpp:1120-1173), with the
//                    cavity mode's upwinding next to walls and continuous blend;
//  * aref_offdiag -- A_ref phi - phi of the implicit diffusion operator
//                    I - dt/Re L (coefficients of solver.cpp:764-847);
//  * alo_coefs    -- low-order predictor operator (the solver.cpp:764-786
//                    second-order row + first-order upwind convection);
//  * p9_coefs     -- the J-weighted flux form of the pressure-correction
//                    operator built from the face velocity of solver.cpp:1402-1436.
#pragma once

#include "cfd_common.h"

namespace cfd {

// Node (i, j) of a level with n0 nodes per row, as an offset from the slot's first
// node: 32-bit, as no level holds more than MAX_LEVEL_NODES nodes (checked where
// cases and packs are created), so the address is one 32-bit multiply-add and a
// widening add, not a 64-bit multiply.
constexpr long long MAX_LEVEL_NODES = 1LL << 28;
__device__ __forceinline__ unsigned idx(int i, int j, int n0) { return (unsigned)j * (unsigned)n0 + (unsigned)i; }

struct KGrid {
    int n0, n1;
    double h1, h2;
};

// reciprocals of the grid constants (host-computed)
template <typename R>
struct RecipT {
    R ih1, ih2;       // 1/h1, 1/h2
    R ih1sq, ih2sq;   // 1/h1^2, 1/h2^2
    R i2h1, i2h2;     // 1/(2 h1), 1/(2 h2)
    R ih1h2, i2h1h2;  // 1/(h1 h2), 1/(2 h1 h2)
};
template <typename R, typename S>
__host__ __device__ inline RecipT<R> recip_cast(const RecipT<S> &a) {
    return RecipT<R>{(R)a.ih1, (R)a.ih2, (R)a.ih1sq, (R)a.ih2sq, (R)a.i2h1, (R)a.i2h2, (R)a.ih1h2, (R)a.i2h1h2};
}
inline RecipT<double> make_recip(double h1, double h2) {
    return RecipT<double>{1.0 / h1, 1.0 / h2, 1.0 / (h1 * h1), 1.0 / (h2 * h2), 1.0 / (2.0 * h1), 1.0 / (2.0 * h2),
                          1.0 / (h1 * h2), 1.0 / (2.0 * h1 * h2)};
}

// conv_deriv with the reciprocal ih = 1/h; the upwind branch divides by vel
// only symbolically: (vel D4 + |vel| D4') / vel = D4 + sign(vel) D4'.
// mode 3 (cavity rows next to a wall): second-order central, first-order upwind
// when |Pe| > 2.  blend > 0 replaces the switch at |Pe| = 2 by a linear blend from
// the central to the upwind form over 2 < |Pe| < 2 + blend (cavity mode).
__device__ __forceinline__ double conv_deriv_r(int mode, double fm2, double fm1, double f0, double fp1,
                                               double fp2, double vel, double pec, double ih, double blend) {
    const bool central = pec <= 2 && pec > -2;
    const double w = blend > 0.0 ? fmin(1.0, (fabs(pec) - 2.0) / blend) : 1.0;
    const double sg = vel > 0.0 ? 1.0 : -1.0;
    if (mode == 3) {
        const double c2 = 0.5 * (fp1 - fm1) * ih;
        if (central) return c2;
        if (blend > 0.0) return c2 + w * sg * (2.0 * f0 - fm1 - fp1) * (0.5 * ih);
        return vel > 0.0 ? (f0 - fm1) * ih : (fp1 - f0) * ih;
    }
    if (mode == 0) return 0.5 * (fp1 - fm1) * ih;
    const double c4 = (1.0 / 12.0) * (8.0 * (fp1 - fm1) - (fp2 - fm2)) * ih;
    if (mode == 1 || central) return c4;
    const double d4p = (fp2 - 4 * fp1 + 6 * f0 - 4 * fm1 + fm2) * (0.25 * ih);
    if (blend > 0.0) return c4 + w * sg * d4p;
    const double d4 = (-fp2 + 8 * fp1 - 8 * fm1 + fm2) * ((1.0 / 12.0) * ih);
    return vel > 0.0 ? d4 + d4p : d4 - d4p;
}

// (A_ref - I) phi at an interior node with dn = dt / nd (nd = Re or Re Pr); F(di, dj)
// reads phi at the neighbour (i + di, j + dj)
template <typename FA>
__device__ __forceinline__ double aref_offdiag_a(FA F, double al, double be, double ga, double pp, double qq,
                                                 const RecipT<double> &r, double dn, bool wide) {
    const double A = al * r.ih1sq, G = ga * r.ih2sq;
    if (!wide) {
        const double Pp = pp * r.i2h1, Qq = qq * r.i2h2;
        const double ce = -dn * (A + Pp), cw = -dn * (A - Pp);
        const double cn = -dn * (G + Qq), cs = -dn * (G - Qq);
        const double cx = dn * be * r.i2h1h2;
        const double cp = 2.0 * dn * (A + G);
        return cp * F(0, 0) + ce * F(1, 0) + cw * F(-1, 0) + cn * F(0, 1) + cs * F(0, -1) +
               cx * (F(1, 1) + F(-1, -1) - F(-1, 1) - F(1, -1));
    }
    const double Pi = pp * r.ih1, Qi = qq * r.ih2, B = be * r.ih1h2;
    const double ce = -dn * ((4.0 / 3.0) * A + (2.0 / 3.0) * Pi);
    const double cw = -dn * ((4.0 / 3.0) * A - (2.0 / 3.0) * Pi);
    const double cn = -dn * ((4.0 / 3.0) * G + (2.0 / 3.0) * Qi);
    const double cs = -dn * ((4.0 / 3.0) * G - (2.0 / 3.0) * Qi);
    const double cne = dn * (8.0 / 9.0) * B;
    const double cnw = -cne;
    const double cnn = dn * ((1.0 / 12.0) * G + (1.0 / 12.0) * Qi);
    const double css = dn * ((1.0 / 12.0) * G - (1.0 / 12.0) * Qi);
    const double cee = dn * ((1.0 / 12.0) * A + (1.0 / 12.0) * Pi);
    const double cww = dn * ((1.0 / 12.0) * A - (1.0 / 12.0) * Pi);
    const double c22 = dn * (1.0 / 72.0) * B;
    const double c2m = -c22;
    const double c12 = -dn * (1.0 / 9.0) * B;
    const double c1m = -c12;
    const double cp = 2.5 * dn * (A + G);
    return cp * F(0, 0) + ce * F(1, 0) + cw * F(-1, 0) + cn * F(0, 1) + cs * F(0, -1) +
           cne * (F(1, 1) + F(-1, -1)) + cnw * (F(-1, 1) + F(1, -1)) +
           cnn * F(0, 2) + css * F(0, -2) + cee * F(2, 0) + cww * F(-2, 0) +
           c22 * (F(2, 2) + F(-2, -2)) + c2m * (F(-2, 2) + F(2, -2)) +
           c12 * (F(1, 2) + F(-1, -2) + F(2, 1) + F(-2, -1)) +
           c1m * (F(-1, 2) + F(1, -2) + F(-2, 1) + F(2, -1));
}

__device__ __forceinline__ double aref_offdiag_r(const double *__restrict__ f, int n0, int i, int j, double al,
                                                 double be, double ga, double pp, double qq,
                                                 const RecipT<double> &r, double dn, bool wide) {
    return aref_offdiag_a([&](int di, int dj) { return f[idx(i + di, j + dj, n0)]; }, al, be, ga, pp, qq, r, dn,
                          wide);
}

// low-order predictor row with reciprocals: dn = dt / nd
template <typename R>
__device__ __forceinline__ void alo_coefs_r(R al, R be, R ga, R pp, R qq, R U, R V, const RecipT<R> &r, R dt, R dn,
                                            R c[9]) {
    const R zero = R(0);
    const R A = al * r.ih1sq, G = ga * r.ih2sq, Pp = pp * r.i2h1, Qq = qq * r.i2h2;
    c[5] = -dn * (A + Pp) + dt * fmin(U, zero) * r.ih1;
    c[3] = -dn * (A - Pp) - dt * fmax(U, zero) * r.ih1;
    c[7] = -dn * (G + Qq) + dt * fmin(V, zero) * r.ih2;
    c[1] = -dn * (G - Qq) - dt * fmax(V, zero) * r.ih2;
    const R cx = dn * be * r.i2h1h2;
    c[8] = cx;
    c[0] = cx;
    c[6] = -cx;
    c[2] = -cx;
    c[4] = R(1) + R(2) * dn * (A + G) + dt * (fabs(U) * r.ih1 + fabs(V) * r.ih2);
}


// mode 2: hybrid, 1: 4th-order central, 0: 2nd-order central (solver.cpp:1124-1172)
__device__ __forceinline__ double conv_deriv(int mode, double fm2, double fm1, double f0, double fp1,
                                             double fp2, double vel, double pec, double h) {
    if (mode == 0) return 0.5 * (fp1 - fm1) / h;
    if (mode == 1 || (pec <= 2 && pec > -2)) {
        const double xpp = 8.0 * (fp1 - fm1);
        const double xnn = fp2 - fm2;
        return (1.0 / 12.0) * (xpp - xnn) / h;
    }
    double ak1 = vel * (-fp2 + 8 * fp1 - 8 * fm1 + fm2) / (12.0 * h);
    const double ak2 = fabs(vel) * (fp2 - 4 * fp1 + 6 * f0 - 4 * fm1 + fm2) / (4.0 * h);
    ak1 = ak1 + ak2;
    return ak1 / vel;
}

// (A_ref - I) phi at interior node (i, j); nd = Re (u, v) or Re*Pr (T)
__device__ __forceinline__ double aref_offdiag(const double *__restrict__ f, int n0, int i, int j,
                                               double al, double be, double ga, double pp, double qq,
                                               double h1, double h2, double dt, double nd, bool wide) {
#define F(di, dj) f[idx(i + (di), j + (dj), n0)]
    if (!wide) {
        const double ce = -dt * (al / (h1 * h1) + pp / (2.0 * h1)) / nd;
        const double cw = -dt * (al / (h1 * h1) - pp / (2.0 * h1)) / nd;
        const double cn = -dt * (ga / (h2 * h2) + qq / (2.0 * h2)) / nd;
        const double cs = -dt * (ga / (h2 * h2) - qq / (2.0 * h2)) / nd;
        const double cx = dt * be / (2.0 * h1 * h2 * nd);
        const double cp = dt * 2.0 * (al / (h1 * h1) + ga / (h2 * h2)) / nd;
        return cp * F(0, 0) + ce * F(1, 0) + cw * F(-1, 0) + cn * F(0, 1) + cs * F(0, -1) +
               cx * (F(1, 1) + F(-1, -1) - F(-1, 1) - F(1, -1));
    }
    const double ce = (-dt) * ((4.0 * al) / (3.0 * h1 * h1) + (2.0 * pp) / (3.0 * h1)) / nd;
    const double cw = (-dt) * ((4.0 * al) / (3.0 * h1 * h1) - (2.0 * pp) / (3.0 * h1)) / nd;
    const double cn = (-dt) * ((4.0 * ga) / (3.0 * h2 * h2) + (2.0 * qq) / (3.0 * h2)) / nd;
    const double cs = (-dt) * ((4.0 * ga) / (3.0 * h2 * h2) - (2.0 * qq) / (3.0 * h2)) / nd;
    const double cne = (-dt) * (-8.0 * be / (9.0 * h1 * h2)) / nd;
    const double cnw = (-dt) * (8.0 * be / (9.0 * h1 * h2)) / nd;
    const double cnn = (-dt) * (-ga / (12.0 * h2 * h2) - qq / (12.0 * h2)) / nd;
    const double css = (-dt) * (-ga / (12.0 * h2 * h2) + qq / (12.0 * h2)) / nd;
    const double cee = (-dt) * (-al / (12.0 * h1 * h1) - pp / (12.0 * h1)) / nd;
    const double cww = (-dt) * (-al / (12.0 * h1 * h1) + pp / (12.0 * h1)) / nd;
    const double c22 = (-dt) * (-be / (72.0 * h1 * h2)) / nd;
    const double c2m = (-dt) * (be / (72.0 * h1 * h2)) / nd;
    const double c12 = (-dt) * (be / (9.0 * h1 * h2)) / nd;
    const double c1m = (-dt) * (-be / (9.0 * h1 * h2)) / nd;
    const double cp = dt * (5.0 * al / (2.0 * h1 * h1) + 5.0 * ga / (2.0 * h2 * h2)) / nd;
    return cp * F(0, 0) + ce * F(1, 0) + cw * F(-1, 0) + cn * F(0, 1) + cs * F(0, -1) +
           cne * (F(1, 1) + F(-1, -1)) + cnw * (F(-1, 1) + F(1, -1)) +
           cnn * F(0, 2) + css * F(0, -2) + cee * F(2, 0) + cww * F(-2, 0) +
           c22 * (F(2, 2) + F(-2, -2)) + c2m * (F(-2, 2) + F(2, -2)) +
           c12 * (F(1, 2) + F(-1, -2) + F(2, 1) + F(-2, -1)) +
           c1m * (F(-1, 2) + F(1, -2) + F(-2, 1) + F(2, -1));
#undef F
}

// Low-order predictor operator row: c[(dj+1)*3+(di+1)], diagonal c[4]
template <typename R>
__device__ __forceinline__ void alo_coefs(R al, R be, R ga, R pp, R qq, R U, R V, R h1, R h2, R dt,
                                          R nd, R c[9]) {
    const R zero = R(0);
    c[5] = -dt * (al / (h1 * h1) + pp / (R(2) * h1)) / nd + dt * fmin(U, zero) / h1;
    c[3] = -dt * (al / (h1 * h1) - pp / (R(2) * h1)) / nd - dt * fmax(U, zero) / h1;
    c[7] = -dt * (ga / (h2 * h2) + qq / (R(2) * h2)) / nd + dt * fmin(V, zero) / h2;
    c[1] = -dt * (ga / (h2 * h2) - qq / (R(2) * h2)) / nd - dt * fmax(V, zero) / h2;
    const R cx = dt * be / (R(2) * h1 * h2 * nd);
    c[8] = cx;
    c[0] = cx;
    c[6] = -cx;
    c[2] = -cx;
    c[4] = R(1) + dt * R(2) * (al / (h1 * h1) + ga / (h2 * h2)) / nd + dt * (fabs(U) / h1 + fabs(V) / h2);
}

// J-weighted flux-form pressure-correction row (see the oracle's p9_coefs).
// Faces touching a wall node carry no flux; neighbours are read clamped into
// the interior by the caller.
template <typename T>
__device__ __forceinline__ void p9_coefs(int n0, int n1, const T *__restrict__ xix,
                                         const T *__restrict__ xiy, const T *__restrict__ etx,
                                         const T *__restrict__ ety, const T *__restrict__ jac,
                                         const T *__restrict__ d, T h1, T h2, int i, int j, T c[9]) {
    const size_t P = idx(i, j, n0), E = idx(i + 1, j, n0), W = idx(i - 1, j, n0);
    const size_t Nn = idx(i, j + 1, n0), S = idx(i, j - 1, n0);
    const bool we = (i + 1 == n0 - 1), ww = (i - 1 == 0), wn = (j + 1 == n1 - 1), ws = (j - 1 == 0);
#pragma unroll
    for (int k = 0; k < 9; k++) c[k] = T(0);
    const T xP = xix[P], yP = xiy[P], eP = etx[P], fP = ety[P], jP = jac[P], dP = d[P];
    const T jxP = jP * xP, jyP = jP * yP, kxP = jP * eP, kyP = jP * fP;
    const T half = T(0.5), four = T(4);
    if (!we) {
        const T xE = xix[E], yE = xiy[E], eE = etx[E], fE = ety[E], jE = jac[E];
        const T mx = half * (jxP + jE * xE), my = half * (jyP + jE * yE);
        const T a = mx * (xP + xE) + my * (yP + yE);
        const T b = mx * (eP + eE) + my * (fP + fE);
        const T de = half * (dP + d[E]);
        const T kn = half * de * a / (h1 * h1), kt = half * de * b / (four * h1 * h2);
        c[5] += kn; c[4] -= kn;
        c[8] += kt; c[7] += kt; c[1] -= kt; c[2] -= kt;
    }
    if (!ww) {
        const T xW = xix[W], yW = xiy[W], eW = etx[W], fW = ety[W], jW = jac[W];
        const T mx = half * (jxP + jW * xW), my = half * (jyP + jW * yW);
        const T a = mx * (xP + xW) + my * (yP + yW);
        const T b = mx * (eP + eW) + my * (fP + fW);
        const T dw = half * (dP + d[W]);
        const T kn = half * dw * a / (h1 * h1), kt = half * dw * b / (four * h1 * h2);
        c[4] -= kn; c[3] += kn;
        c[7] -= kt; c[6] -= kt; c[1] += kt; c[0] += kt;
    }
    if (!wn) {
        const T xN = xix[Nn], yN = xiy[Nn], eN = etx[Nn], fN = ety[Nn], jN = jac[Nn];
        const T mx = half * (kxP + jN * eN), my = half * (kyP + jN * fN);
        const T a = mx * (eP + eN) + my * (fP + fN);
        const T b = mx * (xP + xN) + my * (yP + yN);
        const T dn = half * (dP + d[Nn]);
        const T kn = half * dn * a / (h2 * h2), kt = half * dn * b / (four * h1 * h2);
        c[7] += kn; c[4] -= kn;
        c[8] += kt; c[6] -= kt; c[5] += kt; c[3] -= kt;
    }
    if (!ws) {
        const T xS = xix[S], yS = xiy[S], eS = etx[S], fS = ety[S], jS = jac[S];
        const T mx = half * (kxP + jS * eS), my = half * (kyP + jS * fS);
        const T a = mx * (eP + eS) + my * (fP + fS);
        const T b = mx * (xP + xS) + my * (yP + yS);
        const T ds = half * (dP + d[S]);
        const T kn = half * ds * a / (h2 * h2), kt = half * ds * b / (four * h1 * h2);
        c[4] -= kn; c[1] += kn;
        c[5] -= kt; c[3] += kt; c[2] -= kt; c[0] += kt;
    }
}

// Coefficients (normal kn, tangential kt) of the face between node P=(i, j) and its
// east neighbour (dir 0) or north neighbour (dir 1): the per-face terms of
// p9_coefs, in the same arithmetic.  A face touching a wall node is zero.
template <typename T>
__device__ __forceinline__ void p9_face(int n0, int n1, const T *__restrict__ xix, const T *__restrict__ xiy,
                                        const T *__restrict__ etx, const T *__restrict__ ety,
                                        const T *__restrict__ jac, const T *__restrict__ d, T h1, T h2, int i,
                                        int j, int dir, T &kn, T &kt) {
    kn = T(0);
    kt = T(0);
    const T half = T(0.5), four = T(4);
    const size_t P = idx(i, j, n0);
    const T xP = xix[P], yP = xiy[P], eP = etx[P], fP = ety[P], jP = jac[P], dP = d[P];
    if (dir == 0) {
        if (i < 1 || i > n0 - 3 || j < 1 || j > n1 - 2) return;
        const size_t E = P + 1;
        const T xE = xix[E], yE = xiy[E], eE = etx[E], fE = ety[E], jE = jac[E];
        const T mx = half * (jP * xP + jE * xE), my = half * (jP * yP + jE * yE);
        const T a = mx * (xP + xE) + my * (yP + yE);
        const T b = mx * (eP + eE) + my * (fP + fE);
        const T de = half * (dP + d[E]);
        kn = half * de * a / (h1 * h1);
        kt = half * de * b / (four * h1 * h2);
    } else {
        if (j < 1 || j > n1 - 3 || i < 1 || i > n0 - 2) return;
        const size_t Nn = P + n0;
        const T xN = xix[Nn], yN = xiy[Nn], eN = etx[Nn], fN = ety[Nn], jN = jac[Nn];
        const T mx = half * (jP * eP + jN * eN), my = half * (jP * fP + jN * fN);
        const T a = mx * (eP + eN) + my * (fP + fN);
        const T b = mx * (xP + xN) + my * (yP + yN);
        const T dn = half * (dP + d[Nn]);
        kn = half * dn * a / (h2 * h2);
        kt = half * dn * b / (four * h1 * h2);
    }
}

// The eight face coefficients row (i, j) of the pressure-correction operator
// uses (fe_*: face to the east of a node, fn_*: face to the north): the node's
// own east and north faces and those of its west and south neighbours.
template <typename T>
struct RowFaces {
    T en, et, wn, wt, nn, nt, sn, st;
};

template <typename T>
__device__ __forceinline__ RowFaces<T> load_row_faces(int n0, const T *__restrict__ fen,
                                                      const T *__restrict__ fet, const T *__restrict__ fnn,
                                                      const T *__restrict__ fnt, int i, int j) {
    const size_t P = idx(i, j, n0);
    return {fen[P], fet[P], fen[P - 1], fet[P - 1], fnn[P], fnt[P], fnn[P - n0], fnt[P - n0]};
}

// Row coefficients from the eight faces, accumulated in the order of p9_coefs.
template <typename T>
__device__ __forceinline__ void p9_row_from(const RowFaces<T> &f, T c[9]) {
#pragma unroll
    for (int k = 0; k < 9; k++) c[k] = T(0);
    T kn = f.en, kt = f.et;
    c[5] += kn; c[4] -= kn;
    c[8] += kt; c[7] += kt; c[1] -= kt; c[2] -= kt;
    kn = f.wn; kt = f.wt;
    c[4] -= kn; c[3] += kn;
    c[7] -= kt; c[6] -= kt; c[1] += kt; c[0] += kt;
    kn = f.nn; kt = f.nt;
    c[7] += kn; c[4] -= kn;
    c[8] += kt; c[6] -= kt; c[5] += kt; c[3] -= kt;
    kn = f.sn; kt = f.st;
    c[4] -= kn; c[1] += kn;
    c[5] -= kt; c[3] += kt; c[2] -= kt; c[0] += kt;
}

// Row (i, j) of the pressure-correction operator from stored face coefficients.
template <typename T>
__device__ __forceinline__ void p9_row_faces(int n0, const T *__restrict__ fen, const T *__restrict__ fet,
                                             const T *__restrict__ fnn, const T *__restrict__ fnt, int i, int j,
                                             T c[9]) {
    p9_row_from<T>(load_row_faces<T>(n0, fen, fet, fnn, fnt, i, j), c);
}

}  // namespace cfd
