// Reference-stage kernels (periodic O-grid) behind the cfd_ref_* / cfd_sip9p /
// cfd_gauss C ABI.  Each entry point takes host arrays in the reference's
// layout (converted to j*n0 + i), runs one gfx950 kernel (or a GPU BiCGStab
// for the two solvers) and returns host arrays.  Column n0-1 duplicates
// column 0; neighbour columns wrap with period n0-1 exactly as the reference's
// inn/ipp/inn2/ipp2 maps do (solver.cpp:1076-1102), except that the literal
// column n0-1 is addressed where the reference addresses it.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "cfd_common.h"

namespace cfd {
namespace ref {

__device__ __forceinline__ size_t ix(int i, int j, int n0) { return (size_t)j * n0 + i; }
// neighbour column i+di of column i in [0, n0-2]
__device__ __forceinline__ int pw(int k, int n0) { return k < 0 ? k + n0 - 1 : (k > n0 - 1 ? k - (n0 - 1) : k); }

constexpr int K25(int di, int dj) { return (dj + 2) * 5 + (di + 2); }
constexpr int K9(int di, int dj) { return (dj + 1) * 3 + (di + 1); }

// ---------------------------------------------------------------- a2
__global__ void k_asm_mom(int n0, int n1, double h1, double h2, const double *__restrict__ alph,
                          const double *__restrict__ beta, const double *__restrict__ gamma,
                          const double *__restrict__ p1, const double *__restrict__ q1, double dt, double Re,
                          double Pr, double *__restrict__ AU, double *__restrict__ AT, double *__restrict__ BU,
                          double *__restrict__ BT) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i > n0 - 2 || j < 1 || j > n1 - 2) return;
    const size_t N = (size_t)n0 * n1, id = ix(i, j, n0);
    const double al = alph[id], be = beta[id], ga = gamma[id], pp = p1[id], qq = q1[id];
    double u[25], t[25];
#pragma unroll
    for (int k = 0; k < 25; k++) u[k] = t[k] = 0.0;
    if (j == 1 || j == n1 - 2) {
        u[K25(1, 0)] = -dt * (al / (h1 * h1) + pp / (2.0 * h1)) / Re;
        u[K25(-1, 0)] = -dt * (al / (h1 * h1) - pp / (2.0 * h1)) / Re;
        u[K25(0, 1)] = -dt * (ga / (h2 * h2) + qq / (2.0 * h2)) / Re;
        u[K25(0, -1)] = -dt * (ga / (h2 * h2) - qq / (2.0 * h2)) / Re;
        u[K25(1, 1)] = dt * be / (2.0 * h1 * h2 * Re);
        u[K25(-1, -1)] = u[K25(1, 1)];
        u[K25(-1, 1)] = -dt * be / (2.0 * h1 * h2 * Re);
        u[K25(1, -1)] = u[K25(-1, 1)];
        u[K25(0, 0)] = 1 + dt * 2.0 * (al / (h1 * h1) + ga / (h2 * h2)) / Re;
        t[K25(1, 0)] = -dt * (al / (h1 * h1) + pp / (2.0 * h1)) / (Re * Pr);
        t[K25(-1, 0)] = -dt * (al / (h1 * h1) - pp / (2.0 * h1)) / (Re * Pr);
        t[K25(0, 1)] = -dt * (ga / (h2 * h2) + qq / (2.0 * h2)) / (Re * Pr);
        t[K25(0, -1)] = -dt * (ga / (h2 * h2) - qq / (2.0 * h2)) / (Re * Pr);
        t[K25(1, 1)] = dt * (be / (2.0 * h1 * h2)) / (Re * Pr);
        t[K25(-1, -1)] = t[K25(1, 1)];
        t[K25(-1, 1)] = -dt * (be / (2.0 * h1 * h2)) / (Re * Pr);
        t[K25(1, -1)] = t[K25(-1, 1)];
        t[K25(0, 0)] = 1 + dt * 2.0 * (al / (h1 * h1) + ga / (h2 * h2)) / (Re * Pr);
    } else {
        u[K25(1, 0)] = (-dt) * ((4.0 * al) / (3.0 * h1 * h1) + (2.0 * pp) / (3.0 * h1)) / Re;
        u[K25(-1, 0)] = (-dt) * ((4.0 * al) / (3.0 * h1 * h1) - (2.0 * pp) / (3.0 * h1)) / Re;
        u[K25(0, 1)] = (-dt) * ((4.0 * ga) / (3.0 * h2 * h2) + (2.0 * qq) / (3.0 * h2)) / Re;
        u[K25(0, -1)] = (-dt) * ((4.0 * ga) / (3.0 * h2 * h2) - (2.0 * qq) / (3.0 * h2)) / Re;
        u[K25(1, 1)] = (-dt) * (-8.0 * be / (9.0 * h1 * h2)) / Re;
        u[K25(-1, 1)] = (-dt) * (8.0 * be / (9.0 * h1 * h2)) / Re;
        u[K25(1, -1)] = u[K25(-1, 1)];
        u[K25(-1, -1)] = u[K25(1, 1)];
        u[K25(0, 2)] = (-dt) * (-ga / (12.0 * h2 * h2) - qq / (12.0 * h2)) / Re;
        u[K25(0, -2)] = (-dt) * (-ga / (12.0 * h2 * h2) + qq / (12.0 * h2)) / Re;
        u[K25(2, 0)] = (-dt) * (-al / (12.0 * h1 * h1) - pp / (12.0 * h1)) / Re;
        u[K25(-2, 0)] = (-dt) * (-al / (12.0 * h1 * h1) + pp / (12.0 * h1)) / Re;
        u[K25(2, 2)] = (-dt) * (-be / (72.0 * h1 * h2)) / Re;
        u[K25(-2, 2)] = (-dt) * (be / (72.0 * h1 * h2)) / Re;
        u[K25(2, -2)] = u[K25(-2, 2)];
        u[K25(-2, -2)] = u[K25(2, 2)];
        u[K25(1, 2)] = (-dt) * (be / (9.0 * h1 * h2)) / Re;
        u[K25(-1, 2)] = (-dt) * (-be / (9.0 * h1 * h2)) / Re;
        u[K25(1, -2)] = u[K25(-1, 2)];
        u[K25(-1, -2)] = u[K25(1, 2)];
        u[K25(2, 1)] = u[K25(1, 2)];
        u[K25(-2, 1)] = u[K25(-1, 2)];
        u[K25(2, -1)] = u[K25(-1, 2)];
        u[K25(-2, -1)] = u[K25(1, 2)];
        u[K25(0, 0)] = 1 + dt * (5.0 * al / (2.0 * h1 * h1) + 5.0 * ga / (2.0 * h2 * h2)) / Re;
#pragma unroll
        for (int k = 0; k < 25; k++) t[k] = u[k] / Pr;
        t[K25(0, 0)] = 1 + dt * (5.0 * al / (2.0 * h1 * h1) + 5.0 * ga / (2.0 * h2 * h2)) / (Re * Pr);
    }
    if (j == 1) {
        BU[0 * n0 + i] = u[K25(0, -1)]; BU[1 * n0 + i] = u[K25(1, -1)]; BU[2 * n0 + i] = u[K25(-1, -1)];
        BT[0 * n0 + i] = t[K25(0, -1)]; BT[1 * n0 + i] = t[K25(1, -1)]; BT[2 * n0 + i] = t[K25(-1, -1)];
        u[K25(0, -1)] = u[K25(1, -1)] = u[K25(-1, -1)] = 0;
        t[K25(0, -1)] = t[K25(1, -1)] = t[K25(-1, -1)] = 0;
    }
    if (j == n1 - 2) {
        BU[3 * n0 + i] = u[K25(0, 1)]; BU[4 * n0 + i] = u[K25(1, 1)]; BU[5 * n0 + i] = u[K25(-1, 1)];
        BT[3 * n0 + i] = t[K25(0, 1)]; BT[4 * n0 + i] = t[K25(1, 1)]; BT[5 * n0 + i] = t[K25(-1, 1)];
        u[K25(0, 1)] = u[K25(1, 1)] = u[K25(-1, 1)] = 0;
        t[K25(0, 1)] = t[K25(1, 1)] = t[K25(-1, 1)] = 0;
    }
    const size_t dup = ix(n0 - 1, j, n0);
#pragma unroll
    for (int k = 0; k < 25; k++) {
        AU[(size_t)k * N + id] = u[k];
        AT[(size_t)k * N + id] = t[k];
        if (i == 0) {
            AU[(size_t)k * N + dup] = u[k];
            AT[(size_t)k * N + dup] = t[k];
        }
    }
}

// ---------------------------------------------------------------- a3
__global__ void k_asm_p9(int n0, int n1, double h1, double h2, const double *__restrict__ xix,
                         const double *__restrict__ xiy, const double *__restrict__ etx,
                         const double *__restrict__ ety, double *__restrict__ A9) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i > n0 - 2 || j < 1 || j > n1 - 2) return;
    const size_t N = (size_t)n0 * n1;
    const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1;
    const size_t P = ix(i, j, n0), E = ix(ipp, j, n0), W = ix(inn, j, n0), Nn = ix(i, j + 1, n0), S = ix(i, j - 1, n0);
    const double h11 = 2.0 * h1 * h1, h22 = 2.0 * h2 * h2, h12 = 8.0 * h1 * h2;
    double c[9];
    c[K9(1, 0)] = (xix[P] / h11) * (xix[P] + xix[E]) + (etx[P] / h12) * (xix[Nn] - xix[S]) +
                  (xiy[P] / h11) * (xiy[P] + xiy[E]) + (ety[P] / h12) * (xiy[Nn] - xiy[S]);
    c[K9(-1, 0)] = (xix[P] / h11) * (xix[P] + xix[W]) + (etx[P] / h12) * (xix[S] - xix[Nn]) +
                   (xiy[P] / h11) * (xiy[P] + xiy[W]) + (ety[P] / h12) * (xiy[S] - xiy[Nn]);
    c[K9(0, 1)] = (xix[P] / h12) * (etx[E] - etx[W]) + (etx[P] / h22) * (etx[P] + etx[Nn]) +
                  (xiy[P] / h12) * (ety[E] - ety[W]) + (ety[P] / h22) * (ety[P] + ety[Nn]);
    c[K9(0, -1)] = (xix[P] / h12) * (etx[W] - etx[E]) + (etx[P] / h22) * (etx[P] + etx[S]) +
                   (xiy[P] / h12) * (ety[W] - ety[E]) + (ety[P] / h22) * (ety[P] + ety[S]);
    c[K9(1, 1)] = (xix[P] / h12) * (etx[P] + etx[E]) + (etx[P] / h12) * (xix[P] + xix[Nn]) +
                  (xiy[P] / h12) * (ety[P] + ety[E]) + (ety[P] / h12) * (xiy[P] + xiy[Nn]);
    c[K9(-1, -1)] = (xix[P] / h12) * (etx[P] + etx[W]) + (etx[P] / h12) * (xix[P] + xix[S]) +
                    (xiy[P] / h12) * (ety[P] + ety[W]) + (ety[P] / h12) * (xiy[P] + xiy[S]);
    c[K9(-1, 1)] = -(xix[P] / h12) * (etx[P] + etx[W]) + -(etx[P] / h12) * (xix[P] + xix[Nn]) +
                   -(xiy[P] / h12) * (ety[P] + ety[W]) + -(ety[P] / h12) * (xiy[P] + xiy[Nn]);
    c[K9(1, -1)] = -(xix[P] / h12) * (etx[P] + etx[E]) + -(etx[P] / h12) * (xix[P] + xix[S]) +
                   -(xiy[P] / h12) * (ety[P] + ety[E]) + -(ety[P] / h12) * (xiy[P] + xiy[S]);
    {
        const double pxi = 1.0 / (2. * h1 * h1), pet = 1.0 / (2. * h2 * h2);
        const double ap = -xix[P] * (2. * xix[P] + xix[W] + xix[E]);
        const double bp = -etx[P] * (2. * etx[P] + etx[S] + etx[Nn]);
        const double cp = -xiy[P] * (2. * xiy[P] + xiy[W] + xiy[E]);
        const double dp = -ety[P] * (2. * ety[P] + ety[S] + ety[Nn]);
        c[K9(0, 0)] = ap * pxi + bp * pet + cp * pxi + dp * pet;
    }
    const size_t dup = ix(n0 - 1, j, n0);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        A9[(size_t)k * N + P] = c[k];
        if (i == 0) A9[(size_t)k * N + dup] = c[k];
    }
}

// ---------------------------------------------------------------- a4 + a5
__global__ void k_contra(int n0, int n1, const double *__restrict__ u, const double *__restrict__ v,
                         const double *__restrict__ xix, const double *__restrict__ xiy,
                         const double *__restrict__ etx, const double *__restrict__ ety, double *__restrict__ uxi,
                         double *__restrict__ uet) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= n0 || j >= n1) return;
    const size_t id = ix(i, j, n0);
    uxi[id] = xix[id] * u[id] + xiy[id] * v[id];
    uet[id] = etx[id] * u[id] + ety[id] * v[id];
}

__device__ __forceinline__ double hyb(double fm2, double fm1, double f0, double fp1, double fp2, double vel,
                                      double pec, double h) {
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

struct RhsIn {
    const double *u, *v, *T, *up, *vp, *p, *xix, *xiy, *etx, *ety, *alph, *gamma, *BU, *BT, *uxi, *uet;
};
struct RhsOut {
    double *qu, *qv, *qt, *qup, *qvp;
};

__global__ void k_rhs(int n0, int n1, double h1, double h2, RhsIn a, RhsOut o, double Re, double Pr, double Ri,
                      double dt) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i > n0 - 2 || j < 1 || j > n1 - 2) return;
    const int inn = pw(i - 1, n0), ipp = i + 1, inn2 = pw(i - 2, n0), ipp2 = pw(i + 2, n0);
    const size_t P = ix(i, j, n0);
    const bool interior = (j >= 2 && j <= n1 - 3);
    const double *fl[3] = {a.u, a.v, a.T};
    const double U = a.uxi[P], V = a.uet[P];
    double conv[3];
    for (int k = 0; k < 3; k++) {
        const double *f = fl[k];
        double pec1, pec2;
        if (k <= 1) {
            pec1 = U * Re * h1 / a.alph[P];
            pec2 = V * Re * h2 / a.gamma[P];
        } else {
            pec1 = U * Re * Pr * h1 / a.alph[P];
            pec2 = V * Re * Pr * h2 / a.gamma[P];
        }
        double dxi, det;
        if (interior) {
            dxi = hyb(f[ix(inn2, j, n0)], f[ix(inn, j, n0)], f[P], f[ix(ipp, j, n0)], f[ix(ipp2, j, n0)], U, pec1, h1);
            det = hyb(f[ix(i, j - 2, n0)], f[ix(i, j - 1, n0)], f[P], f[ix(i, j + 1, n0)], f[ix(i, j + 2, n0)], V, pec2,
                      h2);
        } else {
            const double xpp = 8.0 * (f[ix(ipp, j, n0)] - f[ix(inn, j, n0)]);
            const double xnn = f[ix(ipp2, j, n0)] - f[ix(inn2, j, n0)];
            dxi = (1.0 / 12.0) * (xpp - xnn) / h1;
            det = 0.5 * (f[ix(i, j + 1, n0)] - f[ix(i, j - 1, n0)]) / h2;
        }
        conv[k] = U * dxi + V * det;
    }
    const double *p = a.p;
    const double dp_dxi = (p[ix(ipp, j, n0)] - p[ix(inn, j, n0)]) / (2.0 * h1);
    const double dp_de = (p[ix(i, j + 1, n0)] - p[ix(i, j - 1, n0)]) / (2.0 * h2);
    const double dp_dx = a.xix[P] * dp_dxi + a.etx[P] * dp_de;
    const double dp_dy = a.xiy[P] * dp_dxi + a.ety[P] * dp_de;
    double qu = dt * (-conv[0] - dp_dx) + a.u[P];
    double qv = dt * (-conv[1] - dp_dy + Ri * a.T[P]) + a.v[P];
    double qt = -dt * conv[2] + a.T[P];
    double qup = qu + dt * dp_dx;
    double qvp = qv + dt * dp_dy;
    if (j == 1) {
        const size_t s = ix(i, j - 1, n0), se = ix(ipp, j - 1, n0), sw = ix(inn, j - 1, n0);
        const double bs = a.BU[i], bse = a.BU[n0 + i], bsw = a.BU[2 * n0 + i];
        const double ts = a.BT[i], tse = a.BT[n0 + i], tsw = a.BT[2 * n0 + i];
        qu = qu - (bs * a.u[s] + bse * a.u[se] + bsw * a.u[sw]);
        qv = qv - (bs * a.v[s] + bse * a.v[se] + bsw * a.v[sw]);
        qt = qt - (ts * a.T[s] + tse * a.T[se] + tsw * a.T[sw]);
        qup = qup - (bs * a.up[s] + bse * a.up[se] + bsw * a.up[sw]);
        qvp = qvp - (bs * a.vp[s] + bse * a.vp[se] + bsw * a.vp[sw]);
    }
    if (j == n1 - 2) {
        const size_t nn = ix(i, j + 1, n0), ne = ix(ipp, j + 1, n0), nw = ix(inn, j + 1, n0);
        const double bn = a.BU[3 * n0 + i], bne = a.BU[4 * n0 + i], bnw = a.BU[5 * n0 + i];
        const double tn = a.BT[3 * n0 + i], tne = a.BT[4 * n0 + i], tnw = a.BT[5 * n0 + i];
        qu = qu - (bn * a.u[nn] + bne * a.u[ne] + bnw * a.u[nw]);
        qv = qv - (bn * a.v[nn] + bne * a.v[ne] + bnw * a.v[nw]);
        qt = qt - (tn * a.T[nn] + tne * a.T[ne] + tnw * a.T[nw]);
        qup = qup - (bn * a.up[nn] + bne * a.up[ne] + bnw * a.up[nw]);
        qvp = qvp - (bn * a.vp[nn] + bne * a.vp[ne] + bnw * a.vp[nw]);
    }
    o.qu[P] = qu; o.qv[P] = qv; o.qt[P] = qt; o.qup[P] = qup; o.qvp[P] = qvp;
    if (i == 0) {
        const size_t dup = ix(n0 - 1, j, n0);
        o.qu[dup] = qu; o.qv[dup] = qv; o.qt[dup] = qt; o.qup[dup] = qup; o.qvp[dup] = qvp;
    }
}

// ---------------------------------------------------------------- a7
__global__ void k_face_div_ref(int n0, int n1, double h1, double h2, const double *__restrict__ up,
                               const double *__restrict__ vp, const double *__restrict__ p,
                               const double *__restrict__ xix, const double *__restrict__ xiy,
                               const double *__restrict__ etx, const double *__restrict__ ety, double dt,
                               double *__restrict__ q) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i > n0 - 2 || j < 1 || j > n1 - 2) return;
    const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1, jpp = j + 1, jnn = j - 1;
    const size_t P = ix(i, j, n0), E = ix(ipp, j, n0), W = ix(inn, j, n0), Nn = ix(i, jpp, n0), S = ix(i, jnn, n0);
    const double dpdxi_ip = (p[E] - p[P]) / h1;
    const double dpde_ip = (p[ix(ipp, jpp, n0)] + p[Nn] - p[S] - p[ix(ipp, jnn, n0)]) / (4.0 * h2);
    const double dpdxi_in = (p[P] - p[W]) / h1;
    const double dpde_in = (p[Nn] + p[ix(inn, jpp, n0)] - p[S] - p[ix(inn, jnn, n0)]) / (4.0 * h2);
    const double dpdxi_jp = (p[ix(ipp, jpp, n0)] - p[ix(inn, jpp, n0)] + p[E] - p[W]) / (4.0 * h1);
    const double dpde_jp = (p[Nn] - p[P]) / h2;
    const double dpdxi_jn = (p[E] - p[W] + p[ix(ipp, jnn, n0)] - p[ix(inn, jnn, n0)]) / (4.0 * h1);
    const double dpde_jn = (p[P] - p[S]) / h2;
    const double us_ip = 0.5 * (up[P] + up[E]) - 0.5 * dt * ((xix[P] + xix[E]) * dpdxi_ip + (etx[P] + etx[E]) * dpde_ip);
    const double us_in = 0.5 * (up[P] + up[W]) - 0.5 * dt * ((xix[P] + xix[W]) * dpdxi_in + (etx[P] + etx[W]) * dpde_in);
    const double us_jp = 0.5 * (up[P] + up[Nn]) - 0.5 * dt * ((xix[P] + xix[Nn]) * dpdxi_jp + (etx[P] + etx[Nn]) * dpde_jp);
    const double us_jn = 0.5 * (up[P] + up[S]) - 0.5 * dt * ((xix[P] + xix[S]) * dpdxi_jn + (etx[P] + etx[S]) * dpde_jn);
    const double vs_ip = 0.5 * (vp[P] + vp[E]) - 0.5 * dt * ((xiy[P] + xiy[E]) * dpdxi_ip + (ety[P] + ety[E]) * dpde_ip);
    const double vs_in = 0.5 * (vp[P] + vp[W]) - 0.5 * dt * ((xiy[P] + xiy[W]) * dpdxi_in + (ety[P] + ety[W]) * dpde_in);
    const double vs_jp = 0.5 * (vp[P] + vp[Nn]) - 0.5 * dt * ((xiy[P] + xiy[Nn]) * dpdxi_jp + (ety[P] + ety[Nn]) * dpde_jp);
    const double vs_jn = 0.5 * (vp[P] + vp[S]) - 0.5 * dt * ((xiy[P] + xiy[S]) * dpdxi_jn + (ety[P] + ety[S]) * dpde_jn);
    const double qq = (xix[P] * ((us_ip - us_in) / h1)) + (etx[P] * ((us_jp - us_jn) / h2)) +
                      (xiy[P] * ((vs_ip - vs_in) / h1)) + (ety[P] * ((vs_jp - vs_jn) / h2));
    q[P] = qq / dt;
}

// ---------------------------------------------------------------- a9, a10
__global__ void k_pcor_bc(int n0, int n1, double *__restrict__ pc, const double *__restrict__ xnox,
                          const double *__restrict__ xnoy, double uinf, double vinf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n0 - 2) return;
    pc[ix(i, 0, n0)] = pc[ix(i, 1, n0)];
    const int j = n1 - 1;
    const double vnn = uinf * xnox[i] + vinf * xnoy[i];
    const double val = (vnn >= 0) ? pc[ix(i, j - 1, n0)] : 0.0;
    pc[ix(i, j, n0)] = val;
    if (i == 0) {
        pc[ix(n0 - 1, 0, n0)] = pc[ix(0, 1, n0)];
        pc[ix(n0 - 1, j, n0)] = val;
    }
}

__global__ void k_correct_ref(int n0, int n1, double h1, double h2, const double *__restrict__ us,
                              const double *__restrict__ vs, const double *__restrict__ pc,
                              const double *__restrict__ xix, const double *__restrict__ xiy,
                              const double *__restrict__ etx, const double *__restrict__ ety, double dt,
                              double *__restrict__ u, double *__restrict__ v, double *__restrict__ p) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i > n0 - 2 || j < 1 || j > n1 - 2) return;
    const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1;
    const size_t P = ix(i, j, n0);
    const double dpx = 0.5 * (pc[ix(ipp, j, n0)] - pc[ix(inn, j, n0)]) / h1;
    const double dpe = 0.5 * (pc[ix(i, j + 1, n0)] - pc[ix(i, j - 1, n0)]) / h2;
    const double uu = us[P] - dt * (xix[P] * dpx + etx[P] * dpe);
    const double vv = vs[P] - dt * (xiy[P] * dpx + ety[P] * dpe);
    const double pp = p[P] + pc[P];
    u[P] = uu; v[P] = vv; p[P] = pp;
    if (i == 0) {
        const size_t dup = ix(n0 - 1, j, n0);
        u[dup] = uu; v[dup] = vv; p[dup] = pp;
    }
}

// ---------------------------------------------------------------- a11
struct WallRef {
    const double *u, *v, *T, *xix, *xiy, *etx, *ety, *alph, *beta, *gamma, *q1, *ajac, *alc_u, *alc_v;
};

__global__ void k_wall_p_ref(int n0, int n1, double h1, double h2, WallRef w, double *__restrict__ p, double Re,
                             double Ri) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n0 - 2) return;
    const int j = 0;
    const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1;
    const size_t P = ix(i, j, n0);
    const double uxi = w.xix[P] * w.u[P] + w.xiy[P] * w.v[P];
    const double uet = w.etx[P] * w.u[P] + w.ety[P] * w.v[P];
    double dp_dx = 0.0, dp_dy = 0.0;
    for (int k = 0; k < 2; k++) {
        const double *f = k == 0 ? w.u : w.v;
        const double aa = w.alph[P] * (f[ix(ipp, j, n0)] + f[ix(inn, j, n0)] - 2 * f[P]) / (h1 * h1);
        const double gg = w.gamma[P] * (f[ix(i, j + 2, n0)] + f[P] - 2 * f[ix(i, j + 1, n0)]) / (h2 * h2);
        const double bb = w.beta[P] * (f[ix(ipp, j + 1, n0)] + f[ix(inn, j, n0)] - f[ix(inn, j + 1, n0)] -
                                       f[ix(ipp, j, n0)]) / (2 * h1 * h2);
        const double qqq = w.q1[P] * (-3 * f[P] + 4 * f[ix(i, j + 1, n0)] - f[ix(i, j + 2, n0)]) / (2 * h2);
        const double d2u = aa + gg - 2 * bb + qqq;
        double conv = uxi * 0.5 * (f[ix(ipp, j, n0)] - f[ix(inn, j, n0)]) / h1;
        conv = conv + uet * (f[ix(i, j + 1, n0)] - f[P]) / h2;
        if (k == 0) dp_dx = 1.0 * d2u / Re - conv - (w.alc_u ? w.alc_u[i] : 0.0);
        else dp_dy = 1.0 * d2u / Re - conv - (w.alc_v ? w.alc_v[i] : 0.0) + Ri * w.T[P];
    }
    const double val = p[ix(i, j + 1, n0)] - (dp_dx * (-w.xiy[P] * w.ajac[P]) + dp_dy * (w.xix[P] * w.ajac[P])) * h2;
    p[P] = val;
    if (i == 0) p[ix(n0 - 1, j, n0)] = val;
}

// ---------------------------------------------------------------- outer-boundary rows
// Predictor velocities on the outer row j = n1-1 (solver.cpp:1351-1381): the
// current velocity where the free stream enters (vnn >= 0), else the quadratic
// extrapolation (5 f_{j-1} - 4 f_{j-2} + f_{j-3}) / 2.  One thread per column.
__global__ void k_outflow_up_ref(int n0, int n1, double *__restrict__ up, double *__restrict__ vp,
                                 const double *__restrict__ u, const double *__restrict__ v,
                                 const double *__restrict__ xnox, const double *__restrict__ xnoy, double uinf,
                                 double vinf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n0 - 2) return;
    const int j = n1 - 1;
    const size_t P = ix(i, j, n0);
    const double vnn = uinf * xnox[i] + vinf * xnoy[i];
    double a, b;
    if (vnn >= 0) {
        a = u[P];
        b = v[P];
    } else {
        a = (5.0 * up[ix(i, j - 1, n0)] - 4.0 * up[ix(i, j - 2, n0)] + up[ix(i, j - 3, n0)]) / 2.0;
        b = (5.0 * vp[ix(i, j - 1, n0)] - 4.0 * vp[ix(i, j - 2, n0)] + vp[ix(i, j - 3, n0)]) / 2.0;
    }
    up[P] = a;
    vp[P] = b;
    if (i == 0) {
        up[ix(n0 - 1, j, n0)] = a;
        vp[ix(n0 - 1, j, n0)] = b;
    }
}

// Boundary velocities after the correction (solver.cpp:1543-1655), one block:
// the circulation along row n1-2 (a block sum of the trapezoid terms), then per
// outer column the far-field prediction from the radial/tangential velocity of
// row n1-2 -- vr0 cr^2 + vr_inf (1 - cr^2) and vth0 cr^k + vth_inf (1 - cr^k),
// k = 1 when |circulation| > 1e-2 else 2, cr the radius ratio of the two rows
// (cr^2 evaluated as cr * cr) -- where the flow leaves, the free stream (T = 0)
// where it enters, and T convected from the step's start values Told, uet; the
// wall row gets the cylinder velocity (-a y, a x), a = amp * cwt, also in up/vp.
// Rows n1-2 and n1-1 are distinct (n1 >= 5), so the circulation reads no value
// this kernel writes.
constexpr int FF_NT = 256;
struct FarRef {
    const double *x, *y, *Told, *uet, *etx, *ety, *ajac, *xnox, *xnoy;
};
__global__ __launch_bounds__(FF_NT) void k_farfield_ref(int n0, int n1, double h2, FarRef a, double *__restrict__ u,
                                                        double *__restrict__ v, double *__restrict__ T,
                                                        double *__restrict__ up, double *__restrict__ vp,
                                                        double uinf, double vinf, double dt, double amp, double cwt,
                                                        double *__restrict__ circ_out) {
    const int t = threadIdx.x;
    const int jm = n1 - 2, j = n1 - 1;
    const double de = 1.0 / (n0 - 2);
    double part = 0.0;
    for (int i = t; i < n0 - 1; i += FF_NT) {
        const size_t A = ix(i, jm, n0), B = ix(i + 1, jm, n0);
        const double f1 = (u[A] * a.ety[A] - v[A] * a.etx[A]) * fabs(a.ajac[A]);
        const double f2 = (u[B] * a.ety[B] - v[B] * a.etx[B]) * fabs(a.ajac[B]);
        part += de * 0.5 * (f1 + f2);
    }
    __shared__ double s_circ;
    const double circ = block_sum<FF_NT>(part);
    if (t == 0) {
        s_circ = circ;
        if (circ_out) *circ_out = circ;
    }
    __syncthreads();
    const bool k1 = fabs(s_circ) > 1e-2;
    for (int i = t; i < n0 - 1; i += FF_NT) {
        const size_t Q = ix(i, jm, n0), P = ix(i, j, n0);
        const double rq = sqrt(a.x[Q] * a.x[Q] + a.y[Q] * a.y[Q]);
        const double cq = a.x[Q] / rq, sq = a.y[Q] / rq;
        const double vr0 = u[Q] * cq + v[Q] * sq, vt0 = -u[Q] * sq + v[Q] * cq;
        const double rp = sqrt(a.x[P] * a.x[P] + a.y[P] * a.y[P]);
        const double cr = rq / rp, c = a.x[P] / rp, s = a.y[P] / rp;
        const double vrinf = uinf * c + vinf * s, vtinf = -uinf * s + vinf * c;
        const double cr2 = cr * cr, crk = k1 ? cr : cr2;
        const double vr1 = vr0 * cr2 + vrinf * (1 - cr2), vt1 = vt0 * crk + vtinf * (1 - crk);
        const double vnn = uinf * a.xnox[i] + vinf * a.xnoy[i];
        double uu = uinf, vv = vinf, tt = 0.0;
        if (vnn < 0) {
            uu = c * vr1 - s * vt1;
            vv = s * vr1 + c * vt1;
            tt = a.Told[P] - (a.uet[P] * dt / h2) * (a.Told[P] - a.Told[ix(i, j - 1, n0)]);
        }
        u[P] = uu;
        v[P] = vv;
        T[P] = tt;
        if (i == 0) {
            const size_t D = ix(n0 - 1, j, n0);
            u[D] = uu;
            v[D] = vv;
            T[D] = tt;
        }
    }
    for (int i = t; i < n0; i += FF_NT) {
        const size_t P = ix(i, 0, n0);
        const double uu = -amp * cwt * a.y[P], vv = amp * cwt * a.x[P];
        u[P] = uu;
        up[P] = uu;
        v[P] = vv;
        vp[P] = vv;
    }
}

// Outer-row pressure (solver.cpp:1731-1782): the normal momentum equation with
// one-sided eta differences and local acceleration (f - f_old)/dt where the free
// stream enters, Gresho's condition (3 uet_j - 4 uet_{j-1} + uet_{j-2}) / (2 Re h2)
// where it leaves (uxi, uet of the current velocity).  One thread per column.
struct ExitRef {
    const double *u, *v, *T, *uold, *vold, *xix, *xiy, *etx, *ety, *alph, *beta, *gamma, *q1, *ajac, *xnox, *xnoy;
};
__global__ void k_exit_p_ref(int n0, int n1, double h1, double h2, ExitRef w, double *__restrict__ p, double uinf,
                             double vinf, double dt, double Re, double Ri) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n0 - 2) return;
    const int j = n1 - 1;
    const size_t P = ix(i, j, n0), Q = ix(i, j - 1, n0), R = ix(i, j - 2, n0);
    const double vnn = uinf * w.xnox[i] + vinf * w.xnoy[i];
    double val;
    if (vnn >= 0) {
        const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1;
        const double uxi = w.xix[P] * w.u[P] + w.xiy[P] * w.v[P];
        const double uet = w.etx[P] * w.u[P] + w.ety[P] * w.v[P];
        double dp_dx = 0.0, dp_dy = 0.0;
        for (int k = 0; k < 2; k++) {
            const double *f = k == 0 ? w.u : w.v;
            const double *fo = k == 0 ? w.uold : w.vold;
            const double aa = w.alph[P] * (f[ix(ipp, j, n0)] + f[ix(inn, j, n0)] - 2 * f[P]) / (h1 * h1);
            const double gg = w.gamma[P] * (f[P] + f[R] - 2 * f[Q]) / (h2 * h2);
            const double bb = w.beta[P] * (f[ix(ipp, j, n0)] + f[ix(inn, j - 1, n0)] - f[ix(ipp, j - 1, n0)] -
                                           f[ix(inn, j, n0)]) / (2 * h1 * h2);
            const double qqq = w.q1[P] * (3 * f[P] - 4 * f[Q] + f[R]) / (2 * h2);
            const double d2u = aa + gg - 2 * bb + qqq;
            double conv = uxi * 0.5 * (f[ix(ipp, j, n0)] - f[ix(inn, j, n0)]) / h1;
            conv = conv + uet * (3.0 * f[P] - 4 * f[Q] + f[R]) / (2 * h2);
            const double alc = (f[P] - fo[P]) / dt;
            if (k == 0) dp_dx = 1.0 * d2u / Re - conv - alc;
            else dp_dy = 1.0 * d2u / Re - conv - alc + Ri * w.T[P];
        }
        val = p[Q] + (dp_dx * (-w.xiy[P] * w.ajac[P]) + dp_dy * (w.xix[P] * w.ajac[P])) * h2;
    } else {
        const double e0 = w.etx[P] * w.u[P] + w.ety[P] * w.v[P];
        const double e1 = w.etx[Q] * w.u[Q] + w.ety[Q] * w.v[Q];
        const double e2 = w.etx[R] * w.u[R] + w.ety[R] * w.v[R];
        val = 0.5 * (1.0 / Re) * ((3 * e0 - 4 * e1 + e2) / h2);
    }
    p[P] = val;
    if (i == 0) p[ix(n0 - 1, j, n0)] = val;
}

// ---------------------------------------------------------------- post-processing
// Stream function (solver.cpp:1784-1800): along each eta line, the trapezoid sum of
// the J-weighted contravariant flux.  The integral is a prefix sum down each column:
// a 32 x 32 block takes 32 columns,
// each of its 32 row-threads of a column sums a chunk of consecutive rows, the chunk
// totals are scanned in LDS, and every thread writes its rows from its chunk's offset
// (the reference sums the rows one after another; the order of the additions differs
// here, not the terms).
constexpr int SI_W = 32;
__device__ __forceinline__ double si_inc(int n0, double h2, const double *__restrict__ u, const double *__restrict__ v,
                                         const double *__restrict__ xix, const double *__restrict__ xiy,
                                         const double *__restrict__ ajac, int i, int j) {
    const size_t P = ix(i, j, n0), Q = ix(i, j - 1, n0);
    const double ca = (xix[P] * u[P] * fabs(ajac[P]) + xix[Q] * u[Q] * fabs(ajac[Q]));
    const double cb = (xiy[P] * v[P] * fabs(ajac[P]) + xiy[Q] * v[Q] * fabs(ajac[Q]));
    return (ca + cb) * 0.5 * h2;
}
__global__ __launch_bounds__(SI_W * SI_W) void k_si_ref(int n0, int n1, double h2, const double *__restrict__ u,
                                                        const double *__restrict__ v, const double *__restrict__ xix,
                                                        const double *__restrict__ xiy,
                                                        const double *__restrict__ ajac, double *__restrict__ si) {
    __shared__ double tot[SI_W][SI_W + 1];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = blockIdx.x * SI_W + tx;
    const int ch = (n1 - 1 + SI_W - 1) / SI_W;  // rows per chunk
    const int j0 = 1 + ty * ch, j1 = min(n1, j0 + ch);
    double acc = 0.0;
    if (i < n0)
        for (int j = j0; j < j1; j++) acc += si_inc(n0, h2, u, v, xix, xiy, ajac, i, j);
    tot[tx][ty] = acc;
    __syncthreads();
    if (ty == 0) {  // exclusive prefix of the chunk totals of column tx
        double run = 0.0;
        for (int k = 0; k < SI_W; k++) {
            const double t = tot[tx][k];
            tot[tx][k] = run;
            run += t;
        }
    }
    __syncthreads();
    if (i >= n0) return;
    if (ty == 0) si[ix(i, 0, n0)] = 0.0;
    double run = tot[tx][ty];
    for (int j = j0; j < j1; j++) {
        run += si_inc(n0, h2, u, v, xix, xiy, ajac, i, j);
        si[ix(i, j, n0)] = run;
    }
}

// Vorticity dv/dx - du/dy (solver.cpp:1802-1887): central differences inside,
// one-sided eta differences on the wall and outer rows; column n0-1 duplicates 0.
__global__ void k_vort_ref(int n0, int n1, double h1, double h2, const double *__restrict__ u,
                           const double *__restrict__ v, const double *__restrict__ xix,
                           const double *__restrict__ xiy, const double *__restrict__ etx,
                           const double *__restrict__ ety, double *__restrict__ vort) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i > n0 - 2 || j >= n1) return;
    const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1;
    const size_t P = ix(i, j, n0);
    double dv_det, du_det;
    if (j == 0) {
        dv_det = 1.0 / h2 * (v[ix(i, j + 1, n0)] - v[P]);
        du_det = 1.0 / h2 * (u[ix(i, j + 1, n0)] - u[P]);
    } else if (j == n1 - 1) {
        dv_det = 1.0 / h2 * (v[P] - v[ix(i, j - 1, n0)]);
        du_det = 1.0 / h2 * (u[P] - u[ix(i, j - 1, n0)]);
    } else {
        dv_det = 0.5 / h2 * (v[ix(i, j + 1, n0)] - v[ix(i, j - 1, n0)]);
        du_det = 0.5 / h2 * (u[ix(i, j + 1, n0)] - u[ix(i, j - 1, n0)]);
    }
    const double dv_dx = xix[P] * (0.5 / h1 * (v[ix(ipp, j, n0)] - v[ix(inn, j, n0)])) + etx[P] * dv_det;
    const double du_dy = xiy[P] * (0.5 / h1 * (u[ix(ipp, j, n0)] - u[ix(inn, j, n0)])) + ety[P] * du_det;
    const double w = dv_dx - du_dy;
    vort[P] = w;
    if (i == 0) vort[ix(n0 - 1, j, n0)] = w;
}

// Force and moment coefficients and the mean wall Nusselt number from the wall
// row (solver.cpp:1891-1982), one block: the seven trapezoid integrals, then
// out = cl, cd, cm, CL_pr, CD_pr, CL_vor, CD_vor, Nu.  sa, ca: sin and cos of the
// flow angle; nu_den: the ellipse-perimeter factor of Nu (both from the host).
struct ForceRef {
    const double *x, *y, *p, *vort, *T, *ajac, *etx, *ety;
};
__global__ void k_forces_ref(int n0, double h1, double h2, ForceRef a, double Re, double sa, double ca,
                             double nu_den, double *__restrict__ out) {
    double v[7] = {0, 0, 0, 0, 0, 0, 0};  // pr_x, pr_y, vor_x, vor_y, press_i, vor_i, temp_i
    for (int i = threadIdx.x; i < n0 - 1; i += 256) {
        const size_t A = i, B = i + 1;  // row j = 0
        const double PJ1 = a.p[A] * a.ajac[A], pj2 = a.p[B] * a.ajac[B];
        const double VJ1 = a.vort[A] * a.ajac[A], VJ2 = a.vort[B] * a.ajac[B];
        v[0] += 0.5 * h1 * (PJ1 * a.etx[A] + pj2 * a.etx[B]);
        v[1] += 0.5 * h1 * (PJ1 * a.ety[A] + pj2 * a.ety[B]);
        v[2] += 0.5 * h1 * (VJ1 * a.ety[A] + VJ2 * a.ety[B]);
        v[3] += 0.5 * h1 * (VJ1 * a.etx[A] + VJ2 * a.etx[B]);
        const double TJ1 = a.ajac[A] * (a.etx[A] * a.etx[A] + a.ety[A] * a.ety[A]);
        const double TJ2 = a.ajac[B] * (a.etx[B] * a.etx[B] + a.ety[B] * a.ety[B]);
        const double fp1 = PJ1 * (a.x[A] * a.ety[A] - a.y[A] * a.etx[A]);
        const double fp2 = pj2 * (a.x[B] * a.ety[B] - a.y[B] * a.etx[B]);
        const double fv1 = VJ1 * (a.x[A] * a.etx[A] + a.y[A] * a.ety[A]);
        const double fv2 = VJ2 * (a.x[B] * a.etx[B] + a.y[B] * a.ety[B]);
        const double fh1 = TJ1 * (4 * a.T[n0 + A] - 3 * a.T[A] - a.T[2 * (size_t)n0 + A]) / (2 * h2);
        const double fh2 = TJ2 * (4 * a.T[n0 + B] - 3 * a.T[B] - a.T[2 * (size_t)n0 + B]) / (2 * h2);
        v[4] += 0.5 * h1 * (fp1 + fp2);
        v[5] += 0.5 * h1 * (fv1 + fv2);
        v[6] += 0.5 * (fh1 + fh2) * h1;
    }
    block_sums<7, 256>(v);
    if (threadIdx.x != 0) return;
    const double cx = 2 * v[0] + (2.0 / Re) * v[2], cy = 2 * v[1] - (2.0 / Re) * v[3];
    out[0] = cy * sa - cx * ca;
    out[1] = cy * ca + cx * sa;
    out[2] = 2 * v[4] - (2.0 / Re) * v[5];
    out[3] = 2 * v[1] * sa - 2 * v[0] * ca;
    out[4] = 2 * v[1] * ca + 2 * v[0] * sa;
    out[5] = -(2.0 / Re) * v[3] * sa - (2.0 / Re) * v[2] * ca;
    out[6] = -(2.0 / Re) * v[3] * ca + (2.0 / Re) * v[2] * sa;
    out[7] = (2 * v[6]) / nu_den;
}

// ---------------------------------------------------------------- a12, a13
__global__ void k_dil_ref(int n0, int n1, double h1, double h2, const double *__restrict__ u,
                          const double *__restrict__ v, const double *__restrict__ xix,
                          const double *__restrict__ xiy, const double *__restrict__ etx,
                          const double *__restrict__ ety, double *__restrict__ dil, double *__restrict__ part) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    double d = 0.0;
    if (i <= n0 - 2 && j >= 1 && j <= n1 - 2) {
        const int inn = (i == 0) ? n0 - 2 : i - 1, ipp = i + 1;
        const size_t P = ix(i, j, n0);
        d = xix[P] * (u[ix(ipp, j, n0)] - u[ix(inn, j, n0)]) / (2 * h1) +
            etx[P] * (u[ix(i, j + 1, n0)] - u[ix(i, j - 1, n0)]) / (2 * h2) +
            ety[P] * (v[ix(i, j + 1, n0)] - v[ix(i, j - 1, n0)]) / (2 * h2) +
            xiy[P] * (v[ix(ipp, j, n0)] - v[ix(inn, j, n0)]) / (2 * h1);
        dil[P] = d;
        if (i == 0) dil[ix(n0 - 1, j, n0)] = d;
    }
    const double m = block_max<BX * BY>(d);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[blockIdx.x + blockIdx.y * gridDim.x] = m;
}

__global__ void k_max_final(const double *__restrict__ part, int n, double *__restrict__ out) {
    double m = 0.0;  // floor at 0 as solver.cpp:1806 starts dmax at 0
    for (int t = threadIdx.x; t < n; t += blockDim.x) m = fmax(m, part[t]);
    m = block_max<256>(m);
    if (threadIdx.x == 0) *out = m;
}

__global__ void k_nusselt_ref(int n0, double h1, double h2, const double *__restrict__ T,
                              const double *__restrict__ ajac, const double *__restrict__ etx,
                              const double *__restrict__ ety, double ar, double *__restrict__ out,
                              double *__restrict__ dthdn) {
    double acc = 0.0;
    const int j = 0;
    for (int i = threadIdx.x; i < n0 - 1; i += blockDim.x) {
        const int ip = i + 1;
        const size_t A = ix(i, j, n0), B = ix(ip, j, n0);
        const double tj1 = ajac[A] * (etx[A] * etx[A] + ety[A] * ety[A]);
        const double tj2 = ajac[B] * (etx[B] * etx[B] + ety[B] * ety[B]);
        const double fh1 = tj1 * (4 * T[ix(i, j + 1, n0)] - 3 * T[A] - T[ix(i, j + 2, n0)]) / (2 * h2);
        const double fh2 = tj2 * (4 * T[ix(ip, j + 1, n0)] - 3 * T[B] - T[ix(ip, j + 2, n0)]) / (2 * h2);
        acc += 0.5 * (fh1 + fh2) * h1;
    }
    for (int i = threadIdx.x; i < n0; i += blockDim.x) {
        const double d = -(4 * T[ix(i, 1, n0)] - 3 * T[ix(i, 0, n0)] - T[ix(i, 2, n0)]) / (2 * h2);
        dthdn[i] = d * sqrt(etx[ix(i, 0, n0)] * etx[ix(i, 0, n0)] + ety[ix(i, 0, n0)] * ety[ix(i, 0, n0)]);
    }
    acc = block_sum<256>(acc);
    if (threadIdx.x == 0) {
        const double Pi = acos(-1.0);
        *out = (2 * acc) / (Pi * (3 * (1 + (1.0 / ar)) - sqrt((3 + (1.0 / ar)) * ((3.0 / ar) + 1))));
    }
}

// ---------------------------------------------------------------- stored-operator BiCGStab
// Unknowns: columns 0..n0-2, rows 1..n1-2 of the periodic grid.  The operator is
// `npts` stored planes (9 or 25); rows 1 and n1-2 of a 25-point operator use only
// their 9-point planes, as Solver::gauss does (solver.cpp:2331-2353).

__device__ __forceinline__ bool is_unk(int i, int j, int n0, int n1) { return i <= n0 - 2 && j >= 1 && j <= n1 - 2; }

__global__ void k_invdiag(int n0, int n1, int center, const double *__restrict__ A, double *__restrict__ invd) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= n0 || j >= n1) return;
    const size_t P = ix(i, j, n0);
    invd[P] = is_unk(i, j, n0, n1) ? 1.0 / A[(size_t)center * n0 * n1 + P] : 0.0;
}
// duplicate column n0-1 := column 0 on the unknown rows
__global__ void k_dupcol(int n0, int n1, double *__restrict__ x) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= 1 && j <= n1 - 2) x[ix(n0 - 1, j, n0)] = x[ix(0, j, n0)];
}

}  // namespace ref
}  // namespace cfd

// ==========================================================================
using namespace cfd;
using namespace cfd::ref;

namespace {

// RAII bundle of device buffers for one stage call
struct DevBufs {
    std::vector<void *> p;
    ~DevBufs() {
        for (void *q : p) (void)hipFree(q);
    }
    double *get(size_t n) {
        void *q = nullptr;
        if (hipMalloc(&q, n * sizeof(double)) != hipSuccess) return nullptr;
        p.push_back(q);
        return static_cast<double *>(q);
    }
    double *up(const double *h, size_t n) {
        double *d = get(n);
        if (d && h && hipMemcpy(d, h, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        if (d && !h && hipMemset(d, 0, n * sizeof(double)) != hipSuccess) return nullptr;
        return d;
    }
};

#define NEED(x)                          \
    do {                                 \
        if (!(x)) return CFD_E_OOM;      \
    } while (0)

int down(double *h, const double *d, size_t n) {
    if (!h) return CFD_OK;
    CFD_HIP(hipMemcpy(h, d, n * sizeof(double), hipMemcpyDeviceToHost));
    return CFD_OK;
}

bool bad_grid(int n0, int n1) { return n0 < 5 || n1 < 5; }

// Work vectors of the stored-operator BiCGStab on n0 x n1 nodes (device): nine
// node vectors, two rows of block partials and the solver's scalars.
struct SopWork {
    double *r = nullptr, *r0 = nullptr, *pv = nullptr, *vv = nullptr, *s = nullptr, *t = nullptr, *ph = nullptr,
           *sh = nullptr, *invd = nullptr, *part = nullptr, *sc = nullptr;
    int np = 0;  // partials per row
    // the largest number of blocks of a reduction phase: 2-D tiles, 256-node chunks,
    // or columns (the line kernels)
    static int parts(int n0, int n1) {
        const dim3 gr = grid2d(n0, n1);
        const long n1d = ((long)n0 * n1 + 255) / 256;
        return (int)std::max<long>(std::max<long>((long)gr.x * gr.y, n1d), (long)n0);
    }
    static size_t doubles(int n0, int n1) { return 9 * (size_t)n0 * n1 + 2 * (size_t)parts(n0, n1) + 16; }
    static SopWork carve(double *base, int n0, int n1) {
        const size_t N = (size_t)n0 * n1;
        SopWork w;
        double **v[9] = {&w.r, &w.r0, &w.pv, &w.vv, &w.s, &w.t, &w.ph, &w.sh, &w.invd};
        for (int k = 0; k < 9; k++) *v[k] = base + (size_t)k * N;
        w.np = parts(n0, n1);
        w.part = base + 9 * N;
        w.sc = w.part + 2 * (size_t)w.np;
        return w;
    }
};

// the solver's device scalars
enum { Q_RHO, Q_RHO_OLD, Q_ALPHA, Q_OMEGA, Q_BETA, Q_TOL, Q_R0N, Q_RN, Q_SCONV, Q_DONE, Q_NIT, Q_NSC };

// block partials of up to two dot products, row k at part[k * np]
template <int NV, int NT>
__device__ __forceinline__ void sop_partials(const double (&v)[NV], double *part, int np, int b) {
    double t[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) t[k] = v[k];
    block_sums<NV, NT>(t);
    if (threadIdx.x + threadIdx.y * blockDim.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) part[(size_t)k * np + b] = t[k];
    }
}

// y = A x on the unknowns (0 elsewhere) for a 9-point (R = 1) or 25-point (R = 2,
// 9-point on rows 1 and n1-2) operator, columns wrapping with period n0-1
template <int R>
__device__ __forceinline__ double sop_row(int n0, int n1, const double *__restrict__ A, const double *__restrict__ x,
                                          int i, int j) {
    const size_t N = (size_t)n0 * n1, P = ix(i, j, n0);
    const int r = (R == 2 && (j == 1 || j == n1 - 2)) ? 1 : R;
    double s = 0.0;
    for (int dj = -r; dj <= r; dj++)
        for (int di = -r; di <= r; di++) {
            const int k = (R == 2) ? (dj + 2) * 5 + (di + 2) : (dj + 1) * 3 + (di + 1);
            int ii = i + di;
            if (ii < 0) ii += n0 - 1;
            if (ii > n0 - 2) ii -= n0 - 1;
            s += A[(size_t)k * N + P] * x[ix(ii, j + dj, n0)];
        }
    return s;
}

// start: r = q - A x (unknowns), r0 = r, partials of (r, r)
template <int R>
__global__ void k_q_start(int n0, int n1, const double *__restrict__ A, const double *__restrict__ x,
                          const double *__restrict__ q, double *__restrict__ r, double *__restrict__ r0,
                          double *__restrict__ part, int np) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    double v[1] = {0.0};
    if (i < n0 && j < n1) {
        const size_t P = ix(i, j, n0);
        const double rv = is_unk(i, j, n0, n1) ? q[P] - sop_row<R>(n0, n1, A, x, i, j) : 0.0;
        r[P] = rv;
        r0[P] = rv;
        v[0] = rv * rv;
    }
    sop_partials<1, BX * BY>(v, part, np, blockIdx.x + blockIdx.y * gridDim.x);
}

// p = r (first iteration) or r + beta (p - omega v); ph = p / diag
__global__ void k_q_p(size_t n, int first, const double *__restrict__ sc, const double *__restrict__ r,
                      const double *__restrict__ vv, const double *__restrict__ invd, double *__restrict__ pv,
                      double *__restrict__ ph) {
    if (sc[Q_DONE] != 0.0) return;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double p = first ? r[k] : r[k] + sc[Q_BETA] * (pv[k] - sc[Q_OMEGA] * vv[k]);
    pv[k] = p;
    ph[k] = p * invd[k];
}

// y = A x and partials of (w, y); x is ph (v) or sh (t, with (t, s) and (t, t))
template <int R, int NV>
__global__ void k_q_apply(int n0, int n1, const double *__restrict__ sc, const double *__restrict__ A,
                          const double *__restrict__ x, const double *__restrict__ w, double *__restrict__ y,
                          double *__restrict__ part, int np) {
    if (sc[Q_DONE] != 0.0 || (NV == 2 && sc[Q_SCONV] != 0.0)) return;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = 0.0;
    if (i < n0 && j < n1) {
        const size_t P = ix(i, j, n0);
        const double yv = is_unk(i, j, n0, n1) ? sop_row<R>(n0, n1, A, x, i, j) : 0.0;
        y[P] = yv;
        v[0] = w[P] * yv;
        if (NV == 2) v[NV - 1] = yv * yv;
    }
    sop_partials<NV, BX * BY>(v, part, np, blockIdx.x + blockIdx.y * gridDim.x);
}

// s = r - alpha v, sh = s / diag, partials of (s, s)
__global__ void k_q_s(size_t n, const double *__restrict__ sc, const double *__restrict__ r,
                      const double *__restrict__ vv, const double *__restrict__ invd, double *__restrict__ s,
                      double *__restrict__ sh, double *__restrict__ part, int np) {
    if (sc[Q_DONE] != 0.0) return;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (k < n) {
        const double sv = r[k] - sc[Q_ALPHA] * vv[k];
        s[k] = sv;
        sh[k] = sv * invd[k];
        v[0] = sv * sv;
    }
    sop_partials<1, 256>(v, part, np, blockIdx.x);
}

// ---- j-line preconditioner of the 9-point operator: M is the operator's (south,
// centre, north) part, one tridiagonal system per column over the unknown rows
// 1..n1-2 (the radial lines of the O-grid, along which the stretched cells couple
// most strongly), solved exactly by parallel cyclic reduction in one workgroup per
// column: log2(rows) steps, each folding the rows at distance 2^s into a row's
// equation.  Measured on the 350 x 570 ellipse O-grid (Re 1000 step), BiCGStab to
// 0.75e-2 takes 60 iterations with it against 144 with the diagonal alone.
constexpr int LN_MAX = 1024;  // unknown rows per column the line kernels take

// z = M^-1 y on the column of this block (row k = threadIdx.x, node j = k + 1);
// every thread calls it, y valid for k < m
template <int NT>
__device__ __forceinline__ double pcr_column(int n0, int n1, const double *__restrict__ A, int i, double y) {
    __shared__ double sa[2][NT], sb[2][NT], sc_[2][NT], sd[2][NT];
    const int k = threadIdx.x, m = n1 - 2;
    const size_t N = (size_t)n0 * n1;
    double a = 0.0, b = 1.0, c = 0.0, d = 0.0;
    if (k < m) {
        const size_t P = (size_t)(k + 1) * n0 + i;
        a = k > 0 ? A[1 * N + P] : 0.0;       // south neighbour (row 0 is fixed)
        b = A[4 * N + P];
        c = k < m - 1 ? A[7 * N + P] : 0.0;   // north neighbour (row n1-1 is fixed)
        d = y;
    }
    int cur = 0;
    sa[0][k] = a; sb[0][k] = b; sc_[0][k] = c; sd[0][k] = d;
    __syncthreads();
    for (int st = 1; st < m; st <<= 1) {
        if (k < m) {
            double al = 0.0, ga = 0.0, na = 0.0, nc = 0.0, nb = b, ndd = d;
            if (k - st >= 0) {
                al = -a / sb[cur][k - st];
                na = al * sa[cur][k - st];
                nb += al * sc_[cur][k - st];
                ndd += al * sd[cur][k - st];
            }
            if (k + st < m) {
                ga = -c / sb[cur][k + st];
                nc = ga * sc_[cur][k + st];
                nb += ga * sa[cur][k + st];
                ndd += ga * sd[cur][k + st];
            }
            a = na; b = nb; c = nc; d = ndd;
            sa[cur ^ 1][k] = a; sb[cur ^ 1][k] = b; sc_[cur ^ 1][k] = c; sd[cur ^ 1][k] = d;
        }
        cur ^= 1;
        __syncthreads();
    }
    return k < m ? d / b : 0.0;
}

// p = r (first iteration) or r + beta (p - omega v); ph = M^-1 p, per column
template <int NT>
__global__ __launch_bounds__(NT) void k_q_pline(int n0, int n1, int first, const double *__restrict__ sc,
                                                const double *__restrict__ A, const double *__restrict__ r,
                                                const double *__restrict__ vv, double *__restrict__ pv,
                                                double *__restrict__ ph) {
    if (sc[Q_DONE] != 0.0) return;
    const int i = blockIdx.x, k = threadIdx.x, m = n1 - 2;
    const size_t P = (size_t)(k + 1) * n0 + i;
    double p = 0.0;
    if (k < m) {
        p = first ? r[P] : r[P] + sc[Q_BETA] * (pv[P] - sc[Q_OMEGA] * vv[P]);
        pv[P] = p;
    }
    const double z = pcr_column<NT>(n0, n1, A, i, p);
    if (k < m) ph[P] = z;
}

// s = r - alpha v, sh = M^-1 s per column, partials of (s, s) (one per column)
template <int NT>
__global__ __launch_bounds__(NT) void k_q_sline(int n0, int n1, const double *__restrict__ sc,
                                                const double *__restrict__ A, const double *__restrict__ r,
                                                const double *__restrict__ vv, double *__restrict__ s,
                                                double *__restrict__ sh, double *__restrict__ part, int np) {
    if (sc[Q_DONE] != 0.0) return;
    const int i = blockIdx.x, k = threadIdx.x, m = n1 - 2;
    const size_t P = (size_t)(k + 1) * n0 + i;
    double sv = 0.0;
    if (k < m) {
        sv = r[P] - sc[Q_ALPHA] * vv[P];
        s[P] = sv;
    }
    const double z = pcr_column<NT>(n0, n1, A, i, sv);
    if (k < m) sh[P] = z;
    double v[1] = {sv * sv};
    block_sums<1, NT>(v);
    if (k == 0) part[blockIdx.x] = v[0];
}

// x += alpha ph + omega sh (omega taken as 0 after the ||s|| test stopped the
// iteration); r = s - omega t with partials of (r, r) and (r0, r)
__global__ void k_q_xr(size_t n, const double *__restrict__ sc, const double *__restrict__ ph,
                       const double *__restrict__ sh, const double *__restrict__ s, const double *__restrict__ t,
                       const double *__restrict__ r0, double *__restrict__ x, double *__restrict__ r,
                       double *__restrict__ part, int np) {
    if (sc[Q_DONE] != 0.0) return;
    const bool sconv = sc[Q_SCONV] != 0.0;
    const double a = sc[Q_ALPHA], om = sconv ? 0.0 : sc[Q_OMEGA];
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (k < n) {
        x[k] += a * ph[k] + om * sh[k];
        if (!sconv) {
            const double rv = s[k] - om * t[k];
            r[k] = rv;
            v[0] = rv * rv;
            v[1] = r0[k] * rv;
        }
    }
    sop_partials<2, 256>(v, part, np, blockIdx.x);
}

enum { QF_START, QF_V, QF_S, QF_T, QF_XR };
// totals of a phase's partials (one block) and the BiCGStab scalars that follow
// flag (may be null): done, ||r||, ||r0|| and the iteration that ended the solve,
// copied to device memory; hf (may be null): the end-of-phase flag (seq << 1 | done)
// stored into host-mapped memory; it: the iteration this phase belongs to
template <int OP>
__global__ void k_q_fin(const double *__restrict__ part, int np, int nparts, double *__restrict__ sc, double rtol,
                        double *__restrict__ flag, unsigned long long *hf = nullptr, unsigned long long seq = 0,
                        int it = 0) {
    constexpr int NV = (OP == QF_T || OP == QF_XR) ? 2 : 1;
    double tot[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double a = 0.0;
        for (int b = threadIdx.x; b < nparts; b += 256) a += part[(size_t)k * np + b];
        tot[k] = a;
    }
    block_sums<NV, 256>(tot);
    if (threadIdx.x != 0) return;
    if (OP == QF_START) {
        const double rn = sqrt(tot[0]);
        sc[Q_RHO] = tot[0];
        sc[Q_RHO_OLD] = 1.0;
        sc[Q_ALPHA] = 1.0;
        sc[Q_OMEGA] = 1.0;
        sc[Q_BETA] = 0.0;
        sc[Q_R0N] = rn;
        sc[Q_RN] = rn;
        sc[Q_TOL] = rtol * rn;
        sc[Q_SCONV] = 0.0;
        sc[Q_DONE] = rn > 0.0 ? 0.0 : 1.0;
        sc[Q_NIT] = 0.0;
    } else if (sc[Q_DONE] != 0.0) {
        // nothing ran in this phase; the flag below still reports the state
    } else if (OP == QF_V) {
        sc[Q_ALPHA] = sc[Q_RHO] / tot[0];
    } else if (OP == QF_S) {
        const double sn = sqrt(tot[0]);
        if (sn <= sc[Q_TOL]) {
            sc[Q_SCONV] = 1.0;
            sc[Q_RN] = sn;
        }
    } else if (OP == QF_T) {
        if (sc[Q_SCONV] == 0.0) sc[Q_OMEGA] = tot[1] > 0.0 ? tot[0] / tot[1] : 0.0;
    } else {
        if (sc[Q_SCONV] != 0.0) {
            sc[Q_DONE] = 1.0;
            sc[Q_NIT] = it;
        } else {
            const double rn = sqrt(tot[0]);
            sc[Q_RN] = rn;
            if (rn <= sc[Q_TOL] || sc[Q_OMEGA] == 0.0 || tot[1] == 0.0) {
                sc[Q_DONE] = 1.0;
                sc[Q_NIT] = it;
            } else {
                sc[Q_RHO_OLD] = sc[Q_RHO];
                sc[Q_RHO] = tot[1];
                sc[Q_BETA] = (tot[1] / sc[Q_RHO_OLD]) * (sc[Q_ALPHA] / sc[Q_OMEGA]);
            }
        }
    }
    if (flag) {
        flag[0] = sc[Q_DONE];
        flag[1] = sc[Q_RN];
        flag[2] = sc[Q_R0N];
        flag[3] = sc[Q_NIT];
    }
    if (hf)
        __hip_atomic_store(hf, (seq << 1) | (sc[Q_DONE] != 0.0 ? 1ull : 0ull), __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
}

// A host-mapped end-of-iteration flag of the stored-operator BiCGStab: host and
// device views of one pinned, coherent word, and its sequence counter.
struct SopFlag {
    unsigned long long *h = nullptr, *d = nullptr;
    unsigned long long seq = 0;
};

// wait until the flag carries sequence number sq or a later one; done: its done bit.
// Gives up with an error if the stream drains without it.
int sop_wait(const SopFlag &f, unsigned long long sq, hipStream_t st, bool &done) {
    for (long spin = 0;; spin++) {
        const unsigned long long v = __atomic_load_n(f.h, __ATOMIC_ACQUIRE);
        if ((v >> 1) >= sq) {  // this phase's flag or a later one (done stays set)
            done = (v & 1) != 0;
            return CFD_OK;
        }
        if ((spin & 255) == 255) {
            const hipError_t e = hipStreamQuery(st);
            if (e == hipSuccess) {
                const unsigned long long w = __atomic_load_n(f.h, __ATOMIC_ACQUIRE);
                if ((w >> 1) < sq) return CFD_E_HIP;
            } else if (e != hipErrorNotReady) {
                return CFD_E_HIP;
            }
        }
        __builtin_ia32_pause();
    }
}

// GPU BiCGStab on a stored 9/25-point operator A (device planes), x (device,
// boundary rows fixed) updated in place, right-hand side qd (device), on stream st.
// Preconditioner: the j-line solve (9-point operators of at most LN_MAX unknown rows
// per column), else the diagonal.  The recurrence's scalars stay on the device.
// With a mapped flag (fl), iteration k + 1 is queued before the host waits for the
// flag that ends iteration k, so the device never waits for the host and a converged
// solve runs one iteration of no-op launches; without one, the host copies the flag
// (into hflag: 4 doubles) and synchronises the stream once per iteration.
template <int R>
int sop_core(int n0, int n1, const double *A, double *x, const double *qd, const SopWork &w, double rtol, int maxit,
             int *iters, double *relres, hipStream_t st, double *hflag, SopFlag *fl = nullptr) {
    const size_t N = (size_t)n0 * n1;
    const dim3 gr = grid2d(n0, n1), bl(BX, BY);
    const int nb2 = (int)(gr.x * gr.y), lb = 256;
    const int nb1 = (int)((N + lb - 1) / lb);
    const int center = (R == 2) ? 12 : 4;
    const int m0 = n0 - 1, m1 = n1 - 2;
    const bool line = R == 1 && m1 <= LN_MAX;
    const int lnt = m1 <= 256 ? 256 : (m1 <= 512 ? 512 : 1024);
    // the line kernels write the unknowns only: the work vectors start at zero
    CFD_HIP(hipMemsetAsync(w.r, 0, 8 * N * sizeof(double), st));
    k_invdiag<<<gr, bl, 0, st>>>(n0, n1, center, A, w.invd);
    k_q_start<R><<<gr, bl, 0, st>>>(n0, n1, A, x, qd, w.r, w.r0, w.part, w.np);
    k_q_fin<QF_START><<<1, 256, 0, st>>>(w.part, w.np, nb2, w.sc, rtol, w.sc + Q_NSC);
    CFD_LAUNCH_CHECK();
    auto iteration = [&](int it, unsigned long long sq) -> int {
        if (line) {
            if (lnt == 256) k_q_pline<256><<<m0, 256, 0, st>>>(n0, n1, it == 1, w.sc, A, w.r, w.vv, w.pv, w.ph);
            else if (lnt == 512) k_q_pline<512><<<m0, 512, 0, st>>>(n0, n1, it == 1, w.sc, A, w.r, w.vv, w.pv, w.ph);
            else k_q_pline<1024><<<m0, 1024, 0, st>>>(n0, n1, it == 1, w.sc, A, w.r, w.vv, w.pv, w.ph);
        } else {
            k_q_p<<<nb1, lb, 0, st>>>(N, it == 1, w.sc, w.r, w.vv, w.invd, w.pv, w.ph);
        }
        k_q_apply<R, 1><<<gr, bl, 0, st>>>(n0, n1, w.sc, A, w.ph, w.r0, w.vv, w.part, w.np);
        k_q_fin<QF_V><<<1, 256, 0, st>>>(w.part, w.np, nb2, w.sc, rtol, nullptr);
        if (line) {
            if (lnt == 256) k_q_sline<256><<<m0, 256, 0, st>>>(n0, n1, w.sc, A, w.r, w.vv, w.s, w.sh, w.part, w.np);
            else if (lnt == 512)
                k_q_sline<512><<<m0, 512, 0, st>>>(n0, n1, w.sc, A, w.r, w.vv, w.s, w.sh, w.part, w.np);
            else k_q_sline<1024><<<m0, 1024, 0, st>>>(n0, n1, w.sc, A, w.r, w.vv, w.s, w.sh, w.part, w.np);
            k_q_fin<QF_S><<<1, 256, 0, st>>>(w.part, w.np, m0, w.sc, rtol, nullptr);
        } else {
            k_q_s<<<nb1, lb, 0, st>>>(N, w.sc, w.r, w.vv, w.invd, w.s, w.sh, w.part, w.np);
            k_q_fin<QF_S><<<1, 256, 0, st>>>(w.part, w.np, nb1, w.sc, rtol, nullptr);
        }
        k_q_apply<R, 2><<<gr, bl, 0, st>>>(n0, n1, w.sc, A, w.sh, w.s, w.t, w.part, w.np);
        k_q_fin<QF_T><<<1, 256, 0, st>>>(w.part, w.np, nb2, w.sc, rtol, nullptr);
        k_q_xr<<<nb1, lb, 0, st>>>(N, w.sc, w.ph, w.sh, w.s, w.t, w.r0, x, w.r, w.part, w.np);
        k_q_fin<QF_XR><<<1, 256, 0, st>>>(w.part, w.np, nb1, w.sc, rtol, w.sc + Q_NSC, fl ? fl->d : nullptr, sq,
                                          it);
        CFD_LAUNCH_CHECK();
        return CFD_OK;
    };
    int it = 0;
    if (fl) {
        // the start's done state: no iteration when the initial residual is zero
        CFD_HIP(hipMemcpyAsync(hflag, w.sc + Q_NSC, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
        CFD_HIP(hipStreamSynchronize(st));
        if (hflag[0] == 0.0 && maxit > 0) {
            const unsigned long long base = fl->seq;
            fl->seq += (unsigned long long)maxit + 1;
            CFD_RET(iteration(1, base + 1));
            for (it = 1; it <= maxit; it++) {
                if (it < maxit) CFD_RET(iteration(it + 1, base + it + 1));
                bool done = false;
                CFD_RET(sop_wait(*fl, base + it, st, done));
                if (done) break;
            }
            if (it > maxit) it = maxit;
        }
        CFD_HIP(hipMemcpyAsync(hflag, w.sc + Q_NSC, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
        CFD_HIP(hipStreamSynchronize(st));
        if (hflag[0] != 0.0) it = (int)hflag[3];  // the host may have seen a later flag first
    } else {
        CFD_HIP(hipMemcpyAsync(hflag, w.sc + Q_NSC, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
        CFD_HIP(hipStreamSynchronize(st));
        for (it = 1; it <= maxit && hflag[0] == 0.0; it++) {  // none when r0 = 0
            CFD_RET(iteration(it, 0));
            CFD_HIP(hipMemcpyAsync(hflag, w.sc + Q_NSC, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
            CFD_HIP(hipStreamSynchronize(st));
            if (hflag[0] != 0.0) break;
        }
        if (it > maxit) it = maxit;
    }
    if (hflag[2] == 0.0) it = 0;
    k_dupcol<<<(n1 + 255) / 256, 256, 0, st>>>(n0, n1, x);
    CFD_LAUNCH_CHECK();
    if (iters) *iters = it;
    if (relres) *relres = hflag[2] > 0.0 ? hflag[1] / hflag[2] : 0.0;
    return CFD_OK;
}

// a mapped flag word (pinned, coherent host memory with its device address); false
// when it cannot be had (the solve then synchronises once per iteration)
bool flag_alloc(SopFlag &f) {
    void *h = nullptr, *d = nullptr;
    if (hipHostMalloc(&h, sizeof(unsigned long long), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipHostFree(h);
        return false;
    }
    f.h = static_cast<unsigned long long *>(h);
    f.d = static_cast<unsigned long long *>(d);
    *f.h = 0;
    f.seq = 0;
    return true;
}
void flag_free(SopFlag &f) {
    if (f.h) (void)hipHostFree(f.h);
    f = SopFlag{};
}

// The host-array form: uploads the planes, phi and q, solves, downloads phi.
template <int R>
int sop_solve(int n0, int n1, const double *const *planes_host, int npts, double *phi, const double *q,
              double rtol, int maxit, int *iters, double *relres) {
    if (bad_grid(n0, n1) || !phi || !q) return CFD_E_ARG;
    for (int k = 0; k < npts; k++)
        if (!planes_host[k]) return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    double *A = B.get((size_t)npts * N);
    NEED(A);
    for (int k = 0; k < npts; k++)
        CFD_HIP(hipMemcpy(A + (size_t)k * N, planes_host[k], N * sizeof(double), hipMemcpyHostToDevice));
    double *x = B.up(phi, N), *qd = B.up(q, N), *ws = B.get(SopWork::doubles(n0, n1));
    NEED(x && qd && ws);
    double hflag[4] = {0.0, 0.0, 0.0, 0.0};
    SopFlag fl;
    const bool have = flag_alloc(fl);
    const int rc = sop_core<R>(n0, n1, A, x, qd, SopWork::carve(ws, n0, n1), rtol, maxit, iters, relres, nullptr,
                               hflag, have ? &fl : nullptr);
    const hipError_t e = hipDeviceSynchronize();
    flag_free(fl);
    CFD_RET(rc);
    CFD_HIP(e);
    return down(phi, x, N);
}

// ---- the p' solve that allocates nothing: every device buffer of the solve (the
// nine operator planes, phi, q and the BiCGStab work) lives in one file-scope
// array, sized for the reference's own array bounds, np1 x np2 = 350 x 570 nodes
// (solver.cpp:13-14).  Each device holds its own copy, at its own address, found
// once per device with hipGetSymbolAddress.  Calls that overlap on one device
// share the array and race.
constexpr size_t NA_MAX_NODES = (size_t)CFD_SIP9P_NOALLOC_MAX_NODES;
// 9 planes, x, q and SopWork's nine vectors; partials for at most NA_MAX_NODES / 4
// + 1024 tiles per row (a 5-wide grid has one tile per four rows); scalars, flag
constexpr size_t NA_DOUBLES = 20 * NA_MAX_NODES + 2 * (NA_MAX_NODES / 4 + 1024) + 64;
static __device__ __align__(256) double scratch[NA_DOUBLES];

std::mutex na_mu;
double *na_ptr[64] = {nullptr};  // scratch's address on each device

int na_scratch(double **out) {
    int dev = 0;
    CFD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return CFD_E_ARG;
    std::lock_guard<std::mutex> lk(na_mu);
    if (!na_ptr[dev]) {
        void *p = nullptr;
        CFD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        na_ptr[dev] = static_cast<double *>(p);
    }
    *out = na_ptr[dev];
    return CFD_OK;
}

// sop_solve with the buffers carved from scratch
int sop_solve_noalloc(int n0, int n1, const double *const *planes_host, double *phi, const double *q, double rtol,
                      int maxit, int *iters, double *relres) {
    if (bad_grid(n0, n1) || !phi || !q) return CFD_E_ARG;
    for (int k = 0; k < 9; k++)
        if (!planes_host[k]) return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    if (N > NA_MAX_NODES || 11 * N + SopWork::doubles(n0, n1) > NA_DOUBLES) {
        fprintf(stderr, "cfd_sip9p_noalloc: %d x %d nodes exceed its %zu-node scratch\n", n0, n1, NA_MAX_NODES);
        return CFD_E_ARG;
    }
    double *base = nullptr;
    CFD_RET(na_scratch(&base));
    double *A = base, *x = base + 9 * N, *qd = base + 10 * N, *ws = base + 11 * N;
    for (int k = 0; k < 9; k++)
        CFD_HIP(hipMemcpy(A + (size_t)k * N, planes_host[k], N * sizeof(double), hipMemcpyHostToDevice));
    CFD_HIP(hipMemcpy(x, phi, N * sizeof(double), hipMemcpyHostToDevice));
    CFD_HIP(hipMemcpy(qd, q, N * sizeof(double), hipMemcpyHostToDevice));
    double hflag[4] = {0.0, 0.0, 0.0, 0.0};
    CFD_RET(sop_core<1>(n0, n1, A, x, qd, SopWork::carve(ws, n0, n1), rtol, maxit, iters, relres, nullptr, hflag));
    CFD_HIP(hipDeviceSynchronize());
    return down(phi, x, N);
}

// sin and cos of the flow angle and the ellipse-perimeter factor of the Nusselt
// number, evaluated on the host as the reference does (solver.cpp:1927-1982)
struct ForceConst {
    double sa, ca, nu_den;
};
ForceConst force_const(double alpha_deg, double ar) {
    const double Pi = acos(-1.0);
    return ForceConst{sin(alpha_deg * Pi / 180.0), cos(alpha_deg * Pi / 180.0),
                      Pi * (3 * (1 + (1.0 / ar)) - sqrt((3 + (1.0 / ar)) * ((3.0 / ar) + 1)))};
}

// ---------------------------------------------------------------- device-resident time steps
// Local accelerations of the wall row (solver.cpp:1715-1718): a_s = accn_amp sin(2 pi F t).
__global__ void k_alc(int n0, const double *__restrict__ x, const double *__restrict__ y, double a_s,
                      double *__restrict__ alc_u, double *__restrict__ alc_v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n0) return;
    alc_u[i] = a_s * y[i];
    alc_v[i] = -a_s * x[i];
}

// One reference O-grid case resident in HBM: grid, assembled operators, state and
// every stage array of solver.cpp's time step, in one allocation.
struct RefCase {
    int n0 = 0, n1 = 0;
    double dxi = 0, deta = 0;
    cfd_ref_params prm{};
    hipStream_t s = nullptr;
    double *mem = nullptr;
    double *x, *y, *xix, *xiy, *etx, *ety, *alph, *beta, *gamma, *ajac, *p1, *q1, *xnox, *xnoy;
    double *AU, *AT, *BU, *BT, *A9;
    double *u, *v, *T, *p, *up, *vp, *us, *vs, *uold, *vold, *Told, *uxi, *uet, *qu, *qv, *qt, *qup, *qvp, *q,
        *pcor, *dil, *sol, *si, *vort, *alc, *part, *scal, *ws;
    // pinned: circulation, dmax, the eight force coefficients (0-9); the inner
    // solver's flags (12-14)
    double *hscal = nullptr;
    SopFlag flag;  // the inner solver's mapped end-of-iteration flag (null: synchronising)
    bool have_state = false;
};

size_t ref_case_doubles(int n0, int n1) {
    const size_t N = (size_t)n0 * n1;
    const dim3 gr = grid2d(n0, n1);
    return 12 * N + 2 * (size_t)n0 + 59 * N + 12 * (size_t)n0 + 24 * N + 2 * (size_t)n0 + (size_t)gr.x * gr.y + 16 +
           SopWork::doubles(n0, n1) + 64;
}

void ref_case_carve(RefCase *c) {
    const size_t N = (size_t)c->n0 * c->n1, m = (size_t)c->n0;
    double *b = c->mem;
    auto take = [&](size_t n) {
        double *r = b;
        b += n;
        return r;
    };
    double **grid[12] = {&c->x, &c->y, &c->xix, &c->xiy, &c->etx, &c->ety, &c->alph, &c->beta, &c->gamma, &c->ajac,
                         &c->p1, &c->q1};
    for (double **g : grid) *g = take(N);
    c->xnox = take(m);
    c->xnoy = take(m);
    c->AU = take(25 * N);
    c->AT = take(25 * N);
    c->A9 = take(9 * N);
    c->BU = take(6 * m);
    c->BT = take(6 * m);
    double **st[24] = {&c->u,   &c->v,   &c->T,  &c->p,  &c->up, &c->vp,  &c->us,  &c->vs,
                       &c->uold, &c->vold, &c->Told, &c->uxi, &c->uet, &c->qu, &c->qv, &c->qt,
                       &c->qup, &c->qvp, &c->q,  &c->pcor, &c->dil, &c->sol, &c->si, &c->vort};
    for (double **f : st) *f = take(N);
    c->alc = take(2 * m);
    const dim3 gr = grid2d(c->n0, c->n1);
    c->part = take((size_t)gr.x * gr.y);
    c->scal = take(16);
    c->ws = take(SopWork::doubles(c->n0, c->n1));
}

// sol <- its initial value, solve A sol = rhs, then rows 1..n1-2 of sol into dst
// (column n0-1 already duplicates column 0)
template <int R>
int ref_solve_rows(RefCase *c, const double *A, const double *rhs, double *dst, double rtol, int maxit, int *it,
                   double *rr) {
    const size_t N = (size_t)c->n0 * c->n1;
    CFD_RET(sop_core<R>(c->n0, c->n1, A, c->sol, rhs, SopWork::carve(c->ws, c->n0, c->n1), rtol, maxit, it, rr,
                        c->s, c->hscal + 12, c->flag.h ? &c->flag : nullptr));
    CFD_HIP(hipMemcpyAsync(dst + c->n0, c->sol + c->n0, (N - 2 * (size_t)c->n0) * sizeof(double),
                           hipMemcpyDeviceToDevice, c->s));
    return CFD_OK;
}

// One time step: solver.cpp:1057-1852 in order, the Gauss-Seidel and SIP solves
// replaced by the BiCGStab above at the case's tolerances.
int ref_case_step(RefCase *c, cfd_ref_stats *st) {
    const int n0 = c->n0, n1 = c->n1;
    const size_t N = (size_t)n0 * n1, B = N * sizeof(double);
    const dim3 gr = grid2d(n0, n1), bl(BX, BY), g1((n0 + 255) / 256);
    const int nb = (int)(gr.x * gr.y);
    const cfd_ref_params &P = c->prm;
    hipStream_t s = c->s;
    cfd_ref_stats o{};
    if (P.time_accurate) c->prm.time += P.dt;  // 1058, with time a double
    // 1062-1068: contravariant velocities and old T; 1074-1236: right-hand sides
    k_contra<<<gr, bl, 0, s>>>(n0, n1, c->u, c->v, c->xix, c->xiy, c->etx, c->ety, c->uxi, c->uet);
    CFD_HIP(hipMemcpyAsync(c->Told, c->T, B, hipMemcpyDeviceToDevice, s));
    const RhsIn ri{c->u, c->v, c->T, c->up, c->vp, c->p, c->xix, c->xiy, c->etx, c->ety,
                   c->alph, c->gamma, c->BU, c->BT, c->uxi, c->uet};
    const RhsOut ro{c->qu, c->qv, c->qt, c->qup, c->qvp};
    k_rhs<<<gr, bl, 0, s>>>(n0, n1, c->dxi, c->deta, ri, ro, P.Re, P.Pr, P.Ri, P.dt);
    CFD_LAUNCH_CHECK();
    // 1239-1350: u, v and T from their current values, the predictors from zero
    // with their wall row
    const double mt = P.mom_rtol;
    const int mi = P.mom_maxit;
    CFD_HIP(hipMemcpyAsync(c->sol, c->u, B, hipMemcpyDeviceToDevice, s));
    CFD_RET(ref_solve_rows<2>(c, c->AU, c->qu, c->us, mt, mi, &o.iters[0], &o.relres[0]));
    CFD_HIP(hipMemcpyAsync(c->sol, c->v, B, hipMemcpyDeviceToDevice, s));
    CFD_RET(ref_solve_rows<2>(c, c->AU, c->qv, c->vs, mt, mi, &o.iters[1], &o.relres[1]));
    CFD_HIP(hipMemcpyAsync(c->sol, c->T, B, hipMemcpyDeviceToDevice, s));
    CFD_RET(ref_solve_rows<2>(c, c->AT, c->qt, c->T, mt, mi, &o.iters[2], &o.relres[2]));
    double *pf[2] = {c->up, c->vp}, *pq[2] = {c->qup, c->qvp};
    for (int k = 0; k < 2; k++) {
        CFD_HIP(hipMemsetAsync(c->sol, 0, B, s));
        CFD_HIP(hipMemcpyAsync(c->sol, pf[k], (size_t)n0 * sizeof(double), hipMemcpyDeviceToDevice, s));
        CFD_RET(ref_solve_rows<2>(c, c->AU, pq[k], pf[k], mt, mi, &o.iters[3 + k], &o.relres[3 + k]));
    }
    // 1351-1381, 1389-1448
    k_outflow_up_ref<<<g1, 256, 0, s>>>(n0, n1, c->up, c->vp, c->u, c->v, c->xnox, c->xnoy, P.uinf, P.vinf);
    k_face_div_ref<<<gr, bl, 0, s>>>(n0, n1, c->dxi, c->deta, c->up, c->vp, c->p, c->xix, c->xiy, c->etx, c->ety,
                                     P.dt, c->q);
    CFD_LAUNCH_CHECK();
    // 1451-1498: p' from zero, its boundary rows
    CFD_HIP(hipMemsetAsync(c->pcor, 0, B, s));
    CFD_HIP(hipMemcpyAsync(c->uold, c->u, B, hipMemcpyDeviceToDevice, s));
    CFD_HIP(hipMemcpyAsync(c->vold, c->v, B, hipMemcpyDeviceToDevice, s));
    CFD_RET(sop_core<1>(n0, n1, c->A9, c->pcor, c->q, SopWork::carve(c->ws, n0, n1), P.p_rtol, P.p_maxit,
                        &o.iters[5], &o.relres[5], s, c->hscal + 12, c->flag.h ? &c->flag : nullptr));
    k_pcor_bc<<<g1, 256, 0, s>>>(n0, n1, c->pcor, c->xnox, c->xnoy, P.uinf, P.vinf);
    // 1504-1655: correction, far-field and wall velocities
    k_correct_ref<<<gr, bl, 0, s>>>(n0, n1, c->dxi, c->deta, c->us, c->vs, c->pcor, c->xix, c->xiy, c->etx, c->ety,
                                    P.dt, c->u, c->v, c->p);
    const double wt = 2.0 * acos(-1.0) * P.F * P.time;
    const FarRef fr{c->x, c->y, c->Told, c->uet, c->etx, c->ety, c->ajac, c->xnox, c->xnoy};
    k_farfield_ref<<<1, FF_NT, 0, s>>>(n0, n1, c->deta, fr, c->u, c->v, c->T, c->up, c->vp, P.uinf, P.vinf, P.dt,
                                       P.speed_amp, cos(wt), c->scal);
    CFD_LAUNCH_CHECK();
    // 1666-1782: wall and outer-row pressure
    const double a_s = P.accn_amp * sin(wt);
    const double *alc_u = nullptr, *alc_v = nullptr;
    if (a_s != 0.0) {
        k_alc<<<g1, 256, 0, s>>>(n0, c->x, c->y, a_s, c->alc, c->alc + n0);
        alc_u = c->alc;
        alc_v = c->alc + n0;
    }
    const WallRef w{c->u, c->v, c->T, c->xix, c->xiy, c->etx, c->ety, c->alph, c->beta, c->gamma, c->q1, c->ajac,
                    alc_u, alc_v};
    k_wall_p_ref<<<g1, 256, 0, s>>>(n0, n1, c->dxi, c->deta, w, c->p, P.Re, P.Ri);
    const ExitRef e{c->u, c->v, c->T, c->uold, c->vold, c->xix, c->xiy, c->etx,
                    c->ety, c->alph, c->beta, c->gamma, c->q1, c->ajac, c->xnox, c->xnoy};
    k_exit_p_ref<<<g1, 256, 0, s>>>(n0, n1, c->dxi, c->deta, e, c->p, P.uinf, P.vinf, P.dt, P.Re, P.Ri);
    // 1806-1852: dilation and its maximum
    k_dil_ref<<<gr, bl, 0, s>>>(n0, n1, c->dxi, c->deta, c->u, c->v, c->xix, c->xiy, c->etx, c->ety, c->dil, c->part);
    k_max_final<<<1, 256, 0, s>>>(c->part, nb, c->scal + 1);
    // 1784-1982: stream function, vorticity, force coefficients and Nusselt number
    k_si_ref<<<(n0 + SI_W - 1) / SI_W, dim3(SI_W, SI_W), 0, s>>>(n0, n1, c->deta, c->u, c->v, c->xix, c->xiy,
                                                               c->ajac, c->si);
    k_vort_ref<<<gr, bl, 0, s>>>(n0, n1, c->dxi, c->deta, c->u, c->v, c->xix, c->xiy, c->etx, c->ety, c->vort);
    const ForceConst fc = force_const(P.alpha_deg, P.ar);
    k_forces_ref<<<1, 256, 0, s>>>(n0, c->dxi, c->deta,
                                   ForceRef{c->x, c->y, c->p, c->vort, c->T, c->ajac, c->etx, c->ety}, P.Re, fc.sa,
                                   fc.ca, fc.nu_den, c->scal + 2);
    CFD_LAUNCH_CHECK();
    CFD_HIP(hipMemcpyAsync(c->hscal, c->scal, 10 * sizeof(double), hipMemcpyDeviceToHost, s));
    CFD_HIP(hipStreamSynchronize(s));
    o.circ = c->hscal[0];
    o.dmax = c->hscal[1];
    o.cl = c->hscal[2];
    o.cd = c->hscal[3];
    o.cm = c->hscal[4];
    o.cl_pr = c->hscal[5];
    o.cd_pr = c->hscal[6];
    o.cl_vor = c->hscal[7];
    o.cd_vor = c->hscal[8];
    o.nu = c->hscal[9];
    if (st) *st = o;
    if (!std::isfinite(o.dmax)) return CFD_E_NAN;
    return CFD_OK;
}

}  // namespace

extern "C" {

int cfd_ref_assemble_mom(int n0, int n1, double dxi, double deta, const double *alph, const double *beta,
                         const double *gamma, const double *p1, const double *q1, double dt, double Re, double Pr,
                         double *AU, double *AT, double *BU, double *BT) {
    if (bad_grid(n0, n1) || !alph || !beta || !gamma || !AU || !AT || !BU || !BT) return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    double *a = B.up(alph, N), *b = B.up(beta, N), *g = B.up(gamma, N), *pp = B.up(p1, N), *qq = B.up(q1, N);
    double *au = B.up(nullptr, 25 * N), *at = B.up(nullptr, 25 * N), *bu = B.up(nullptr, 6 * (size_t)n0),
           *bt = B.up(nullptr, 6 * (size_t)n0);
    NEED(a && b && g && pp && qq && au && at && bu && bt);
    k_asm_mom<<<grid2d(n0, n1), dim3(BX, BY)>>>(n0, n1, dxi, deta, a, b, g, pp, qq, dt, Re, Pr, au, at, bu, bt);
    CFD_LAUNCH_CHECK();
    CFD_RET(down(AU, au, 25 * N));
    CFD_RET(down(AT, at, 25 * N));
    CFD_RET(down(BU, bu, 6 * (size_t)n0));
    return down(BT, bt, 6 * (size_t)n0);
}

int cfd_ref_assemble_p9(int n0, int n1, double dxi, double deta, const double *xix, const double *xiy,
                        const double *etx, const double *ety, double *A9) {
    if (bad_grid(n0, n1) || !xix || !xiy || !etx || !ety || !A9) return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    double *a = B.up(xix, N), *b = B.up(xiy, N), *c = B.up(etx, N), *d = B.up(ety, N), *o = B.up(nullptr, 9 * N);
    NEED(a && b && c && d && o);
    k_asm_p9<<<grid2d(n0, n1), dim3(BX, BY)>>>(n0, n1, dxi, deta, a, b, c, d, o);
    CFD_LAUNCH_CHECK();
    return down(A9, o, 9 * N);
}

int cfd_ref_rhs(int n0, int n1, double dxi, double deta, const double *u, const double *v, const double *T,
                const double *up, const double *vp, const double *p, const double *xix, const double *xiy,
                const double *etx, const double *ety, const double *alph, const double *gamma, const double *BU,
                const double *BT, double Re, double Pr, double Ri, double dt, double *uxi, double *uet, double *qu,
                double *qv, double *qt, double *qup, double *qvp) {
    if (bad_grid(n0, n1) || !u || !v || !T || !up || !vp || !p || !xix || !xiy || !etx || !ety || !alph ||
        !gamma || !BU || !BT)
        return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    RhsIn a{B.up(u, N), B.up(v, N), B.up(T, N), B.up(up, N), B.up(vp, N), B.up(p, N), B.up(xix, N),
            B.up(xiy, N), B.up(etx, N), B.up(ety, N), B.up(alph, N), B.up(gamma, N), B.up(BU, 6 * (size_t)n0),
            B.up(BT, 6 * (size_t)n0), B.up(nullptr, N), B.up(nullptr, N)};
    RhsOut o{B.up(nullptr, N), B.up(nullptr, N), B.up(nullptr, N), B.up(nullptr, N), B.up(nullptr, N)};
    NEED(a.u && a.v && a.T && a.up && a.vp && a.p && a.xix && a.xiy && a.etx && a.ety && a.alph && a.gamma &&
         a.BU && a.BT && a.uxi && a.uet && o.qu && o.qv && o.qt && o.qup && o.qvp);
    const dim3 gr = grid2d(n0, n1), bl(BX, BY);
    k_contra<<<gr, bl>>>(n0, n1, a.u, a.v, a.xix, a.xiy, a.etx, a.ety, (double *)a.uxi, (double *)a.uet);
    k_rhs<<<gr, bl>>>(n0, n1, dxi, deta, a, o, Re, Pr, Ri, dt);
    CFD_LAUNCH_CHECK();
    CFD_RET(down(uxi, a.uxi, N));
    CFD_RET(down(uet, a.uet, N));
    CFD_RET(down(qu, o.qu, N));
    CFD_RET(down(qv, o.qv, N));
    CFD_RET(down(qt, o.qt, N));
    CFD_RET(down(qup, o.qup, N));
    return down(qvp, o.qvp, N);
}

int cfd_gauss(int n0, int n1, const double *ap, const double *ae, const double *as, const double *an,
              const double *aw, const double *ase, const double *asw, const double *ane, const double *anw,
              const double *ass, const double *assee, const double *assww, const double *asse, const double *assw,
              const double *asee, const double *asww, const double *ann, const double *annee,
              const double *annww, const double *anne, const double *annw, const double *anee, const double *anww,
              const double *aee, const double *aww, double *phi, const double *q, double rtol, int maxit,
              int *iters, double *relres) {
    // planes in (dj+2)*5 + (di+2) order
    const double *pl[25] = {assww, assw, ass, asse, assee,
                            asww, asw, as, ase, asee,
                            aww, aw, ap, ae, aee,
                            anww, anw, an, ane, anee,
                            annww, annw, ann, anne, annee};
    return sop_solve<2>(n0, n1, pl, 25, phi, q, rtol, maxit, iters, relres);
}

int cfd_sip9p(int n0, int n1, const double *ap, const double *ae, const double *as, const double *an,
              const double *aw, const double *ase, const double *asw, const double *ane, const double *anw,
              double *phi, const double *q, double rtol, int maxit, int *iters, double *relres) {
    const double *pl[9] = {asw, as, ase, aw, ap, ae, anw, an, ane};
    return sop_solve<1>(n0, n1, pl, 9, phi, q, rtol, maxit, iters, relres);
}

int cfd_sip9p_noalloc(int n0, int n1, const double *ap, const double *ae, const double *as, const double *an,
                      const double *aw, const double *ase, const double *asw, const double *ane, const double *anw,
                      double *phi, const double *q, double rtol, int maxit, int *iters, double *relres) {
    const double *pl[9] = {asw, as, ase, aw, ap, ae, anw, an, ane};
    return sop_solve_noalloc(n0, n1, pl, phi, q, rtol, maxit, iters, relres);
}

int cfd_ref_face_div(int n0, int n1, double dxi, double deta, const double *up, const double *vp, const double *p,
                     const double *xix, const double *xiy, const double *etx, const double *ety, double dt,
                     double *q) {
    if (bad_grid(n0, n1) || !up || !vp || !p || !xix || !xiy || !etx || !ety || !q) return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    double *a = B.up(up, N), *b = B.up(vp, N), *c = B.up(p, N), *d = B.up(xix, N), *e = B.up(xiy, N),
           *f = B.up(etx, N), *g = B.up(ety, N), *o = B.up(q, N);
    NEED(a && b && c && d && e && f && g && o);
    k_face_div_ref<<<grid2d(n0, n1), dim3(BX, BY)>>>(n0, n1, dxi, deta, a, b, c, d, e, f, g, dt, o);
    CFD_LAUNCH_CHECK();
    return down(q, o, N);
}

int cfd_ref_pcor_bc(int n0, int n1, double *pcor, const double *xnox, const double *xnoy, double uinf,
                    double vinf) {
    if (bad_grid(n0, n1) || !pcor || !xnox || !xnoy) return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    double *pc = B.up(pcor, N), *a = B.up(xnox, n0), *b = B.up(xnoy, n0);
    NEED(pc && a && b);
    k_pcor_bc<<<(n0 + 255) / 256, 256>>>(n0, n1, pc, a, b, uinf, vinf);
    CFD_LAUNCH_CHECK();
    return down(pcor, pc, N);
}

int cfd_ref_correct(int n0, int n1, double dxi, double deta, const double *us, const double *vs,
                    const double *pcor, const double *xix, const double *xiy, const double *etx, const double *ety,
                    double dt, double *u, double *v, double *p) {
    if (bad_grid(n0, n1) || !us || !vs || !pcor || !xix || !xiy || !etx || !ety || !u || !v || !p)
        return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    double *a = B.up(us, N), *b = B.up(vs, N), *c = B.up(pcor, N), *d = B.up(xix, N), *e = B.up(xiy, N),
           *f = B.up(etx, N), *g = B.up(ety, N), *uu = B.up(u, N), *vv = B.up(v, N), *pp = B.up(p, N);
    NEED(a && b && c && d && e && f && g && uu && vv && pp);
    k_correct_ref<<<grid2d(n0, n1), dim3(BX, BY)>>>(n0, n1, dxi, deta, a, b, c, d, e, f, g, dt, uu, vv, pp);
    CFD_LAUNCH_CHECK();
    CFD_RET(down(u, uu, N));
    CFD_RET(down(v, vv, N));
    return down(p, pp, N);
}

int cfd_ref_wall_pressure(int n0, int n1, double dxi, double deta, const double *u, const double *v,
                          const double *T, double *p, const double *xix, const double *xiy, const double *etx,
                          const double *ety, const double *alph, const double *beta, const double *gamma,
                          const double *q1, const double *ajac, const double *alc_u, const double *alc_v, double Re,
                          double Ri) {
    if (bad_grid(n0, n1) || !u || !v || !T || !p || !xix || !xiy || !etx || !ety || !alph || !beta || !gamma ||
        !ajac)
        return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    WallRef w{B.up(u, N), B.up(v, N), B.up(T, N), B.up(xix, N), B.up(xiy, N), B.up(etx, N), B.up(ety, N),
              B.up(alph, N), B.up(beta, N), B.up(gamma, N), B.up(q1, N), B.up(ajac, N),
              alc_u ? B.up(alc_u, n0) : nullptr, alc_v ? B.up(alc_v, n0) : nullptr};
    double *pp = B.up(p, N);
    NEED(w.u && w.v && w.T && w.xix && w.xiy && w.etx && w.ety && w.alph && w.beta && w.gamma && w.q1 && w.ajac && pp);
    k_wall_p_ref<<<(n0 + 255) / 256, 256>>>(n0, n1, dxi, deta, w, pp, Re, Ri);
    CFD_LAUNCH_CHECK();
    return down(p, pp, N);
}

int cfd_ref_outflow_up(int n0, int n1, double *up, double *vp, const double *u, const double *v,
                       const double *xnox, const double *xnoy, double uinf, double vinf) {
    if (bad_grid(n0, n1) || !up || !vp || !u || !v || !xnox || !xnoy) return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    double *a = B.up(up, N), *b = B.up(vp, N), *c = B.up(u, N), *d = B.up(v, N), *e = B.up(xnox, n0),
           *f = B.up(xnoy, n0);
    NEED(a && b && c && d && e && f);
    k_outflow_up_ref<<<(n0 + 255) / 256, 256>>>(n0, n1, a, b, c, d, e, f, uinf, vinf);
    CFD_LAUNCH_CHECK();
    CFD_RET(down(up, a, N));
    return down(vp, b, N);
}

int cfd_ref_farfield(int n0, int n1, double deta, const double *x, const double *y, double *u, double *v, double *T,
                     double *up, double *vp, const double *Told, const double *uet, const double *etx,
                     const double *ety, const double *ajac, const double *xnox, const double *xnoy, double uinf,
                     double vinf, double dt, double speed_amp, double F, double time, double *circ) {
    if (bad_grid(n0, n1) || !x || !y || !u || !v || !T || !up || !vp || !Told || !uet || !etx || !ety || !ajac ||
        !xnox || !xnoy)
        return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    const FarRef a{B.up(x, N), B.up(y, N), B.up(Told, N), B.up(uet, N), B.up(etx, N),
                   B.up(ety, N), B.up(ajac, N), B.up(xnox, n0), B.up(xnoy, n0)};
    double *uu = B.up(u, N), *vv = B.up(v, N), *tt = B.up(T, N), *pu = B.up(up, N), *pv = B.up(vp, N),
           *ci = B.get(1);
    NEED(a.x && a.y && a.Told && a.uet && a.etx && a.ety && a.ajac && a.xnox && a.xnoy && uu && vv && tt && pu &&
         pv && ci);
    const double cwt = cos(2.0 * acos(-1.0) * F * time);
    k_farfield_ref<<<1, FF_NT>>>(n0, n1, deta, a, uu, vv, tt, pu, pv, uinf, vinf, dt, speed_amp, cwt, ci);
    CFD_LAUNCH_CHECK();
    CFD_RET(down(u, uu, N));
    CFD_RET(down(v, vv, N));
    CFD_RET(down(T, tt, N));
    CFD_RET(down(up, pu, N));
    CFD_RET(down(vp, pv, N));
    return down(circ, ci, 1);
}

int cfd_ref_exit_pressure(int n0, int n1, double dxi, double deta, const double *u, const double *v,
                          const double *T, const double *uold, const double *vold, double *p, const double *xix,
                          const double *xiy, const double *etx, const double *ety, const double *alph,
                          const double *beta, const double *gamma, const double *q1, const double *ajac,
                          const double *xnox, const double *xnoy, double uinf, double vinf, double dt, double Re,
                          double Ri) {
    if (bad_grid(n0, n1) || !u || !v || !T || !uold || !vold || !p || !xix || !xiy || !etx || !ety || !alph ||
        !beta || !gamma || !ajac || !xnox || !xnoy)
        return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    const ExitRef w{B.up(u, N),     B.up(v, N),     B.up(T, N),     B.up(uold, N),  B.up(vold, N),  B.up(xix, N),
                    B.up(xiy, N),   B.up(etx, N),   B.up(ety, N),   B.up(alph, N),  B.up(beta, N),  B.up(gamma, N),
                    B.up(q1, N),    B.up(ajac, N),  B.up(xnox, n0), B.up(xnoy, n0)};
    double *pp = B.up(p, N);
    NEED(w.u && w.v && w.T && w.uold && w.vold && w.xix && w.xiy && w.etx && w.ety && w.alph && w.beta && w.gamma &&
         w.q1 && w.ajac && w.xnox && w.xnoy && pp);
    k_exit_p_ref<<<(n0 + 255) / 256, 256>>>(n0, n1, dxi, deta, w, pp, uinf, vinf, dt, Re, Ri);
    CFD_LAUNCH_CHECK();
    return down(p, pp, N);
}

int cfd_ref_dilation(int n0, int n1, double dxi, double deta, const double *u, const double *v, const double *xix,
                     const double *xiy, const double *etx, const double *ety, double *dil, double *dmax) {
    if (bad_grid(n0, n1) || !u || !v || !xix || !xiy || !etx || !ety) return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    const dim3 gr = grid2d(n0, n1);
    const int nb = (int)(gr.x * gr.y);
    DevBufs B;
    double *a = B.up(u, N), *b = B.up(v, N), *c = B.up(xix, N), *d = B.up(xiy, N), *e = B.up(etx, N),
           *f = B.up(ety, N), *o = B.up(dil, N), *part = B.get((size_t)nb), *m = B.get(1);
    NEED(a && b && c && d && e && f && o && part && m);
    k_dil_ref<<<gr, dim3(BX, BY)>>>(n0, n1, dxi, deta, a, b, c, d, e, f, o, part);
    k_max_final<<<1, 256>>>(part, nb, m);
    CFD_LAUNCH_CHECK();
    CFD_RET(down(dil, o, N));
    return down(dmax, m, 1);
}

int cfd_ref_nusselt(int n0, int n1, double dxi, double deta, const double *T, const double *ajac, const double *etx,
                    const double *ety, double ar, double *nu, double *dthdn) {
    if (bad_grid(n0, n1) || !T || !ajac || !etx || !ety || !nu) return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    double *a = B.up(T, N), *b = B.up(ajac, N), *c = B.up(etx, N), *d = B.up(ety, N), *o = B.get(1),
           *prof = B.get((size_t)n0);
    NEED(a && b && c && d && o && prof);
    k_nusselt_ref<<<1, 256>>>(n0, dxi, deta, a, b, c, d, ar, o, prof);
    CFD_LAUNCH_CHECK();
    CFD_RET(down(nu, o, 1));
    return down(dthdn, prof, (size_t)n0);
}

int cfd_ref_case_create(int n0, int n1, double dxi, double deta, const double *x, const double *y, const double *xix,
                        const double *xiy, const double *etx, const double *ety, const double *alph,
                        const double *beta, const double *gamma, const double *ajac, const double *p1,
                        const double *q1, const double *xnox, const double *xnoy, const cfd_ref_params *prm,
                        cfd_ref_case *out) {
    if (!out) return CFD_E_ARG;
    *out = nullptr;
    if (bad_grid(n0, n1) || !x || !y || !xix || !xiy || !etx || !ety || !alph || !beta || !gamma || !ajac ||
        !xnox || !xnoy || !prm || !(prm->dt > 0) || !(prm->Re > 0) || !(prm->Pr > 0) || !(prm->mom_rtol > 0) ||
        !(prm->p_rtol > 0) || prm->mom_maxit < 1 || prm->p_maxit < 1 || !(prm->ar > 0))
        return CFD_E_ARG;
    RefCase *c = new RefCase;
    c->n0 = n0;
    c->n1 = n1;
    c->dxi = dxi;
    c->deta = deta;
    c->prm = *prm;
    const size_t nd = ref_case_doubles(n0, n1), N = (size_t)n0 * n1;
    auto fail = [&](int rc) {
        cfd_ref_case_destroy(reinterpret_cast<cfd_ref_case>(c));
        return rc;
    };
    if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) return fail(CFD_E_HIP);
    if (hipMalloc(&c->mem, nd * sizeof(double)) != hipSuccess) return fail(CFD_E_OOM);
    if (hipHostMalloc(&c->hscal, 16 * sizeof(double), hipHostMallocDefault) != hipSuccess) return fail(CFD_E_OOM);
    (void)flag_alloc(c->flag);
    ref_case_carve(c);
    if (hipMemsetAsync(c->mem, 0, nd * sizeof(double), c->s) != hipSuccess) return fail(CFD_E_HIP);
    const double *src[12] = {x, y, xix, xiy, etx, ety, alph, beta, gamma, ajac, p1, q1};
    double *dst[12] = {c->x, c->y, c->xix, c->xiy, c->etx, c->ety, c->alph, c->beta, c->gamma, c->ajac, c->p1, c->q1};
    for (int k = 0; k < 12; k++)
        if (src[k] && hipMemcpyAsync(dst[k], src[k], N * sizeof(double), hipMemcpyHostToDevice, c->s) != hipSuccess)
            return fail(CFD_E_HIP);
    if (hipMemcpyAsync(c->xnox, xnox, n0 * sizeof(double), hipMemcpyHostToDevice, c->s) != hipSuccess ||
        hipMemcpyAsync(c->xnoy, xnoy, n0 * sizeof(double), hipMemcpyHostToDevice, c->s) != hipSuccess)
        return fail(CFD_E_HIP);
    // operators of the reference's constructor (solver.cpp:751-1039); the temperature
    // solve passes the (-1,+1) plane in the (-1,+2) slot (solver.cpp:1289)
    const dim3 gr = grid2d(n0, n1), bl(BX, BY);
    k_asm_mom<<<gr, bl, 0, c->s>>>(n0, n1, dxi, deta, c->alph, c->beta, c->gamma, c->p1, c->q1, prm->dt, prm->Re,
                                   prm->Pr, c->AU, c->AT, c->BU, c->BT);
    k_asm_p9<<<gr, bl, 0, c->s>>>(n0, n1, dxi, deta, c->xix, c->xiy, c->etx, c->ety, c->A9);
    if (hipGetLastError() != hipSuccess) return fail(CFD_E_HIP);
    if (hipMemcpyAsync(c->AT + 21 * N, c->AT + 16 * N, N * sizeof(double), hipMemcpyDeviceToDevice, c->s) !=
            hipSuccess ||
        hipStreamSynchronize(c->s) != hipSuccess)
        return fail(CFD_E_HIP);
    *out = reinterpret_cast<cfd_ref_case>(c);
    return CFD_OK;
}

int cfd_ref_case_set_state(cfd_ref_case h, const double *u, const double *v, const double *T, const double *p,
                           const double *up, const double *vp) {
    RefCase *c = reinterpret_cast<RefCase *>(h);
    if (!c || !u || !v || !T || !p || !up || !vp) return CFD_E_ARG;
    const size_t B = (size_t)c->n0 * c->n1 * sizeof(double);
    const double *src[6] = {u, v, T, p, up, vp};
    double *dst[6] = {c->u, c->v, c->T, c->p, c->up, c->vp};
    for (int k = 0; k < 6; k++) CFD_HIP(hipMemcpyAsync(dst[k], src[k], B, hipMemcpyHostToDevice, c->s));
    CFD_HIP(hipStreamSynchronize(c->s));
    c->have_state = true;
    return CFD_OK;
}

int cfd_ref_case_get_state(cfd_ref_case h, double *u, double *v, double *T, double *p, double *up, double *vp) {
    RefCase *c = reinterpret_cast<RefCase *>(h);
    if (!c) return CFD_E_ARG;
    const size_t B = (size_t)c->n0 * c->n1 * sizeof(double);
    double *dst[6] = {u, v, T, p, up, vp};
    const double *src[6] = {c->u, c->v, c->T, c->p, c->up, c->vp};
    for (int k = 0; k < 6; k++)
        if (dst[k]) CFD_HIP(hipMemcpyAsync(dst[k], src[k], B, hipMemcpyDeviceToHost, c->s));
    CFD_HIP(hipStreamSynchronize(c->s));
    return CFD_OK;
}

int cfd_ref_case_step(cfd_ref_case h, int nsteps, cfd_ref_stats *last) {
    RefCase *c = reinterpret_cast<RefCase *>(h);
    if (!c || nsteps < 0) return CFD_E_ARG;
    if (!c->have_state) return CFD_E_STATE;
    for (int k = 0; k < nsteps; k++) CFD_RET(ref_case_step(c, last));
    return CFD_OK;
}

int cfd_ref_case_get_post(cfd_ref_case h, double *si, double *vort) {
    RefCase *c = reinterpret_cast<RefCase *>(h);
    if (!c) return CFD_E_ARG;
    const size_t B = (size_t)c->n0 * c->n1 * sizeof(double);
    if (si) CFD_HIP(hipMemcpyAsync(si, c->si, B, hipMemcpyDeviceToHost, c->s));
    if (vort) CFD_HIP(hipMemcpyAsync(vort, c->vort, B, hipMemcpyDeviceToHost, c->s));
    CFD_HIP(hipStreamSynchronize(c->s));
    return CFD_OK;
}

int cfd_ref_stream_function(int n0, int n1, double deta, const double *u, const double *v, const double *xix,
                            const double *xiy, const double *ajac, double *si) {
    if (bad_grid(n0, n1) || !u || !v || !xix || !xiy || !ajac || !si) return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    double *a = B.up(u, N), *b = B.up(v, N), *c = B.up(xix, N), *d = B.up(xiy, N), *e = B.up(ajac, N), *o = B.get(N);
    NEED(a && b && c && d && e && o);
    k_si_ref<<<(n0 + SI_W - 1) / SI_W, dim3(SI_W, SI_W)>>>(n0, n1, deta, a, b, c, d, e, o);
    CFD_LAUNCH_CHECK();
    return down(si, o, N);
}

int cfd_ref_vorticity(int n0, int n1, double dxi, double deta, const double *u, const double *v, const double *xix,
                      const double *xiy, const double *etx, const double *ety, double *vort) {
    if (bad_grid(n0, n1) || !u || !v || !xix || !xiy || !etx || !ety || !vort) return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    double *a = B.up(u, N), *b = B.up(v, N), *c = B.up(xix, N), *d = B.up(xiy, N), *e = B.up(etx, N),
           *f = B.up(ety, N), *o = B.up(vort, N);
    NEED(a && b && c && d && e && f && o);
    k_vort_ref<<<grid2d(n0, n1), dim3(BX, BY)>>>(n0, n1, dxi, deta, a, b, c, d, e, f, o);
    CFD_LAUNCH_CHECK();
    return down(vort, o, N);
}

int cfd_ref_forces(int n0, int n1, double dxi, double deta, const double *x, const double *y, const double *p,
                   const double *vort, const double *T, const double *ajac, const double *etx, const double *ety,
                   double Re, double alpha_deg, double ar, double *coef) {
    if (bad_grid(n0, n1) || !x || !y || !p || !vort || !T || !ajac || !etx || !ety || !coef || !(ar > 0))
        return CFD_E_ARG;
    const size_t N = (size_t)n0 * n1;
    DevBufs B;
    const ForceRef a{B.up(x, N), B.up(y, N), B.up(p, N), B.up(vort, N), B.up(T, N), B.up(ajac, N), B.up(etx, N),
                     B.up(ety, N)};
    double *o = B.get(8);
    NEED(a.x && a.y && a.p && a.vort && a.T && a.ajac && a.etx && a.ety && o);
    const ForceConst fc = force_const(alpha_deg, ar);
    k_forces_ref<<<1, 256>>>(n0, dxi, deta, a, Re, fc.sa, fc.ca, fc.nu_den, o);
    CFD_LAUNCH_CHECK();
    return down(coef, o, 8);
}

int cfd_ref_case_time(cfd_ref_case h, double *time) {
    RefCase *c = reinterpret_cast<RefCase *>(h);
    if (!c || !time) return CFD_E_ARG;
    *time = c->prm.time;
    return CFD_OK;
}

void cfd_ref_case_destroy(cfd_ref_case h) {
    RefCase *c = reinterpret_cast<RefCase *>(h);
    if (!c) return;
    if (c->s) (void)hipStreamSynchronize(c->s);
    if (c->mem) (void)hipFree(c->mem);
    if (c->hscal) (void)hipHostFree(c->hscal);
    flag_free(c->flag);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
}

const char *cfd_version(void) { return "cfd_gpu 0.1 (gfx950)"; }

int cfd_device_count(int *count) {
    if (!count) return CFD_E_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return CFD_OK;
}

}  // extern "C"
