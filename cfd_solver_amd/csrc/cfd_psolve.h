// Pressure-correction solver on gfx950: face-coefficient operator, geometric
// multigrid V-cycle and right-preconditioned BiCGStab, templated on the
// precision of the vectors and stored coefficients (double, or float for the
// inner solve of the mixed path).  Dot products accumulate in double, scalars
// live in device memory as double.
//
// The operator is the J-weighted flux form of the face velocity of
// solver.cpp:1402-1436 (see p9_coefs).  Per level and per outer iteration the
// four face coefficients (east/north, normal/tangential) are computed once in
// FP64 from the level's metrics and d and stored in T; a row is assembled from
// the faces of P, W and S, so an operator application reads 4 coefficient
// arrays and the vector instead of 5 metric arrays, d and the vector.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "cfd_cavity_kernels.h"

namespace cfd {

constexpr int PS_MAXLEV = 16;
constexpr int PS_NT = BX * BY;

// device scalar slots used by the solver (pairs reduced together are adjacent)
enum {
    S_RHO = 0, S_RHO_OLD, S_ALPHA, S_OMEGA, S_RV, S_TS, S_TT, S_SS, S_BB, S_RR, S_SUM, S_R0R,
    S_MASS, S_ENERGY, S_MOM, S_DMAX, S_NUH, S_NUC,
    // device-side control of the captured (graph) solve
    S_DONE, S_SCONV, S_TOL, S_IT, S_RN, S_BN, S_RELRES, S_RTOL,
    // mixed-precision refinement (FP64 defect)
    S_BB64, S_RR64, S_TOL64, S_RN64, S_BN64, S_RDONE, S_ITTOT, S_RELRES64,
    S_NSCAL
};

// FP64 geometry of one multigrid level (level 0 aliases the case's arrays)
struct Geom {
    int n0 = 0, n1 = 0;
    double h1 = 0, h2 = 0;
    double *xix = nullptr, *xiy = nullptr, *etx = nullptr, *ety = nullptr, *jac = nullptr, *d = nullptr;
};

template <typename T>
struct OpLevel {
    int n0 = 0, n1 = 0;
    T *fen = nullptr, *fet = nullptr, *fnn = nullptr, *fnt = nullptr, *invd = nullptr;
    T *x = nullptr, *b = nullptr, *r = nullptr, *t = nullptr;
};

template <typename T>
struct FArgs {
    int n0, n1;
    const T *fen, *fet, *fnn, *fnt;
};
template <typename T>
FArgs<T> fargs(const OpLevel<T> &L) {
    return FArgs<T>{L.n0, L.n1, L.fen, L.fet, L.fnn, L.fnt};
}

__device__ __forceinline__ int blin() { return blockIdx.x + blockIdx.y * gridDim.x; }

// face coefficients of a level, computed in FP64 and stored as T
template <typename T>
__global__ __launch_bounds__(PS_NT) void k_p_faces(int n0, int n1, double h1, double h2, const double *__restrict__ xix,
                                                   const double *__restrict__ xiy, const double *__restrict__ etx,
                                                   const double *__restrict__ ety, const double *__restrict__ jac,
                                                   const double *__restrict__ d, T *__restrict__ fen,
                                                   T *__restrict__ fet, T *__restrict__ fnn, T *__restrict__ fnt) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= n0 || j >= n1) return;
    const size_t P = idx(i, j, n0);
    double kn, kt;
    if (i < n0 - 1) {
        p9_face<double>(n0, n1, xix, xiy, etx, ety, jac, d, h1, h2, i, j, 0, kn, kt);
    } else {
        kn = kt = 0.0;
    }
    fen[P] = (T)kn;
    fet[P] = (T)kt;
    if (j < n1 - 1) {
        p9_face<double>(n0, n1, xix, xiy, etx, ety, jac, d, h1, h2, i, j, 1, kn, kt);
    } else {
        kn = kt = 0.0;
    }
    fnn[P] = (T)kn;
    fnt[P] = (T)kt;
}

template <typename T>
__device__ __forceinline__ T frow(const FArgs<T> &a, const T *__restrict__ x, int i, int j) {
    T c[9];
    p9_row_faces<T>(a.n0, a.fen, a.fet, a.fnn, a.fnt, i, j, c);
    T s = T(0);
#pragma unroll
    for (int dj = -1; dj <= 1; dj++)
#pragma unroll
        for (int di = -1; di <= 1; di++) {
            const int ii = clampi(i + di, 1, a.n0 - 2), jj = clampi(j + dj, 1, a.n1 - 2);
            s += c[(dj + 1) * 3 + (di + 1)] * x[idx(ii, jj, a.n0)];
        }
    return s;
}

template <typename T>
__global__ __launch_bounds__(PS_NT) void k_fapply(const double *gd, FArgs<T> a, const T *__restrict__ x, T *__restrict__ y) {
    if (gd && *gd != 0.0) return;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.n0 || j >= a.n1) return;
    const size_t P = idx(i, j, a.n0);
    if (i == 0 || j == 0 || i == a.n0 - 1 || j == a.n1 - 1) { y[P] = T(0); return; }
    y[P] = frow<T>(a, x, i, j);
}

template <typename T>
__global__ __launch_bounds__(PS_NT) void k_fresid(const double *gd, FArgs<T> a, const T *__restrict__ x, const T *__restrict__ b,
                                                  T *__restrict__ r) {
    if (gd && *gd != 0.0) return;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.n0 || j >= a.n1) return;
    const size_t P = idx(i, j, a.n0);
    if (i == 0 || j == 0 || i == a.n0 - 1 || j == a.n1 - 1) { r[P] = T(0); return; }
    r[P] = b[P] - frow<T>(a, x, i, j);
}

template <typename T>
__global__ __launch_bounds__(PS_NT) void k_finvdiag(FArgs<T> a, T *__restrict__ invd) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i < 1 || i > a.n0 - 2 || j < 1 || j > a.n1 - 2) return;
    T c[9];
    p9_row_faces<T>(a.n0, a.fen, a.fet, a.fnn, a.fnt, i, j, c);
    T s = T(0);
    for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++)
            if (clampi(i + di, 1, a.n0 - 2) == i && clampi(j + dj, 1, a.n1 - 2) == j) s += c[(dj + 1) * 3 + (di + 1)];
    invd[idx(i, j, a.n0)] = T(1) / s;
}

// Jacobi sweep: xo = xi + w (b - L xi) invd (FIRST: xo = w b invd)
template <typename T, bool FIRST>
__global__ __launch_bounds__(PS_NT) void k_fsmooth(const double *gd, FArgs<T> a, T w, const T *__restrict__ xi, const T *__restrict__ b,
                                                   const T *__restrict__ invd, T *__restrict__ xo) {
    if (gd && *gd != 0.0) return;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.n0 || j >= a.n1) return;
    const size_t P = idx(i, j, a.n0);
    if (i == 0 || j == 0 || i == a.n0 - 1 || j == a.n1 - 1) { xo[P] = T(0); return; }
    if (FIRST) xo[P] = w * b[P] * invd[P];
    else xo[P] = xi[P] + w * (b[P] - frow<T>(a, xi, i, j)) * invd[P];
}

template <typename T>
__global__ __launch_bounds__(PS_NT) void k_restrict(const double *gd, int n0, const T *__restrict__ r, int m0, int m1, T *__restrict__ bc) {
    if (gd && *gd != 0.0) return;
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
    if (I >= m0 || J >= m1) return;
    const size_t Pc = idx(I, J, m0);
    if (I == 0 || J == 0 || I == m0 - 1 || J == m1 - 1) { bc[Pc] = T(0); return; }
    const int i = 2 * I, j = 2 * J;
    bc[Pc] = (T(4) * r[idx(i, j, n0)] +
              T(2) * (r[idx(i - 1, j, n0)] + r[idx(i + 1, j, n0)] + r[idx(i, j - 1, n0)] + r[idx(i, j + 1, n0)]) +
              (r[idx(i - 1, j - 1, n0)] + r[idx(i + 1, j - 1, n0)] + r[idx(i - 1, j + 1, n0)] +
               r[idx(i + 1, j + 1, n0)])) /
             T(16);
}

template <typename T>
__global__ __launch_bounds__(PS_NT) void k_prolong_add(const double *gd, int n0, int n1, T *__restrict__ x, int m0, int m1,
                                                       const T *__restrict__ xc) {
    if (gd && *gd != 0.0) return;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i < 1 || i > n0 - 2 || j < 1 || j > n1 - 2) return;
    const int I = i >> 1, J = j >> 1;
#define XC(a_, b_) xc[idx(clampi((a_), 1, m0 - 2), clampi((b_), 1, m1 - 2), m0)]
    T e;
    if ((i & 1) == 0 && (j & 1) == 0) e = XC(I, J);
    else if ((j & 1) == 0) e = T(0.5) * (XC(I, J) + XC(I + 1, J));
    else if ((i & 1) == 0) e = T(0.5) * (XC(I, J) + XC(I, J + 1));
    else e = T(0.25) * (XC(I, J) + XC(I + 1, J) + XC(I, J + 1) + XC(I + 1, J + 1));
#undef XC
    x[idx(i, j, n0)] += e;
}

template <typename T>
__global__ __launch_bounds__(PS_NT) void k_sum_int(const double *gd, int n0, int n1, const T *__restrict__ v, double *__restrict__ part) {
    if (gd && *gd != 0.0) return;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    double s = 0.0;
    if (i >= 1 && i <= n0 - 2 && j >= 1 && j <= n1 - 2) s = (double)v[idx(i, j, n0)];
    s = block_sum<PS_NT>(s);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[blin()] = s;
}

template <typename T>
__global__ __launch_bounds__(PS_NT) void k_sub_mean(const double *gd, int n0, int n1, T *__restrict__ v, const double *__restrict__ scal,
                                                    int slot) {
    if (gd && *gd != 0.0) return;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i < 1 || i > n0 - 2 || j < 1 || j > n1 - 2) return;
    const double m = scal[slot] / ((double)(n0 - 2) * (double)(n1 - 2));
    v[idx(i, j, n0)] = (T)((double)v[idx(i, j, n0)] - m);
}

__global__ void k_inject(int n0, const double *__restrict__ f, int m0, int m1, double *__restrict__ c) {
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
    if (I >= m0 || J >= m1) return;
    c[idx(I, J, m0)] = f[idx(2 * I, 2 * J, n0)];
}

template <typename A, typename B>
__global__ void k_cvt(const double *gd, size_t n, const A *__restrict__ a, B *__restrict__ b) {
    if (gd && *gd != 0.0) return;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = (B)a[k];
}

// acc += (double) e on every node
template <typename T>
__global__ void k_acc(const double *gd, size_t n, const T *__restrict__ e, double *__restrict__ acc) {
    if (gd && *gd != 0.0) return;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) acc[k] += (double)e[k];
}

// --- BiCGStab vector kernels
template <typename T>
__global__ __launch_bounds__(PS_NT) void k_dot2(const double *gd, int n0, int n1, const T *__restrict__ a, const T *__restrict__ b,
                                                const T *__restrict__ c, const T *__restrict__ d2,
                                                double *__restrict__ part, int pstride) {
    if (gd && *gd != 0.0) return;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    double s1 = 0.0, s2 = 0.0;
    if (i >= 1 && i <= n0 - 2 && j >= 1 && j <= n1 - 2) {
        const size_t P = idx(i, j, n0);
        s1 = (double)a[P] * (double)b[P];
        if (c) s2 = (double)c[P] * (double)d2[P];
    }
    s1 = block_sum<PS_NT>(s1);
    s2 = block_sum<PS_NT>(s2);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        part[blin()] = s1;
        part[pstride + blin()] = s2;
    }
}

template <typename T>
__global__ __launch_bounds__(PS_NT) void k_bicg_p(const double *gd, int n0, int n1, bool first, const double *__restrict__ scal,
                                                  const T *__restrict__ r, const T *__restrict__ vv, T *__restrict__ pv) {
    if (gd && *gd != 0.0) return;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= n0 || j >= n1) return;
    const size_t P = idx(i, j, n0);
    if (first) { pv[P] = r[P]; return; }
    const double beta = (scal[S_RHO] / scal[S_RHO_OLD]) * (scal[S_ALPHA] / scal[S_OMEGA]);
    pv[P] = (T)((double)r[P] + beta * ((double)pv[P] - scal[S_OMEGA] * (double)vv[P]));
}

template <typename T>
__global__ __launch_bounds__(PS_NT) void k_bicg_s(const double *gd, int n0, int n1, double *__restrict__ scal, const T *__restrict__ r,
                                                  const T *__restrict__ vv, T *__restrict__ s, double *__restrict__ part) {
    if (gd && *gd != 0.0) return;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const double alpha = scal[S_RHO] / scal[S_RV];
    double ss = 0.0;
    if (i < n0 && j < n1) {
        const size_t P = idx(i, j, n0);
        const T sv = (T)((double)r[P] - alpha * (double)vv[P]);
        s[P] = sv;
        if (i >= 1 && i <= n0 - 2 && j >= 1 && j <= n1 - 2) ss = (double)sv * (double)sv;
    }
    ss = block_sum<PS_NT>(ss);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[blin()] = ss;
    if (blin() == 0 && threadIdx.x == 0 && threadIdx.y == 0) scal[S_ALPHA] = alpha;
}

template <typename T>
__global__ __launch_bounds__(PS_NT) void k_bicg_x1(int n0, int n1, const double *__restrict__ scal,
                                                   const T *__restrict__ ph, T *__restrict__ x) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= n0 || j >= n1) return;
    const size_t P = idx(i, j, n0);
    x[P] = (T)((double)x[P] + scal[S_ALPHA] * (double)ph[P]);
}

template <typename T>
__global__ __launch_bounds__(PS_NT) void k_bicg_xr(const double *gd, int n0, int n1, const double *__restrict__ scal,
                                                   const T *__restrict__ ph, const T *__restrict__ sh,
                                                   const T *__restrict__ s, const T *__restrict__ t,
                                                   const T *__restrict__ r0, T *__restrict__ x, T *__restrict__ r,
                                                   double *__restrict__ part, int pstride) {
    if (gd && *gd != 0.0) return;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const double tt = scal[S_TT];
    const double omega = (tt > 0.0) ? scal[S_TS] / tt : 0.0;
    const double alpha = scal[S_ALPHA];
    double rr = 0.0, r0r = 0.0;
    if (i < n0 && j < n1) {
        const size_t P = idx(i, j, n0);
        x[P] = (T)((double)x[P] + alpha * (double)ph[P] + omega * (double)sh[P]);
        const T rv = (T)((double)s[P] - omega * (double)t[P]);
        r[P] = rv;
        if (i >= 1 && i <= n0 - 2 && j >= 1 && j <= n1 - 2) {
            rr = (double)rv * (double)rv;
            r0r = (double)r0[P] * (double)rv;
        }
    }
    rr = block_sum<PS_NT>(rr);
    r0r = block_sum<PS_NT>(r0r);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        part[blin()] = rr;
        part[pstride + blin()] = r0r;
    }
}

__global__ void k_set_omega(const double *gd, double *__restrict__ scal) {
    if (gd && *gd != 0.0) return;
    const double tt = scal[S_TT];
    scal[S_OMEGA] = (tt > 0.0) ? scal[S_TS] / tt : 0.0;
}
__global__ void k_rotate_rho(double *__restrict__ scal) {
    scal[S_RHO_OLD] = scal[S_RHO];
    scal[S_RHO] = scal[S_R0R];
}
__global__ void k_init_scal(double *__restrict__ scal) {
    scal[S_RHO] = scal[S_RR];
    scal[S_RHO_OLD] = 1.0;
    scal[S_ALPHA] = 1.0;
    scal[S_OMEGA] = 1.0;
}

// sum of `nparts` partials for each of `nv` arrays (stride pstride) -> scal[dst + k]
__global__ void k_reduce_sum(const double *gd, const double *__restrict__ part, int pstride, int nparts, int nv, double *__restrict__ scal,
                             int dst) {
    if (gd && *gd != 0.0) return;
    const int k = blockIdx.x;
    if (k >= nv) return;
    double s = 0.0;
    for (int t = threadIdx.x; t < nparts; t += blockDim.x) s += part[(size_t)k * pstride + t];
    s = block_sum<256>(s);
    if (threadIdx.x == 0) scal[dst + k] = s;
}

// ---------------------------------------------------------------------------
// Coarse-level V-cycle in one workgroup.  Levels at or below a few thousand
// nodes are latency-bound: one launch per sweep would cost more than the sweep.
// This kernel runs the whole V-cycle from level l0 down to the coarsest and
// back in a single 1024-thread workgroup, separating the phases with block
// barriers (which also order the global-memory traffic of the workgroup).
// The arithmetic per node is the same device code as the per-level kernels.

constexpr int CW = 1024;  // threads of the coarse workgroup

template <typename T>
struct CLev {
    int n0, n1;
    T *fen, *fet, *fnn, *fnt, *invd, *x, *b, *r, *t;
};
template <typename T>
struct CoarseArgs {
    int nlev;
    int pre, post, coarse;
    T omega;
    CLev<T> L[PS_MAXLEV];
};

template <typename T>
__device__ __forceinline__ void cw_smooth(const CLev<T> &L, const T *xi, T *xo, T w, bool first) {
    const FArgs<T> a{L.n0, L.n1, L.fen, L.fet, L.fnn, L.fnt};
    const int n = L.n0 * L.n1;
    for (int k = threadIdx.x; k < n; k += CW) {
        const int i = k % L.n0, j = k / L.n0;
        if (i == 0 || j == 0 || i == L.n0 - 1 || j == L.n1 - 1) { xo[k] = T(0); continue; }
        if (first) xo[k] = w * L.b[k] * L.invd[k];
        else xo[k] = xi[k] + w * (L.b[k] - frow<T>(a, xi, i, j)) * L.invd[k];
    }
    __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(CW) void k_mg_coarse(const double *gd, CoarseArgs<T> A, int l0) {
    if (gd && *gd != 0.0) return;
    __shared__ double red[CW / 64];
    T *xc[PS_MAXLEV], *tc[PS_MAXLEV];
    for (int l = 0; l < A.nlev; l++) { xc[l] = A.L[l].x; tc[l] = A.L[l].t; }
    const T w = A.omega;
    const int last = A.nlev - 1;
    // sweeps on level l starting from zero or from xc[l]; result in xc[l]
    auto sweeps = [&](int l, int count, bool from_zero) {
        for (int s = 0; s < count; s++) {
            if (from_zero && s == 0) {
                cw_smooth<T>(A.L[l], xc[l], xc[l], w, true);
            } else {
                cw_smooth<T>(A.L[l], xc[l], tc[l], w, false);
                T *tmp = xc[l]; xc[l] = tc[l]; tc[l] = tmp;
            }
        }
    };
    for (int l = l0; l < last; l++) {
        const CLev<T> &L = A.L[l];
        const int n = L.n0 * L.n1;
        if (A.pre > 0) {
            sweeps(l, A.pre, true);
        } else {
            for (int k = threadIdx.x; k < n; k += CW) xc[l][k] = T(0);
            __syncthreads();
        }
        // residual
        const FArgs<T> a{L.n0, L.n1, L.fen, L.fet, L.fnn, L.fnt};
        for (int k = threadIdx.x; k < n; k += CW) {
            const int i = k % L.n0, j = k / L.n0;
            if (i == 0 || j == 0 || i == L.n0 - 1 || j == L.n1 - 1) { L.r[k] = T(0); continue; }
            L.r[k] = L.b[k] - frow<T>(a, xc[l], i, j);
        }
        __syncthreads();
        // full weighting onto level l+1 and the interior sum for the mean projection
        const CLev<T> &C = A.L[l + 1];
        const int m = C.n0 * C.n1;
        double part = 0.0;
        for (int k = threadIdx.x; k < m; k += CW) {
            const int I = k % C.n0, J = k / C.n0;
            if (I == 0 || J == 0 || I == C.n0 - 1 || J == C.n1 - 1) { C.b[k] = T(0); continue; }
            const int i = 2 * I, j = 2 * J, f = L.n0;
            const T *r = L.r;
            const T v = (T(4) * r[j * f + i] +
                         T(2) * (r[j * f + i - 1] + r[j * f + i + 1] + r[(j - 1) * f + i] + r[(j + 1) * f + i]) +
                         (r[(j - 1) * f + i - 1] + r[(j - 1) * f + i + 1] + r[(j + 1) * f + i - 1] +
                          r[(j + 1) * f + i + 1])) /
                        T(16);
            C.b[k] = v;
            part += (double)v;
        }
        // block sum (same shape as block_sum, sized for CW threads)
        part = wave_sum(part);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
        __syncthreads();
        double tot = 0.0;
        for (int q = 0; q < CW / 64; q++) tot += red[q];
        __syncthreads();
        const double mean = tot / ((double)(C.n0 - 2) * (double)(C.n1 - 2));
        for (int k = threadIdx.x; k < m; k += CW) {
            const int I = k % C.n0, J = k / C.n0;
            if (I == 0 || J == 0 || I == C.n0 - 1 || J == C.n1 - 1) continue;
            C.b[k] = (T)((double)C.b[k] - mean);
        }
        __syncthreads();
    }
    // coarsest level
    if (A.coarse > 0) {
        sweeps(last, A.coarse, true);
    } else {
        const int n = A.L[last].n0 * A.L[last].n1;
        for (int k = threadIdx.x; k < n; k += CW) xc[last][k] = T(0);
        __syncthreads();
    }
    for (int l = last - 1; l >= l0; l--) {
        const CLev<T> &L = A.L[l];
        const CLev<T> &C = A.L[l + 1];
        const int n = L.n0 * L.n1;
        const T *xcc = xc[l + 1];
        for (int k = threadIdx.x; k < n; k += CW) {
            const int i = k % L.n0, j = k / L.n0;
            if (i < 1 || i > L.n0 - 2 || j < 1 || j > L.n1 - 2) continue;
            const int I = i >> 1, J = j >> 1, m0 = C.n0, m1 = C.n1;
#define XC(a_, b_) xcc[clampi((b_), 1, m1 - 2) * m0 + clampi((a_), 1, m0 - 2)]
            T e;
            if ((i & 1) == 0 && (j & 1) == 0) e = XC(I, J);
            else if ((j & 1) == 0) e = T(0.5) * (XC(I, J) + XC(I + 1, J));
            else if ((i & 1) == 0) e = T(0.5) * (XC(I, J) + XC(I, J + 1));
            else e = T(0.25) * (XC(I, J) + XC(I + 1, J) + XC(I, J + 1) + XC(I + 1, J + 1));
#undef XC
            xc[l][k] += e;
        }
        __syncthreads();
        sweeps(l, A.post, false);
    }
    // hand the level-l0 result back in its x buffer
    if (xc[l0] != A.L[l0].x) {
        const int n = A.L[l0].n0 * A.L[l0].n1;
        for (int k = threadIdx.x; k < n; k += CW) A.L[l0].x[k] = xc[l0][k];
    }
}

// --- small guarded helpers and the device-side control of the captured solve
template <typename T>
__global__ void k_fill(const double *gd, size_t n, T *__restrict__ a, T v) {
    if (gd && *gd != 0.0) return;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = v;
}
template <typename T>
__global__ void k_copy(const double *gd, size_t n, const T *__restrict__ a, T *__restrict__ b) {
    if (gd && *gd != 0.0) return;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = a[k];
}
// x += alpha ph, only when this iteration stopped at the ||s|| test
template <typename T>
__global__ __launch_bounds__(PS_NT) void k_bicg_x1g(int n0, int n1, const double *scal, const T *__restrict__ ph,
                                                    T *__restrict__ x) {
    if (scal[S_SCONV] == 0.0) return;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= n0 || j >= n1) return;
    const size_t P = idx(i, j, n0);
    x[P] = (T)((double)x[P] + scal[S_ALPHA] * (double)ph[P]);
}
// after (b,b), (r,r): tolerance, initial test and start values; an outer guard
// (the refinement's done flag) marks the whole solve as done
__global__ void k_g_init(double *scal, double rtol, int rtol_slot, const double *outer) {
    const double bn = sqrt(scal[S_BB]);
    const double rt = rtol_slot >= 0 ? scal[rtol_slot] : rtol;
    const double tol = rt * bn;
    const double rn = sqrt(scal[S_RR]);
    scal[S_BN] = bn;
    scal[S_TOL] = tol;
    scal[S_RN] = rn;
    scal[S_IT] = 0.0;
    scal[S_SCONV] = 0.0;
    scal[S_RHO] = scal[S_RR];
    scal[S_RHO_OLD] = 1.0;
    scal[S_ALPHA] = 1.0;
    scal[S_OMEGA] = 1.0;
    scal[S_DONE] = (bn == 0.0 || rn <= tol || (outer && *outer != 0.0)) ? 1.0 : 0.0;
}
__global__ void k_g_begin(double *scal, int k) {
    if (scal[S_DONE] != 0.0) return;
    scal[S_IT] = (double)k;
    if (scal[S_RHO] == 0.0) scal[S_DONE] = 1.0;
}
__global__ void k_g_check_s(double *scal) {
    if (scal[S_DONE] != 0.0) { scal[S_SCONV] = 0.0; return; }
    const double sn = sqrt(scal[S_SS]);
    if (sn <= scal[S_TOL]) {
        scal[S_SCONV] = 1.0;
        scal[S_DONE] = 1.0;
        scal[S_RN] = sn;
    } else {
        scal[S_SCONV] = 0.0;
    }
}
__global__ void k_g_check_r(double *scal) {
    if (scal[S_DONE] != 0.0) return;
    const double rn = sqrt(scal[S_RR]);
    scal[S_RN] = rn;
    if (rn <= scal[S_TOL] || scal[S_OMEGA] == 0.0) {
        scal[S_DONE] = 1.0;
    } else {
        scal[S_RHO_OLD] = scal[S_RHO];
        scal[S_RHO] = scal[S_R0R];
    }
}
__global__ void k_g_final(double *scal) { scal[S_RELRES] = scal[S_BN] > 0.0 ? scal[S_RN] / scal[S_BN] : 0.0; }

// ---------------------------------------------------------------------------
// host side

// what the solver needs from its owner (stream, reductions, timing hooks)
struct PsCtx {
    hipStream_t s;
    double *part;
    int pstride;
    double *scal;
    double *hscal;
    void (*prof_b)(void *, int);
    void (*prof_e)(void *);
    void (*drain)(void *);
    void *owner;
    std::vector<void *> *allocs;
    size_t *bytes;
};

inline int ps_alloc(PsCtx &cx, void **p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) return CFD_E_OOM;
    if (hipMemsetAsync(*p, 0, bytes, cx.s) != hipSuccess) return CFD_E_HIP;
    cx.allocs->push_back(*p);
    *cx.bytes += bytes;
    return CFD_OK;
}

inline int nblk(int n0, int n1) {
    const dim3 g = grid2d(n0, n1);
    return (int)(g.x * g.y);
}

inline int ps_reduce(PsCtx &cx, int nparts, int nv, int dst, const double *gd = nullptr) {
    k_reduce_sum<<<nv, 256, 0, cx.s>>>(gd, cx.part, cx.pstride, nparts, nv, cx.scal, dst);
    CFD_LAUNCH_CHECK();
    return CFD_OK;
}

inline int ps_sync(PsCtx &cx) {
    CFD_HIP(hipMemcpyAsync(cx.hscal, cx.scal, S_NSCAL * sizeof(double), hipMemcpyDeviceToHost, cx.s));
    CFD_HIP(hipStreamSynchronize(cx.s));
    if (cx.drain) cx.drain(cx.owner);
    return CFD_OK;
}

template <typename T>
int ps_project_mean(PsCtx &cx, int n0, int n1, T *v, const double *gd = nullptr) {
    k_sum_int<T><<<grid2d(n0, n1), dim3(BX, BY), 0, cx.s>>>(gd, n0, n1, v, cx.part);
    CFD_LAUNCH_CHECK();
    CFD_RET(ps_reduce(cx, nblk(n0, n1), 1, S_SUM, gd));
    k_sub_mean<T><<<grid2d(n0, n1), dim3(BX, BY), 0, cx.s>>>(gd, n0, n1, v, cx.scal, S_SUM);
    CFD_LAUNCH_CHECK();
    return CFD_OK;
}

template <typename T>
int ps_fill(PsCtx &cx, T *a, size_t n, const double *gd = nullptr) {
    k_fill<T><<<(unsigned)((n + 255) / 256), 256, 0, cx.s>>>(gd, n, a, T(0));
    CFD_LAUNCH_CHECK();
    return CFD_OK;
}
template <typename T>
int ps_copy(PsCtx &cx, const T *a, T *b, size_t n, const double *gd = nullptr) {
    k_copy<T><<<(unsigned)((n + 255) / 256), 256, 0, cx.s>>>(gd, n, a, b);
    CFD_LAUNCH_CHECK();
    return CFD_OK;
}

// Multigrid-preconditioned BiCGStab on the face-coefficient operator in precision T.
template <typename T>
struct PSolver {
    OpLevel<T> lev[PS_MAXLEV];
    int nlev = 0;
    T *bi[8] = {nullptr};  // r r0 pv vv s t ph sh (level-0 size)
    int pre = 2, post = 2, coarse = 30;
    double omega = 0.8;
    int lc = 0;                  // first level handled by the single-workgroup kernel
    CoarseArgs<T> cargs{};
    static constexpr int COARSE_NODES = 4225;  // levels with at most 65 x 65 nodes

    int init(PsCtx &cx, const Geom *g, int nlev_, const cfd_params &prm) {
        nlev = nlev_;
        pre = prm.mg_pre;
        post = prm.mg_post;
        coarse = prm.mg_coarse;
        omega = prm.mg_omega;
        for (int l = 0; l < nlev; l++) {
            OpLevel<T> &L = lev[l];
            L.n0 = g[l].n0;
            L.n1 = g[l].n1;
            const size_t M = (size_t)L.n0 * L.n1 * sizeof(T);
            T **arr[] = {&L.fen, &L.fet, &L.fnn, &L.fnt, &L.invd, &L.x, &L.b, &L.r, &L.t};
            for (T **a : arr) CFD_RET(ps_alloc(cx, (void **)a, M));
        }
        const size_t M0 = (size_t)lev[0].n0 * lev[0].n1 * sizeof(T);
        for (int k = 0; k < 8; k++) CFD_RET(ps_alloc(cx, (void **)&bi[k], M0));
        lc = nlev - 1;
        while (lc > 0 && (size_t)lev[lc - 1].n0 * lev[lc - 1].n1 <= COARSE_NODES) lc--;
        cargs.nlev = nlev;
        cargs.pre = pre;
        cargs.post = post;
        cargs.coarse = coarse;
        cargs.omega = (T)omega;
        for (int l = 0; l < nlev; l++) {
            const OpLevel<T> &L = lev[l];
            cargs.L[l] = CLev<T>{L.n0, L.n1, L.fen, L.fet, L.fnn, L.fnt, L.invd, L.x, L.b, L.r, L.t};
        }
        return CFD_OK;
    }

    // per outer iteration: face coefficients and inverse diagonals of every level
    int setup(PsCtx &cx, const Geom *g) {
        const dim3 bl(BX, BY);
        for (int l = 0; l < nlev; l++) {
            OpLevel<T> &L = lev[l];
            const dim3 gr = grid2d(L.n0, L.n1);
            k_p_faces<T><<<gr, bl, 0, cx.s>>>(L.n0, L.n1, g[l].h1, g[l].h2, g[l].xix, g[l].xiy, g[l].etx, g[l].ety,
                                              g[l].jac, g[l].d, L.fen, L.fet, L.fnn, L.fnt);
            k_finvdiag<T><<<gr, bl, 0, cx.s>>>(fargs(L), L.invd);
            CFD_LAUNCH_CHECK();
        }
        return CFD_OK;
    }

    // sweeps on level l; result in L.x (host-side ping-pong of x and t)
    int smooth(PsCtx &cx, int l, int sweeps, bool from_zero, const double *gd) {
        OpLevel<T> &L = lev[l];
        const dim3 gr = grid2d(L.n0, L.n1), bl(BX, BY);
        const FArgs<T> a = fargs(L);
        const T w = (T)omega;
        for (int s = 0; s < sweeps; s++) {
            if (from_zero && s == 0) {
                k_fsmooth<T, true><<<gr, bl, 0, cx.s>>>(gd, a, w, L.x, L.b, L.invd, L.x);
            } else {
                if (l == 0) cx.prof_b(cx.owner, 4);
                k_fsmooth<T, false><<<gr, bl, 0, cx.s>>>(gd, a, w, L.x, L.b, L.invd, L.t);
                if (l == 0) cx.prof_e(cx.owner);
                std::swap(L.x, L.t);
            }
            CFD_LAUNCH_CHECK();
        }
        return CFD_OK;
    }

    int vcycle(PsCtx &cx, int l, const double *gd) {
        OpLevel<T> &L = lev[l];
        if (l == lc) {
            // the rest of the cycle in one workgroup; the level's x/t roles are fixed there
            k_mg_coarse<T><<<1, CW, 0, cx.s>>>(gd, cargs, l);
            CFD_LAUNCH_CHECK();
            return CFD_OK;
        }
        const size_t M = (size_t)L.n0 * L.n1;
        const dim3 gr = grid2d(L.n0, L.n1), bl(BX, BY);
        if (pre > 0) {
            CFD_RET(smooth(cx, l, pre, true, gd));
        } else {
            CFD_RET(ps_fill<T>(cx, L.x, M, gd));
        }
        k_fresid<T><<<gr, bl, 0, cx.s>>>(gd, fargs(L), L.x, L.b, L.r);
        CFD_LAUNCH_CHECK();
        OpLevel<T> &C = lev[l + 1];
        k_restrict<T><<<grid2d(C.n0, C.n1), bl, 0, cx.s>>>(gd, L.n0, L.r, C.n0, C.n1, C.b);
        CFD_LAUNCH_CHECK();
        CFD_RET(ps_project_mean<T>(cx, C.n0, C.n1, C.b, gd));
        CFD_RET(vcycle(cx, l + 1, gd));
        k_prolong_add<T><<<gr, bl, 0, cx.s>>>(gd, L.n0, L.n1, L.x, C.n0, C.n1, C.x);
        CFD_LAUNCH_CHECK();
        return smooth(cx, l, post, false, gd);
    }

    int precond(PsCtx &cx, const T *y, T *z, const double *gd) {
        OpLevel<T> &L = lev[0];
        const size_t M = (size_t)L.n0 * L.n1;
        CFD_RET(ps_copy<T>(cx, y, L.b, M, gd));
        CFD_RET(vcycle(cx, 0, gd));
        CFD_RET(ps_copy<T>(cx, L.x, z, M, gd));
        return CFD_OK;
    }

    // r = b - A x (x = 0 on entry), projected; r0 = r; (b,b) and (r,r) into S_BB, S_RR
    int start(PsCtx &cx, const T *b, T *x, const double *gd) {
        OpLevel<T> &L0 = lev[0];
        const int n0 = L0.n0, n1 = L0.n1;
        const size_t M = (size_t)n0 * n1;
        const dim3 gr = grid2d(n0, n1), bl(BX, BY);
        T *r = bi[0], *r0 = bi[1], *pv = bi[2], *vv = bi[3];
        k_fresid<T><<<gr, bl, 0, cx.s>>>(gd, fargs(L0), x, b, r);
        CFD_LAUNCH_CHECK();
        CFD_RET(ps_project_mean<T>(cx, n0, n1, r, gd));
        CFD_RET(ps_copy<T>(cx, r, r0, M, gd));
        CFD_RET(ps_fill<T>(cx, pv, M, gd));
        CFD_RET(ps_fill<T>(cx, vv, M, gd));
        k_dot2<T><<<gr, bl, 0, cx.s>>>(gd, n0, n1, b, b, r, r, cx.part, cx.pstride);
        CFD_LAUNCH_CHECK();
        return ps_reduce(cx, nblk(n0, n1), 2, S_BB, gd);
    }

    // Synchronous solve: L x = b to ||r|| <= rtol ||b|| with host-side tests,
    // iterate for iterate the oracle's bicgstab(); x must be zero on entry.
    int solve(PsCtx &cx, const T *b, T *x, double rtol, int maxit, int *its_out, double *relres) {
        OpLevel<T> &L0 = lev[0];
        const int n0 = L0.n0, n1 = L0.n1;
        const dim3 gr = grid2d(n0, n1), bl(BX, BY);
        const int nb = nblk(n0, n1);
        T *r = bi[0], *r0 = bi[1], *pv = bi[2], *vv = bi[3], *s = bi[4], *t = bi[5], *ph = bi[6], *sh = bi[7];
        const FArgs<T> a0 = fargs(L0);
        CFD_RET(start(cx, b, x, nullptr));
        k_g_init<<<1, 1, 0, cx.s>>>(cx.scal, rtol, -1, nullptr);
        CFD_LAUNCH_CHECK();
        CFD_RET(ps_sync(cx));
        const double bnorm = std::sqrt(cx.hscal[S_BB]);
        double rnorm = std::sqrt(cx.hscal[S_RR]);
        *its_out = 0;
        if (bnorm == 0.0) { *relres = 0.0; return CFD_OK; }
        const double tol = rtol * bnorm;
        if (rnorm <= tol) { *relres = rnorm / bnorm; return CFD_OK; }
        double rho = cx.hscal[S_RR];
        int it;
        for (it = 1; it <= maxit; it++) {
            if (rho == 0.0) break;
            k_bicg_p<T><<<gr, bl, 0, cx.s>>>(nullptr, n0, n1, it == 1, cx.scal, r, vv, pv);
            CFD_LAUNCH_CHECK();
            CFD_RET(precond(cx, pv, ph, nullptr));
            cx.prof_b(cx.owner, 0);
            k_fapply<T><<<gr, bl, 0, cx.s>>>(nullptr, a0, ph, vv);
            cx.prof_e(cx.owner);
            CFD_LAUNCH_CHECK();
            k_dot2<T><<<gr, bl, 0, cx.s>>>(nullptr, n0, n1, r0, vv, nullptr, nullptr, cx.part, cx.pstride);
            CFD_LAUNCH_CHECK();
            CFD_RET(ps_reduce(cx, nb, 1, S_RV));
            k_bicg_s<T><<<gr, bl, 0, cx.s>>>(nullptr, n0, n1, cx.scal, r, vv, s, cx.part);
            CFD_LAUNCH_CHECK();
            CFD_RET(ps_reduce(cx, nb, 1, S_SS));
            CFD_RET(ps_sync(cx));
            const double snorm = std::sqrt(cx.hscal[S_SS]);
            if (snorm <= tol) {
                k_bicg_x1<T><<<gr, bl, 0, cx.s>>>(n0, n1, cx.scal, ph, x);
                CFD_LAUNCH_CHECK();
                rnorm = snorm;
                break;
            }
            CFD_RET(precond(cx, s, sh, nullptr));
            cx.prof_b(cx.owner, 0);
            k_fapply<T><<<gr, bl, 0, cx.s>>>(nullptr, a0, sh, t);
            cx.prof_e(cx.owner);
            CFD_LAUNCH_CHECK();
            k_dot2<T><<<gr, bl, 0, cx.s>>>(nullptr, n0, n1, t, s, t, t, cx.part, cx.pstride);
            CFD_LAUNCH_CHECK();
            CFD_RET(ps_reduce(cx, nb, 2, S_TS));
            k_set_omega<<<1, 1, 0, cx.s>>>(nullptr, cx.scal);
            CFD_LAUNCH_CHECK();
            k_bicg_xr<T><<<gr, bl, 0, cx.s>>>(nullptr, n0, n1, cx.scal, ph, sh, s, t, r0, x, r, cx.part, cx.pstride);
            CFD_LAUNCH_CHECK();
            CFD_RET(ps_reduce(cx, nb, 1, S_RR));
            k_reduce_sum<<<1, 256, 0, cx.s>>>(nullptr, cx.part + cx.pstride, cx.pstride, nb, 1, cx.scal, S_R0R);
            CFD_LAUNCH_CHECK();
            k_rotate_rho<<<1, 1, 0, cx.s>>>(cx.scal);
            CFD_LAUNCH_CHECK();
            CFD_RET(ps_sync(cx));
            rnorm = std::sqrt(cx.hscal[S_RR]);
            const double omega_h = cx.hscal[S_OMEGA];
            if (rnorm <= tol || omega_h == 0.0) break;
            rho = cx.hscal[S_RHO];
        }
        *its_out = it > maxit ? maxit : it;
        *relres = rnorm / bnorm;
        return CFD_OK;
    }

    // Capturable solve: exactly `maxit` unrolled iterations whose kernels turn into
    // no-ops once the device-side tests (the same ones as solve()) have fired; no
    // host synchronisation.  Result: iterations in S_IT, relative residual in
    // S_RELRES.  rtol comes from `rtol` or, if rtol_slot >= 0, from scal[rtol_slot];
    // `outer` (may be null) skips the whole solve when nonzero.
    int solve_graph(PsCtx &cx, const T *b, T *x, double rtol, int rtol_slot, int maxit, const double *outer) {
        OpLevel<T> &L0 = lev[0];
        const int n0 = L0.n0, n1 = L0.n1;
        const dim3 gr = grid2d(n0, n1), bl(BX, BY);
        const int nb = nblk(n0, n1);
        T *r = bi[0], *r0 = bi[1], *pv = bi[2], *vv = bi[3], *s = bi[4], *t = bi[5], *ph = bi[6], *sh = bi[7];
        const FArgs<T> a0 = fargs(L0);
        const double *dn = cx.scal + S_DONE;
        CFD_RET(start(cx, b, x, outer));
        k_g_init<<<1, 1, 0, cx.s>>>(cx.scal, rtol, rtol_slot, outer);
        CFD_LAUNCH_CHECK();
        for (int it = 1; it <= maxit; it++) {
            k_g_begin<<<1, 1, 0, cx.s>>>(cx.scal, it);
            k_bicg_p<T><<<gr, bl, 0, cx.s>>>(dn, n0, n1, it == 1, cx.scal, r, vv, pv);
            CFD_LAUNCH_CHECK();
            CFD_RET(precond(cx, pv, ph, dn));
            cx.prof_b(cx.owner, 0);
            k_fapply<T><<<gr, bl, 0, cx.s>>>(dn, a0, ph, vv);
            cx.prof_e(cx.owner);
            k_dot2<T><<<gr, bl, 0, cx.s>>>(dn, n0, n1, r0, vv, nullptr, nullptr, cx.part, cx.pstride);
            CFD_LAUNCH_CHECK();
            CFD_RET(ps_reduce(cx, nb, 1, S_RV, dn));
            k_bicg_s<T><<<gr, bl, 0, cx.s>>>(dn, n0, n1, cx.scal, r, vv, s, cx.part);
            CFD_LAUNCH_CHECK();
            CFD_RET(ps_reduce(cx, nb, 1, S_SS, dn));
            k_g_check_s<<<1, 1, 0, cx.s>>>(cx.scal);
            k_bicg_x1g<T><<<gr, bl, 0, cx.s>>>(n0, n1, cx.scal, ph, x);
            CFD_LAUNCH_CHECK();
            CFD_RET(precond(cx, s, sh, dn));
            cx.prof_b(cx.owner, 0);
            k_fapply<T><<<gr, bl, 0, cx.s>>>(dn, a0, sh, t);
            cx.prof_e(cx.owner);
            k_dot2<T><<<gr, bl, 0, cx.s>>>(dn, n0, n1, t, s, t, t, cx.part, cx.pstride);
            CFD_LAUNCH_CHECK();
            CFD_RET(ps_reduce(cx, nb, 2, S_TS, dn));
            k_set_omega<<<1, 1, 0, cx.s>>>(dn, cx.scal);
            k_bicg_xr<T><<<gr, bl, 0, cx.s>>>(dn, n0, n1, cx.scal, ph, sh, s, t, r0, x, r, cx.part, cx.pstride);
            CFD_LAUNCH_CHECK();
            CFD_RET(ps_reduce(cx, nb, 1, S_RR, dn));
            k_reduce_sum<<<1, 256, 0, cx.s>>>(dn, cx.part + cx.pstride, cx.pstride, nb, 1, cx.scal, S_R0R);
            k_g_check_r<<<1, 1, 0, cx.s>>>(cx.scal);
            CFD_LAUNCH_CHECK();
        }
        k_g_final<<<1, 1, 0, cx.s>>>(cx.scal);
        CFD_LAUNCH_CHECK();
        return CFD_OK;
    }
};

}  // namespace cfd
