// Pressure-correction solver on gfx950: face-coefficient operator, geometric
// multigrid V-cycle and right-preconditioned BiCGStab, templated on the
// precision of the vectors and stored coefficients (double, or float for the
// inner solve of the mixed path).  Dot products accumulate in double, scalars
// live in device memory as double.
//
// The operator is the J-weighted flux form of the face velocity of
// solver.cpp:1402-1436 (see p9_coefs).  Per outer iteration the four face
// coefficients (east/north, normal/tangential) of every level are computed once
// in FP64 from the level's metrics and d and stored in T; a row is assembled from
// the faces of P, W and S, so an operator application reads 4 coefficient
// arrays and the vector instead of 5 metric arrays, d and the vector.
//
// Launch structure (one BiCGStab iteration: 5 + 2 x (2 x fine levels + 1) launches,
// two more where the operator application is not fused into the finest up leg):
//  * a fine multigrid level costs two launches: k_mg_down (pre-smoothing from
//    zero, residual, full-weighting restriction and the coarse interior sum, on
//    an LDS tile) and k_mg_up (bilinear prolongation and post-smoothing on an
//    LDS tile); levels of at most 33 x 33 nodes run in one single-workgroup
//    kernel resident in LDS (k_mg_coarse);
//  * the ||s|| and end-of-iteration reductions are finished by one-block kernels
//    (k_bi_fin) that also evaluate the solver's scalar recurrences and
//    convergence tests; alpha and omega, and the coarse-level means of the
//    V-cycle, are totalled from the block partials by the kernels that consume
//    them;
//  * the convergence state lives in device scalars: the same kernel sequence
//    runs eagerly (with host reads of the flags between half-iterations) or
//    captured in a graph (all iterations unrolled; kernels after convergence
//    return at their first instruction).
//
// Slots: every kernel runs blockIdx.z over the slots of a batch of equally sized
// cases (grid2d(...).z = slots).  Slot z's arrays start at z x (the array's size
// per slot), its device scalars at z x SS; a block of a slot that is inactive,
// or whose guard flag is set, returns at once, so slots converge and are
// refilled independently.  A single case is a batch of one slot.
#pragma once

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <vector>

#include "cfd_cavity_kernels.h"

namespace cfd {

constexpr int PS_MAXLEV = 16;
constexpr int PS_NT = BX * BY;

// device scalar slots
enum {
    S_RHO = 0, S_RHO_OLD, S_ALPHA, S_OMEGA, S_RV, S_TS, S_TT, S_SS, S_BB, S_RR, S_SUM, S_R0R,
    S_MASS, S_ENERGY, S_MOM, S_DMAX, S_NUH, S_NUC,
    // BiCGStab control: done flag, iteration that stopped at the ||s|| test,
    // tolerance, iteration count, norms, relative residual, tolerance source
    S_DONE, S_SCONV, S_TOL, S_IT, S_RN, S_BN, S_RELRES, S_RTOL,
    // mixed-precision refinement (FP64 defect)
    S_BB64, S_RR64, S_TOL64, S_RN64, S_BN64, S_RDONE, S_ITTOT, S_RELRES64,
    S_PSUM,   // interior sum of p' (its mean is removed by the correction)
    S_ACTIVE, // slot holds a case (set by the host)
    S_MASS0, S_ENERGY0,  // residual norms of the case's first SIMPLE iteration
    S_FRESH,  // 1 until the statistics of the case's first SIMPLE iteration are finished (set by the host)
    S_PMAX,   // > 0: the slot's BiCGStab iteration cap below the launched count; < 0: the
              // slot's p' is one V-cycle, p^ of the first iteration (set by the host)
    S_NSCAL
};
constexpr int SS = 64;  // device scalars per slot
static_assert(S_NSCAL <= SS, "scalar slots");

__device__ __forceinline__ size_t zslot() { return (size_t)blockIdx.z; }
// slot's scalars; true when the block has nothing to do (inactive slot, or guard set)
__device__ __forceinline__ bool slot_skip(const double *sc, int gs) {
    return sc[S_ACTIVE] == 0.0 || (gs >= 0 && sc[gs] != 0.0);
}

// FP64 geometry of one multigrid level (level 0 aliases the case's arrays)
struct Geom {
    int n0 = 0, n1 = 0;
    double h1 = 0, h2 = 0;
    double *xix = nullptr, *xiy = nullptr, *etx = nullptr, *ety = nullptr, *jac = nullptr;
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
template <typename T>
__device__ __forceinline__ FArgs<T> fz(FArgs<T> a, size_t o) {
    a.fen += o; a.fet += o; a.fnn += o; a.fnt += o;
    return a;
}

__device__ __forceinline__ int blin() { return blockIdx.x + blockIdx.y * gridDim.x; }
__device__ __forceinline__ int tlin() { return threadIdx.x + threadIdx.y * blockDim.x; }
__device__ __forceinline__ bool inner_node(int i, int j, int n0, int n1) {
    return i >= 1 && i <= n0 - 2 && j >= 1 && j <= n1 - 2;
}

// The elementwise kernels take the nodes of a slot in memory order: FLAT_NPT
// nodes per thread, each pass over a contiguous PS_NT-node chunk, so a launch has
// ceil(n0 n1 / (FLAT_NPT PS_NT)) blocks along x (1027 at 1025^2, fewer than the
// resident-block capacity; 5 at 65^2) and every block reads and writes one
// contiguous range, with the loads of all its nodes issued together and one block
// reduction for all of them (measured: one node per thread 9.1 us per k_bi_xr at
// 1025^2, two 8.0-8.4 us, four 7.3 us).
constexpr int FLAT_NPT = 4;
struct FlatNode {
    int i, j;
    size_t P;
    bool in;  // the node exists
};
// the m-th node (0 <= m < FLAT_NPT) of this thread
__device__ __forceinline__ FlatNode flat_node(int n0, int n1, int m) {
    const unsigned q = (blockIdx.x * (unsigned)FLAT_NPT + (unsigned)m) * (unsigned)PS_NT + (unsigned)tlin();
    FlatNode f;
    f.in = q < (unsigned)n0 * (unsigned)n1;
    const unsigned j = f.in ? q / (unsigned)n0 : 0u;
    f.j = (int)j;
    f.i = f.in ? (int)(q - j * (unsigned)n0) : 0;
    f.P = f.in ? (size_t)q : 0;
    return f;
}
inline int nflat(int n0, int n1) { return (int)(((long)n0 * n1 + FLAT_NPT * PS_NT - 1) / (FLAT_NPT * PS_NT)); }

// ---------------------------------------------------------------------------
// Reductions in two launches: a kernel stores one partial per block and value
// (part[k * pstride + block]), a one-block kernel adds them in block order (a
// deterministic result) and evaluates what depends on the totals.  (A last-block
// finish inside the first kernel needs a device-scope fence per block, which
// costs more on gfx950 than the extra launch.)
template <int NV, int NT = PS_NT>
__device__ __forceinline__ void block_partials(double (&v)[NV], double *part, int pstride) {
    block_sums<NV, NT>(v);
    if (tlin() == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) part[(size_t)k * pstride + blin()] = v[k];
    }
}

// totals of NV partial arrays, summed by the first VS threads of the NT-thread
// block in a fixed order: each of them keeps eight running sums per value, and
// every load of a pass over 8 * VS partials is issued before any is added (one
// memory round trip per pass instead of one per partial); valid in thread 0.
// The other threads add zeros, so for a given VS the bits do not depend on NT.
template <int NV, int NT = PS_NT, int VS = NT>
__device__ __forceinline__ void sum_partials(const double *part, int pstride, int nparts, double (&tot)[NV]) {
    static_assert(VS <= NT && VS % 64 == 0, "summing lanes");
    constexpr int U = 8;
    const int t = tlin();
    const bool lane = t < VS;
    const int last = max(nparts - 1, 0);
    double s[NV][U];
    {  // the first pass, outside the loop: straight-line loads the compiler issues early
        double v[NV][U];
#pragma unroll
        for (int k = 0; k < NV; k++)
#pragma unroll
            for (int u = 0; u < U; u++) v[k][u] = part[(size_t)k * pstride + min(u * VS + t, last)];
#pragma unroll
        for (int k = 0; k < NV; k++)
#pragma unroll
            for (int u = 0; u < U; u++) s[k][u] = lane && u * VS + t < nparts ? v[k][u] : 0.0;
    }
    for (int t0 = U * VS; t0 < nparts; t0 += U * VS) {
        double v[NV][U];
#pragma unroll
        for (int k = 0; k < NV; k++)
#pragma unroll
            for (int u = 0; u < U; u++) v[k][u] = part[(size_t)k * pstride + min(t0 + u * VS + t, nparts - 1)];
#pragma unroll
        for (int k = 0; k < NV; k++)
#pragma unroll
            for (int u = 0; u < U; u++) s[k][u] += lane && t0 + u * VS + t < nparts ? v[k][u] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < NV; k++)
        tot[k] = ((s[k][0] + s[k][1]) + (s[k][2] + s[k][3])) + ((s[k][4] + s[k][5]) + (s[k][6] + s[k][7]));
    block_sums<NV, NT>(tot);
}

// threads of the one-block kernels that finish a reduction
constexpr int FIN_NT = 1024;
// the BiCGStab finishes: at 1025^2 their phases have at most 2048 partials per
// value (one pass of eight loads per thread at 256 threads; larger grids take
// more passes), and four wavefronts reduce and start faster than sixteen
constexpr int BI_FIN_NT = 256;
// blocks x partials above which alpha and omega get their own finish kernels
constexpr long BI_FUSE_MAX = 8L << 20;

// ---------------------------------------------------------------------------
// operator rows

// Row (i, j) of the stored operator applied to x, with x read through `xat`
// at the neighbours clamped into the interior (Neumann ghost).
template <typename T>
__device__ __forceinline__ RowFaces<T> row_faces(const FArgs<T> &a, int i, int j) {
    return load_row_faces<T>(a.n0, a.fen, a.fet, a.fnn, a.fnt, i, j);
}

// the same row from faces already loaded
template <typename T, typename F>
__device__ __forceinline__ T frow_v(const RowFaces<T> &f, int n0, int n1, int i, int j, F xat) {
    T c[9];
    p9_row_from<T>(f, c);
    T s = T(0);
#pragma unroll
    for (int dj = -1; dj <= 1; dj++)
#pragma unroll
        for (int di = -1; di <= 1; di++) {
            const int ii = clampi(i + di, 1, n0 - 2), jj = clampi(j + dj, 1, n1 - 2);
            s += c[(dj + 1) * 3 + (di + 1)] * xat(ii, jj);
        }
    return s;
}

template <typename T, typename F>
__device__ __forceinline__ T frow_f(const FArgs<T> &a, int i, int j, F xat) {
    return frow_v<T>(row_faces<T>(a, i, j), a.n0, a.n1, i, j, xat);
}

// The same row over an LDS tile of row stride S whose every entry holds the value
// of the node it covers clamped into the interior (the Neumann ghost already in
// place): the neighbours are read at fixed offsets from `c`, the entry of the
// row's own node, with the terms added in frow_v's order, so the result has
// frow_v's bits.
template <typename T, int S>
__device__ __forceinline__ T frow_tile(const RowFaces<T> &f, const T *c) {
    T w[9];
    p9_row_from<T>(f, w);
    T s = T(0);
#pragma unroll
    for (int dj = -1; dj <= 1; dj++)
#pragma unroll
        for (int di = -1; di <= 1; di++) s += w[(dj + 1) * 3 + (di + 1)] * c[dj * S + di];
    return s;
}

// entry of node (gi, gj) in an LDS tile of row stride S whose first entry is node
// (fi0 - H, fj0 - H)
template <int H, int S>
__device__ __forceinline__ int tile_at(int gi, int gj, int fi0, int fj0) {
    return (gj - fj0 + H) * S + (gi - fi0 + H);
}

// Loads of a strided block loop issued together: element k = tid + m * NT of a
// COUNT-element region gets load(k) for every m before any is used, so a staging
// phase costs one memory round trip instead of one per iteration.  The loads are
// unconditional (no branch between them): load(k) must read valid addresses for
// every k < COUNT (clamped ones where the element is unused), and k past the
// region is clamped to COUNT - 1.  body(k, r) then runs for k < COUNT in order.
template <typename R, int NIT>
struct Batch {
    R r[NIT];
};

template <int COUNT, int NT, typename L>
__device__ __forceinline__ auto batch_load(int tid, L load) {
    constexpr int NIT = (COUNT + NT - 1) / NT;
    Batch<decltype(load(0)), NIT> b;
#pragma unroll
    for (int m = 0; m < NIT; m++) b.r[m] = load(min(tid + m * NT, COUNT - 1));
    return b;
}

template <int COUNT, int NT, typename R, int NIT, typename B>
__device__ __forceinline__ void batch_run(int tid, const Batch<R, NIT> &b, B body) {
#pragma unroll
    for (int m = 0; m < NIT; m++) {
        const int k = tid + m * NT;
        if (k < COUNT) body(k, b.r[m]);
    }
}

template <typename T>
__device__ __forceinline__ T frow(const FArgs<T> &a, const T *__restrict__ x, int i, int j) {
    return frow_f<T>(a, i, j, [&](int ii, int jj) { return x[idx(ii, jj, a.n0)]; });
}

// y = L x (interior), 0 on the border
template <typename T>
__global__ __launch_bounds__(PS_NT) void k_fapply(FArgs<T> a, const T *__restrict__ x, T *__restrict__ y) {
    const size_t o = zslot() * a.n0 * a.n1;
    a = fz(a, o); x += o; y += o;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.n0 || j >= a.n1) return;
    y[idx(i, j, a.n0)] = inner_node(i, j, a.n0, a.n1) ? frow<T>(a, x, i, j) : T(0);
}

// ---------------------------------------------------------------------------
// per-outer-iteration set-up of every level in one launch each

template <typename T>
struct LevTab {
    int nlev;
    int bofs[PS_MAXLEV + 1];
    int bx[PS_MAXLEV];
    int n0[PS_MAXLEV], n1[PS_MAXLEV];
    double h1[PS_MAXLEV], h2[PS_MAXLEV];
    double ih1sq[PS_MAXLEV], ih2sq[PS_MAXLEV], i4h1h2[PS_MAXLEV];  // 1/h1^2, 1/h2^2, 1/(4 h1 h2)
    const double *xix[PS_MAXLEV], *xiy[PS_MAXLEV], *etx[PS_MAXLEV], *ety[PS_MAXLEV], *jac[PS_MAXLEV];
    T *fen[PS_MAXLEV], *fet[PS_MAXLEV], *fnn[PS_MAXLEV], *fnt[PS_MAXLEV], *invd[PS_MAXLEV];
    const double *d0;
    int dn0;
    double *f64[4];  // optional FP64 copy of level 0's faces (en, et, nn, nt)
};

template <typename T>
__device__ __forceinline__ bool lev_of_block(const LevTab<T> &A, int &l, int &i, int &j) {
    const int b = blockIdx.x;
    l = 0;
    while (l + 1 < A.nlev && A.bofs[l + 1] <= b) l++;
    const int lb = xcd_order(b - A.bofs[l], A.bofs[l + 1] - A.bofs[l]);  // neighbours on one XCD
    i = (lb % A.bx[l]) * BX + (threadIdx.x % BX);
    j = (lb / A.bx[l]) * BY + (threadIdx.x / BX);
    return i < A.n0[l] && j < A.n1[l];
}

// Metric terms and d at one node, loaded once and shared by the faces that touch it.
struct FNode {
    double x, y, e, f, j, d;  // xi_x, xi_y, eta_x, eta_y, J, d
};
__device__ __forceinline__ FNode fnode_load(const double *__restrict__ xix, const double *__restrict__ xiy,
                                            const double *__restrict__ etx, const double *__restrict__ ety,
                                            const double *__restrict__ jac, const double *__restrict__ d0,
                                            int n0, int n1, int dn0, int lv, int i, int j) {
    const int ic = clampi(i, 0, n0 - 1), jc = clampi(j, 0, n1 - 1);
    const unsigned P = (unsigned)jc * (unsigned)n0 + (unsigned)ic;
    return FNode{xix[P], xiy[P], etx[P], ety[P], jac[P],
                 d0[((unsigned)jc << lv) * (unsigned)dn0 + ((unsigned)ic << lv)]};
}

// Face (normal kn, tangential kt) between node P and its east (dir 0) or north
// (dir 1) neighbour Q: the arithmetic of p9_face with the grid divisions taken as
// host-computed reciprocals (ih1sq = 1/h1^2, ih2sq = 1/h2^2, i4 = 1/(4 h1 h2)); d
// is level 0's d injected at (i << lv, j << lv).  The caller zeroes faces that
// touch a wall node.
// The face's metric factors (a: normal, b: tangential; d does not enter them).
__device__ __forceinline__ void p9_face_ab(const FNode &P, const FNode &Q, int dir, double &a, double &b) {
    const double half = 0.5;
    if (dir == 0) {
        const double mx = half * (P.j * P.x + Q.j * Q.x), my = half * (P.j * P.y + Q.j * Q.y);
        a = mx * (P.x + Q.x) + my * (P.y + Q.y);
        b = mx * (P.e + Q.e) + my * (P.f + Q.f);
    } else {
        const double mx = half * (P.j * P.e + Q.j * Q.e), my = half * (P.j * P.f + Q.j * Q.f);
        a = mx * (P.e + Q.e) + my * (P.f + Q.f);
        b = mx * (P.x + Q.x) + my * (P.y + Q.y);
    }
}
// the face from its metric factors and the d of its two nodes
__device__ __forceinline__ void p9_face_d(double a, double b, double dP, double dQ, int dir, double ih1sq,
                                          double ih2sq, double i4, double &kn, double &kt) {
    const double half = 0.5;
    const double de = half * (dP + dQ);
    kn = half * de * a * (dir == 0 ? ih1sq : ih2sq);
    kt = half * de * b * i4;
}
// (p9_face_ab then p9_face_d, written out)
__device__ __forceinline__ void p9_face_of(const FNode &P, const FNode &Q, int dir, double ih1sq, double ih2sq,
                                           double i4, double &kn, double &kt) {
    const double half = 0.5;
    if (dir == 0) {
        const double mx = half * (P.j * P.x + Q.j * Q.x), my = half * (P.j * P.y + Q.j * Q.y);
        const double a = mx * (P.x + Q.x) + my * (P.y + Q.y);
        const double b = mx * (P.e + Q.e) + my * (P.f + Q.f);
        const double de = half * (P.d + Q.d);
        kn = half * de * a * ih1sq;
        kt = half * de * b * i4;
    } else {
        const double mx = half * (P.j * P.e + Q.j * Q.e), my = half * (P.j * P.f + Q.j * Q.f);
        const double a = mx * (P.e + Q.e) + my * (P.f + Q.f);
        const double b = mx * (P.x + Q.x) + my * (P.y + Q.y);
        const double dn = half * (P.d + Q.d);
        kn = half * dn * a * ih2sq;
        kt = half * dn * b * i4;
    }
}

// Face coefficients of every node of every level, and the inverse diagonal of
// the clamped operator (every coefficient whose neighbour clamps onto the node)
// at interior nodes: the thread also evaluates its west neighbour's east face and
// its south neighbour's north face, so the row needs no second launch.  A block's
// BX x BY tile and a one-node ring have their metric terms and d staged in LDS
// (loads clamped into the grid, all issued before any is used); a face touching a
// wall node is zeroed after its arithmetic.
template <typename T>
__global__ __launch_bounds__(PS_NT) void k_faces_all(const double *scal, LevTab<T> A) {
    if (slot_skip(scal + zslot() * SS, -1)) return;
    int l, i, j;
    const bool own = lev_of_block(A, l, i, j);
    const int n0 = A.n0[l], n1 = A.n1[l];
    const size_t o = zslot() * n0 * n1, o0 = zslot() * A.dn0 * A.n1[0];
    const size_t P = o + idx(min(i, n0 - 1), min(j, n1 - 1), n0);
    const double *d0 = A.d0 + o0;
    const double *xix = A.xix[l] + o, *xiy = A.xiy[l] + o, *etx = A.etx[l] + o, *ety = A.ety[l] + o,
                 *jac = A.jac[l] + o;
    const double r1 = A.ih1sq[l], r2 = A.ih2sq[l], r12 = A.i4h1h2[l];
    const int dn0 = A.dn0;
    constexpr int SX = BX + 2, SY = BY + 2;
    __shared__ FNode s_n[SX * SY];
    {
        const int tid = threadIdx.x;
        const int i0 = i - tid % BX, j0 = j - tid / BX;  // the tile's origin
        const auto ln = batch_load<SX * SY, PS_NT>(tid, [&](int k) {
            return fnode_load(xix, xiy, etx, ety, jac, d0, n0, n1, dn0, l, i0 - 1 + k % SX, j0 - 1 + k / SX);
        });
        batch_run<SX * SY, PS_NT>(tid, ln, [&](int k, const FNode &v) { s_n[k] = v; });
    }
    __syncthreads();
    if (!own) return;
    const int sc = (threadIdx.x / BX + 1) * SX + threadIdx.x % BX + 1;
    const FNode nP = s_n[sc], nE = s_n[sc + 1], nW = s_n[sc - 1], nN = s_n[sc + SX], nS = s_n[sc - SX];
    double en, et, nn, nt, wn, wt, sn, st;
    p9_face_of(nP, nE, 0, r1, r2, r12, en, et);
    p9_face_of(nP, nN, 1, r1, r2, r12, nn, nt);
    p9_face_of(nW, nP, 0, r1, r2, r12, wn, wt);
    p9_face_of(nS, nP, 1, r1, r2, r12, sn, st);
    // east face of (i, j) valid for 1 <= i <= n0 - 3, 1 <= j <= n1 - 2; north face
    // for 1 <= j <= n1 - 3, 1 <= i <= n0 - 2 (zero on faces touching a wall node)
    auto ev = [&](int ii, int jj) { return ii >= 1 && ii <= n0 - 3 && jj >= 1 && jj <= n1 - 2; };
    auto nv = [&](int ii, int jj) { return jj >= 1 && jj <= n1 - 3 && ii >= 1 && ii <= n0 - 2; };
    if (!ev(i, j)) en = et = 0.0;
    if (!nv(i, j)) nn = nt = 0.0;
    if (!ev(i - 1, j)) wn = wt = 0.0;
    if (!nv(i, j - 1)) sn = st = 0.0;
    A.fen[l][P] = (T)en;
    A.fet[l][P] = (T)et;
    A.fnn[l][P] = (T)nn;
    A.fnt[l][P] = (T)nt;
    if (l == 0 && A.f64[0]) {
        A.f64[0][P] = en;
        A.f64[1][P] = et;
        A.f64[2][P] = nn;
        A.f64[3][P] = nt;
    }
    if (!inner_node(i, j, n0, n1)) return;
    T c[9];
    p9_row_from<T>(RowFaces<T>{(T)en, (T)et, (T)wn, (T)wt, (T)nn, (T)nt, (T)sn, (T)st}, c);
    T s = T(0);
    for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++)
            if (clampi(i + di, 1, n0 - 2) == i && clampi(j + dj, 1, n1 - 2) == j) s += c[(dj + 1) * 3 + (di + 1)];
    A.invd[l][P] = T(1) / s;
}

// ---------------------------------------------------------------------------
// Fine multigrid levels.  Both kernels work on a 64 x 16 tile of fine nodes per
// 256-thread block and stage their intermediate sweeps in LDS with the halo the
// next stencil needs, so a level's pre-smoothing, residual and restriction read
// the level's arrays once, and so do its prolongation and post-smoothing.
// Jacobi sweep: x <- x + w (b - L x) / diag; the first sweep from zero is
// x = w b / diag (no neighbours).

constexpr int MG_FX = 64, MG_FY = 8;   // fine tile
constexpr int MG_NT_WIDE = 1024;       // threads per tile on levels with few tiles
constexpr int MG_NT_NARROW = 512;      // threads per tile elsewhere
constexpr int MG_NT_SLIM = 256;        // a third size the autotuner may choose
// block-size classes of a multigrid launch (index into the sizes above)
enum { MGB_SLIM = 0, MGB_NARROW = 1, MGB_WIDE = 2, MGB_COUNT = 3 };
// calls f(std::integral_constant<int, threads>) for block-size class nb
template <typename F>
inline void by_block_size(int nb, F f) {
    switch (nb) {
    case MGB_SLIM: f(std::integral_constant<int, MG_NT_SLIM>{}); break;
    case MGB_WIDE: f(std::integral_constant<int, MG_NT_WIDE>{}); break;
    default: f(std::integral_constant<int, MG_NT_NARROW>{}); break;
    }
}
// Lanes that sum a tile's reductions: every tile runs at least this many threads,
// and the sums use exactly these lanes in a fixed order, so a tile's partials have
// the same bits whichever block size runs it.
constexpr int MG_SUM_LANES = 256;
constexpr long MG_WIDE_BELOW = 256;    // tiles per launch below which tiles get MG_NT_WIDE threads
constexpr int MG_CX = MG_FX / 2, MG_CY = MG_FY / 2;

// Tiles of a multigrid level cover nodes 0 .. n - 2 in each direction
// (tiles_inner): 16 x 128 tiles at 1025^2 instead of 17 x 129.  The wall nodes that
// fall outside every tile (column n0 - 1 when the tile width divides n0 - 1, row
// n1 - 1 likewise) are set by the tiles next to them (edge_walls).

// Calls f(node index) for the east-wall nodes of a TX x TY tile at (i0, j0) when the
// tile's east edge is the column before the wall, for the north-wall nodes likewise,
// and for the corner when both hold.
template <int TX, int TY, typename F>
__device__ __forceinline__ void edge_walls(int tid, int nt, int i0, int j0, int n0, int n1, F f) {
    const bool e = i0 + TX == n0 - 1, n = j0 + TY == n1 - 1;
    if (!e && !n) return;
    for (int k = tid; k < TY + TX + 1; k += nt) {
        int i, j;
        if (k < TY) {
            if (!e) continue;
            i = n0 - 1;
            j = j0 + k;
        } else if (k < TY + TX) {
            if (!n) continue;
            i = i0 + k - TY;
            j = n1 - 1;
        } else {
            if (!(e && n)) continue;
            i = n0 - 1;
            j = n1 - 1;
        }
        if (i < n0 && j < n1) f(idx(i, j, n0));
    }
}

template <typename T>
struct DownArgs {
    FArgs<T> a;
    const T *invd;
    const T *bsrc;       // right-hand side (restricted, or the preconditioner input on level 0)
    const double *bpart; // block partials of bsrc's interior sum, whose mean is removed first (null: none)
    int nbp;             // number of those partials
    T *bout;             // bsrc with its mean removed, kept for the post-smoothing (null: not kept)
    T *x;                // smoothed x
    int m0, m1;          // coarse level
    T *craw;             // restricted residual on the coarse level (mean not removed)
    double *cpart;       // block partials of craw's interior sum (one per block)
    T w;
    size_t lps;          // per-slot stride of the partial arrays
    const double *scal;  // device scalars (guard)
    int gs;              // guard slot (-1: none)
};

template <int NPRE>
struct DownTile {
    static constexpr int HB = NPRE + 1;  // halo of the staged right-hand side
    static constexpr int BWX = MG_FX + 2 * HB, BWY = MG_FY + 2 * HB;
    static constexpr int SX = MG_FX + 4, SY = MG_FY + 4;  // smoothed x (halo 2)
    static constexpr int RX = MG_FX + 2, RY = MG_FY + 2;  // residual (halo 1)
    static constexpr int NB = BWX * BWY;
    static constexpr int NX1 = NPRE == 2 ? NB : 0;        // first sweep (reused for the residual)
    static constexpr int NXS = NPRE >= 1 ? SX * SY : 0;
    static constexpr int NR = NPRE == 2 ? 0 : (NPRE == 1 ? RX * RY : 0);
    static constexpr int TOTAL = NB + NX1 + NXS + NR;
};

// FP32 tiles issue the smoothing loads together with the right-hand side's (one
// memory round trip per launch); FP64 tiles, which would need twice the registers,
// issue them after the first phase.
template <typename T>
constexpr bool mg_early() {
    return sizeof(T) == 4;
}

// MEAN: the right-hand side's interior mean is removed first (A.bpart holds the
// partial sums); false on level 0, where it is not, so no registers are held for it
template <typename T, int NPRE, int NTB, bool MEAN>
__global__ __launch_bounds__(NTB) void k_mg_down(DownArgs<T> A) {
    const double *guard = A.scal + zslot() * SS;
    {
        const size_t of = zslot() * A.a.n0 * A.a.n1, oc = zslot() * A.m0 * A.m1, op = zslot() * A.lps;
        A.a = fz(A.a, of);
        A.invd += of;
        A.bsrc += of;
        if (A.bout) A.bout += of;
        A.x += of;
        A.craw += oc;
        if (MEAN) A.bpart += op;
        A.cpart += op;
    }
    using TL = DownTile<NPRE>;
    __shared__ T lds[TL::TOTAL];
    T *s_b = lds;
    T *s_x1 = lds + TL::NB;
    T *s_xs = s_x1 + TL::NX1;
    T *s_r = NPRE == 2 ? s_x1 : (NPRE == 1 ? s_xs + TL::NXS : s_b);
    const int tid = tlin();
    const int n0 = A.a.n0, n1 = A.a.n1;
    const int2 tl = xcd_tile();
    const int fi0 = tl.x * MG_FX, fj0 = tl.y * MG_FY;
    const T w = A.w;
    // right-hand side (and 1/diag for a first sweep staged here), halo HB
    struct BI {
        T b, id;
    };
    const auto lb = batch_load<TL::NB, NTB>(tid, [&](int k) {
        const int gi = fi0 - TL::HB + k % TL::BWX, gj = fj0 - TL::HB + k / TL::BWX;
        const size_t P = idx(clampi(gi, 1, n0 - 2), clampi(gj, 1, n1 - 2), n0);
        return BI{A.bsrc[P], NPRE == 2 ? A.invd[P] : T(0)};
    });
    // faces and 1/diag of the smoothed region (halo 2), also used by the residual
    struct XR {
        RowFaces<T> f;
        T id;
    };
    constexpr int NXR = NPRE >= 1 ? TL::SX * TL::SY : 1;
    auto xs_load = [&]() {
        return batch_load<NXR, NTB>(tid, [&](int k) {
            const int gi = clampi(fi0 - 2 + k % TL::SX, 1, n0 - 2), gj = clampi(fj0 - 2 + k / TL::SX, 1, n1 - 2);
            return XR{row_faces<T>(A.a, gi, gj), A.invd[idx(gi, gj, n0)]};
        });
    };
    decltype(xs_load()) lx;
    if (NPRE >= 1 && mg_early<T>()) lx = xs_load();
    if (slot_skip(guard, A.gs)) return;
    double mean = 0.0;
    if (MEAN) {
        double tot[1];
        sum_partials<1, NTB, MG_SUM_LANES>(A.bpart, 0, A.nbp, tot);
        __shared__ double s_mean;
        if (tid == 0) s_mean = tot[0] / ((double)(n0 - 2) * (double)(n1 - 2));
        __syncthreads();
        mean = s_mean;
    }
    // Every entry of s_b, s_x1, s_xs and s_r holds the value of its node clamped into
    // the interior (the loads are at those nodes), so the sweeps and the residual
    // read their neighbours at fixed offsets (frow_tile); the restriction reads
    // interior nodes only, and the wall values written to memory stay zero.
    // mean-free right-hand side; first sweep from zero (no neighbours) in place
    batch_run<TL::NB, NTB>(tid, lb, [&](int k, const BI &r) {
        const T bv = MEAN ? (T)((double)r.b - mean) : r.b;
        s_b[k] = bv;
        if (NPRE == 2) s_x1[k] = w * bv * r.id;
    });
    if (NPRE >= 1 && !mg_early<T>()) lx = xs_load();
    __syncthreads();
    if (NPRE >= 1) {
        constexpr int HB = TL::HB, BWX = TL::BWX;
        batch_run<NXR, NTB>(tid, lx, [&](int k, const XR &r) {
            const int gi = clampi(fi0 - 2 + k % TL::SX, 1, n0 - 2), gj = clampi(fj0 - 2 + k / TL::SX, 1, n1 - 2);
            const int kb = tile_at<HB, BWX>(gi, gj, fi0, fj0);
            T v;
            if (NPRE == 1) {
                v = w * s_b[kb] * r.id;
            } else {
                const T *c = s_x1 + kb;
                v = c[0] + w * (s_b[kb] - frow_tile<T, BWX>(r.f, c)) * r.id;
            }
            s_xs[k] = v;
        });
        __syncthreads();
        // residual (halo 1) on the same elements, the ring of the region left out
        batch_run<NXR, NTB>(tid, lx, [&](int k, const XR &r) {
            const int li = k % TL::SX, lj = k / TL::SX;
            if (li < 1 || li > TL::SX - 2 || lj < 1 || lj > TL::SY - 2) return;
            const int gi = clampi(fi0 - 2 + li, 1, n0 - 2), gj = clampi(fj0 - 2 + lj, 1, n1 - 2);
            const T ax = frow_tile<T, TL::SX>(r.f, s_xs + tile_at<2, TL::SX>(gi, gj, fi0, fj0));
            s_r[(lj - 1) * TL::RX + (li - 1)] = s_b[tile_at<HB, BWX>(gi, gj, fi0, fj0)] - ax;
        });
    }
    // owned nodes: smoothed x and the mean-free right-hand side
    for (int k = tid; k < MG_FX * MG_FY; k += NTB) {
        const int li = k % MG_FX, lj = k / MG_FX;
        const int gi = fi0 + li, gj = fj0 + lj;
        if (gi >= n0 || gj >= n1) continue;
        const size_t P = idx(gi, gj, n0);
        const bool in = inner_node(gi, gj, n0, n1);
        A.x[P] = NPRE >= 1 && in ? s_xs[(lj + 2) * TL::SX + (li + 2)] : T(0);
        if (A.bout) A.bout[P] = in ? s_b[(lj + TL::HB) * TL::BWX + (li + TL::HB)] : T(0);
    }
    edge_walls<MG_FX, MG_FY>(tid, NTB, fi0, fj0, n0, n1, [&](size_t P) {
        A.x[P] = T(0);
        if (A.bout) A.bout[P] = T(0);
    });
    edge_walls<MG_CX, MG_CY>(tid, NTB, tl.x * MG_CX, tl.y * MG_CY, A.m0, A.m1, [&](size_t P) { A.craw[P] = T(0); });
    __syncthreads();
    // full weighting onto the coarse tile, one coarse node per thread (threads
    // beyond the tile's MG_CX * MG_CY coarse nodes have none)
    double cs[1] = {0.0};
    if (NTB <= MG_CX * MG_CY || tid < MG_CX * MG_CY) {
        const int ci = tid % MG_CX, cj = tid / MG_CX;
        const int I = tl.x * MG_CX + ci, J = tl.y * MG_CY + cj;
        if (I < A.m0 && J < A.m1) {
            T v = T(0);
            if (inner_node(I, J, A.m0, A.m1)) {
                // residual-region index of fine node (2I + di, 2J + dj)
                auto R = [&](int di, int dj) {
                    const int li = 2 * ci + 1 + di, lj = 2 * cj + 1 + dj;
                    return NPRE == 0 ? s_b[(lj + TL::HB - 1) * TL::BWX + (li + TL::HB - 1)] : s_r[lj * TL::RX + li];
                };
                v = (T(4) * R(0, 0) + T(2) * (R(-1, 0) + R(1, 0) + R(0, -1) + R(0, 1)) +
                     (R(-1, -1) + R(1, -1) + R(-1, 1) + R(1, 1))) /
                    T(16);
                cs[0] = (double)v;
            }
            A.craw[idx(I, J, A.m0)] = v;
        }
    }
    block_partials<1, NTB>(cs, A.cpart, 0);
}

template <typename T>
struct UpArgs {
    FArgs<T> a;
    const T *invd;
    const T *b;    // right-hand side of the level (mean removed)
    const T *x;    // pre-smoothed x
    int m0, m1;
    const T *xc;   // coarse-level correction
    T *out;        // x + prolongation, post-smoothed
    T w;
    const double *scal;  // device scalars (guard)
    int gs;              // guard slot (-1: none)
};

template <typename T, int NPOST, int NTB>
__global__ __launch_bounds__(NTB) void k_mg_up(UpArgs<T> A) {
    const double *guard = A.scal + zslot() * SS;
    {
        const size_t of = zslot() * A.a.n0 * A.a.n1, oc = zslot() * A.m0 * A.m1;
        A.a = fz(A.a, of);
        A.invd += of;
        A.b += of;
        A.x += of;
        A.out += of;
        A.xc += oc;
    }
    constexpr int YX = MG_FX + 2 * NPOST, YY = MG_FY + 2 * NPOST;  // x + e, halo NPOST
    // first post-sweep region (halo NPOST - 1): its elements also produce the output
    constexpr int QX = NPOST >= 1 ? MG_FX + 2 * (NPOST - 1) : 1, QY = NPOST >= 1 ? MG_FY + 2 * (NPOST - 1) : 1;
    __shared__ T s_y[YX * YY];
    __shared__ T s_z[NPOST == 2 ? QX * QY : 1];
    const int tid = tlin();
    const int n0 = A.a.n0, n1 = A.a.n1, m0 = A.m0, m1 = A.m1;
    const int2 tl = xcd_tile();
    const int fi0 = tl.x * MG_FX, fj0 = tl.y * MG_FY;
    const T w = A.w;
    const T *xc = A.xc;
    // x and the coarse values its bilinear interpolation uses, halo NPOST
    struct YR {
        T x, c00, c10, c01, c11;
    };
#define XC(a_, b_) xc[idx(clampi((a_), 1, m0 - 2), clampi((b_), 1, m1 - 2), m0)]
    const auto ly = batch_load<YX * YY, NTB>(tid, [&](int k) {
        const int gi = clampi(fi0 - NPOST + k % YX, 1, n0 - 2), gj = clampi(fj0 - NPOST + k / YX, 1, n1 - 2);
        const int I = gi >> 1, J = gj >> 1;
        return YR{A.x[idx(gi, gj, n0)], XC(I, J), XC(I + 1, J), XC(I, J + 1), XC(I + 1, J + 1)};
    });
#undef XC
    // faces, right-hand side and 1/diag of the first post-sweep region
    struct PR {
        RowFaces<T> f;
        T b, id;
    };
    auto q_load = [&]() {
        return batch_load<QX * QY, NTB>(tid, [&](int k) {
            const int gi = clampi(fi0 - (NPOST - 1) + k % QX, 1, n0 - 2);
            const int gj = clampi(fj0 - (NPOST - 1) + k / QX, 1, n1 - 2);
            const size_t P = idx(gi, gj, n0);
            return PR{row_faces<T>(A.a, gi, gj), A.b[P], A.invd[P]};
        });
    };
    decltype(q_load()) lq;
    if (NPOST >= 1 && mg_early<T>()) lq = q_load();
    if (slot_skip(guard, A.gs)) return;
    edge_walls<MG_FX, MG_FY>(tid, NTB, fi0, fj0, n0, n1, [&](size_t P) { A.out[P] = T(0); });
    // Every entry of s_y and s_z holds the value of its node clamped into the
    // interior (the loads are at those nodes), so the sweeps read their neighbours
    // at fixed offsets (frow_tile); the wall values written to memory stay zero.
    batch_run<YX * YY, NTB>(tid, ly, [&](int k, const YR &r) {
        const int ui = fi0 - NPOST + k % YX, uj = fj0 - NPOST + k / YX;
        const int gi = clampi(ui, 1, n0 - 2), gj = clampi(uj, 1, n1 - 2);
        T e;
        if ((gi & 1) == 0 && (gj & 1) == 0) e = r.c00;
        else if ((gj & 1) == 0) e = T(0.5) * (r.c00 + r.c10);
        else if ((gi & 1) == 0) e = T(0.5) * (r.c00 + r.c01);
        else e = T(0.25) * (r.c00 + r.c10 + r.c01 + r.c11);
        const T v = r.x + e;
        if (NPOST == 0) {
            if (ui < n0 && uj < n1) A.out[idx(ui, uj, n0)] = inner_node(ui, uj, n0, n1) ? v : T(0);
        } else {
            s_y[k] = v;
        }
    });
    if (NPOST == 0) return;
    if (!mg_early<T>()) lq = q_load();
    __syncthreads();
    // first post sweep on the Q region (NPOST == 2), output on its inner part
    if (NPOST == 2) {
        batch_run<QX * QY, NTB>(tid, lq, [&](int k, const PR &r) {
            const int gi = clampi(fi0 - 1 + k % QX, 1, n0 - 2), gj = clampi(fj0 - 1 + k / QX, 1, n1 - 2);
            const T *c = s_y + tile_at<2, YX>(gi, gj, fi0, fj0);
            s_z[k] = c[0] + w * (r.b - frow_tile<T, YX>(r.f, c)) * r.id;
        });
        __syncthreads();
    }
    batch_run<QX * QY, NTB>(tid, lq, [&](int k, const PR &r) {
        constexpr int H = NPOST - 1;  // ring of the Q region outside the tile
        const int li = k % QX - H, lj = k / QX - H;
        if (li < 0 || li >= MG_FX || lj < 0 || lj >= MG_FY) return;
        const int gi = fi0 + li, gj = fj0 + lj;
        if (gi >= n0 || gj >= n1) return;
        T v = T(0);
        if (inner_node(gi, gj, n0, n1)) {
            const T *c = NPOST == 1 ? s_y + tile_at<1, YX>(gi, gj, fi0, fj0) : s_z + tile_at<1, QX>(gi, gj, fi0, fj0);
            const T ax = NPOST == 1 ? frow_tile<T, YX>(r.f, c) : frow_tile<T, QX>(r.f, c);
            v = c[0] + w * (r.b - ax) * r.id;
        }
        A.out[idx(gi, gj, n0)] = v;
    });
}

// ---------------------------------------------------------------------------
// Coarse-level V-cycle in one workgroup.  Levels of at most 33 x 33 nodes are
// latency-bound: a launch per sweep would cost more than the sweep, and so would
// a global-memory round trip per sweep.  Every thread owns one interior node of
// each level (interior node q = tid; at most 31 x 31 = 961 of them) and keeps
// that node's operator row, 1/diag and right-hand side in registers for the whole
// V-cycle; LDS holds only each level's iterate and its ping-pong buffer, whose
// wall entries stay zero.  The row is the stored operator's with the Neumann
// ghosts (neighbours clamped into the interior) folded into the interior
// coefficients, so a sweep is nine LDS reads and nine multiply-adds per node.
// Levels of at most 64 interior nodes are owned by wavefront 0 alone, whose
// sweeps need no block barrier.  The set-up issues every global load (faces and
// 1/diag of every level, level l0's right-hand side) before any is used.

constexpr int CW = 1024;          // threads of the coarse workgroup
constexpr int CW_NODES = 1089;    // first level handled: at most 33 x 33 nodes
constexpr int CN_MAX = 1536;      // nodes of all LDS-resident levels together
constexpr int CW_ARR = 2;         // LDS arrays per level: x and its ping-pong buffer
constexpr int CW_LEVS = 4;        // at most this many levels in the kernel (a 33 x 33 level has four)
constexpr int CW_WAVE = 64;       // levels with at most this many interior nodes run in wavefront 0

template <typename T>
struct CLev {
    int n0, n1;
    const T *fen, *fet, *fnn, *fnt, *invd;
};
template <typename T>
struct CoarseArgs {
    int nlev;
    int pre, post, coarse;
    T omega;
    CLev<T> L[PS_MAXLEV];
    int off[PS_MAXLEV];  // LDS offset of each level's arrays (levels >= l0)
};

// total of v over a workgroup of NTH threads, in every thread
template <int NTH = CW>
__device__ __forceinline__ double cw_sum(double v) {
    __shared__ double red[NTH / 64];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double tot = 0.0;
    for (int q = 0; q < NTH / 64; q++) tot += red[q];
    __syncthreads();
    return tot;
}

// the interior node of one level a thread owns
template <typename T>
struct CNode {
    int k;     // node index j * n0 + i; -1: none
    int ij;    // i | j << 16
    T c[9];    // operator row, ghosts folded in
    T id;      // 1 / diag
    T b;       // right-hand side
    __device__ int i() const { return ij & 0xffff; }
    __device__ int j() const { return ij >> 16; }
};

// row (i, j) from its faces with every clamped neighbour's coefficient moved onto
// the interior node it is clamped to: next to the west (east) wall the west (east)
// column folds onto the centre column, next to the south (north) wall the south
// (north) row onto the centre row.  Static indices only (the row stays in
// registers); an interior row is unchanged.
template <typename T>
__device__ __forceinline__ void fold_row(const RowFaces<T> &f, int n0, int n1, int i, int j, T (&c)[9]) {
    p9_row_from<T>(f, c);
    const bool w_ = i == 1, e_ = i == n0 - 2, s_ = j == 1, n_ = j == n1 - 2;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        T &cw = c[3 * r], &cc = c[3 * r + 1], &cx = c[3 * r + 2];
        cc = w_ ? cc + cw : cc;
        cw = w_ ? T(0) : cw;
        cc = e_ ? cc + cx : cc;
        cx = e_ ? T(0) : cx;
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
        T &cs = c[q], &cc = c[3 + q], &cn = c[6 + q];
        cc = s_ ? cc + cs : cc;
        cs = s_ ? T(0) : cs;
        cc = n_ ? cc + cn : cc;
        cn = n_ ? T(0) : cn;
    }
}

template <typename T>
__device__ __forceinline__ T cw_row(const CNode<T> &c, const T *x, int n0) {
    T s = T(0);
#pragma unroll
    for (int dj = -1; dj <= 1; dj++)
#pragma unroll
        for (int di = -1; di <= 1; di++) s += c.c[(dj + 1) * 3 + (di + 1)] * x[c.k + dj * n0 + di];
    return s;
}

// The owned interior node of every level l0 + L of a single-workgroup V-cycle: its
// index, the loads of its row faces and 1/diag (slot blockIdx.z of the level
// arrays in A), b = 0; levels past the last reload the last one and own no node.
template <typename T>
__device__ __forceinline__ void cw_row_loads(const CoarseArgs<T> &A, int l0, CNode<T> (&nd)[CW_LEVS],
                                             RowFaces<T> (&rf)[CW_LEVS]) {
    const int tid = threadIdx.x, last = A.nlev - 1;
#pragma unroll
    for (int L = 0; L < CW_LEVS; L++) {
        const int l = min(l0 + L, last);
        const CLev<T> &cl = A.L[l];
        const int m0 = cl.n0 - 2, m1 = cl.n1 - 2;
        const int q = min(tid, m0 * m1 - 1);
        const int i = 1 + q % m0, j = 1 + q / m0;
        const size_t o = zslot() * cl.n0 * cl.n1;
        nd[L].ij = i | j << 16;
        nd[L].k = (l0 + L <= last && tid < m0 * m1) ? j * cl.n0 + i : -1;
        rf[L] = load_row_faces<T>(cl.n0, cl.fen + o, cl.fet + o, cl.fnn + o, cl.fnt + o, i, j);
        nd[L].id = cl.invd[o + (size_t)j * cl.n0 + i];
        nd[L].b = T(0);
    }
}
// the rows from the loaded faces, Neumann ghosts folded in
template <typename T>
__device__ __forceinline__ void cw_row_fold(const CoarseArgs<T> &A, int l0, CNode<T> (&nd)[CW_LEVS],
                                            const RowFaces<T> (&rf)[CW_LEVS]) {
    const int last = A.nlev - 1;
#pragma unroll
    for (int L = 0; L < CW_LEVS; L++) {
        const int l = min(l0 + L, last);
        fold_row<T>(rf[L], A.L[l].n0, A.L[l].n1, nd[L].i(), nd[L].j(), nd[L].c);
    }
}

// The V-cycle of levels l0 .. A.nlev - 1 in one workgroup of NTH threads (at least
// the interior nodes of level l0), from the rows, inverse diagonals and level-l0
// right-hand side in nd[] (nd[L].b of the coarser levels is set here), iterates in
// the LDS arrays at sh + A.off[l]; the result is level l0's iterate, complete and
// visible to every thread on return.
template <typename T, int NTH = CW>
__device__ __forceinline__ void cw_vcycle(const CoarseArgs<T> &A, int l0, CNode<T> (&nd)[CW_LEVS], T *sh) {
    const int tid = threadIdx.x;
    const T w = A.omega;
    const int last = A.nlev - 1;
    const int nl = last - l0 + 1;
    auto NN = [&](int l) { return A.L[l].n0 * A.L[l].n1; };
    auto XA = [&](int L) { return sh + A.off[l0 + L]; };             // iterate of level l0 + L
    auto TA = [&](int L) { return sh + A.off[l0 + L] + NN(l0 + L); }; // its ping-pong buffer
    auto INNER = [&](int L) { return (A.L[l0 + L].n0 - 2) * (A.L[l0 + L].n1 - 2); };
    // Levels of at most CW_WAVE interior nodes ("wave levels", from LW on) are owned
    // by wavefront 0 alone: their down leg, the coarsest solve and their up leg run
    // in that wavefront, ordered by wavefront barriers (its LDS writes precede its
    // next reads), while the other wavefronts wait at one block barrier.
    int LW = nl;
    for (int L = nl - 1; L >= 0 && INNER(L) <= CW_WAVE; L--) LW = L;
    auto wsync = [&]() __attribute__((always_inline)) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    // `count` Jacobi sweeps of level L (from zero when from_zero), the result left in
    // the level's iterate; WAVE: run by wavefront 0 with wavefront barriers, else by
    // the block (ending with a block barrier)
    auto sweeps = [&](auto wave_tag, const CNode<T> &c, int L, int count, bool from_zero)
                      __attribute__((always_inline)) {
        constexpr bool WAVE = decltype(wave_tag)::value;
        if (count <= 0) return;
        const int n0 = A.L[l0 + L].n0;
        T *x = XA(L), *t = TA(L);
        // output of sweep s: the buffers alternate so that the last one is x
        // (from zero); otherwise sweep 0 reads x and writes t
        auto out_of = [&](int s) __attribute__((always_inline)) {
            return from_zero ? (((count - 1 - s) & 1) ? t : x) : ((s & 1) ? x : t);
        };
        for (int s = 0; s < count; s++) {
            if (c.k >= 0) {
                T *out = out_of(s);
                if (from_zero && s == 0) {
                    out[c.k] = w * c.b * c.id;
                } else {
                    const T *in = (out == x) ? t : x;
                    out[c.k] = in[c.k] + w * (c.b - cw_row<T>(c, in, n0)) * c.id;
                }
            }
            if (WAVE) wsync();
            else if (s + 1 < count) __syncthreads();
        }
        if (!from_zero && (count & 1) && c.k >= 0) x[c.k] = t[c.k];  // the last sweep wrote t
        if (WAVE) wsync();
        else __syncthreads();
    };
    using Block = std::integral_constant<bool, false>;
    using Wave = std::integral_constant<bool, true>;
    // full weighting of level L's residual r onto the owned node of level L + 1
    auto restrict_to = [&](const T *r, int n0, const CNode<T> &cc) __attribute__((always_inline)) {
        T v = T(0);
        if (cc.k >= 0) {
            const int k = 2 * cc.j() * n0 + 2 * cc.i();
            v = (T(4) * r[k] + T(2) * (r[k - 1] + r[k + 1] + r[k - n0] + r[k + n0]) +
                 (r[k - n0 - 1] + r[k - n0 + 1] + r[k + n0 - 1] + r[k + n0 + 1])) /
                T(16);
        }
        return v;
    };
    // bilinear prolongation of level L + 1's correction added to level L's iterate
    auto prolong = [&](const CNode<T> &c, int L) __attribute__((always_inline)) {
        if (c.k < 0) return;
        const int m0 = A.L[l0 + L + 1].n0, m1 = A.L[l0 + L + 1].n1;
        const T *xcc = XA(L + 1);
        T *x = XA(L);
        const int i = c.i(), j = c.j(), I = i >> 1, J = j >> 1;
        // the four coarse values read unconditionally (clamped into the interior) and
        // the node's parity case selected, so that a wavefront does not branch
        const int I0 = clampi(I, 1, m0 - 2), I1 = clampi(I + 1, 1, m0 - 2);
        const int J0 = clampi(J, 1, m1 - 2) * m0, J1 = clampi(J + 1, 1, m1 - 2) * m0;
        const T a = xcc[J0 + I0], b = xcc[J0 + I1], cn = xcc[J1 + I0], d = xcc[J1 + I1];
        const bool ie = (i & 1) == 0, je = (j & 1) == 0;
        const T e = ie ? (je ? a : T(0.5) * (a + cn)) : (je ? T(0.5) * (a + b) : T(0.25) * (a + b + cn + d));
        x[c.k] += e;
    };

    // down, block levels: smoothing, residual, full weighting onto the next level
    // (its interior mean removed)
#pragma unroll
    for (int L = 0; L < CW_LEVS - 1; L++) {
        if (L < nl - 1 && L < LW) {
            const int n0 = A.L[l0 + L].n0;
            const CNode<T> &c = nd[L];
            T *x = XA(L), *r = TA(L);
            if (A.pre > 0) sweeps(Block{}, c, L, A.pre, true);  // with pre == 0 x stays zero
            if (c.k >= 0) r[c.k] = c.b - cw_row<T>(c, x, n0);
            __syncthreads();
            CNode<T> &cc = nd[L + 1];
            const T v = restrict_to(r, n0, cc);
            const double mean = cw_sum<NTH>(cc.k >= 0 ? (double)v : 0.0) / (double)INNER(L + 1);
            cc.b = (T)((double)v - mean);
        }
    }
    // wave levels, wavefront 0
    if (LW < nl) {
        if (tid < 64) {
#pragma unroll
            for (int L = 0; L < CW_LEVS - 1; L++) {
                if (L >= LW && L < nl - 1) {
                    const int n0 = A.L[l0 + L].n0;
                    const CNode<T> &c = nd[L];
                    T *x = XA(L), *r = TA(L);
                    if (A.pre > 0) sweeps(Wave{}, c, L, A.pre, true);
                    if (c.k >= 0) r[c.k] = c.b - cw_row<T>(c, x, n0);
                    wsync();
                    CNode<T> &cc = nd[L + 1];
                    const T v = restrict_to(r, n0, cc);
                    const double tot = wave_sum(cc.k >= 0 ? (double)v : 0.0);
                    cc.b = (T)((double)v - tot / (double)INNER(L + 1));
                }
            }
#pragma unroll
            for (int L = 0; L < CW_LEVS; L++)
                if (L == nl - 1) sweeps(Wave{}, nd[L], L, A.coarse, true);  // with coarse == 0 x stays zero
#pragma unroll
            for (int L = CW_LEVS - 2; L >= 0; L--) {
                if (L >= LW && L < nl - 1) {
                    prolong(nd[L], L);
                    wsync();
                    sweeps(Wave{}, nd[L], L, A.post, false);
                }
            }
        }
        __syncthreads();
    } else {
        // the coarsest level is a block level
#pragma unroll
        for (int L = 0; L < CW_LEVS; L++)
            if (L == nl - 1) sweeps(Block{}, nd[L], L, A.coarse, true);
    }
    // up, block levels: bilinear prolongation of the correction, post-smoothing
#pragma unroll
    for (int L = CW_LEVS - 2; L >= 0; L--) {
        if (L < nl - 1 && L < LW) {
            prolong(nd[L], L);
            __syncthreads();
            sweeps(Block{}, nd[L], L, A.post, false);
        }
    }
}

// b0: right-hand side of level l0 (b0part: nb0p block partials of its interior sum,
// whose mean is removed; null: used as is); the level-l0 result is written to x0out.
template <typename T>
__global__ __launch_bounds__(CW) void k_mg_coarse(const double *scal, int gs, CoarseArgs<T> A, int l0, const T *b0,
                                                  const double *b0part, int nb0p, size_t lps, T *x0out) {
    if (slot_skip(scal + zslot() * SS, gs)) return;
    {
        const size_t o0 = zslot() * A.L[l0].n0 * A.L[l0].n1;
        b0 += o0;
        x0out += o0;
        if (b0part) b0part += zslot() * lps;
    }
    __shared__ T sh[CN_MAX * CW_ARR];
    const int tid = threadIdx.x;
    const int last = A.nlev - 1;
    auto NN = [&](int l) { return A.L[l].n0 * A.L[l].n1; };
    auto INNER = [&](int L) { return (A.L[l0 + L].n0 - 2) * (A.L[l0 + L].n1 - 2); };

    // set-up loads: row faces and 1/diag of the owned node of every level
    CNode<T> nd[CW_LEVS];
    RowFaces<T> rf[CW_LEVS];
    cw_row_loads<T>(A, l0, nd, rf);
    const T braw = b0[nd[0].j() * A.L[l0].n0 + nd[0].i()];
    double part = 0.0;
    if (b0part)
        for (int t = tid; t < nb0p; t += CW) part += b0part[t];
    {
        int tot = 0;
        for (int l = l0; l <= last; l++) tot += CW_ARR * NN(l);
        for (int k = tid; k < tot; k += CW) sh[k] = T(0);
    }
    cw_row_fold<T>(A, l0, nd, rf);
    {
        const double mean = cw_sum(part) / (double)INNER(0);  // also orders the zero fill
        nd[0].b = b0part ? (T)((double)braw - mean) : braw;
    }

    cw_vcycle<T>(A, l0, nd, sh);
    {
        const T *x = sh + A.off[l0];
        for (int k = tid; k < NN(l0); k += CW) x0out[k] = x[k];
    }
}

// ---------------------------------------------------------------------------
// BiCGStab (right preconditioned): one kernel per vector phase writing block
// partials of its dot products, then k_bi_fin<phase> with the totals and the
// scalar recurrence / convergence test of that phase.

struct RedArgs {
    double *part;   // partials: per slot NV x pstride, slot stride psz
    int pstride;
    size_t psz;
    double *scal;   // device scalars, slot stride SS
};
// Partial rows PS_APPLY_ROW and PS_APPLY_ROW + 1 hold the operator-application
// phases of BiCGStab ((r0, v); (t, s) and (t, t)), which the next vector kernel
// totals itself while it writes its own partials to rows 0-1; the owner of the
// buffers allocates at least PS_APPLY_ROW + 2 rows.
constexpr int PS_APPLY_ROW = 6;
// the block's slot view of the reduction buffers
__device__ __forceinline__ RedArgs rz(RedArgs ra) {
    ra.part += zslot() * ra.psz;
    ra.scal += zslot() * SS;
    return ra;
}

// x (and, if acc != null, acc += x first) <- 0; r = b (interior); partials of
// sum(r), (b, b).  outer != null and set: the solve is skipped (S_DONE = 1),
// x is still cleared.
template <typename T>
__global__ __launch_bounds__(PS_NT) void k_bi_start1(int og, int n0, int n1, const double *__restrict__ b,
                                                     T *__restrict__ x, double *__restrict__ acc, T *__restrict__ r,
                                                     RedArgs ra) {
    ra = rz(ra);
    if (slot_skip(ra.scal, -1)) return;
    {
        const size_t o = zslot() * n0 * n1;
        b += o; x += o; r += o;
        if (acc) acc += o;
    }
    const bool skip = og >= 0 && ra.scal[og] != 0.0;
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int m = 0; m < FLAT_NPT; m++) {
        const FlatNode fn = flat_node(n0, n1, m);
        if (!fn.in) continue;
        const size_t P = fn.P;
        if (acc) acc[P] += (double)x[P];
        x[P] = T(0);
        if (skip) continue;
        T rv = T(0);
        if (inner_node(fn.i, fn.j, n0, n1)) {
            rv = (T)b[P];
            v[0] += (double)rv;
            v[1] += (double)rv * (double)rv;
        }
        r[P] = rv;
    }
    if (skip) {
        if (blin() == 0 && tlin() == 0) ra.scal[S_DONE] = 1.0;
        return;
    }
    block_partials<2>(v, ra.part, ra.pstride);
}

// r -= interior mean (S_SUM), r0 = r; partials of (r, r)
template <typename T>
__global__ __launch_bounds__(PS_NT) void k_bi_start2(int og, int n0, int n1, T *__restrict__ r, T *__restrict__ r0,
                                                     RedArgs ra) {
    ra = rz(ra);
    if (slot_skip(ra.scal, og)) return;
    {
        const size_t o = zslot() * n0 * n1;
        r += o; r0 += o;
    }
    const double mean = ra.scal[S_SUM] / ((double)(n0 - 2) * (double)(n1 - 2));
    double v[1] = {0.0};
#pragma unroll
    for (int m = 0; m < FLAT_NPT; m++) {
        const FlatNode fn = flat_node(n0, n1, m);
        if (!fn.in) continue;
        const size_t P = fn.P;
        T rv = T(0);
        if (inner_node(fn.i, fn.j, n0, n1)) {
            rv = (T)((double)r[P] - mean);
            r[P] = rv;
            v[0] += (double)rv * (double)rv;
        }
        r0[P] = rv;
    }
    block_partials<1>(v, ra.part, ra.pstride);
}

// The start when the right-hand side's interior sum and sum of squares are already
// in S_SUM / S_BB (and S_DONE, S_IT, S_SCONV cleared), e.g. from the pass that
// produced b: x = 0, r = b - interior mean, r0 = r, partials of (r, r).  With og
// set the solve is skipped (S_DONE = 1) after clearing x, as in k_bi_start1.
template <typename T>
__global__ __launch_bounds__(PS_NT) void k_bi_start_m(int og, int n0, int n1, const double *__restrict__ b,
                                                      T *__restrict__ x, T *__restrict__ r, T *__restrict__ r0,
                                                      RedArgs ra) {
    ra = rz(ra);
    if (slot_skip(ra.scal, -1)) return;
    {
        const size_t o = zslot() * n0 * n1;
        b += o; x += o; r += o; r0 += o;
    }
    const bool skip = og >= 0 && ra.scal[og] != 0.0;
    const double mean = ra.scal[S_SUM] / ((double)(n0 - 2) * (double)(n1 - 2));
    double v[1] = {0.0};
#pragma unroll
    for (int m = 0; m < FLAT_NPT; m++) {
        const FlatNode fn = flat_node(n0, n1, m);
        if (!fn.in) continue;
        const size_t P = fn.P;
        x[P] = T(0);
        if (skip) continue;
        T rv = T(0);
        if (inner_node(fn.i, fn.j, n0, n1)) {
            rv = (T)(b[P] - mean);
            v[0] += (double)rv * (double)rv;
        }
        r[P] = rv;
        r0[P] = rv;
    }
    if (skip) {
        if (blin() == 0 && tlin() == 0) ra.scal[S_DONE] = 1.0;
        return;
    }
    block_partials<1>(v, ra.part, ra.pstride);
}

// p = r (first iteration) or r + beta (p - omega v); stops the solve when rho = 0
template <typename T>
__global__ __launch_bounds__(PS_NT) void k_bi_p(int n0, int n1, int it, double *scal, const T *__restrict__ r,
                                                const T *__restrict__ vv, T *__restrict__ pv) {
    scal += zslot() * SS;
    if (slot_skip(scal, S_DONE)) return;
    {
        const size_t o = zslot() * n0 * n1;
        r += o; vv += o; pv += o;
    }
    const double rho = scal[S_RHO];
    if (blin() == 0 && tlin() == 0) {
        scal[S_IT] = (double)it;
        if (rho == 0.0) scal[S_DONE] = 1.0;
    }
    if (rho == 0.0) return;
    double beta = 0.0, omega = 0.0;  // read only after the first iteration
    if (it != 1) {
        beta = (rho / scal[S_RHO_OLD]) * (scal[S_ALPHA] / scal[S_OMEGA]);
        omega = scal[S_OMEGA];
    }
#pragma unroll
    for (int m = 0; m < FLAT_NPT; m++) {
        const FlatNode fn = flat_node(n0, n1, m);
        if (!fn.in) continue;
        const size_t P = fn.P;
        if (it == 1) pv[P] = r[P];
        else pv[P] = (T)((double)r[P] + beta * ((double)pv[P] - omega * (double)vv[P]));
    }
}

// The finest level's up leg (k_mg_up with two post-sweeps) fused with the
// operator application that follows the preconditioner in BiCGStab: the
// prolongation and the two post-sweeps run on a region one node wider than
// k_mg_up's (halos 3, 2 and 1), so that the output z (ph or sh) is known on the
// tile and the ring around it; then y = L z on the tile, with the partials of
// k_bi_v (MGA_V: (r0, y)) or k_bi_t (MGA_T: (y, s), (y, y); s is the level's
// right-hand side).  Writes z and y.
enum { MGA_V, MGA_T };
template <typename T, int NTB, int MODE>
__global__ __launch_bounds__(NTB) void k_mg_up_apply(UpArgs<T> A, T *__restrict__ yout, const T *__restrict__ r0,
                                                     RedArgs ra) {
    ra = rz(ra);
    const double *guard = A.scal + zslot() * SS;
    {
        const size_t of = zslot() * A.a.n0 * A.a.n1, oc = zslot() * A.m0 * A.m1;
        A.a = fz(A.a, of);
        A.invd += of;
        A.b += of;
        A.x += of;
        A.out += of;
        A.xc += oc;
        yout += of;
        if (MODE == MGA_V) r0 += of;
    }
    constexpr int YX = MG_FX + 6, YY = MG_FY + 6;  // x + e, halo 3
    constexpr int QX = MG_FX + 4, QY = MG_FY + 4;  // first post-sweep, halo 2
    constexpr int ZX = MG_FX + 2, ZY = MG_FY + 2;  // second post-sweep (z), halo 1
    constexpr int NV = MODE == MGA_V ? 1 : 2;
    __shared__ T s_y[YX * YY];
    __shared__ T s_q[QX * QY];
    __shared__ T s_o[ZX * ZY];
    __shared__ double s_dp[NV][MG_FX * MG_FY];  // each tile node's dot-product terms
    const int tid = tlin();
    const int n0 = A.a.n0, n1 = A.a.n1, m0 = A.m0, m1 = A.m1;
    const int2 tl = xcd_tile();
    const int fi0 = tl.x * MG_FX, fj0 = tl.y * MG_FY;
    const T w = A.w;
    const T *xc = A.xc;
    struct YR {
        T x, c00, c10, c01, c11;
    };
#define XC(a_, b_) xc[idx(clampi((a_), 1, m0 - 2), clampi((b_), 1, m1 - 2), m0)]
    const auto ly = batch_load<YX * YY, NTB>(tid, [&](int k) {
        const int gi = clampi(fi0 - 3 + k % YX, 1, n0 - 2), gj = clampi(fj0 - 3 + k / YX, 1, n1 - 2);
        const int I = gi >> 1, J = gj >> 1;
        return YR{A.x[idx(gi, gj, n0)], XC(I, J), XC(I + 1, J), XC(I, J + 1), XC(I + 1, J + 1)};
    });
#undef XC
    // faces, right-hand side, 1/diag (and r0) of the first post-sweep region
    struct PR {
        RowFaces<T> f;
        T b, id, r0;
    };
    auto q_load = [&]() {
        return batch_load<QX * QY, NTB>(tid, [&](int k) {
            const int gi = clampi(fi0 - 2 + k % QX, 1, n0 - 2);
            const int gj = clampi(fj0 - 2 + k / QX, 1, n1 - 2);
            const size_t P = idx(gi, gj, n0);
            return PR{row_faces<T>(A.a, gi, gj), A.b[P], A.invd[P], MODE == MGA_V ? r0[P] : T(0)};
        });
    };
    decltype(q_load()) lq;
    if (mg_early<T>()) lq = q_load();
    if (slot_skip(guard, A.gs)) return;
    edge_walls<MG_FX, MG_FY>(tid, NTB, fi0, fj0, n0, n1, [&](size_t P) {
        A.out[P] = T(0);
        yout[P] = T(0);
    });
    // Every entry of s_y, s_q and s_o holds the value of its node clamped into the
    // interior (the loads above are at those nodes), so the sweeps and the operator
    // read their neighbours at fixed offsets (frow_tile); the wall values written
    // to memory stay zero.
    batch_run<YX * YY, NTB>(tid, ly, [&](int k, const YR &r) {
        const int gi = clampi(fi0 - 3 + k % YX, 1, n0 - 2), gj = clampi(fj0 - 3 + k / YX, 1, n1 - 2);
        T e;
        if ((gi & 1) == 0 && (gj & 1) == 0) e = r.c00;
        else if ((gj & 1) == 0) e = T(0.5) * (r.c00 + r.c10);
        else if ((gi & 1) == 0) e = T(0.5) * (r.c00 + r.c01);
        else e = T(0.25) * (r.c00 + r.c10 + r.c01 + r.c11);
        s_y[k] = r.x + e;
    });
    if (!mg_early<T>()) lq = q_load();
    __syncthreads();
    // first post-sweep on the halo-2 region
    batch_run<QX * QY, NTB>(tid, lq, [&](int k, const PR &r) {
        const int gi = clampi(fi0 - 2 + k % QX, 1, n0 - 2), gj = clampi(fj0 - 2 + k / QX, 1, n1 - 2);
        const T *c = s_y + tile_at<3, YX>(gi, gj, fi0, fj0);
        s_q[k] = c[0] + w * (r.b - frow_tile<T, YX>(r.f, c)) * r.id;
    });
    __syncthreads();
    // second post-sweep on the halo-1 region: z, written on the tile
    batch_run<QX * QY, NTB>(tid, lq, [&](int k, const PR &r) {
        const int li = k % QX - 1, lj = k / QX - 1;
        if (li < 0 || li >= ZX || lj < 0 || lj >= ZY) return;
        const int ui = fi0 - 1 + li, uj = fj0 - 1 + lj;
        const int gi = clampi(ui, 1, n0 - 2), gj = clampi(uj, 1, n1 - 2);
        const T *c = s_q + tile_at<2, QX>(gi, gj, fi0, fj0);
        const T v = c[0] + w * (r.b - frow_tile<T, QX>(r.f, c)) * r.id;
        s_o[lj * ZX + li] = v;
        if (li >= 1 && li <= MG_FX && lj >= 1 && lj <= MG_FY && ui < n0 && uj < n1)
            A.out[idx(ui, uj, n0)] = inner_node(ui, uj, n0, n1) ? v : T(0);
    });
    __syncthreads();
    // y = L z on the tile and the dot products: each node's terms go to LDS and the
    // first MG_SUM_LANES lanes add them in a fixed order (the same bits for any NTB)
    batch_run<QX * QY, NTB>(tid, lq, [&](int k, const PR &r) {
        const int li = k % QX - 2, lj = k / QX - 2;
        if (li < 0 || li >= MG_FX || lj < 0 || lj >= MG_FY) return;
        const int gi = fi0 + li, gj = fj0 + lj;
        double d0 = 0.0, d1 = 0.0;
        if (gi < n0 && gj < n1) {
            T y = T(0);
            if (inner_node(gi, gj, n0, n1)) {
                y = frow_tile<T, ZX>(r.f, s_o + tile_at<1, ZX>(gi, gj, fi0, fj0));
                d0 = MODE == MGA_V ? (double)r.r0 * (double)y : (double)y * (double)r.b;
                d1 = (double)y * (double)y;
            }
            yout[idx(gi, gj, n0)] = y;
        }
        s_dp[0][lj * MG_FX + li] = d0;
        if (NV == 2) s_dp[NV - 1][lj * MG_FX + li] = d1;
    });
    __syncthreads();
    static_assert(MG_FX * MG_FY == 2 * MG_SUM_LANES && NTB >= MG_SUM_LANES, "tile sum shape");
    double dv[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) dv[q] = tid < MG_SUM_LANES ? s_dp[q][tid] + s_dp[q][tid + MG_SUM_LANES] : 0.0;
    block_partials<NV, NTB>(dv, ra.part, ra.pstride);
}

// v = L ph; partials of (r0, v)
template <typename T>
__global__ __launch_bounds__(PS_NT) void k_bi_v(FArgs<T> a, const T *__restrict__ ph, T *__restrict__ vv,
                                                const T *__restrict__ r0, RedArgs ra) {
    ra = rz(ra);
    if (slot_skip(ra.scal, S_DONE)) return;
    {
        const size_t o = zslot() * a.n0 * a.n1;
        a = fz(a, o); ph += o; vv += o; r0 += o;
    }
    const int2 tl = xcd_tile();
    const int i = tl.x * BX + threadIdx.x, j = tl.y * BY + threadIdx.y;
    double v[1] = {0.0};
    if (i < a.n0 && j < a.n1) {
        const size_t P = idx(i, j, a.n0);
        T y = T(0);
        if (inner_node(i, j, a.n0, a.n1)) {
            y = frow<T>(a, ph, i, j);
            v[0] = (double)r0[P] * (double)y;
        }
        vv[P] = y;
    }
    block_partials<1>(v, ra.part, ra.pstride);
}

// s = r - alpha v; partials of (s, s)
// alpha = rho / (r0, v) from the nvp partials of the v phase (vpart, row
// PS_APPLY_ROW), totalled by every block in the fixed order of a one-block finish
// (the same bits) while its own nodes load; block 0 keeps (r0, v) and alpha in the
// device scalars for the kernels that follow.  nvp = 0: alpha is already in the
// scalars (k_bi_fin<FIN_V>, where every block re-reading the partials would cost
// more than the finish launch)
template <typename T>
__global__ __launch_bounds__(PS_NT) void k_bi_s(int n0, int n1, const T *__restrict__ r, const T *__restrict__ vv,
                                                T *__restrict__ s, RedArgs ra, const double *__restrict__ vpart,
                                                int nvp) {
    ra = rz(ra);
    if (slot_skip(ra.scal, S_DONE)) return;
    if (ra.scal[S_PMAX] < 0.0) {
        // a one-V-cycle slot: x = 0 + 1 p^ by k_bi_xr's ||s||-test path, nothing
        // else of the iteration runs
        if (blin() == 0 && tlin() == 0) {
            ra.scal[S_ALPHA] = 1.0;
            ra.scal[S_SCONV] = 1.0;
            ra.scal[S_DONE] = 1.0;
        }
        return;
    }
    vpart += zslot() * ra.psz;
    {
        const size_t o = zslot() * n0 * n1;
        r += o; vv += o; s += o;
    }
    FlatNode fn[FLAT_NPT];
    T rl[FLAT_NPT], vl[FLAT_NPT];
#pragma unroll
    for (int m = 0; m < FLAT_NPT; m++) {
        fn[m] = flat_node(n0, n1, m);
        rl[m] = r[fn[m].P];  // P = 0 for a missing node: in bounds, unused
        vl[m] = vv[fn[m].P];
    }
    double alpha;
    if (nvp > 0) {
        double tv[1];
        sum_partials<1, PS_NT>(vpart, ra.pstride, nvp, tv);
        __shared__ double s_alpha;
        if (tlin() == 0) {
            const double a = ra.scal[S_RHO] / tv[0];
            s_alpha = a;
            if (blin() == 0) {
                ra.scal[S_RV] = tv[0];
                ra.scal[S_ALPHA] = a;
            }
        }
        __syncthreads();
        alpha = s_alpha;
    } else {
        alpha = ra.scal[S_ALPHA];
    }
    double v[1] = {0.0};
#pragma unroll
    for (int m = 0; m < FLAT_NPT; m++) {
        if (!fn[m].in) continue;
        const T sv = (T)((double)rl[m] - alpha * (double)vl[m]);
        s[fn[m].P] = sv;
        if (inner_node(fn[m].i, fn[m].j, n0, n1)) v[0] += (double)sv * (double)sv;
    }
    block_partials<1>(v, ra.part, ra.pstride);
}

// t = L sh; partials of (t, s), (t, t)
template <typename T>
__global__ __launch_bounds__(PS_NT) void k_bi_t(FArgs<T> a, const T *__restrict__ sh, T *__restrict__ t,
                                                const T *__restrict__ s, RedArgs ra) {
    ra = rz(ra);
    if (slot_skip(ra.scal, S_DONE)) return;
    {
        const size_t o = zslot() * a.n0 * a.n1;
        a = fz(a, o); sh += o; t += o; s += o;
    }
    const int2 tl = xcd_tile();
    const int i = tl.x * BX + threadIdx.x, j = tl.y * BY + threadIdx.y;
    double v[2] = {0.0, 0.0};
    if (i < a.n0 && j < a.n1) {
        const size_t P = idx(i, j, a.n0);
        T y = T(0);
        if (inner_node(i, j, a.n0, a.n1)) {
            y = frow<T>(a, sh, i, j);
            v[0] = (double)y * (double)s[P];
            v[1] = (double)y * (double)y;
        }
        t[P] = y;
    }
    block_partials<2>(v, ra.part, ra.pstride);
}

// x += alpha ph + omega sh, r = s - omega t, partials of (r, r), (r0, r); in the
// iteration that stopped at the ||s|| test only x += alpha ph.  omega = (t, s) /
// (t, t) from the ntp partials of the t phase (tpart, rows PS_APPLY_ROW and
// PS_APPLY_ROW + 1), totalled by every block as k_bi_s totals alpha's; block 0
// keeps the totals and omega in the device scalars.  ntp = 0: omega is already in
// the scalars (k_bi_fin<FIN_T>)
template <typename T>
__global__ __launch_bounds__(PS_NT) void k_bi_xr(int n0, int n1, int it, const T *__restrict__ ph,
                                                 const T *__restrict__ sh, const T *__restrict__ s,
                                                 const T *__restrict__ t, const T *__restrict__ r0, T *__restrict__ x,
                                                 T *__restrict__ r, RedArgs ra, const double *__restrict__ tpart,
                                                 int ntp) {
    ra = rz(ra);
    double *sc = ra.scal;
    if (sc[S_ACTIVE] == 0.0) return;
    const bool sconv = sc[S_SCONV] == (double)it;
    if (!sconv && sc[S_DONE] != 0.0) return;
    tpart += zslot() * ra.psz;
    {
        const size_t o = zslot() * n0 * n1;
        ph += o; sh += o; s += o; t += o; r0 += o; x += o; r += o;
    }
    const double alpha = sc[S_ALPHA];
    if (sconv) {
#pragma unroll
        for (int m = 0; m < FLAT_NPT; m++) {
            const FlatNode fn = flat_node(n0, n1, m);
            if (fn.in) x[fn.P] = (T)((double)x[fn.P] + alpha * (double)ph[fn.P]);
        }
        return;
    }
    FlatNode fn[FLAT_NPT];
    T xl[FLAT_NPT], pl[FLAT_NPT], hl[FLAT_NPT], sl[FLAT_NPT], tl[FLAT_NPT], ol[FLAT_NPT];
#pragma unroll
    for (int m = 0; m < FLAT_NPT; m++) {
        fn[m] = flat_node(n0, n1, m);
        const size_t P = fn[m].P;  // 0 for a missing node: in bounds, unused
        xl[m] = x[P];
        pl[m] = ph[P];
        hl[m] = sh[P];
        sl[m] = s[P];
        tl[m] = t[P];
        ol[m] = r0[P];
    }
    double omega;
    if (ntp > 0) {
        double tt[2];
        sum_partials<2, PS_NT>(tpart, ra.pstride, ntp, tt);
        __shared__ double s_omega;
        if (tlin() == 0) {
            const double om = tt[1] > 0.0 ? tt[0] / tt[1] : 0.0;
            s_omega = om;
            if (blin() == 0) {
                sc[S_TS] = tt[0];
                sc[S_TT] = tt[1];
                sc[S_OMEGA] = om;
            }
        }
        __syncthreads();
        omega = s_omega;
    } else {
        omega = sc[S_OMEGA];
    }
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int m = 0; m < FLAT_NPT; m++) {
        if (!fn[m].in) continue;
        const size_t P = fn[m].P;
        x[P] = (T)((double)xl[m] + alpha * (double)pl[m] + omega * (double)hl[m]);
        const T rv = (T)((double)sl[m] - omega * (double)tl[m]);
        r[P] = rv;
        if (inner_node(fn[m].i, fn[m].j, n0, n1)) {
            v[0] += (double)rv * (double)rv;
            v[1] += (double)ol[m] * (double)rv;
        }
    }
    block_partials<2>(v, ra.part, ra.pstride);
}

// (the v and t phases, FIN_V and FIN_T, are totalled inside k_bi_s and k_bi_xr,
// or by their own finish on large grids)
enum { FIN_START1, FIN_START2, FIN_V, FIN_S, FIN_T, FIN_XR };

// the scalar recurrence / convergence test of phase OP from the totals
template <int OP, int NV>
__device__ __forceinline__ void bi_fin_scalars(double *s, const double (&tot)[NV], int it, double rtol,
                                                    int rtol_slot) {
    if (OP == FIN_START1) {
        s[S_SUM] = tot[0];
        s[S_BB] = tot[NV - 1];
        s[S_DONE] = 0.0;
        s[S_IT] = 0.0;
        s[S_SCONV] = 0.0;
    } else if (OP == FIN_START2) {
        const double bn = sqrt(s[S_BB]);
        const double rt = rtol_slot >= 0 ? s[rtol_slot] : rtol;
        const double tol = rt * bn;
        const double rn = sqrt(tot[0]);
        s[S_RR] = tot[0];
        s[S_BN] = bn;
        s[S_TOL] = tol;
        s[S_RN] = rn;
        s[S_RELRES] = bn > 0.0 ? rn / bn : 0.0;
        s[S_RHO] = tot[0];
        s[S_RHO_OLD] = 1.0;
        s[S_ALPHA] = 1.0;
        s[S_OMEGA] = 1.0;
        s[S_DONE] = (bn == 0.0 || (rn <= tol && s[S_PMAX] >= 0.0)) ? 1.0 : 0.0;
    } else if (OP == FIN_V) {
        s[S_RV] = tot[0];
        s[S_ALPHA] = s[S_RHO] / tot[0];
    } else if (OP == FIN_T) {
        s[S_TS] = tot[0];
        s[S_TT] = tot[NV - 1];
        s[S_OMEGA] = tot[NV - 1] > 0.0 ? tot[0] / tot[NV - 1] : 0.0;
    } else if (OP == FIN_S) {
        s[S_SS] = tot[0];
        const double sn = sqrt(tot[0]);
        if (sn <= s[S_TOL]) {
            s[S_SCONV] = (double)it;
            s[S_DONE] = 1.0;
            s[S_RN] = sn;
            s[S_RELRES] = s[S_BN] > 0.0 ? sn / s[S_BN] : 0.0;
        }
    } else {
        s[S_RR] = tot[0];
        s[S_R0R] = tot[NV - 1];
        const double rn = sqrt(tot[0]);
        s[S_RN] = rn;
        s[S_RELRES] = s[S_BN] > 0.0 ? rn / s[S_BN] : 0.0;
        if (rn <= s[S_TOL] || s[S_OMEGA] == 0.0) {
            s[S_DONE] = 1.0;
        } else {
            s[S_RHO_OLD] = s[S_RHO];
            s[S_RHO] = tot[NV - 1];
            // the slot's own cap: its remaining launched iterations are no-ops
            if (s[S_PMAX] > 0.0 && (double)it >= s[S_PMAX]) s[S_DONE] = 1.0;
        }
    }
}

// totals of a phase's partials and what follows from them; skipped while *guard
// is set (the phase's kernel did not run)
// FIN_XR with flag != null also writes the slot's end-of-iteration flag into host
// memory: (seq << 1) | done, done when the slot is skipped or has converged.
template <int OP>
__global__ __launch_bounds__(BI_FIN_NT) void k_bi_fin(int gs, RedArgs ra, int nparts, int it, double rtol, int rtol_slot,
                                                   unsigned long long seq = 0, unsigned long long *flag = nullptr) {
    ra = rz(ra);
    const bool skip = slot_skip(ra.scal, gs);  // read while the partials load
    constexpr int NV = (OP == FIN_START1 || OP == FIN_XR || OP == FIN_T) ? 2 : 1;
    double tot[NV];
    sum_partials<NV, BI_FIN_NT>(ra.part, ra.pstride, nparts, tot);
    if (tlin() != 0) return;
    if (!skip) bi_fin_scalars<OP>(ra.scal, tot, it, rtol, rtol_slot);
    if (OP == FIN_XR && flag) {
        const unsigned long long done = (skip || ra.scal[S_DONE] != 0.0) ? 1ull : 0ull;
        __threadfence_system();
        __hip_atomic_store(flag + zslot(), (seq << 1) | done, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

template <typename A, typename B>
__global__ void k_cvt(size_t n, const A *__restrict__ a, B *__restrict__ b) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = (B)a[k];
}
template <typename T>
__global__ void k_fill(size_t n, T *__restrict__ a, T v) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = v;
}
// injection of the fine level's values onto the coarse level, slot z = blockIdx.z
__global__ void k_inject(int n0, int n1, const double *__restrict__ f, int m0, int m1, double *__restrict__ c) {
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
    if (I >= m0 || J >= m1) return;
    c[zslot() * m0 * m1 + idx(I, J, m0)] = f[zslot() * n0 * n1 + idx(2 * I, 2 * J, n0)];
}

// ---------------------------------------------------------------------------
// host side

// what the solver needs from its owner (stream, reductions, timing hooks)
struct PsCtx {
    hipStream_t s;
    int B;          // slots
    double *part;   // B x psz
    int pstride;
    size_t psz;
    double *scal;   // B x SS
    double *hscal;  // pinned host mirror of scal
    Timed (*prof)(void *, int);  // events timing the next launch of a kind (none: untimed)
    void (*drain)(void *);
    void *owner;
    std::vector<void *> *allocs;
    size_t *bytes;
    bool dry;       // count bytes only (no device allocation)
    // per-slot end-of-iteration flags in pinned, coherent host memory (device and
    // host views) and the owner's flag sequence counter; null: not used
    unsigned long long *flag_d = nullptr;
    unsigned long long *flag_h = nullptr;
    unsigned long long *seq = nullptr;
    RedArgs ra() const { return RedArgs{part, pstride, psz, scal}; }
    // every active slot has its S_DONE flag set (host mirror)
    bool all_done() const {
        for (int z = 0; z < B; z++) {
            const double *h = hscal + (size_t)z * SS;
            if (h[S_ACTIVE] != 0.0 && h[S_DONE] == 0.0) return false;
        }
        return true;
    }
};

// live-timing kinds of the solver's fine-level kernels
enum { PK_APPLY = 0, PK_MG_DOWN = 4, PK_MG_UP = 5 };

inline int ps_alloc(PsCtx &cx, void **p, size_t bytes) {
    if (cx.dry) {
        *p = nullptr;
        *cx.bytes += bytes;
        return CFD_OK;
    }
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

// Wait until every slot's end-of-iteration flag carries sequence number `sq`
// (written by k_bi_fin<FIN_XR> straight into host memory), then report whether
// every slot is done.  Cheaper than copying the scalars back and synchronising
// the stream.  If the stream drains and a flag is still missing the protocol is
// broken: CFD_E_HIP, never an endless wait.
inline int ps_wait(PsCtx &cx, unsigned long long sq, bool &all_done) {
    for (long spin = 0;; spin++) {
        bool seen = true, done = true;
        for (int z = 0; z < cx.B && seen; z++) {
            const unsigned long long v = __atomic_load_n(cx.flag_h + z, __ATOMIC_ACQUIRE);
            if ((v >> 1) != sq) seen = false;
            else if ((v & 1) == 0) done = false;
        }
        if (seen) {
            all_done = done;
            break;
        }
        if ((spin & 255) == 255) {
            const hipError_t e = hipStreamQuery(cx.s);
            if (e == hipSuccess) {  // drained: the flags are final now
                for (int z = 0; z < cx.B; z++)
                    if ((__atomic_load_n(cx.flag_h + z, __ATOMIC_ACQUIRE) >> 1) != sq) return CFD_E_HIP;
            } else if (e != hipErrorNotReady) {
                return CFD_E_HIP;
            }
        }
        __builtin_ia32_pause();
    }
    if (cx.drain) cx.drain(cx.owner);
    return CFD_OK;
}

inline int ps_sync(PsCtx &cx) {
    CFD_HIP(hipMemcpyAsync(cx.hscal, cx.scal, (size_t)cx.B * SS * sizeof(double), hipMemcpyDeviceToHost, cx.s));
    CFD_HIP(hipStreamSynchronize(cx.s));
    if (cx.drain) cx.drain(cx.owner);
    return CFD_OK;
}

// Multigrid-preconditioned BiCGStab on the face-coefficient operator in precision T.
template <typename T>
struct PSolver {
    OpLevel<T> lev[PS_MAXLEV];
    int nlev = 0;
    T *bi[8] = {nullptr};  // r r0 pv vv s t ph sh (level-0 size)
    int pre = 2, post = 2, coarse = 10;
    double omega = 0.8;
    int lc = 0;                  // first level handled by the single-workgroup kernel
    CoarseArgs<T> cargs{};
    LevTab<T> tab{};
    int tab_blocks = 0;
    double *lpart = nullptr;       // block partials of the restricted right-hand sides' sums
    int lofs[PS_MAXLEV] = {0};     // level l's partials at lpart + lofs[l] (l >= 1), per slot
    int lnp[PS_MAXLEV] = {0};      // and their number
    size_t lps = 0;                // per-slot stride of lpart
    int B = 1;                     // slots
    // per fine level: the block-size class (MGB_*) of its down (up) kernel, from
    // the launch-size rule or timed by tune()
    int wdn[PS_MAXLEV] = {}, wup[PS_MAXLEV] = {};
    bool tuned = false;

    int init(PsCtx &cx, const Geom *g, int nlev_, const cfd_params &prm) {
        nlev = nlev_;
        pre = prm.mg_pre;
        post = prm.mg_post;
        coarse = prm.mg_coarse;
        omega = prm.mg_omega;
        if (pre < 0 || pre > 2 || post < 0 || post > 2 || coarse < 0) return CFD_E_ARG;
        // the LDS-resident coarse kernel takes every level from the first one with
        // at most CW_NODES nodes whose levels together fit its LDS arrays
        lc = nlev - 1;
        while (lc > 0 && g[lc - 1].n0 * g[lc - 1].n1 <= CW_NODES) lc--;
        for (;;) {
            int tot = 0;
            for (int l = lc; l < nlev; l++) tot += g[l].n0 * g[l].n1;
            if (tot <= CN_MAX) break;
            if (lc + 1 >= nlev) return CFD_E_ARG;
            lc++;
        }
        if (nlev - lc > CW_LEVS) return CFD_E_ARG;
        B = cx.B;
        for (int l = 0; l < nlev; l++) {
            OpLevel<T> &L = lev[l];
            L.n0 = g[l].n0;
            L.n1 = g[l].n1;
            const size_t M = (size_t)B * L.n0 * L.n1 * sizeof(T);
            T **arr[] = {&L.fen, &L.fet, &L.fnn, &L.fnt, &L.invd, &L.x, &L.b, &L.t};
            for (T **a : arr) CFD_RET(ps_alloc(cx, (void **)a, M));
        }
        const size_t M0 = (size_t)B * lev[0].n0 * lev[0].n1 * sizeof(T);
        for (int k = 0; k < 8; k++) CFD_RET(ps_alloc(cx, (void **)&bi[k], M0));
        int np = 0;
        for (int l = 1; l <= lc; l++) {
            lofs[l] = np;
            lnp[l] = down_grid(l - 1).x * down_grid(l - 1).y;
            np += lnp[l];
        }
        lps = (size_t)np + 1;
        CFD_RET(ps_alloc(cx, (void **)&lpart, (size_t)B * lps * sizeof(double)));
        cargs.nlev = nlev;
        cargs.pre = pre;
        cargs.post = post;
        cargs.coarse = coarse;
        cargs.omega = (T)omega;
        int off = 0;
        for (int l = 0; l < nlev; l++) {
            const OpLevel<T> &L = lev[l];
            cargs.L[l] = CLev<T>{L.n0, L.n1, L.fen, L.fet, L.fnn, L.fnt, L.invd};
            cargs.off[l] = off;
            if (l >= lc) off += CW_ARR * L.n0 * L.n1;
        }
        tab.nlev = nlev;
        int b = 0;
        for (int l = 0; l < nlev; l++) {
            const dim3 gr = grid2d(g[l].n0, g[l].n1);
            tab.bofs[l] = b;
            tab.bx[l] = (int)gr.x;
            b += (int)(gr.x * gr.y);
            tab.n0[l] = g[l].n0;
            tab.n1[l] = g[l].n1;
            tab.h1[l] = g[l].h1;
            tab.h2[l] = g[l].h2;
            tab.ih1sq[l] = 1.0 / (g[l].h1 * g[l].h1);
            tab.ih2sq[l] = 1.0 / (g[l].h2 * g[l].h2);
            tab.i4h1h2[l] = 1.0 / (4.0 * g[l].h1 * g[l].h2);
            tab.xix[l] = g[l].xix;
            tab.xiy[l] = g[l].xiy;
            tab.etx[l] = g[l].etx;
            tab.ety[l] = g[l].ety;
            tab.jac[l] = g[l].jac;
            tab.fen[l] = lev[l].fen;
            tab.fet[l] = lev[l].fet;
            tab.fnn[l] = lev[l].fnn;
            tab.fnt[l] = lev[l].fnt;
            tab.invd[l] = lev[l].invd;
        }
        tab.bofs[nlev] = b;
        tab_blocks = b;
        tab.dn0 = g[0].n0;
        size_rule();
        return CFD_OK;
    }

    // per outer iteration: face coefficients and inverse diagonals of every level
    // from level 0's d; f64 (optional) receives level 0's faces in FP64
    int setup(PsCtx &cx, const double *d0, double *const f64[4]) {
        LevTab<T> t = tab;
        t.d0 = d0;
        for (int k = 0; k < 4; k++) t.f64[k] = f64 ? f64[k] : nullptr;
        const dim3 gr(tab_blocks, 1, B);
        k_faces_all<T><<<gr, PS_NT, 0, cx.s>>>(cx.scal, t);
        CFD_LAUNCH_CHECK();
        return CFD_OK;
    }

    // k_mg_down<T, pre, threads, mean> on grid gr
    template <int NPRE, int NTB>
    static void launch_down_n(const dim3 &gr, hipStream_t st, const DownArgs<T> &a, bool mean, const Timed &tm) {
        if (mean) launch_timed(tm, k_mg_down<T, NPRE, NTB, true>, gr, NTB, st, a);
        else launch_timed(tm, k_mg_down<T, NPRE, NTB, false>, gr, NTB, st, a);
    }
    template <int NPRE>
    static void launch_down_p(const dim3 &gr, hipStream_t st, const DownArgs<T> &a, int nb, bool mean,
                              const Timed &tm) {
        by_block_size(nb, [&](auto n) { launch_down_n<NPRE, decltype(n)::value>(gr, st, a, mean, tm); });
    }
    static void launch_down(const dim3 &gr, hipStream_t st, const DownArgs<T> &a, int pre, int nb, bool mean,
                            const Timed &tm = Timed{}) {
        switch (pre) {
        case 0: launch_down_p<0>(gr, st, a, nb, mean, tm); break;
        case 1: launch_down_p<1>(gr, st, a, nb, mean, tm); break;
        default: launch_down_p<2>(gr, st, a, nb, mean, tm); break;
        }
    }
    template <int NPOST>
    static void launch_up_p(const dim3 &gr, hipStream_t st, const UpArgs<T> &a, int nb, const Timed &tm) {
        by_block_size(nb, [&](auto n) {
            constexpr int NTB = decltype(n)::value;
            launch_timed(tm, k_mg_up<T, NPOST, NTB>, gr, NTB, st, a);
        });
    }

    // The rule without tuning: a launch with fewer tiles (over all slots) than CUs
    // runs them with 1024 threads each (with one wavefront per SIMD a tile's phases
    // are bound by that wavefront's serial instruction latency, so four times the
    // threads cut it about fourfold); larger launches use 512.  tune() times 256,
    // 512 and 1024 instead: 256 threads fit all 2048 level-0 tiles of a 1025^2 down
    // leg in one round of resident blocks, and win for the many small tiles of a
    // batch, while the levels below favour 512 or 1024.  A tile's sums use
    // MG_SUM_LANES lanes whatever the block size, so a case computes the same bits
    // alone and in a batch, tuned or not.
    static bool wide_blocks(const dim3 &gr) { return (long)gr.x * gr.y * gr.z < MG_WIDE_BELOW; }

    dim3 up_grid(int l) const {
        return dim3(tiles_inner(lev[l].n0, MG_FX), tiles_inner(lev[l].n1, MG_FY), B);
    }
    void size_rule() {
        for (int l = 0; l < lc; l++) {
            dim3 gd = down_grid(l);
            gd.z = B;
            wdn[l] = wide_blocks(gd) ? MGB_WIDE : MGB_NARROW;
            wup[l] = wide_blocks(up_grid(l)) ? MGB_WIDE : MGB_NARROW;
        }
    }

    // Autotuning of the multigrid block sizes: each fine level's down and up
    // kernels are timed with MG_NT_NARROW and MG_NT_WIDE threads per tile on the
    // solver's own buffers (three launches of each, alternating, HIP events on the
    // solver's stream) and the faster is kept.  A tile's sums use the same lanes
    // whatever the block size, so the choice changes timing only, never a result.
    // Runs once; not while the stream is being captured.
    int tune(PsCtx &cx) {
        if (tuned) return CFD_OK;
        tuned = true;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        CFD_HIP(hipEventCreate(&e0));
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            return CFD_E_HIP;
        }
        int rc = CFD_OK;
        // one untimed launch of each block size (a kernel's first launch carries
        // one-off costs), then the size with the smallest minimum over five rounds of
        // alternating timed launches (a minimum ignores launches slowed by other work)
        auto time_sizes = [&](auto launch, int &pick) {
            for (int k = 0; k < MGB_COUNT && rc == CFD_OK; k++) {
                launch(k);
                if (hipGetLastError() != hipSuccess) rc = CFD_E_HIP;
            }
            float best[MGB_COUNT];
            for (int k = 0; k < MGB_COUNT; k++) best[k] = 3.0e38f;
            for (int rep = 0; rep < 5 && rc == CFD_OK; rep++)
                for (int k = 0; k < MGB_COUNT && rc == CFD_OK; k++) {
                    float ms = 0.f;
                    if (hipEventRecord(e0, cx.s) != hipSuccess) rc = CFD_E_HIP;
                    launch(k);
                    if (hipGetLastError() != hipSuccess || hipEventRecord(e1, cx.s) != hipSuccess ||
                        hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                        rc = CFD_E_HIP;
                    best[k] = std::min(best[k], ms);
                }
            if (rc == CFD_OK) {
                int k0 = 0;
                for (int k = 1; k < MGB_COUNT; k++)
                    if (best[k] < best[k0]) k0 = k;
                pick = k0;
            }
        };
        const T w = (T)omega;
        const RedArgs ra = cx.ra();
        for (int l = 0; l < lc && rc == CFD_OK; l++) {
            OpLevel<T> &L = lev[l];
            OpLevel<T> &C = lev[l + 1];
            const DownArgs<T> a{fargs(L), L.invd, L.t, l == 0 ? nullptr : lpart + lofs[l], lnp[l],
                                l == 0 ? nullptr : L.b, L.x, C.n0, C.n1, C.t, lpart + lofs[l + 1], w, lps, cx.scal,
                                -1};
            dim3 gd = down_grid(l);
            gd.z = B;
            time_sizes([&](int nb) { launch_down(gd, cx.s, a, pre, nb, a.bpart != nullptr); }, wdn[l]);
            const UpArgs<T> u{fargs(L), L.invd, L.b, L.x, C.n0, C.n1, C.t, L.t, w, cx.scal, -1};
            const dim3 gu = up_grid(l);
            if (l == 0 && fused_tail()) {
                RedArgs rap = ra;
                rap.part += (size_t)PS_APPLY_ROW * ra.pstride;
                const Tail tl{MGA_T, bi[5], nullptr, rap};
                time_sizes([&](int nb) {
                    by_block_size(nb, [&](auto n) { launch_up_apply<decltype(n)::value>(gu, cx.s, u, tl, Timed{}); });
                }, wup[l]);
            } else {
                time_sizes([&](int nb) {
                    switch (post) {
                    case 0: launch_up_p<0>(gu, cx.s, u, nb, Timed{}); break;
                    case 1: launch_up_p<1>(gu, cx.s, u, nb, Timed{}); break;
                    default: launch_up_p<2>(gu, cx.s, u, nb, Timed{}); break;
                    }
                }, wup[l]);
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc;
    }

    // launch shape of k_mg_down on fine level l (one block per coarse 32 x 4 tile)
    dim3 down_grid(int l) const {
        return dim3(tiles_inner(lev[l + 1].n0, MG_CX), tiles_inner(lev[l + 1].n1, MG_CY));
    }

    // z = M^-1 y: one V-cycle from zero with right-hand side y (slots whose
    // scalar `gs` is set are skipped)
    // The operator application that follows a preconditioner call in BiCGStab,
    // fused into the finest up leg when there is a fine level and two post-sweeps:
    // y = L z with the partials of k_bi_v (MGA_V, with r0) or k_bi_t (MGA_T).
    struct Tail {
        int mode;
        T *y;
        const T *r0;
        RedArgs ra;
    };
    bool fused_tail() const { return lc > 0 && post == 2; }
    // blocks (partials) of the fused launch: level 0's 64 x 8 tiles
    int tail_parts() const { return tiles_inner(lev[0].n0, MG_FX) * tiles_inner(lev[0].n1, MG_FY); }
    template <int NTB>
    static void launch_up_apply(const dim3 &gr, hipStream_t st, const UpArgs<T> &a, const Tail &t, const Timed &tm) {
        if (t.mode == MGA_V) launch_timed(tm, k_mg_up_apply<T, NTB, MGA_V>, gr, NTB, st, a, t.y, t.r0, t.ra);
        else launch_timed(tm, k_mg_up_apply<T, NTB, MGA_T>, gr, NTB, st, a, t.y, t.r0, t.ra);
    }

    int precond(PsCtx &cx, const T *y, T *z, int gs, const Tail *tail = nullptr) {
        const T w = (T)omega;
        for (int l = 0; l < lc; l++) {
            OpLevel<T> &L = lev[l];
            OpLevel<T> &C = lev[l + 1];
            DownArgs<T> a{fargs(L), L.invd, l == 0 ? y : L.t, l == 0 ? nullptr : lpart + lofs[l], lnp[l],
                          l == 0 ? nullptr : L.b, L.x, C.n0, C.n1, C.t, lpart + lofs[l + 1], w, lps, cx.scal, gs};
            dim3 gr = down_grid(l);
            gr.z = B;
            launch_down(gr, cx.s, a, pre, wdn[l], a.bpart != nullptr,
                        l == 0 ? cx.prof(cx.owner, PK_MG_DOWN) : Timed{});
            CFD_LAUNCH_CHECK();
        }
        T *cout = lc == 0 ? z : lev[lc].x;
        k_mg_coarse<T><<<dim3(1, 1, B), CW, 0, cx.s>>>(cx.scal, gs, cargs, lc, lc == 0 ? y : lev[lc].t,
                                                       lc == 0 ? nullptr : lpart + lofs[lc], lnp[lc], lps, cout);
        CFD_LAUNCH_CHECK();
        for (int l = lc - 1; l >= 0; l--) {
            OpLevel<T> &L = lev[l];
            OpLevel<T> &C = lev[l + 1];
            UpArgs<T> a{fargs(L), L.invd, l == 0 ? y : L.b, L.x, C.n0, C.n1, l + 1 == lc ? C.x : C.t,
                        l == 0 ? z : L.t, w, cx.scal, gs};
            const dim3 gr = up_grid(l);
            const Timed tm = l == 0 ? cx.prof(cx.owner, PK_MG_UP) : Timed{};
            if (l == 0 && tail) {
                by_block_size(wup[l], [&](auto n) { launch_up_apply<decltype(n)::value>(gr, cx.s, a, *tail, tm); });
                CFD_LAUNCH_CHECK();
                continue;
            }
            switch (post) {
            case 0: launch_up_p<0>(gr, cx.s, a, wup[l], tm); break;
            case 1: launch_up_p<1>(gr, cx.s, a, wup[l], tm); break;
            default: launch_up_p<2>(gr, cx.s, a, wup[l], tm); break;
            }
            CFD_LAUNCH_CHECK();
        }
        return CFD_OK;
    }

    // L x = b to ||r|| <= rtol ||b|| (rtol from scal[rtol_slot] when rtol_slot >= 0),
    // per slot, x cleared first (acc += x before clearing when acc != null);
    // skipped for slots whose scalar `og` (>= 0) is set.  Iteration for iteration
    // the oracle's bicgstab().  graph: no host synchronisation, `maxit` iterations
    // launched; otherwise the host reads the device flags at the end of each
    // iteration and stops launching once every slot is done (a slot that stops
    // at the start or at the ||s|| test turns the rest of its iteration's
    // launches into no-ops).  The iteration count and relative residual are
    // left in S_IT / S_RELRES of each slot.
    // sums_ready: S_SUM / S_BB already hold b's interior sum and sum of squares and
    // acc is null -- one start pass instead of two.
    int solve(PsCtx &cx, const double *b, T *x, double *acc, double rtol, int rtol_slot, int maxit, int og,
              bool graph, bool sums_ready = false) {
        OpLevel<T> &L0 = lev[0];
        const int n0 = L0.n0, n1 = L0.n1;
        dim3 gr = grid2d(n0, n1);  // k_bi_v, k_bi_t: 64 x 4 tiles
        gr.z = B;
        const dim3 gf(nflat(n0, n1), 1, B);  // the elementwise kernels
        const dim3 bl(BX, BY), g1(1, 1, B);
        const int nb = nflat(n0, n1);
        const bool fz = fused_tail();
        const int nbt = fz ? tail_parts() : nblk(n0, n1);  // partials of the v and t phases
        // alpha and omega totalled by every block of k_bi_s / k_bi_xr while that reads
        // little (measured: 1027 blocks x 2048 partials at 1025^2 is faster fused), by
        // a one-block finish on larger grids (the reads grow as blocks x partials)
        const bool sep = (long)nb * nbt > BI_FUSE_MAX;
        T *r = bi[0], *r0 = bi[1], *pv = bi[2], *vv = bi[3], *s = bi[4], *t = bi[5], *ph = bi[6], *sh = bi[7];
        const FArgs<T> a0 = fargs(L0);
        const RedArgs ra = cx.ra();
        RedArgs rap = ra;  // rows of the operator-application phases
        rap.part += (size_t)PS_APPLY_ROW * ra.pstride;
        if (sums_ready && !acc) {
            k_bi_start_m<T><<<gf, bl, 0, cx.s>>>(og, n0, n1, b, x, r, r0, ra);
        } else {
            k_bi_start1<T><<<gf, bl, 0, cx.s>>>(og, n0, n1, b, x, acc, r, ra);
            k_bi_fin<FIN_START1><<<g1, BI_FIN_NT, 0, cx.s>>>(og, ra, nb, 0, 0.0, -1);
            k_bi_start2<T><<<gf, bl, 0, cx.s>>>(og, n0, n1, r, r0, ra);
        }
        k_bi_fin<FIN_START2><<<g1, BI_FIN_NT, 0, cx.s>>>(og, ra, nb, 0, rtol, rtol_slot);
        CFD_LAUNCH_CHECK();
        for (int it = 1; it <= maxit; it++) {
            k_bi_p<T><<<gf, bl, 0, cx.s>>>(n0, n1, it, cx.scal, r, vv, pv);
            CFD_LAUNCH_CHECK();
            const Tail tv{MGA_V, vv, r0, rap};
            CFD_RET(precond(cx, pv, ph, S_DONE, fz ? &tv : nullptr));
            if (!fz)
                launch_timed(cx.prof(cx.owner, PK_APPLY), k_bi_v<T>, gr, bl, cx.s, a0, (const T *)ph, vv,
                             (const T *)r0, rap);
            if (sep) k_bi_fin<FIN_V><<<g1, BI_FIN_NT, 0, cx.s>>>(S_DONE, rap, nbt, it, 0.0, -1);
            k_bi_s<T><<<gf, bl, 0, cx.s>>>(n0, n1, r, vv, s, ra, (const double *)rap.part, sep ? 0 : nbt);
            k_bi_fin<FIN_S><<<g1, BI_FIN_NT, 0, cx.s>>>(S_DONE, ra, nb, it, 0.0, -1);
            CFD_LAUNCH_CHECK();
            const Tail tt{MGA_T, t, nullptr, rap};
            CFD_RET(precond(cx, s, sh, S_DONE, fz ? &tt : nullptr));
            if (!fz)
                launch_timed(cx.prof(cx.owner, PK_APPLY), k_bi_t<T>, gr, bl, cx.s, a0, (const T *)sh, t,
                             (const T *)s, rap);
            if (sep) k_bi_fin<FIN_T><<<g1, BI_FIN_NT, 0, cx.s>>>(S_DONE, rap, nbt, it, 0.0, -1);
            k_bi_xr<T><<<gf, bl, 0, cx.s>>>(n0, n1, it, ph, sh, s, t, r0, x, r, ra, (const double *)rap.part,
                                          sep ? 0 : nbt);
            // eager: the iteration's end flags go straight to host memory and the host
            // waits for them (no copy, no stream synchronisation)
            const bool flags = !graph && cx.flag_d && cx.seq;
            const unsigned long long sq = flags ? ++*cx.seq : 0;
            k_bi_fin<FIN_XR><<<g1, BI_FIN_NT, 0, cx.s>>>(S_DONE, ra, nb, it, 0.0, -1, sq, flags ? cx.flag_d : nullptr);
            CFD_LAUNCH_CHECK();
            if (flags) {
                bool done = false;
                CFD_RET(ps_wait(cx, sq, done));
                if (done) return CFD_OK;
            } else if (!graph) {
                CFD_RET(ps_sync(cx));
                if (cx.all_done()) return CFD_OK;
            }
        }
        return CFD_OK;
    }

    // x = one V-cycle from zero on b with its interior mean removed, no Krylov
    // iteration and no tolerance (the oracle's vcycle_solve()); S_SUM / S_BB hold
    // b's interior sums when sums_ready, as for solve().  A slot whose scalar og is
    // set gets x = 0.
    int vsolve(PsCtx &cx, const double *b, T *x, int og, bool sums_ready) {
        const int n0 = lev[0].n0, n1 = lev[0].n1;
        const dim3 gf(nflat(n0, n1), 1, B), bl(BX, BY), g1(1, 1, B);
        T *r = bi[0], *r0 = bi[1];
        const RedArgs ra = cx.ra();
        if (sums_ready) {
            k_bi_start_m<T><<<gf, bl, 0, cx.s>>>(og, n0, n1, b, x, r, r0, ra);
        } else {
            k_bi_start1<T><<<gf, bl, 0, cx.s>>>(og, n0, n1, b, x, nullptr, r, ra);
            k_bi_fin<FIN_START1><<<g1, BI_FIN_NT, 0, cx.s>>>(og, ra, nflat(n0, n1), 0, 0.0, -1);
            k_bi_start2<T><<<gf, bl, 0, cx.s>>>(og, n0, n1, r, r0, ra);
        }
        CFD_LAUNCH_CHECK();
        return precond(cx, r, x, og);
    }
};

}  // namespace cfd
