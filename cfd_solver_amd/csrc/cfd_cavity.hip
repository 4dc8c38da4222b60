// Cavity SIMPLE path on gfx950: device-resident batches of equally sized cases,
// per-stage kernels, BiCGStab + geometric multigrid for the pressure correction,
// and the outer-iteration driver behind the cfd_batch_* / cfd_case_* C ABI.
//
// A batch holds `slots` cases of n0 x n1 nodes.  Every kernel runs blockIdx.z
// over the slots: slot z's arrays start at z x n0 x n1, its physics (Re, Ri, Pr,
// relaxation, wall values) is SlotP z, its device scalars start at z x SS.  A
// slot that holds no case (SlotP::active = 0) costs one early return per block.
// A single case (cfd_case_*) is a batch of one slot.
//
// One outer iteration (all on the batch's stream):
//   k_residual      reference RHS (hybrid convection, grad p, buoyancy) minus the
//                   reference implicit operator applied to the current fields
//                   -> R_u, R_v, R_T, the d factors, U, V      (solver.cpp:1062-1236, 751-848)
//   k_jacobin       (A_lo/alpha) delta = R, Jacobi sweeps for u, v, T (four per launch)
//   k_face_div      u* = u + du, u-hat = u* + d grad p, and q = J-weighted divergence
//                   of the face velocities (solver.cpp:1389-1448)
//   p' solve        BiCGStab, right-preconditioned by a multigrid V-cycle
//   k_correct       u = u* - d grad p', p += alpha_p p', T += dT (solver.cpp:1504-1541)
//   BC kernels      adiabatic sides, wall pressure from normal momentum (solver.cpp:1667-1728)
//   reductions      sum|q|, sum|R_T|, dmax (solver.cpp:1806-1852), wall Nusselt (1951-1982)
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cfd_psolve.h"

namespace cfd {

constexpr int MAXLEV = PS_MAXLEV;
constexpr int CFD_PROF_KINDS = 8;
constexpr int PROF_EVENTS = 4096;
constexpr int NT = BX * BY;  // threads per 2-D block

// per-slot physics and wall data (device array, one entry per slot)
struct SlotP {
    double Re, Ri, Pr, dt, alpha_u, alpha_t, alpha_p;
    double lid, t_hot, t_cold, len_hot, len_cold;
    double active;  // 1: the slot holds a case
    double dnu, dnt;  // dt / Re, dt / (Re Pr)
    double scheme;   // convection next to walls (cfd_params::wall_upwind) on this slot's level
    double fas_src;  // 1: the slot is a coarse level of a FAS hierarchy (applies the level's sources)
    double deep;     // 1: the slot is the coarsest level of its case's hierarchy
    double pmaxit;   // > 0: the slot's p' BiCGStab iteration cap; < 0: its p' is one V-cycle
                     // (0: the level's p_maxit)
};

// grid shape shared by the slots, and the slots' physics
struct CGrid {
    int n0, n1;
    double h1, h2;
    size_t N;          // nodes per slot
    const SlotP *sp;   // sp[blockIdx.z]
    RecipT<double> r;  // reciprocals of the grid constants
    double blend;      // width of the blended hybrid switch (0: the reference's step)
    __host__ __device__ KGrid kg() const { return KGrid{n0, n1, h1, h2}; }
};

// Source terms of a coarse FAS level (null on a level without them): u, v, T of
// the momentum and energy residuals, m of the mass residual; on the border, T
// holds the side-wall and m the wall-pressure sources.  A slot that is not a coarse
// level (SlotP::fas_src = 0: the top level of a case packed beside coarse levels of
// larger ones) has none.
struct Src {
    const double *u, *v, *T, *m;
};
__device__ __forceinline__ Src src_at(Src s, size_t o, const SlotP &ph) {
    if (ph.fas_src == 0.0) return Src{nullptr, nullptr, nullptr, nullptr};
    if (s.u) {
        s.u += o; s.v += o; s.T += o; s.m += o;
    }
    return s;
}

struct Case {
    int n0 = 0, n1 = 0;
    int B = 1;           // slots
    size_t N = 0;        // nodes per slot
    double h1 = 0, h2 = 0;
    cfd_params prm{};
    bool mixed = false;
    hipStream_t s = nullptr;
    int device = 0;
    // per-slot physics (device copy, host copy) and host-side per-slot state
    SlotP *spd = nullptr;
    // the same table with only the slots that have a coarser level (spd_nd) or that
    // are a coarse level (spd_cr) active; both equal spd in a hierarchy of equal
    // slots, separate arrays in a pack's slab (cfd_pack)
    SlotP *spd_nd = nullptr, *spd_cr = nullptr;
    std::vector<SlotP> sph;
    std::vector<double> mass0, energy0;
    std::vector<int> iter;
    size_t scr = 0;      // scratch per slot
    // grid (every array: slots x N)
    double *xg, *yg, *xix, *xiy, *etx, *ety, *alph, *beta, *gamma, *ajac, *p1, *q1;
    // state and FP64 work
    double *u, *v, *T, *p;
    double *d;
    double *uh, *vh, *q, *pc, *scratch;
    // predictor arrays in the predictor precision (double, or float in the mixed path)
    void *Ru, *Rv, *RT, *U, *V;
    void *jb[6];                           // Jacobi ping-pong: du dv dT | tu tv tT
    float *a32[5] = {nullptr};             // alph, beta, gamma, p1, q1 in FP32 (mixed)
    // pressure correction
    Geom geom[MAXLEV];
    int nlev = 0;
    PSolver<double> *ps64 = nullptr;
    PSolver<float> *ps32 = nullptr;
    float *ef = nullptr;                   // mixed: FP32 correction of one refinement step
    double *r64 = nullptr;                 // mixed: FP64 defect
    double *f64[4] = {nullptr};            // mixed: level-0 face coefficients in FP64 (defect)
    double *part = nullptr;  // reduction partials: per slot PS_SLOTS rows of pstride
    int pstride = 0;
    double *scal = nullptr;  // device scalars, SS per slot
    double *hscal = nullptr; // pinned host mirror
    // per-slot end-of-iteration flags of the eager p' solve: pinned, coherent host
    // memory (host and device views) and their sequence counter
    unsigned long long *hflag = nullptr, *dflag = nullptr, flag_seq = 0;
    size_t bytes = 0;
    bool dry = false;    // size estimate only: no device work
    std::vector<void *> allocs;
    // live kernel timing: event pairs around fine-level launches of the hot kernels
    bool prof = false;
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_kind;   // kind of each pending pair
    int ev_used = 0;            // events recorded since the last drain
    double prof_ms[CFD_PROF_KINDS] = {0};
    long prof_n[CFD_PROF_KINDS] = {0};
    // residual tile rows (4 or 8) of this level, and whether the tuner has chosen them
    int res_ty = 8;
    bool res_tuned = false;
    // captured outer iteration
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool capturing = false;
    // FAS hierarchy: the next coarser level (same slots, same stream), this level's
    // residuals at its current state (FP64: u, v, T, and the wall pressure the wall
    // update would set), and on a coarse level its sources and restricted state
    Case *coarse = nullptr;
    bool own_stream = true;
    double *fr[4] = {nullptr};
    double *src[4] = {nullptr};
    double *x0[4] = {nullptr};
    bool use_src = false;
    // a coarse level run by k_simple_small: its p' levels' FP64 face coefficients
    // (east/north, normal/tangential) and inverse diagonals
    double *sl_tab[CW_LEVS][5] = {};
    struct SmallArgs *sl_args = nullptr;  // its arguments (device memory)
};

// ------------------------------------------------------------------------
// reductions

// ------------------------------------------------------------------------
// residual + d factors (oracle: or_cav_residuals)

// the grid arrays of one case, offset to a slot
struct GridPtrs {
    const double *xix, *xiy, *etx, *ety, *alph, *beta, *gamma, *p1, *q1;
};

// Tiled: a 64 x 8 tile of nodes per 64 x 4 block (two nodes per thread).  u, v, T
// (halo 2) and p (halo 1) are staged in LDS with coalesced loads issued together,
// as are the metric and operator coefficients of the thread's own nodes, so the
// 25-point rows read LDS instead of issuing a global load per stencil point.
// Halo entries beyond the grid hold clamped copies that no interior row reads.
// Tiles are RT_X wide and TY = 4 or 8 rows high (one or two nodes per thread):
// the autotuner (tune_res) times both on each level and keeps the faster (the
// taller tile re-reads fewer halo rows, the shorter one gives small levels more
// blocks and a shorter chain per thread); a node's arithmetic does not depend on it.
constexpr int RT_X = 64;
inline dim3 grid_res(int n0, int n1, int B, int ty) { return dim3(tiles_inner(n0, RT_X), tiles_inner(n1, ty), B); }

// metric and operator coefficients of one node
struct NC {
    double xix, xiy, etx, ety, al, be, ga, pp, qq;
};
// residuals, contravariant velocities and d of one interior node
struct ResOut {
    double ru, rv, rt, U, V, d;
};

// Residual at interior node (i, j) of grid g (oracle: or_cav_residuals): at(f, di, dj)
// reads field f (one of s_u, s_v, s_t, the arrays of u, v, T in the caller's layout)
// at the neighbour (i + di, j + dj), pat(di, dj) reads p there; P indexes the node's
// FAS sources.
template <typename FA, typename PA>
__device__ __forceinline__ ResOut residual_node(const CGrid &g, const SlotP &ph, const NC &c, int i, int j, size_t P,
                                                const double *s_u, const double *s_v, const double *s_t, FA at,
                                                PA pat, const Src &src) {
    const int n0 = g.n0, n1 = g.n1;
    const double h1 = g.h1, h2 = g.h2;
    const bool fit_i = (i >= 2 && i <= n0 - 3), fit_j = (j >= 2 && j <= n1 - 3);
    const bool hyb = fit_i && fit_j;
    int mode_i = hyb ? 2 : (fit_i ? 1 : 0);
    int mode_j = hyb ? 2 : (fit_j ? 1 : 0);
    const int scheme = (int)ph.scheme;
    if (scheme == 1) {
        mode_i = fit_i ? 2 : 3;
        mode_j = fit_j ? 2 : 3;
    } else if (scheme == 2) {
        mode_i = mode_j = 3;
    }
    const double al = c.al, be = c.be, ga = c.ga, pp = c.pp, qq = c.qq;
    const double UP = c.xix * at(s_u, 0, 0) + c.xiy * at(s_v, 0, 0);
    const double VP = c.etx * at(s_u, 0, 0) + c.ety * at(s_v, 0, 0);
    const double *fl[3] = {s_u, s_v, s_t};
    double conv[3];
    const double ial = 1.0 / al, iga = 1.0 / ga;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double *f = fl[k];
        const double sc = (k <= 1) ? ph.Re : ph.Re * ph.Pr;
        const double pec1 = UP * sc * h1 * ial;
        const double pec2 = VP * sc * h2 * iga;
        const double fm2i = fit_i ? at(f, -2, 0) : 0.0, fp2i = fit_i ? at(f, 2, 0) : 0.0;
        const double fm2j = fit_j ? at(f, 0, -2) : 0.0, fp2j = fit_j ? at(f, 0, 2) : 0.0;
        const double dxi = conv_deriv_r(mode_i, fm2i, at(f, -1, 0), at(f, 0, 0), at(f, 1, 0), fp2i, UP, pec1,
                                        g.r.ih1, g.blend);
        const double det = conv_deriv_r(mode_j, fm2j, at(f, 0, -1), at(f, 0, 0), at(f, 0, 1), fp2j, VP, pec2,
                                        g.r.ih2, g.blend);
        conv[k] = UP * dxi + VP * det;
    }
    const double dp_dxi = (pat(1, 0) - pat(-1, 0)) * g.r.i2h1;
    const double dp_de = (pat(0, 1) - pat(0, -1)) * g.r.i2h2;
    const double dp_dx = c.xix * dp_dxi + c.etx * dp_de;
    const double dp_dy = c.xiy * dp_dxi + c.ety * dp_de;
    const double dt = ph.dt;
    auto A_ = [&](const double *f, double dn) {
        return aref_offdiag_a([&](int di, int dj) { return at(f, di, dj); }, al, be, ga, pp, qq, g.r, dn, hyb);
    };
    double ru = dt * (-conv[0] - dp_dx) - A_(s_u, ph.dnu);
    double rv = dt * (-conv[1] - dp_dy + ph.Ri * at(s_t, 0, 0)) - A_(s_v, ph.dnu);
    double rt = -dt * conv[2] - A_(s_t, ph.dnt);
    if (src.u) {  // FAS sources of a coarse level
        ru += dt * src.u[P];
        rv += dt * src.v[P];
        rt += dt * src.T[P];
    }
    double cl[9];
    alo_coefs_r<double>(al, be, ga, pp, qq, UP, VP, g.r, dt, ph.dnu, cl);
    return ResOut{ru, rv, rt, UP, VP, dt * ph.alpha_u / cl[4]};
}

template <typename TR, int TY>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3))) void k_residual(
    CGrid g, const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ T,
    const double *__restrict__ p, GridPtrs gp, TR *__restrict__ Ru, TR *__restrict__ Rv, TR *__restrict__ RT,
    double *__restrict__ dd, TR *__restrict__ Ua, TR *__restrict__ Va, RedArgs ra, Src src) {
    const SlotP ph = g.sp[blockIdx.z];
    ra = rz(ra);
    const size_t o = zslot() * g.N;
    u += o; v += o; T += o; p += o; Ru += o; Rv += o; RT += o; dd += o; Ua += o; Va += o;
    src = src_at(src, o, ph);
    constexpr int FX = RT_X + 4, FY = TY + 4, PX = RT_X + 2, PY = TY + 2;
    constexpr int NODES = TY / BY;  // nodes per thread
    static_assert(TY % BY == 0, "tile rows");
    __shared__ double s_u[FX * FY], s_v[FX * FY], s_t[FX * FY], s_p[PX * PY];
    const int tid = tlin();
    const int n0 = g.n0, n1 = g.n1;
    const int2 tl = xcd_tile();
    const int i0 = tl.x * RT_X, j0 = tl.y * TY;
    struct F3 {
        double u, v, t;
    };
    const auto lf = batch_load<FX * FY, NT>(tid, [&](int k) {
        const size_t Q = idx(clampi(i0 - 2 + k % FX, 0, n0 - 1), clampi(j0 - 2 + k / FX, 0, n1 - 1), n0);
        return F3{u[Q], v[Q], T[Q]};
    });
    const auto lp = batch_load<PX * PY, NT>(tid, [&](int k) {
        return p[idx(clampi(i0 - 1 + k % PX, 0, n0 - 1), clampi(j0 - 1 + k / PX, 0, n1 - 1), n0)];
    });
    NC nc[NODES];
#pragma unroll
    for (int m = 0; m < NODES; m++) {
        const size_t Q = o + idx(clampi(i0 + (int)threadIdx.x, 0, n0 - 1),
                                 clampi(j0 + (int)threadIdx.y + m * BY, 0, n1 - 1), n0);
        nc[m] = NC{gp.xix[Q], gp.xiy[Q], gp.etx[Q], gp.ety[Q], gp.alph[Q],
                   gp.beta[Q], gp.gamma[Q], gp.p1[Q], gp.q1[Q]};
    }
    if (ph.active == 0.0) return;
    batch_run<FX * FY, NT>(tid, lf, [&](int k, const F3 &f) {
        s_u[k] = f.u;
        s_v[k] = f.v;
        s_t[k] = f.t;
    });
    batch_run<PX * PY, NT>(tid, lp, [&](int k, double pv) { s_p[k] = pv; });
    __syncthreads();
    double e_abs = 0.0, m_abs = 0.0;
#pragma unroll
    for (int m = 0; m < NODES; m++) {
        const int li = threadIdx.x, lj = threadIdx.y + m * BY;
        const int i = i0 + li, j = j0 + lj;
        if (i < 1 || i > n0 - 2 || j < 1 || j > n1 - 2) continue;
        const size_t P = idx(i, j, n0);
        auto at = [&](const double *sf, int di, int dj) { return sf[(lj + 2 + dj) * FX + (li + 2 + di)]; };
        auto pat = [&](int di, int dj) { return s_p[(lj + 1 + dj) * PX + (li + 1 + di)]; };
        const ResOut ro = residual_node(g, ph, nc[m], i, j, P, s_u, s_v, s_t, at, pat, src);
        Ru[P] = (TR)ro.ru;
        Rv[P] = (TR)ro.rv;
        RT[P] = (TR)ro.rt;
        Ua[P] = (TR)ro.U;
        Va[P] = (TR)ro.V;
        dd[P] = ro.d;
        e_abs += fabs(ro.rt);
        m_abs += fabs(ro.ru) + fabs(ro.rv);
    }
    double sums[2] = {e_abs, m_abs};
    block_partials<2>(sums, ra.part, ra.pstride);
}

// d on the border = nearest interior value (corners from the diagonal)
__global__ void k_d_border(int n0, int n1, double *__restrict__ d) {
    d += zslot() * n0 * n1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n0) {
        const int i = t;
        if (i >= 1 && i <= n0 - 2) {
            d[idx(i, 0, n0)] = d[idx(i, 1, n0)];
            d[idx(i, n1 - 1, n0)] = d[idx(i, n1 - 2, n0)];
        }
    }
    if (t < n1) {
        const int j = t;
        const int jj = j == 0 ? 1 : (j == n1 - 1 ? n1 - 2 : j);
        d[idx(0, j, n0)] = d[idx(1, jj, n0)];
        d[idx(n0 - 1, j, n0)] = d[idx(n0 - 2, jj, n0)];
    }
}

// one Jacobi sweep of (A_lo/alpha) delta = R for u, v (shared operator) and T,
// in the predictor precision TR
template <typename TR, bool FIRST>
__global__ __launch_bounds__(NT) void k_jacobi(
    CGrid g, const TR *__restrict__ Ru, const TR *__restrict__ Rv, const TR *__restrict__ RT,
    const TR *__restrict__ Ua, const TR *__restrict__ Va, const TR *__restrict__ alph, const TR *__restrict__ beta,
    const TR *__restrict__ gamma, const TR *__restrict__ p1, const TR *__restrict__ q1, const TR *__restrict__ du,
    const TR *__restrict__ dv, const TR *__restrict__ dT, TR *__restrict__ nu, TR *__restrict__ nv,
    TR *__restrict__ nT) {
    const SlotP ph = g.sp[blockIdx.z];
    if (ph.active == 0.0) return;
    {
        const size_t o = zslot() * g.N;
        Ru += o; Rv += o; RT += o; Ua += o; Va += o; alph += o; beta += o; gamma += o; p1 += o; q1 += o;
        du += o; dv += o; dT += o; nu += o; nv += o; nT += o;
    }
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const int n0 = g.n0, n1 = g.n1;
    if (i < 1 || i > n0 - 2 || j < 1 || j > n1 - 2) return;
    const size_t P = idx(i, j, n0);
    const TR dt = (TR)ph.dt;
    const RecipT<TR> r = recip_cast<TR>(g.r);
    TR cu[9], ct[9];
    alo_coefs_r<TR>(alph[P], beta[P], gamma[P], p1[P], q1[P], Ua[P], Va[P], r, dt, (TR)ph.dnu, cu);
    alo_coefs_r<TR>(alph[P], beta[P], gamma[P], p1[P], q1[P], Ua[P], Va[P], r, dt, (TR)ph.dnt, ct);
    TR su = Ru[P], sv = Rv[P], st = RT[P];
    if (!FIRST) {
#pragma unroll
        for (int dj = -1; dj <= 1; dj++)
#pragma unroll
            for (int di = -1; di <= 1; di++) {
                if (di == 0 && dj == 0) continue;
                const int k = (dj + 1) * 3 + (di + 1);
                const size_t Q = idx(i + di, j + dj, n0);
                su -= cu[k] * du[Q];
                sv -= cu[k] * dv[Q];
                st -= ct[k] * dT[Q];
            }
    }
    const TR iu = (TR)ph.alpha_u / cu[4], it = (TR)ph.alpha_t / ct[4];
    nu[P] = su * iu;
    nv[P] = sv * iu;
    nT[P] = st * it;
}

// Several Jacobi sweeps per launch (k_jacobin): arguments, and the tile.
template <typename TR>
struct JArgs {
    const TR *Ru, *Rv, *RT, *Ua, *Va, *alph, *beta, *gamma, *p1, *q1, *du, *dv, *dT;
    TR *nu, *nv, *nT;
};

constexpr int JT_X = 64, JT_Y = 8;   // output tile of k_jacobin

// the node's coefficient inputs and right-hand sides
template <typename TR>
struct JNode {
    TR al, be, ga, pp, qq, U, V, ru, rv, rt;
};
template <typename TR>
__device__ __forceinline__ JNode<TR> jac_load(const JArgs<TR> &a, size_t P) {
    return JNode<TR>{a.alph[P], a.beta[P], a.gamma[P], a.p1[P], a.q1[P], a.Ua[P], a.Va[P], a.Ru[P], a.Rv[P], a.RT[P]};
}

// one sweep at interior node (i, j); `dat(f, i, j)` reads the old correction of
// field f (0 u, 1 v, 2 T) at a neighbour
// the node's A_lo rows (u and v; T) and relaxed inverse diagonals
template <typename TR>
struct JRow {
    TR cu[9], ct[9], iu, it;
};
template <typename TR>
__device__ __forceinline__ JRow<TR> jac_row(const JNode<TR> &c, const SlotP &ph, const RecipT<TR> &r) {
    JRow<TR> w;
    const TR dt = (TR)ph.dt;
    alo_coefs_r<TR>(c.al, c.be, c.ga, c.pp, c.qq, c.U, c.V, r, dt, (TR)ph.dnu, w.cu);
    alo_coefs_r<TR>(c.al, c.be, c.ga, c.pp, c.qq, c.U, c.V, r, dt, (TR)ph.dnt, w.ct);
    w.iu = (TR)ph.alpha_u / w.cu[4];
    w.it = (TR)ph.alpha_t / w.ct[4];
    return w;
}
// one sweep at interior node (i, j) from the node's rows
template <typename TR, bool FIRST, typename F>
__device__ __forceinline__ void jac_node_w(const JNode<TR> &c, const JRow<TR> &w, int i, int j, F dat, TR &ou,
                                           TR &ov, TR &ot) {
    TR su = c.ru, sv = c.rv, st = c.rt;
    if (!FIRST) {
#pragma unroll
        for (int dj = -1; dj <= 1; dj++)
#pragma unroll
            for (int di = -1; di <= 1; di++) {
                if (di == 0 && dj == 0) continue;
                const int k = (dj + 1) * 3 + (di + 1);
                su -= w.cu[k] * dat(0, i + di, j + dj);
                sv -= w.cu[k] * dat(1, i + di, j + dj);
                st -= w.ct[k] * dat(2, i + di, j + dj);
            }
    }
    ou = su * w.iu;
    ov = sv * w.iu;
    ot = st * w.it;
}

template <typename TR, bool FIRST, typename F>
__device__ __forceinline__ void jac_node(const JNode<TR> &c, const SlotP &ph, const RecipT<TR> &r, int i, int j,
                                         F dat, TR &ou, TR &ov, TR &ot) {
    const TR dt = (TR)ph.dt;
    TR cu[9], ct[9];
    alo_coefs_r<TR>(c.al, c.be, c.ga, c.pp, c.qq, c.U, c.V, r, dt, (TR)ph.dnu, cu);
    alo_coefs_r<TR>(c.al, c.be, c.ga, c.pp, c.qq, c.U, c.V, r, dt, (TR)ph.dnt, ct);
    TR su = c.ru, sv = c.rv, st = c.rt;
    if (!FIRST) {
#pragma unroll
        for (int dj = -1; dj <= 1; dj++)
#pragma unroll
            for (int di = -1; di <= 1; di++) {
                if (di == 0 && dj == 0) continue;
                const int k = (dj + 1) * 3 + (di + 1);
                su -= cu[k] * dat(0, i + di, j + dj);
                sv -= cu[k] * dat(1, i + di, j + dj);
                st -= ct[k] * dat(2, i + di, j + dj);
            }
    }
    const TR iu = (TR)ph.alpha_u / cu[4], it = (TR)ph.alpha_t / ct[4];
    ou = su * iu;
    ov = sv * iu;
    ot = st * it;
}

// NS sweeps per launch on a JT_X x JT_Y tile (temporal blocking): the old
// correction (halo NS) is staged in LDS, and the coefficients of the first
// sweep's region (halo NS - 1) are loaded into registers, all loads issued
// together.  Sweep s covers the region with halo NS - 1 - s, on the inner part of
// the same elements, so a node's coefficients are read once per launch; the
// intermediate sweeps ping-pong between two LDS buffers (zero on wall nodes, like
// the arrays) and the last writes the tile.  Node for node the arithmetic of
// k_jacobi.  Halo entries beyond the grid hold clamped copies that no interior
// node reads.
// FP32 sweeps after the first run four waves per SIMD (at 128 VGPRs, with a few
// spilled): 5 % faster per launch than three waves; the first sweep, which stages
// no old correction, is faster at its natural three.
template <typename TR, bool FIRST, int NS>
__global__ __launch_bounds__(NT, sizeof(TR) == 4 && !FIRST ? 4 : 1) void k_jacobin(CGrid g, JArgs<TR> a) {
    const SlotP ph = g.sp[blockIdx.z];
    {
        const size_t o = zslot() * g.N;
        a.Ru += o; a.Rv += o; a.RT += o; a.Ua += o; a.Va += o; a.alph += o; a.beta += o; a.gamma += o;
        a.p1 += o; a.q1 += o; a.du += o; a.dv += o; a.dT += o; a.nu += o; a.nv += o; a.nT += o;
    }
    constexpr int AX = JT_X + 2 * NS, AY = JT_Y + 2 * NS;              // old correction, halo NS
    constexpr int MX = JT_X + 2 * (NS - 1), MY = JT_Y + 2 * (NS - 1);  // first sweep, halo NS - 1
    constexpr int NA = FIRST ? 1 : AX * AY, NM = MX * MY;
    __shared__ TR s_old[3 * NA];
    __shared__ TR s_buf[2][3 * NM];
    const int tid = tlin();
    const int n0 = g.n0, n1 = g.n1;
    const int2 tl = xcd_tile();
    const int fi0 = tl.x * JT_X, fj0 = tl.y * JT_Y;
    const RecipT<TR> r = recip_cast<TR>(g.r);
    struct O3 {
        TR u, v, t;
    };
    const auto lo = batch_load<NA, NT>(tid, [&](int k) {
        if (FIRST) return O3{TR(0), TR(0), TR(0)};
        const size_t Q = idx(clampi(fi0 - NS + k % AX, 0, n0 - 1), clampi(fj0 - NS + k / AX, 0, n1 - 1), n0);
        return O3{a.du[Q], a.dv[Q], a.dT[Q]};
    });
    const auto lc = batch_load<NM, NT>(tid, [&](int k) {
        return jac_load<TR>(a, idx(clampi(fi0 - (NS - 1) + k % MX, 0, n0 - 1),
                                   clampi(fj0 - (NS - 1) + k / MX, 0, n1 - 1), n0));
    });
    if (ph.active == 0.0) return;
    if (!FIRST) {
        batch_run<NA, NT>(tid, lo, [&](int k, const O3 &v) {
            s_old[k] = v.u;
            s_old[NA + k] = v.v;
            s_old[2 * NA + k] = v.t;
        });
        __syncthreads();
    }
    auto old_at = [&](int f, int ii, int jj) { return s_old[f * NA + (jj - fj0 + NS) * AX + (ii - fi0 + NS)]; };
#pragma unroll
    for (int sw = 0; sw < NS; sw++) {
        const TR *in = s_buf[(sw + 1) & 1];  // the previous sweep's output (sw >= 1)
        TR *out = s_buf[sw & 1];
        auto mid_at = [&](int f, int ii, int jj) {
            return in[f * NM + (jj - fj0 + NS - 1) * MX + (ii - fi0 + NS - 1)];
        };
        batch_run<NM, NT>(tid, lc, [&](int k, const JNode<TR> &c) {
            const int li = k % MX, lj = k / MX;
            if (li < sw || li >= MX - sw || lj < sw || lj >= MY - sw) return;  // outside this sweep's region
            const int gi = fi0 - (NS - 1) + li, gj = fj0 - (NS - 1) + lj;
            const bool in_ = inner_node(gi, gj, n0, n1);
            if (sw == NS - 1) {  // the tile itself: to global memory
                if (!in_) return;
                const size_t P = idx(gi, gj, n0);
                TR ou, ov, ot;
                if (sw == 0) jac_node<TR, FIRST>(c, ph, r, gi, gj, old_at, ou, ov, ot);
                else jac_node<TR, false>(c, ph, r, gi, gj, mid_at, ou, ov, ot);
                a.nu[P] = ou;
                a.nv[P] = ov;
                a.nT[P] = ot;
                return;
            }
            TR ou = TR(0), ov = TR(0), ot = TR(0);
            if (in_) {
                if (sw == 0) jac_node<TR, FIRST>(c, ph, r, gi, gj, old_at, ou, ov, ot);
                else jac_node<TR, false>(c, ph, r, gi, gj, mid_at, ou, ov, ot);
            }
            out[k] = ou;
            out[NM + k] = ov;
            out[2 * NM + k] = ot;
        });
        if (sw < NS - 1) __syncthreads();
    }
}

// arrays a face-divergence node reads at its neighbours: u-hat, v-hat, d, xi_x, xi_y,
// eta_x, eta_y, J
enum { FD_UH, FD_VH, FD_D, FD_X, FD_Y, FD_E, FD_F, FD_J, FD_N };

// J-weighted divergence of the face velocities at interior node (i, j) (oracle:
// face_div, without its source): L(a, di, dj) reads array a (FD_*) at the neighbour
// (i + di, j + dj), P_(di, dj) reads p there.  Every face is evaluated, and the
// faces touching a wall node are then replaced by zero, so wall values read here
// need only be finite or not at all: they are selected away.
template <typename LA, typename PA>
__device__ __forceinline__ double face_div_node(const CGrid &g, int i, int j, LA L, PA P_) {
    const int n0 = g.n0, n1 = g.n1;
    // face between (i, j) and its neighbour (di, dj) (one of them zero): the
    // face-normal velocity times the J-weighted metric of direction xi (di != 0)
    // or eta (dj != 0)
    auto face = [&](int di, int dj, double dpdxi, double dpde) {
        const double dm = 0.5 * (L(FD_D, 0, 0) + L(FD_D, di, dj));
        const double uf = 0.5 * (L(FD_UH, 0, 0) + L(FD_UH, di, dj)) -
                          0.5 * dm * ((L(FD_X, 0, 0) + L(FD_X, di, dj)) * dpdxi +
                                      (L(FD_E, 0, 0) + L(FD_E, di, dj)) * dpde);
        const double vf = 0.5 * (L(FD_VH, 0, 0) + L(FD_VH, di, dj)) -
                          0.5 * dm * ((L(FD_Y, 0, 0) + L(FD_Y, di, dj)) * dpdxi +
                                      (L(FD_F, 0, 0) + L(FD_F, di, dj)) * dpde);
        const int mx = di != 0 ? FD_X : FD_E, my = di != 0 ? FD_Y : FD_F;
        return 0.5 * (L(FD_J, 0, 0) * L(mx, 0, 0) + L(FD_J, di, dj) * L(mx, di, dj)) * uf +
               0.5 * (L(FD_J, 0, 0) * L(my, 0, 0) + L(FD_J, di, dj) * L(my, di, dj)) * vf;
    };
    double fe = face(1, 0, (P_(1, 0) - P_(0, 0)) * g.r.ih1,
                     (P_(1, 1) + P_(0, 1) - P_(0, -1) - P_(1, -1)) * (0.25 * g.r.ih2));
    double fw = face(-1, 0, (P_(0, 0) - P_(-1, 0)) * g.r.ih1,
                     (P_(0, 1) + P_(-1, 1) - P_(0, -1) - P_(-1, -1)) * (0.25 * g.r.ih2));
    double gn = face(0, 1, (P_(1, 1) - P_(-1, 1) + P_(1, 0) - P_(-1, 0)) * (0.25 * g.r.ih1),
                     (P_(0, 1) - P_(0, 0)) * g.r.ih2);
    double gs = face(0, -1, (P_(1, 0) - P_(-1, 0) + P_(1, -1) - P_(-1, -1)) * (0.25 * g.r.ih1),
                     (P_(0, 0) - P_(0, -1)) * g.r.ih2);
    fe = i + 1 != n0 - 1 ? fe : 0.0;
    fw = i - 1 != 0 ? fw : 0.0;
    gn = j + 1 != n1 - 1 ? gn : 0.0;
    gs = j - 1 != 0 ? gs : 0.0;
    return (fe - fw) * g.r.ih1 + (gn - gs) * g.r.ih2;
}

// Face velocities and their divergence in one pass.  On a 64 x 4 tile plus a
// one-node halo, û = u* + d grad p with u* = u + du (the pressure-free velocity
// the face interpolation starts from) is evaluated into LDS from the tile's
// inputs (p with a two-node halo); then q, the J-weighted flux divergence of the
// face velocities, with block partials of Σ|q| (mass residual, rmass) and of Σq,
// Σq² (the FP64 norms the mixed p' solve starts from, rq).  All loads of the tile
// are issued together.  Halo entries beyond the grid hold clamped copies, and û
// on a wall node is evaluated like an interior one: both only enter faces that
// touch a wall, which are replaced by zero.
template <typename TR>
__global__ __launch_bounds__(NT) void k_face_div(
    CGrid g, const double *__restrict__ u, const double *__restrict__ v, const TR *__restrict__ du,
    const TR *__restrict__ dv, const double *__restrict__ p, const double *__restrict__ d,
    const double *__restrict__ xix, const double *__restrict__ xiy, const double *__restrict__ etx,
    const double *__restrict__ ety, const double *__restrict__ jac, double *__restrict__ q, RedArgs rmass,
    RedArgs rq, const double *__restrict__ sm) {
    const bool active = g.sp[blockIdx.z].active != 0.0;
    if (g.sp[blockIdx.z].fas_src == 0.0) sm = nullptr;
    rmass = rz(rmass);
    rq = rz(rq);
    const size_t o = zslot() * g.N;
    q += o;
    constexpr int HX = BX + 2, HY = BY + 2, PX = BX + 4, PY = BY + 4;
    enum { UH = FD_UH, VH = FD_VH, DD = FD_D, XX = FD_X, XY = FD_Y, EX = FD_E, EY = FD_F, JJ = FD_J, NA = FD_N };
    __shared__ double s[NA][HX * HY];
    __shared__ double s_p[PX * PY];
    const int n0 = g.n0, n1 = g.n1;
    const int tid = tlin();
    const int2 tl = xcd_tile();
    const int i0 = tl.x * BX, j0 = tl.y * BY;
    struct H {
        double u, v, du, dv, d, x, y, e, f, j;
    };
    const auto lh = batch_load<HX * HY, NT>(tid, [&](int k) {
        const size_t Q = o + idx(clampi(i0 - 1 + k % HX, 0, n0 - 1), clampi(j0 - 1 + k / HX, 0, n1 - 1), n0);
        return H{u[Q], v[Q], du ? (double)du[Q] : 0.0, dv ? (double)dv[Q] : 0.0, d[Q], xix[Q], xiy[Q], etx[Q], ety[Q],
                 jac[Q]};
    });
    const auto lp = batch_load<PX * PY, NT>(tid, [&](int k) {
        return p[o + idx(clampi(i0 - 2 + k % PX, 0, n0 - 1), clampi(j0 - 2 + k / PX, 0, n1 - 1), n0)];
    });
    if (!active) return;
    batch_run<PX * PY, NT>(tid, lp, [&](int k, double pv) { s_p[k] = pv; });
    batch_run<HX * HY, NT>(tid, lh, [&](int k, const H &h) {
        s[DD][k] = h.d;
        s[XX][k] = h.x;
        s[XY][k] = h.y;
        s[EX][k] = h.e;
        s[EY][k] = h.f;
        s[JJ][k] = h.j;
    });
    __syncthreads();
    // û, v̂ at every node of the halo-1 region (the arithmetic of u* and û)
    batch_run<HX * HY, NT>(tid, lh, [&](int k, const H &h) {
        const int ki = k % HX, kj = k / HX;
        auto P_ = [&](int di, int dj) { return s_p[(kj + 1 + dj) * PX + (ki + 1 + di)]; };
        const double a = h.u + h.du, b = h.v + h.dv;
        const double dp_dxi = (P_(1, 0) - P_(-1, 0)) * g.r.i2h1;
        const double dp_de = (P_(0, 1) - P_(0, -1)) * g.r.i2h2;
        s[UH][k] = a + h.d * (h.x * dp_dxi + h.e * dp_de);
        s[VH][k] = b + h.d * (h.y * dp_dxi + h.f * dp_de);
    });
    __syncthreads();
    const int li = threadIdx.x, lj = threadIdx.y;
    const int i = i0 + li, j = j0 + lj;
    double aq[1] = {0.0}, nq[2] = {0.0, 0.0};
    if (i >= 1 && i <= n0 - 2 && j >= 1 && j <= n1 - 2) {
        // array a at the node offset (di, dj) from (i, j)
        auto L = [&](int a, int di, int dj) { return s[a][(lj + 1 + dj) * HX + (li + 1 + di)]; };
        auto P_ = [&](int di, int dj) { return s_p[(lj + 2 + dj) * PX + (li + 2 + di)]; };
        double qq = face_div_node(g, i, j, L, P_);
        if (sm) qq -= sm[o + idx(i, j, n0)];
        q[idx(i, j, n0)] = qq;
        aq[0] = fabs(qq);
        nq[0] = qq;
        nq[1] = qq * qq;
    }
    double all[3] = {aq[0], nq[0], nq[1]};
    block_sums<3, NT>(all);
    if (tid == 0) {
        rmass.part[blin()] = all[0];
        rq.part[blin()] = all[1];
        rq.part[(size_t)rq.pstride + blin()] = all[2];
    }
}

// u = u* - d grad p' with u* = u + du (p' read clamped = wall copy), p += alpha_p p'
// with p' taken with its interior mean (scal[S_PSUM] / interior count) removed,
// T += dT (interior); then the signed max of the nodal dilation of the corrected
// velocities (solver.cpp:1821-1850) as block partials (rd).  A 64 x 4 tile computes
// the corrected u, v on itself plus a one-node halo into LDS (p' staged with a
// two-node halo; wall nodes keep their stored velocities, which no correction
// changes), writes its own nodes and takes the dilation from LDS.  All loads of
// the tile are issued together; halo entries beyond the grid are clamped copies
// that no interior node uses.
template <typename TR>
__global__ __launch_bounds__(NT) void k_correct(
    CGrid g, const TR *__restrict__ du, const TR *__restrict__ dv, const TR *__restrict__ dT,
    const double *__restrict__ pc, const double *__restrict__ d, const double *__restrict__ xix,
    const double *__restrict__ xiy, const double *__restrict__ etx, const double *__restrict__ ety,
    const double *__restrict__ scal, double *__restrict__ u, double *__restrict__ v, double *__restrict__ p,
    double *__restrict__ T, RedArgs rd) {
    const SlotP &sp = g.sp[blockIdx.z];
    const size_t o = zslot() * g.N;
    rd = rz(rd);
    scal += zslot() * SS;
    constexpr int HX = BX + 2, HY = BY + 2, PX = BX + 4, PY = BY + 4;
    __shared__ double s_pc[PX * PY];
    __shared__ double s_u[HX * HY], s_v[HX * HY];
    const int n0 = g.n0, n1 = g.n1;
    const int tid = tlin();
    const int2 tl = xcd_tile();
    const int i0 = tl.x * BX, j0 = tl.y * BY;
    struct H {
        double u, v, du, dv, d, x, y, e, f;
    };
    const auto lh = batch_load<HX * HY, NT>(tid, [&](int k) {
        const size_t Q = o + idx(clampi(i0 - 1 + k % HX, 0, n0 - 1), clampi(j0 - 1 + k / HX, 0, n1 - 1), n0);
        return H{u[Q], v[Q], (double)du[Q], (double)dv[Q], d[Q], xix[Q], xiy[Q], etx[Q], ety[Q]};
    });
    const auto lp = batch_load<PX * PY, NT>(tid, [&](int k) {
        return pc[o + idx(clampi(i0 - 2 + k % PX, 0, n0 - 1), clampi(j0 - 2 + k / PX, 0, n1 - 1), n0)];
    });
    const int li = threadIdx.x, lj = threadIdx.y;
    const int i = i0 + li, j = j0 + lj;
    const bool own = i >= 1 && i <= n0 - 2 && j >= 1 && j <= n1 - 2;
    const size_t P = o + idx(clampi(i, 0, n0 - 1), clampi(j, 0, n1 - 1), n0);
    const double p_own = p[P], T_own = T[P], dT_own = (double)dT[P];
    const double m_x = xix[P], m_y = xiy[P], m_e = etx[P], m_f = ety[P];
    if (sp.active == 0.0) return;
    const double alpha_p = sp.alpha_p;
    const double m = scal[S_PSUM] / ((double)(n0 - 2) * (double)(n1 - 2));
    batch_run<PX * PY, NT>(tid, lp, [&](int k, double pv) { s_pc[k] = pv; });
    __syncthreads();
    // corrected velocities of the halo-1 region
    batch_run<HX * HY, NT>(tid, lh, [&](int k, const H &h) {
        const int gi = i0 - 1 + k % HX, gj = j0 - 1 + k / HX;
        double un = h.u, vn = h.v;
        if (gi >= 1 && gi <= n0 - 2 && gj >= 1 && gj <= n1 - 2) {
            auto PC = [&](int ii, int jj) { return s_pc[(jj - j0 + 2) * PX + (ii - i0 + 2)]; };
            const double pe = PC(clampi(gi + 1, 1, n0 - 2), gj) - m, pw = PC(clampi(gi - 1, 1, n0 - 2), gj) - m;
            const double pn = PC(gi, clampi(gj + 1, 1, n1 - 2)) - m, ps = PC(gi, clampi(gj - 1, 1, n1 - 2)) - m;
            const double dpx = 0.5 * (pe - pw) * g.r.ih1;
            const double dpe = 0.5 * (pn - ps) * g.r.ih2;
            const double us = h.u + h.du, vs = h.v + h.dv;
            un = us - h.d * (h.x * dpx + h.e * dpe);
            vn = vs - h.d * (h.y * dpx + h.f * dpe);
        }
        s_u[k] = un;
        s_v[k] = vn;
    });
    __syncthreads();
    double dm = 0.0;
    if (own) {
        auto U = [&](int di, int dj) { return s_u[(lj + 1 + dj) * HX + (li + 1 + di)]; };
        auto V = [&](int di, int dj) { return s_v[(lj + 1 + dj) * HX + (li + 1 + di)]; };
        u[P] = U(0, 0);
        v[P] = V(0, 0);
        p[P] = p_own + alpha_p * (s_pc[(lj + 2) * PX + (li + 2)] - m);
        T[P] = T_own + dT_own;
        dm = m_x * (U(1, 0) - U(-1, 0)) / (2 * g.h1) + m_e * (U(0, 1) - U(0, -1)) / (2 * g.h2) +
             m_f * (V(0, 1) - V(0, -1)) / (2 * g.h2) + m_y * (V(1, 0) - V(-1, 0)) / (2 * g.h1);
    }
    dm = block_max<NT>(dm);
    if (tid == 0) rd.part[blin()] = dm;
}

// New temperature of side-wall node (0, j) (east: (n0 - 1, j)) from the old values
// T(i, j) and the node's alpha, beta (oracle: temperature_sides, without its source)
template <typename TA>
__device__ __forceinline__ double t_side_value(const KGrid &g, TA T, double alph, double beta, bool east, int j) {
    const int n0 = g.n0, i = east ? n0 - 1 : 0;
    const double te = (T(i, j + 1) - T(i, j - 1)) / (2.0 * g.h2);
    const double txi = beta / alph * te;
    return east ? (4.0 * T(n0 - 2, j) - T(n0 - 3, j) + 2.0 * g.h1 * txi) / 3.0
                : (4.0 * T(1, j) - T(2, j) - 2.0 * g.h1 * txi) / 3.0;
}

// Adiabatic side walls spread over the chip (the column-strided wall accesses of
// one block would be bound by that CU's outstanding misses): the new wall values,
// computed from the old ones, go to scratch; k_p_walls uses and writes them.
constexpr int WB = 64;  // threads per block of the wall kernels
__global__ __launch_bounds__(WB) void k_t_sides(CGrid g, const double *__restrict__ T,
                                                const double *__restrict__ alph, const double *__restrict__ beta,
                                                double *__restrict__ scratch, size_t scr,
                                                const double *__restrict__ sT) {
    if (g.sp[blockIdx.z].active == 0.0) return;
    if (g.sp[blockIdx.z].fas_src == 0.0) sT = nullptr;
    {
        const size_t o = zslot() * g.N;
        T += o; alph += o; beta += o;
        if (sT) sT += o;
        scratch += zslot() * scr;
    }
    const int n0 = g.n0, n1 = g.n1;
    const int j = 1 + (int)(blockIdx.x * WB + threadIdx.x);
    if (j > n1 - 2) return;
    auto Tat = [&](int ii, int jj) { return T[idx(ii, jj, n0)]; };
    scratch[j] = t_side_value(g.kg(), Tat, alph[idx(0, j, n0)], beta[idx(0, j, n0)], false, j);
    scratch[n1 + j] = t_side_value(g.kg(), Tat, alph[idx(n0 - 1, j, n0)], beta[idx(n0 - 1, j, n0)], true, j);
    if (sT) {  // FAS boundary sources of a coarse level
        scratch[j] += sT[idx(0, j, n0)];
        scratch[n1 + j] += sT[idx(n0 - 1, j, n0)];
    }
}
// wall pressure from the normal momentum equation (oracle: wall_p_value)
struct WallIn {
    const double *u, *v, *T, *p, *xix, *xiy, *etx, *ety, *alph, *beta, *gamma, *p1, *q1, *ajac;
};
__device__ __forceinline__ WallIn wall_at(WallIn w, size_t o) {
    w.u += o; w.v += o; w.T += o; w.p += o; w.xix += o; w.xiy += o; w.etx += o; w.ety += o;
    w.alph += o; w.beta += o; w.gamma += o; w.p1 += o; w.q1 += o; w.ajac += o;
    return w;
}

// Component k (0: x, from u; 1: y, from v, with the buoyancy term) of the pressure
// gradient the momentum equation gives at wall node (i, j) of wall `side`.  Tw: the
// wall node's temperature (the side walls' new values are passed in before they
// are written back)
__device__ __forceinline__ double wall_dp(const WallIn &w, KGrid g, double Re, double Ri, int side, int i, int j,
                                          double Tw, int k) {
    const int n0 = g.n0;
    const double h1 = g.h1, h2 = g.h2;
    const size_t P = idx(i, j, n0);
    const double U = w.xix[P] * w.u[P] + w.xiy[P] * w.v[P];
    const double V = w.etx[P] * w.u[P] + w.ety[P] * w.v[P];
    {
        const double *f = k == 0 ? w.u : w.v;
#define F(ii, jj) f[idx(ii, jj, n0)]
        double d2u, conv;
        if (side <= 1) {
            const int s = (side == 0) ? 1 : -1;
            const double aa = w.alph[P] * (F(i + 1, j) + F(i - 1, j) - 2 * F(i, j)) / (h1 * h1);
            const double gg = w.gamma[P] * (F(i, j + 2 * s) + F(i, j) - 2 * F(i, j + s)) / (h2 * h2);
            double bb;
            if (side == 0)
                bb = w.beta[P] * (F(i + 1, j + 1) + F(i - 1, j) - F(i - 1, j + 1) - F(i + 1, j)) / (2 * h1 * h2);
            else
                bb = w.beta[P] * (F(i + 1, j) + F(i - 1, j - 1) - F(i + 1, j - 1) - F(i - 1, j)) / (2 * h1 * h2);
            const double qqq = w.q1[P] * s * (-3 * F(i, j) + 4 * F(i, j + s) - F(i, j + 2 * s)) / (2 * h2);
            d2u = aa + gg - 2 * bb + qqq;
            conv = U * 0.5 * (F(i + 1, j) - F(i - 1, j)) / h1;
            conv = conv + V * s * (-3 * F(i, j) + 4 * F(i, j + s) - F(i, j + 2 * s)) / (2 * h2);
        } else {
            const int s = (side == 2) ? 1 : -1;
            const double aa = w.gamma[P] * (F(i, j + 1) + F(i, j - 1) - 2 * F(i, j)) / (h2 * h2);
            const double gg = w.alph[P] * (F(i + 2 * s, j) + F(i, j) - 2 * F(i + s, j)) / (h1 * h1);
            double bb;
            if (side == 2)
                bb = w.beta[P] * (F(i + 1, j + 1) + F(i, j - 1) - F(i + 1, j - 1) - F(i, j + 1)) / (2 * h1 * h2);
            else
                bb = w.beta[P] * (F(i, j + 1) + F(i - 1, j - 1) - F(i, j - 1) - F(i - 1, j + 1)) / (2 * h1 * h2);
            const double ppp = w.p1[P] * s * (-3 * F(i, j) + 4 * F(i + s, j) - F(i + 2 * s, j)) / (2 * h1);
            d2u = aa + gg - 2 * bb + ppp;
            conv = V * 0.5 * (F(i, j + 1) - F(i, j - 1)) / h2;
            conv = conv + U * s * (-3 * F(i, j) + 4 * F(i + s, j) - F(i + 2 * s, j)) / (2 * h1);
        }
#undef F
        return k == 0 ? d2u / Re - conv : d2u / Re - conv + Ri * Tw;
    }
}
// the wall value from the two components: p at the next node inward minus the
// normal derivative times the spacing
__device__ __forceinline__ double wall_p_from(const WallIn &w, KGrid g, int side, int i, int j, double dpx,
                                              double dpy) {
    const int n0 = g.n0;
    const double h1 = g.h1, h2 = g.h2;
    const size_t P = idx(i, j, n0);
    const double J = w.ajac[P];
    if (side <= 1) {
        const double tx = -w.xiy[P] * J, ty = w.xix[P] * J;
        const double dpn = (dpx * tx + dpy * ty) * h2;
        return side == 0 ? w.p[idx(i, j + 1, n0)] - dpn : w.p[idx(i, j - 1, n0)] + dpn;
    }
    const double tx = w.ety[P] * J, ty = -w.etx[P] * J;
    const double dpn = (dpx * tx + dpy * ty) * h1;
    return side == 2 ? w.p[idx(i + 1, j, n0)] - dpn : w.p[idx(i - 1, j, n0)] + dpn;
}
__device__ double wall_p_value(const WallIn &w, KGrid g, double Re, double Ri, int side, int i, int j, double Tw) {
    const double dpx = wall_dp(w, g, Re, Ri, side, i, j, Tw, 0), dpy = wall_dp(w, g, Re, Ri, side, i, j, Tw, 1);
    return wall_p_from(w, g, side, i, j, dpx, dpy);
}

// Wall pressure written in place: wall_p_value reads p only at interior nodes, so
// no thread reads a value another one writes.  Block row y < 4 holds the nodes of
// wall y (one per thread); row 4 the corners, each the mean of the two adjacent
// wall values, computed by a lane pair.  The side walls' new temperatures come from
// scratch (k_t_sides): each side-wall thread uses its node's value and writes it
// into T; no thread here reads side-wall T from T.
// sm (may be null): FAS boundary sources added to every wall value after the
// corners are formed; pout (may be null): the new values go there instead of into p
// and T is left alone (the wall residual of a FAS restriction).
__global__ __launch_bounds__(WB) void k_p_walls(CGrid cg, WallIn w, double *__restrict__ p,
                                                double *__restrict__ T, const double *__restrict__ scratch,
                                                size_t scr, const double *__restrict__ sm, double *__restrict__ pout) {
    const SlotP &sp = cg.sp[blockIdx.z];
    if (sp.active == 0.0) return;
    const double Re = sp.Re, Ri = sp.Ri;
    const KGrid g = cg.kg();
    w = wall_at(w, zslot() * cg.N);
    p += zslot() * cg.N;
    T += zslot() * cg.N;
    if (sp.fas_src == 0.0) sm = nullptr;
    if (sm) sm += zslot() * cg.N;
    double *const pw = pout ? pout + zslot() * cg.N : p;
    scratch += zslot() * scr;
    const int side = blockIdx.y;
    // a lane pair per wall node: lane part 0 forms the x component of the pressure
    // gradient, lane part 1 the y component (two chains of half the length), and
    // part 0 combines them
    const int t = (int)(blockIdx.x * WB + threadIdx.x) >> 1, part = threadIdx.x & 1;
    const int n0 = g.n0, n1 = g.n1;
    // temperature of wall node (i, j) of wall `s`: new side-wall value or T
    auto Tw = [&](int s, int i, int j) {
        return s == 2 ? scratch[j] : (s == 3 ? scratch[n1 + j] : T[idx(i, j, n0)]);
    };
    if (side < 4) {
        const int len = side <= 1 ? n0 : n1;
        if (t < 1 || t > len - 2) return;
        const int i = side == 2 ? 0 : (side == 3 ? n0 - 1 : t);
        const int j = side == 0 ? 0 : (side == 1 ? n1 - 1 : t);
        const double tw = Tw(side, i, j);
        const size_t P = idx(i, j, n0);
        const double dp = wall_dp(w, g, Re, Ri, side, i, j, tw, part);
        const double dq = __shfl_xor(dp, 1, 64);  // the pair's other component
        if (part == 0) {
            pw[P] = wall_p_from(w, g, side, i, j, dp, dq) + (sm ? sm[P] : 0.0);
            if (side >= 2 && !pout) T[P] = tw;
        }
        return;
    }
    if (blockIdx.x != 0 || threadIdx.x >= 16) return;
    // corner c = threadIdx.x / 4 (bit 0: east, bit 1: north); lanes h = 0 of it take
    // the bottom/top wall value next to it, h = 1 the side wall value, each value
    // formed by a lane pair as above
    const int c = threadIdx.x >> 2, h = (threadIdx.x >> 1) & 1;
    const bool east = c & 1, north = c & 2;
    const int ci = east ? n0 - 1 : 0, cj = north ? n1 - 1 : 0;
    const int ws = h == 0 ? (north ? 1 : 0) : (east ? 3 : 2);
    const int wi = h == 0 ? (east ? n0 - 2 : 1) : ci;
    const int wj = h == 0 ? cj : (north ? n1 - 2 : 1);
    const double dp = wall_dp(w, g, Re, Ri, ws, wi, wj, Tw(ws, wi, wj), part);
    const double dq = __shfl_xor(dp, 1, 64);
    const double v = wall_p_from(w, g, ws, wi, wj, part == 0 ? dp : dq, part == 0 ? dq : dp);
    const double o = __shfl_xor(v, 2, 64);
    const size_t C = idx(ci, cj, n0);
    if (h == 0 && part == 0) pw[C] = 0.5 * (v + o) + (sm ? sm[C] : 0.0);
}

// velocity and temperature wall values (oracle: velocity_temperature_bcs)
__global__ void k_uvt_walls(CGrid g, double *__restrict__ u, double *__restrict__ v, double *__restrict__ T) {
    const SlotP &sp = g.sp[blockIdx.z];
    if (sp.active == 0.0) return;
    const double lid = sp.lid, th = sp.t_hot, tc = sp.t_cold;
    {
        const size_t o = zslot() * g.N;
        u += o; v += o; T += o;
    }
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n0 = g.n0, n1 = g.n1;
    if (t < n0) {
        u[idx(t, 0, n0)] = 0.0;
        v[idx(t, 0, n0)] = 0.0;
        u[idx(t, n1 - 1, n0)] = (t == 0 || t == n0 - 1) ? 0.0 : lid;
        v[idx(t, n1 - 1, n0)] = 0.0;
        T[idx(t, 0, n0)] = th;
        T[idx(t, n1 - 1, n0)] = tc;
    }
    if (t < n1) {
        u[idx(0, t, n0)] = 0.0;
        v[idx(0, t, n0)] = 0.0;
        u[idx(n0 - 1, t, n0)] = 0.0;
        v[idx(n0 - 1, t, n0)] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// SIMPLE iterations of a small FAS level in one workgroup (oracle: or_cav_outer
// without its statistics).  A coarse level of at most 33 x 33 nodes runs its
// iterations (the coarsest level fas_coarse of them per cycle) at a few
// microseconds of work per stage, which the multi-kernel iteration spreads over
// ~50 launches.  Here one 1024-thread workgroup per slot runs `iters` iterations
// with block barriers between the stages: the level's state, d, the predictor
// corrections and u-hat live in LDS; each thread owns one interior node and keeps
// its row of every p' level in registers (as k_mg_coarse does) and its p' solve
// vectors; the BiCGStab p' solve is preconditioned by the same single-workgroup
// V-cycle (cw_vcycle).  FP64 in both precision modes (a coarse level steers the
// fine level's iterates and sets no result of its own).  The per-node arithmetic
// is the stage kernels' (residual_node, jac_node, face_div_node, t_side_value,
// wall_p_value, p9_face_of).
constexpr int SL_NT = 1024;
constexpr int SL_NODES = 1089;  // at most 33 x 33 nodes (and at most SL_NT interior ones)
enum { SL_U, SL_V, SL_T, SL_P, SL_D, SL_A0, SL_A1, SL_A2, SL_B0, SL_B1, SL_B2, SL_NARR };

struct SmallArgs {
    CGrid g;
    GridPtrs gp;
    const double *jac;
    double *u, *v, *T, *p;
    Src src;
    int iters, sweeps, p_maxit;
    double p_rtol;
    CoarseArgs<double> mg;  // the p' levels' shapes, face tables, V-cycle sweeps and LDS offsets
    Geom lev[CW_LEVS];      // the p' levels' metrics (level 0 is this level)
};

// totals of NV values over the workgroup of NT threads, in every thread
template <int NT, int NV>
__device__ __forceinline__ void sl_sums(double (&v)[NV]) {
    __shared__ double red[NV][NT / 64];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) red[k][threadIdx.x >> 6] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double t = 0.0;
        for (int q = 0; q < NT / 64; q++) t += red[k][q];
        v[k] = t;
    }
    __syncthreads();
}

// border node b (0 <= b < 2 (n0 + n1) - 4) of an n0 x n1 grid, as a node index
__device__ __forceinline__ int sl_border(int b, int n0, int n1) {
    if (b < n0) return b;                          // j = 0
    b -= n0;
    if (b < n0) return (n1 - 1) * n0 + b;          // j = n1 - 1
    b -= n0;
    if (b < n1 - 2) return (b + 1) * n0;           // i = 0
    return (b - (n1 - 2) + 1) * n0 + n0 - 1;       // i = n0 - 1
}

// LDS of k_simple_small: the level's arrays, and the V-cycle's iterates
__shared__ double sl_sa[SL_NARR * SL_NODES];
__shared__ double sl_mg[CN_MAX * CW_ARR];
__device__ __forceinline__ double *sl_arr(int k) { return sl_sa + k * SL_NODES; }
// The 256-thread kernel (at most SL_FF_MAX nodes on all its p' levels together)
// also keeps the metric factors of every node's east and north faces of every p'
// level (p9_face_ab: four arrays, the levels one after the other), which the
// level's iterations do not change, so each iteration's face tables read them
// and d from LDS instead of five metric arrays at five nodes from global memory.
constexpr int SL_FF_MAX = 768;
__shared__ double sl_ff[4 * SL_FF_MAX];

// a thread's node: interior node (i, j) = index P when own
struct SlNode {
    int i, j, P;
    bool own;
};

// Stages 1-4 of an iteration: residuals and d at the own node (d into SL_D), the
// predictor's Jacobi sweeps (ping-pong between the A and B arrays, wall entries
// zero), u* = u + du into SL_U / SL_V (no stage before the correction reads the
// old u at a neighbour), T += dT, and u-hat, v-hat into SL_A0 / SL_A1.  Ends with
// a block barrier.
template <int NT>
__device__ __forceinline__ void sl_predict(const SmallArgs *__restrict__ a, SlNode t) {
    const CGrid &g = a->g;
    const SlotP ph = g.sp[blockIdx.z];
    const int n0 = g.n0, n1 = g.n1, P = t.P, tid = threadIdx.x;
    const size_t o = zslot() * g.N;
    double *const U_ = sl_arr(SL_U), *const V_ = sl_arr(SL_V), *const T_ = sl_arr(SL_T), *const P_ = sl_arr(SL_P);
    double *const D_ = sl_arr(SL_D);
    const Src src = src_at(a->src, o, ph);
    const GridPtrs &gq = a->gp;
    const size_t G = o + P;
    const NC c{gq.xix[G], gq.xiy[G], gq.etx[G], gq.ety[G], gq.alph[G], gq.beta[G], gq.gamma[G], gq.p1[G], gq.q1[G]};
    ResOut ro{};
    if (t.own) {
        auto at = [&](const double *f, int di, int dj) { return f[P + dj * n0 + di]; };
        auto pat = [&](int di, int dj) { return P_[P + dj * n0 + di]; };
        ro = residual_node(g, ph, c, t.i, t.j, P, U_, V_, T_, at, pat, src);
        D_[P] = ro.d;
    }
    double *da[3] = {sl_arr(SL_A0), sl_arr(SL_A1), sl_arr(SL_A2)};
    double *db[3] = {sl_arr(SL_B0), sl_arr(SL_B1), sl_arr(SL_B2)};
    for (int b = tid; b < 2 * (n0 + n1) - 4; b += NT) {
        const int k = sl_border(b, n0, n1);
#pragma unroll
        for (int f = 0; f < 3; f++) da[f][k] = db[f][k] = 0.0;
    }
    const int sweeps = a->sweeps;
    if (sweeps <= 0 && t.own) da[0][P] = da[1][P] = da[2][P] = 0.0;
    const JNode<double> jn{c.al, c.be, c.ga, c.pp, c.qq, ro.U, ro.V, ro.ru, ro.rv, ro.rt};
    const JRow<double> jw = jac_row<double>(jn, ph, g.r);  // the rows, once for all sweeps
    for (int sw = 0; sw < sweeps; sw++) {
        if (t.own) {
            auto dat = [&](int f, int ii, int jj) { return da[f][jj * n0 + ii]; };
            double ou, ov, ot;
            if (sw == 0) jac_node_w<double, true>(jn, jw, t.i, t.j, dat, ou, ov, ot);
            else jac_node_w<double, false>(jn, jw, t.i, t.j, dat, ou, ov, ot);
            db[0][P] = ou;
            db[1][P] = ov;
            db[2][P] = ot;
        }
        __syncthreads();
#pragma unroll
        for (int f = 0; f < 3; f++) {
            double *s = da[f];
            da[f] = db[f];
            db[f] = s;
        }
    }
    if (t.own) {
        const double us = U_[P] + da[0][P], vs = V_[P] + da[1][P];
        T_[P] = T_[P] + da[2][P];
        U_[P] = us;
        V_[P] = vs;
        const double dp_dxi = (P_[P + 1] - P_[P - 1]) * g.r.i2h1;
        const double dp_de = (P_[P + n0] - P_[P - n0]) * g.r.i2h2;
        sl_arr(SL_A0)[P] = us + ro.d * (c.xix * dp_dxi + c.etx * dp_de);
        sl_arr(SL_A1)[P] = vs + ro.d * (c.xiy * dp_dxi + c.ety * dp_de);
    }
    __syncthreads();
}

// Stage 5: the face divergence at the own node (with the level's mass source) into
// SL_B0
__device__ __forceinline__ void sl_face_div(const SmallArgs *__restrict__ a, SlNode t) {
    if (!t.own) return;
    const CGrid &g = a->g;
    const int n0 = g.n0, P = t.P;
    const size_t o = zslot() * g.N;
    const GridPtrs &gq = a->gp;
    const double *const uh = sl_arr(SL_A0), *const vh = sl_arr(SL_A1), *const D_ = sl_arr(SL_D);
    const double *const P_ = sl_arr(SL_P);
    auto L = [&](int k, int di, int dj) -> double {
        const int Q = P + dj * n0 + di;
        switch (k) {
        case FD_UH: return uh[Q];
        case FD_VH: return vh[Q];
        case FD_D: return D_[Q];
        case FD_X: return gq.xix[o + Q];
        case FD_Y: return gq.xiy[o + Q];
        case FD_E: return gq.etx[o + Q];
        case FD_F: return gq.ety[o + Q];
        default: return a->jac[o + Q];
        }
    };
    auto pq = [&](int di, int dj) { return P_[P + dj * n0 + di]; };
    double qv = face_div_node(g, t.i, t.j, L, pq);
    if (a->src.m && g.sp[blockIdx.z].fas_src != 0.0) qv -= a->src.m[o + P];
    sl_arr(SL_B0)[P] = qv;
}

// node k = (li, lj) of a p' level: the faces touching a wall node zeroed, the four
// face tables written, and at an interior node 1 / diag of the clamped operator
__device__ __forceinline__ void sl_face_store(const CLev<double> &cl, size_t ol, int k, int li, int lj, int ln0,
                                              int ln1, double en, double et, double nn, double nt, double wn,
                                              double wt, double sn, double st) {
    auto ev = [&](int ii, int jj) { return ii >= 1 && ii <= ln0 - 3 && jj >= 1 && jj <= ln1 - 2; };
    auto nv = [&](int ii, int jj) { return jj >= 1 && jj <= ln1 - 3 && ii >= 1 && ii <= ln0 - 2; };
    if (!ev(li, lj)) en = et = 0.0;
    if (!nv(li, lj)) nn = nt = 0.0;
    if (!ev(li - 1, lj)) wn = wt = 0.0;
    if (!nv(li, lj - 1)) sn = st = 0.0;
    const_cast<double *>(cl.fen)[ol + k] = en;
    const_cast<double *>(cl.fet)[ol + k] = et;
    const_cast<double *>(cl.fnn)[ol + k] = nn;
    const_cast<double *>(cl.fnt)[ol + k] = nt;
    if (!inner_node(li, lj, ln0, ln1)) return;
    double cr[9];
    p9_row_from<double>(RowFaces<double>{en, et, wn, wt, nn, nt, sn, st}, cr);
    double dg = 0.0;
    for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++)
            if (clampi(li + di, 1, ln0 - 2) == li && clampi(lj + dj, 1, ln1 - 2) == lj) dg += cr[(dj + 1) * 3 + (di + 1)];
    const_cast<double *>(cl.invd)[ol + k] = 1.0 / dg;
}

// the metric factors of the p' levels' faces into sl_ff (the 256-thread kernel, once
// per launch; p9_face_ab of (P, E) and (P, N), neighbours clamped into the grid as
// in the face tables)
template <int NT>
__device__ __forceinline__ void sl_face_factors(const SmallArgs *__restrict__ a) {
    const int last = a->mg.nlev - 1;
    int fo = 0;  // the level's first entry
#pragma unroll
    for (int L = 0; L < CW_LEVS; L++) {
        if (L > last) break;
        const Geom &gl = a->lev[L];
        const int ln0 = gl.n0, ln1 = gl.n1, nn = ln0 * ln1;
        const size_t ol = zslot() * (size_t)nn;
        for (int k = threadIdx.x; k < nn; k += NT) {
            const int li = k % ln0, lj = k / ln0;
            auto fn = [&](int ii, int jj) {
                const int ic = clampi(ii, 0, ln0 - 1), jc = clampi(jj, 0, ln1 - 1);
                const size_t Q = ol + (size_t)jc * ln0 + ic;
                return FNode{gl.xix[Q], gl.xiy[Q], gl.etx[Q], gl.ety[Q], gl.jac[Q], 0.0};
            };
            const FNode nP = fn(li, lj), nE = fn(li + 1, lj), nN = fn(li, lj + 1);
            p9_face_ab(nP, nE, 0, sl_ff[fo + k], sl_ff[SL_FF_MAX + fo + k]);
            p9_face_ab(nP, nN, 1, sl_ff[2 * SL_FF_MAX + fo + k], sl_ff[3 * SL_FF_MAX + fo + k]);
        }
        fo += nn;
    }
}

// Stage 6a: the face coefficients and inverse diagonals of every p' level from the
// level's metrics and d injected from this level (k_faces_all's arithmetic), into
// the level tables (global memory).  Ends with a block barrier.  The 256-thread
// kernel takes each face's metric factors from sl_ff (sl_face_tables_ff).
template <int NT>
__device__ __forceinline__ void sl_face_tables(const SmallArgs *__restrict__ a) {
    const double *const D_ = sl_arr(SL_D);
    const int n0 = a->g.n0, last = a->mg.nlev - 1;
#pragma unroll
    for (int L = 0; L < CW_LEVS; L++) {
        if (L > last) break;
        const Geom &gl = a->lev[L];
        const CLev<double> &cl = a->mg.L[L];
        const int ln0 = gl.n0, ln1 = gl.n1;
        const size_t ol = zslot() * (size_t)ln0 * ln1;
        const double r1 = 1.0 / (gl.h1 * gl.h1), r2 = 1.0 / (gl.h2 * gl.h2), r12 = 1.0 / (4.0 * gl.h1 * gl.h2);
        for (int k = threadIdx.x; k < ln0 * ln1; k += NT) {
            const int li = k % ln0, lj = k / ln0;
            auto fn = [&](int ii, int jj) {
                const int ic = clampi(ii, 0, ln0 - 1), jc = clampi(jj, 0, ln1 - 1);
                const size_t Q = ol + (size_t)jc * ln0 + ic;
                return FNode{gl.xix[Q], gl.xiy[Q], gl.etx[Q], gl.ety[Q], gl.jac[Q], D_[(jc << L) * n0 + (ic << L)]};
            };
            const FNode nP = fn(li, lj), nE = fn(li + 1, lj), nW = fn(li - 1, lj), nN = fn(li, lj + 1),
                        nS = fn(li, lj - 1);
            double en, et, nn, nt, wn, wt, sn, st;
            p9_face_of(nP, nE, 0, r1, r2, r12, en, et);
            p9_face_of(nP, nN, 1, r1, r2, r12, nn, nt);
            p9_face_of(nW, nP, 0, r1, r2, r12, wn, wt);
            p9_face_of(nS, nP, 1, r1, r2, r12, sn, st);
            sl_face_store(cl, ol, k, li, lj, ln0, ln1, en, et, nn, nt, wn, wt, sn, st);
        }
    }
    __syncthreads();
}
template <int NT>
__device__ __forceinline__ void sl_face_tables_ff(const SmallArgs *__restrict__ a) {
    const double *const D_ = sl_arr(SL_D);
    const int n0 = a->g.n0, last = a->mg.nlev - 1;
    int fo = 0;  // the level's first entry in sl_ff
#pragma unroll
    for (int L = 0; L < CW_LEVS; L++) {
        if (L > last) break;
        const Geom &gl = a->lev[L];
        const CLev<double> &cl = a->mg.L[L];
        const int ln0 = gl.n0, ln1 = gl.n1;
        const size_t ol = zslot() * (size_t)ln0 * ln1;
        const double r1 = 1.0 / (gl.h1 * gl.h1), r2 = 1.0 / (gl.h2 * gl.h2), r12 = 1.0 / (4.0 * gl.h1 * gl.h2);
        const double *const f = sl_ff + fo;
        for (int k = threadIdx.x; k < ln0 * ln1; k += NT) {
            const int li = k % ln0, lj = k / ln0;
            // d at the clamped nodes; a clamped west or south neighbour's face is not
            // this node's, but every such face touches a wall node and is zeroed
            auto dat = [&](int ii, int jj) {
                const int ic = clampi(ii, 0, ln0 - 1), jc = clampi(jj, 0, ln1 - 1);
                return D_[(jc << L) * n0 + (ic << L)];
            };
            const int kw = lj * ln0 + clampi(li - 1, 0, ln0 - 1), ks = clampi(lj - 1, 0, ln1 - 1) * ln0 + li;
            const double dP = dat(li, lj);
            double en, et, nn, nt, wn, wt, sn, st;
            p9_face_d(f[k], f[SL_FF_MAX + k], dP, dat(li + 1, lj), 0, r1, r2, r12, en, et);
            p9_face_d(f[2 * SL_FF_MAX + k], f[3 * SL_FF_MAX + k], dP, dat(li, lj + 1), 1, r1, r2, r12, nn, nt);
            p9_face_d(f[kw], f[SL_FF_MAX + kw], dat(li - 1, lj), dP, 0, r1, r2, r12, wn, wt);
            p9_face_d(f[2 * SL_FF_MAX + ks], f[3 * SL_FF_MAX + ks], dat(li, lj - 1), dP, 1, r1, r2, r12, sn, st);
            sl_face_store(cl, ol, k, li, lj, ln0, ln1, en, et, nn, nt, wn, wt, sn, st);
        }
        fo += ln0 * ln1;
    }
    __syncthreads();
}

// Stage 6b: L_d p' = q by BiCGStab, right-preconditioned by the single-workgroup
// V-cycle, from p' = 0 (oracle: bicgstab), with the rows of the own node of every
// level in registers (as k_mg_coarse holds them) and the vectors at the own node in
// the free work arrays; p' with its interior mean removed goes to SL_A1.  Ends with
// a block barrier.
template <int NT>
__device__ __forceinline__ void sl_pprime(const SmallArgs *__restrict__ a, SlNode t) {
    const int n0 = a->g.n0, P = t.P;
    const bool own = t.own;
    const double inv_m = 1.0 / ((double)(n0 - 2) * (double)(a->g.n1 - 2));
    CNode<double> nd[CW_LEVS];
    {
        RowFaces<double> rf[CW_LEVS];
        cw_row_loads<double>(a->mg, 0, nd, rf);
        cw_row_fold<double>(a->mg, 0, nd, rf);
    }
    double *const R = sl_arr(SL_B0), *const R0 = sl_arr(SL_B1), *const PV = sl_arr(SL_B2);
    double *const VV = sl_arr(SL_A2), *const X = sl_arr(SL_A0);
    const double bq = own ? R[P] : 0.0;  // the face divergence
    double ms[1] = {bq};
    sl_sums<NT, 1>(ms);
    const double r = own ? bq - ms[0] * inv_m : 0.0;
    if (own) {
        R[P] = r;
        R0[P] = r;
        X[P] = 0.0;
    }
    double nb[2] = {bq * bq, r * r};
    sl_sums<NT, 2>(nb);
    const double bnorm = sqrt(nb[0]), tol = a->p_rtol * bnorm;
    const double *z = sl_mg + a->mg.off[0];  // the V-cycle's result on this level
    const double pm = a->g.sp[blockIdx.z].pmaxit;
    const int maxit = pm != 0.0 ? (int)pm : a->p_maxit;  // < 0: one V-cycle
    if (maxit < 0) {
        if (bnorm != 0.0) {
            nd[0].b = r;
            cw_vcycle<double, NT>(a->mg, 0, nd, sl_mg);
            if (own) X[P] = z[P];
        }
    } else if (bnorm != 0.0 && sqrt(nb[1]) > tol) {
        double rho_old = 1.0, alpha = 1.0, omega = 1.0;
        for (int it = 1; it <= maxit; it++) {
            double t1[1] = {own ? R0[P] * R[P] : 0.0};
            sl_sums<NT, 1>(t1);
            const double rho = t1[0];
            if (rho == 0.0) break;
            double pv = 0.0;
            if (own) {
                pv = it == 1 ? R[P] : R[P] + ((rho / rho_old) * (alpha / omega)) * (PV[P] - omega * VV[P]);
                PV[P] = pv;
            }
            nd[0].b = pv;
            cw_vcycle<double, NT>(a->mg, 0, nd, sl_mg);
            const double phv = own ? z[P] : 0.0;
            const double vv = own ? cw_row<double>(nd[0], z, n0) : 0.0;
            if (own) VV[P] = vv;
            double t2[1] = {own ? R0[P] * vv : 0.0};
            sl_sums<NT, 1>(t2);
            alpha = rho / t2[0];
            const double sv = own ? R[P] - alpha * vv : 0.0;
            double t3[1] = {sv * sv};
            sl_sums<NT, 1>(t3);
            if (sqrt(t3[0]) <= tol) {
                if (own) X[P] = X[P] + alpha * phv;
                break;
            }
            nd[0].b = sv;
            cw_vcycle<double, NT>(a->mg, 0, nd, sl_mg);
            const double shv = own ? z[P] : 0.0;
            const double tv = own ? cw_row<double>(nd[0], z, n0) : 0.0;
            double t4[2] = {tv * tv, tv * sv};
            sl_sums<NT, 2>(t4);
            omega = t4[0] > 0.0 ? t4[1] / t4[0] : 0.0;
            const double rn = sv - omega * tv;
            if (own) {
                X[P] = X[P] + (alpha * phv + omega * shv);
                R[P] = rn;
            }
            double t5[1] = {rn * rn};
            sl_sums<NT, 1>(t5);
            if (sqrt(t5[0]) <= tol || omega == 0.0) break;
            rho_old = rho;
        }
    }
    const double xv = own ? X[P] : 0.0;
    double xs[1] = {xv};
    sl_sums<NT, 1>(xs);
    if (own) sl_arr(SL_A1)[P] = xv - xs[0] * inv_m;
    __syncthreads();
}

// Stages 7-9: correction u = u* - d grad p' (p' read clamped into the interior:
// the wall copies), p += alpha_p p'; the adiabatic side walls (new values from the
// old ones); the wall pressure (reads p at interior nodes only), the corners from
// the new wall values (kept without sources in SL_B0), then the sources.  Ends
// with a block barrier.
template <int NT>
__device__ __forceinline__ void sl_correct_walls(const SmallArgs *__restrict__ a, SlNode t) {
    const CGrid &g = a->g;
    const SlotP *const ph = g.sp + blockIdx.z;
    const int n0 = g.n0, n1 = g.n1, m0 = n0 - 2, m1 = n1 - 2, P = t.P, tid = threadIdx.x;
    const size_t o = zslot() * g.N;
    double *const U_ = sl_arr(SL_U), *const V_ = sl_arr(SL_V), *const T_ = sl_arr(SL_T), *const P_ = sl_arr(SL_P);
    const double *const D_ = sl_arr(SL_D), *const PC = sl_arr(SL_A1);
    const GridPtrs gp{a->gp.xix + o, a->gp.xiy + o, a->gp.etx + o, a->gp.ety + o, a->gp.alph + o,
                      a->gp.beta + o, a->gp.gamma + o, a->gp.p1 + o, a->gp.q1 + o};
    const Src src = src_at(a->src, o, *ph);
    if (t.own) {
        const int i = t.i, j = t.j;
        const double pe = PC[clampi(i + 1, 1, n0 - 2) + j * n0], pw = PC[clampi(i - 1, 1, n0 - 2) + j * n0];
        const double pn = PC[i + clampi(j + 1, 1, n1 - 2) * n0], ps = PC[i + clampi(j - 1, 1, n1 - 2) * n0];
        const double dpx = 0.5 * (pe - pw) * g.r.ih1;
        const double dpe = 0.5 * (pn - ps) * g.r.ih2;
        const double dP = D_[P];
        U_[P] = U_[P] - dP * (gp.xix[P] * dpx + gp.etx[P] * dpe);
        V_[P] = V_[P] - dP * (gp.xiy[P] * dpx + gp.ety[P] * dpe);
        P_[P] = P_[P] + ph->alpha_p * PC[P];
    }
    __syncthreads();
    const KGrid kg = g.kg();
    double tw = 0.0;
    int Pt = -1;
    if (tid < 2 * m1) {
        const bool east = tid >= m1;
        const int jw = 1 + (east ? tid - m1 : tid);
        Pt = jw * n0 + (east ? n0 - 1 : 0);
        auto Tat = [&](int ii, int jj) { return T_[jj * n0 + ii]; };
        tw = t_side_value(kg, Tat, gp.alph[Pt], gp.beta[Pt], east, jw);
        if (src.T) tw += src.T[Pt];
    }
    __syncthreads();
    if (Pt >= 0) T_[Pt] = tw;
    __syncthreads();
    double *const PW = sl_arr(SL_B0);
    if (tid < 2 * (m0 + m1)) {
        int side, iw, jw;
        if (tid < 2 * m0) {
            side = tid < m0 ? 0 : 1;
            iw = 1 + (tid < m0 ? tid : tid - m0);
            jw = side == 0 ? 0 : n1 - 1;
        } else {
            const int q = tid - 2 * m0;
            side = q < m1 ? 2 : 3;
            jw = 1 + (q < m1 ? q : q - m1);
            iw = side == 2 ? 0 : n0 - 1;
        }
        const WallIn w{U_, V_, T_, P_, gp.xix, gp.xiy, gp.etx, gp.ety, gp.alph, gp.beta, gp.gamma,
                       gp.p1, gp.q1, a->jac + o};
        const int Pw = jw * n0 + iw;
        const double pw = wall_p_value(w, kg, ph->Re, ph->Ri, side, iw, jw, T_[Pw]);
        PW[Pw] = pw;
        P_[Pw] = pw + (src.m ? src.m[Pw] : 0.0);
    }
    __syncthreads();
    if (tid < 4) {
        const bool east = tid & 1, north = tid & 2;
        const int ci = east ? n0 - 1 : 0, cj = north ? n1 - 1 : 0;
        const int C = cj * n0 + ci;
        const double v1 = PW[cj * n0 + (east ? n0 - 2 : 1)], v2 = PW[(north ? n1 - 2 : 1) * n0 + ci];
        P_[C] = 0.5 * (v1 + v2) + (src.m ? src.m[C] : 0.0);
    }
    __syncthreads();
}

// a: the level's arguments in device memory (scalar loads, no per-lane copy);
// iters SIMPLE iterations, iters_deep on a slot that is its case's coarsest level
// NT threads per workgroup (at least the level's interior nodes, and 2 (n0 + n1))
template <int NT>
__global__ __launch_bounds__(NT) void k_simple_small(const SmallArgs *__restrict__ a, int iters, int iters_deep) {
    const SlotP &sp = a->g.sp[blockIdx.z];
    if (sp.active == 0.0) return;
    if (sp.deep != 0.0) iters = iters_deep;
    if (iters <= 0) return;
    const int n0 = a->g.n0, n1 = a->g.n1, N = n0 * n1;
    const size_t o = zslot() * a->g.N;
    const int tid = threadIdx.x;
    const int m0 = n0 - 2, m1 = n1 - 2;
    SlNode t;
    t.own = tid < m0 * m1;
    const int qn = t.own ? tid : 0;
    t.i = 1 + qn % m0;
    t.j = 1 + qn / m0;
    t.P = t.j * n0 + t.i;
    for (int k = tid; k < N; k += NT) {
        sl_arr(SL_U)[k] = a->u[o + k];
        sl_arr(SL_V)[k] = a->v[o + k];
        sl_arr(SL_T)[k] = a->T[o + k];
        sl_arr(SL_P)[k] = a->p[o + k];
    }
    for (int k = tid; k < CN_MAX * CW_ARR; k += NT) sl_mg[k] = 0.0;  // wall entries stay zero
    if constexpr (NT == 256) sl_face_factors<NT>(a);
    __syncthreads();
    // The stages run as the cases of one loop: a stage's values cannot stay live into
    // the next one, so the register demand is the largest stage's, not their sum.
    for (int k = 0; k < 5 * iters; k++) {
        switch (k % 5) {
        case 0: sl_predict<NT>(a, t); break;
        case 1: sl_face_div(a, t); break;  // the face tables' barrier orders it before the solve
        case 2:
            if constexpr (NT == 256) sl_face_tables_ff<NT>(a);
            else sl_face_tables<NT>(a);
            break;
        case 3: sl_pprime<NT>(a, t); break;
        default: sl_correct_walls<NT>(a, t); break;
        }
    }
    for (int k = tid; k < N; k += NT) {
        a->u[o + k] = sl_arr(SL_U)[k];
        a->v[o + k] = sl_arr(SL_V)[k];
        a->T[o + k] = sl_arr(SL_T)[k];
        a->p[o + k] = sl_arr(SL_P)[k];
    }
}

// Partial-sum slots (pstride each) per case slot: 0-2 for reductions finished
// right away, 3-4 for the energy/momentum residual norms and 5 for the mass
// residual, which stay there until the end-of-iteration finish below.
enum { PS_RES = 3, PS_MASS = 5, PS_SLOTS = 8 };  // rows 6-7: the p' solver's operator-application phases
static_assert(PS_SLOTS >= PS_APPLY_ROW + 2 && PS_MASS < PS_APPLY_ROW, "partial rows");

// Wall-Nusselt integrand of segment [i, i + 1] of the bottom (s = 1, j = 0) or top
// (s = -1, j = n1 - 1) wall from its twelve loaded values (oracle: nusselt).
struct NuSeg {
    double ja, xa, ya, jb, xb, yb, t0a, t1a, t2a, t0b, t1b, t2b;
};
__device__ __forceinline__ NuSeg nu_load(const KGrid &g, const double *__restrict__ T,
                                         const double *__restrict__ ajac, const double *__restrict__ etx,
                                         const double *__restrict__ ety, int i, int j, int s) {
    const int n0 = g.n0;
    const size_t A = idx(i, j, n0), B = idx(i + 1, j, n0);
    return NuSeg{ajac[A], etx[A], ety[A], ajac[B], etx[B], ety[B], T[A], T[idx(i, j + s, n0)],
                 T[idx(i, j + 2 * s, n0)], T[B], T[idx(i + 1, j + s, n0)], T[idx(i + 1, j + 2 * s, n0)]};
}
__device__ __forceinline__ double nu_term(const KGrid &g, const NuSeg &q) {
    const double tj1 = q.ja * (q.xa * q.xa + q.ya * q.ya);
    const double tj2 = q.jb * (q.xb * q.xb + q.yb * q.yb);
    const double fh1 = tj1 * (4 * q.t1a - 3 * q.t0a - q.t2a) / (2 * g.h2);
    const double fh2 = tj2 * (4 * q.t1b - 3 * q.t0b - q.t2b) / (2 * g.h2);
    return 0.5 * (fh1 + fh2) * g.h1;
}

// one block: dmax from the partials, floored at 0 (solver.cpp:1806), both wall
// Nusselt numbers, and the energy, momentum (nres partials) and mass (nparts
// partials) residual norms of the iteration.  Every thread first issues the loads
// of its share of all five quantities (one memory round trip for grids up to
// 5 * FIN_NT partials and FIN_NT wall segments), then one block reduction.
__global__ __launch_bounds__(FIN_NT) void k_dmax_nu_fin(CGrid cg, RedArgs ra, int nparts, int nres,
                                                        const double *__restrict__ T,
                                                        const double *__restrict__ ajac,
                                                        const double *__restrict__ etx,
                                                        const double *__restrict__ ety) {
    const SlotP &sp = cg.sp[blockIdx.z];
    if (sp.active == 0.0) return;
    ra = rz(ra);
    const double *part = ra.part;
    double *scal = ra.scal;
    const double len_hot = sp.len_hot, len_cold = sp.len_cold, dT = sp.t_hot - sp.t_cold;
    const KGrid g = cg.kg();
    {
        const size_t o = zslot() * cg.N;
        T += o; ajac += o; etx += o; ety += o;
    }
    constexpr int U = 5;
    const int t = tlin();
    const size_t ps = ra.pstride;
    const double *pr0 = part + (size_t)PS_RES * ps, *pr1 = pr0 + ps, *pm = part + (size_t)PS_MASS * ps;
    const int lr = max(nres - 1, 0), lp = max(nparts - 1, 0), ns = g.n0 - 1;
    // first pass: every load before any use
    double r0[U], r1[U], mm[U], dx[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        r0[u] = pr0[min(u * FIN_NT + t, lr)];
        r1[u] = pr1[min(u * FIN_NT + t, lr)];
        mm[u] = pm[min(u * FIN_NT + t, lp)];
        dx[u] = part[min(u * FIN_NT + t, lp)];
    }
    const int i0 = min(t, max(ns - 1, 0));
    const NuSeg qh = nu_load(g, T, ajac, etx, ety, i0, 0, 1);
    const NuSeg qc = nu_load(g, T, ajac, etx, ety, i0, g.n1 - 1, -1);
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // energy, momentum, mass, Nu hot, Nu cold
    double m = -1.0e300;
#pragma unroll
    for (int u = 0; u < U; u++) {
        if (u * FIN_NT + t < nres) {
            v[0] += r0[u];
            v[1] += r1[u];
        }
        if (u * FIN_NT + t < nparts) {
            v[2] += mm[u];
            m = fmax(m, dx[u]);
        }
    }
    if (t < ns) {
        v[3] += nu_term(g, qh);
        v[4] += nu_term(g, qc);
    }
    // larger grids: the rest, pass by pass
    for (int k = U * FIN_NT + t; k < nres; k += FIN_NT) {
        v[0] += pr0[k];
        v[1] += pr1[k];
    }
    for (int k = U * FIN_NT + t; k < nparts; k += FIN_NT) {
        v[2] += pm[k];
        m = fmax(m, part[k]);
    }
    for (int i = FIN_NT + t; i < ns; i += FIN_NT) {
        v[3] += nu_term(g, nu_load(g, T, ajac, etx, ety, i, 0, 1));
        v[4] += nu_term(g, nu_load(g, T, ajac, etx, ety, i, g.n1 - 1, -1));
    }
    // one block reduction of the five sums and the maximum
    __shared__ double s_v[6][FIN_NT / 64];
#pragma unroll
    for (int q = 0; q < 5; q++) v[q] = wave_sum(v[q]);
    m = wave_max(m);
    if ((t & 63) == 0) {
#pragma unroll
        for (int q = 0; q < 5; q++) s_v[q][t >> 6] = v[q];
        s_v[5][t >> 6] = m;
    }
    __syncthreads();
    if (t >= 64) return;
#pragma unroll
    for (int q = 0; q < 5; q++) v[q] = wave_sum(t < FIN_NT / 64 ? s_v[q][t] : 0.0);
    m = wave_max(t < FIN_NT / 64 ? s_v[5][t] : -1.0e300);
    const double res[2] = {v[0], v[1]}, mass[1] = {v[2]};
    const double nh = v[3], nc = v[4];
    if (t == 0) {
        const double den = dT != 0.0 ? dT : 1.0;
        scal[S_ENERGY] = res[0];
        scal[S_MOM] = res[1];
        scal[S_MASS] = mass[0];
        scal[S_DMAX] = fmax(m, 0.0);
        scal[S_NUH] = -nh / (len_hot * den);
        scal[S_NUC] = nc / (len_cold * den);
        if (scal[S_FRESH] != 0.0) {  // the case's first iteration: the normalisers of the residual ratios
            scal[S_MASS0] = mass[0];
            scal[S_ENERGY0] = res[0];
            scal[S_FRESH] = 0.0;
        }
    }
}

// ------------------------------------------------------------------------
// FAS multigrid outer cycle (oracle: fas_cycle).  Levels are batches of the same
// slots on the same stream; level l+1 takes every other node of level l.

// The transfers run blockIdx.z over the fine level's slots (sp: the fine level's
// slot table, in which a slot without a coarse level is inactive); fine slot z's
// coarse level is slot cmap[z] of the coarse level (z itself when cmap is null:
// a hierarchy of equal slots).
__device__ __forceinline__ size_t cslot(const int *cmap) {
    return cmap ? (size_t)cmap[blockIdx.z] : zslot();
}

// coarse node (I, J) <- fine node (2I, 2J) for u, v, T, p, kept in x0 as well
struct FasInj {
    int n0, m0, m1;
    size_t Nf, Nc;
    const SlotP *sp;
    const int *cmap;
    const double *f[4];
    double *c[4], *c0[4];
};
__global__ void k_fas_inject(FasInj a) {
    if (a.sp[blockIdx.z].active == 0.0) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.m0 * a.m1) return;
    const int I = k % a.m0, J = k / a.m0;
    const size_t F = zslot() * a.Nf + idx(2 * I, 2 * J, a.n0), C = cslot(a.cmap) * a.Nc + (size_t)k;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const double v = a.f[q][F];
        a.c[q][C] = v;
        a.c0[q][C] = v;
    }
}

// Sources of the coarse level.  Interior: s_u = FW(R_f / dt) - R_c / dt (v, T
// likewise; R_f with the fine level's own sources, R_c of the restricted state
// without), s_m = q_c - FW(q_f); side walls: s_T = the fine side-wall update's
// change injected minus the coarse one's.  Every other entry is zeroed (the wall
// pressure sources follow in k_fas_src_p).
struct FasSrc {
    int n0, n1, m0, m1;
    size_t Nf, Nc, scrf, scrc;
    const SlotP *sp;
    const int *cmap;
    const double *fr[3], *fq, *fT, *fscr;
    const double *cr[3], *cq, *cT, *cscr;
    double *s[4];
};
__global__ void k_fas_src(FasSrc a) {
    const SlotP &ph = a.sp[blockIdx.z];
    if (ph.active == 0.0) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int m0 = a.m0, m1 = a.m1, n0 = a.n0;
    if (k >= m0 * m1) return;
    const int I = k % m0, J = k / m0;
    const size_t zf = zslot() * a.Nf, zc = cslot(a.cmap) * a.Nc, C = zc + (size_t)k;
    const double idt = 1.0 / ph.dt;
    double sv[4] = {0.0, 0.0, 0.0, 0.0};
    if (I >= 1 && I <= m0 - 2 && J >= 1 && J <= m1 - 2) {
        const int i = 2 * I, j = 2 * J;
        auto fw = [&](const double *r, double sc) {
            auto R = [&](int di, int dj) { return r[zf + idx(i + di, j + dj, n0)] * sc; };
            return (4.0 * R(0, 0) + 2.0 * (R(-1, 0) + R(1, 0) + R(0, -1) + R(0, 1)) +
                    (R(-1, -1) + R(1, -1) + R(-1, 1) + R(1, 1))) / 16.0;
        };
#pragma unroll
        for (int q = 0; q < 3; q++) sv[q] = fw(a.fr[q], idt) - a.cr[q][C] * idt;
        sv[3] = a.cq[C] - fw(a.fq, 1.0);
    } else if ((I == 0 || I == m0 - 1) && J >= 1 && J <= m1 - 2) {
        const int i = I == 0 ? 0 : a.n0 - 1, j = 2 * J;
        const double *fs = a.fscr + zslot() * a.scrf, *cs = a.cscr + cslot(a.cmap) * a.scrc;
        const double wf = (I == 0 ? fs[j] : fs[a.n1 + j]) - a.fT[zf + idx(i, j, n0)];
        const double wc = (I == 0 ? cs[J] : cs[m1 + J]) - a.cT[C];
        sv[2] = wf - wc;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) a.s[q][C] = sv[q];
}

// wall-pressure sources of the coarse level's border: the fine wall update's
// change injected minus the coarse one's (after the coarse side-wall update with
// its source)
struct FasSrcP {
    int n0, m0, m1;
    size_t Nf, Nc;
    const SlotP *sp;
    const int *cmap;
    const double *fpw, *fp, *cpw, *cp;
    double *sm;
};
__global__ void k_fas_src_p(FasSrcP a) {
    if (a.sp[blockIdx.z].active == 0.0) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int m0 = a.m0, m1 = a.m1;
    const int nb = 2 * m0 + 2 * (m1 - 2);  // border nodes
    if (k >= nb) return;
    int I, J;
    if (k < m0) { I = k; J = 0; }
    else if (k < 2 * m0) { I = k - m0; J = m1 - 1; }
    else if (k < 2 * m0 + m1 - 2) { I = 0; J = 1 + k - 2 * m0; }
    else { I = m0 - 1; J = 1 + k - 2 * m0 - (m1 - 2); }
    const size_t F = zslot() * a.Nf + idx(2 * I, 2 * J, a.n0), C = cslot(a.cmap) * a.Nc + idx(I, J, m0);
    a.sm[C] = (a.fpw[F] - a.fp[F]) - (a.cpw[C] - a.cp[C]);
}

// fine interior += bilinear interpolation of the coarse change x - x0 (u, v, T, p)
struct FasPro {
    int n0, n1, m0;
    size_t Nf, Nc;
    const SlotP *sp;
    const int *cmap;
    const double *c[4], *c0[4];
    double *f[4];
};
__global__ void k_fas_prolong(FasPro a) {
    if (a.sp[blockIdx.z].active == 0.0) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int n0 = a.n0, n1 = a.n1, m0 = a.m0;
    if (k >= n0 * n1) return;
    const int i = k % n0, j = k / n0;
    if (i < 1 || i > n0 - 2 || j < 1 || j > n1 - 2) return;
    const size_t zc = cslot(a.cmap) * a.Nc;
    const int I = i >> 1, J = j >> 1;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        auto E = [&](int ii, int jj) {
            const size_t C = zc + idx(ii, jj, m0);
            return a.c[q][C] - a.c0[q][C];
        };
        double e;
        if ((i & 1) == 0 && (j & 1) == 0) e = E(I, J);
        else if ((j & 1) == 0) e = 0.5 * (E(I, J) + E(I + 1, J));
        else if ((i & 1) == 0) e = 0.5 * (E(I, J) + E(I, J + 1));
        else e = 0.25 * (E(I, J) + E(I + 1, J) + E(I, J + 1) + E(I + 1, J + 1));
        a.f[q][zslot() * a.Nf + (size_t)k] += e;
    }
}

// ------------------------------------------------------------------------
// FP64 pressure-correction operator straight from the metrics (the defect of
// the mixed path's iterative refinement, and cfd_case_p_apply)

struct MArgs {
    int n0, n1;
    double h1, h2;
    const double *xix, *xiy, *etx, *ety, *jac, *d;
};

__device__ __forceinline__ double mrow(const MArgs &a, const double *__restrict__ x, int i, int j) {
    double c[9];
    p9_coefs<double>(a.n0, a.n1, a.xix, a.xiy, a.etx, a.ety, a.jac, a.d, a.h1, a.h2, i, j, c);
    double s = 0.0;
#pragma unroll
    for (int dj = -1; dj <= 1; dj++)
#pragma unroll
        for (int di = -1; di <= 1; di++) {
            const int ii = clampi(i + di, 1, a.n0 - 2), jj = clampi(j + dj, 1, a.n1 - 2);
            s += c[(dj + 1) * 3 + (di + 1)] * x[idx(ii, jj, a.n0)];
        }
    return s;
}

// y = L x, or y = b - L x when b != nullptr; 0 on the border
__global__ __launch_bounds__(NT) void k_mapply(MArgs a, const double *__restrict__ x,
                                               const double *__restrict__ b, double *__restrict__ y) {
    {
        const size_t o = zslot() * a.n0 * a.n1;
        a.xix += o; a.xiy += o; a.etx += o; a.ety += o; a.jac += o; a.d += o;
        x += o; y += o;
        if (b) b += o;
    }
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.n0 || j >= a.n1) return;
    const size_t P = idx(i, j, a.n0);
    if (i == 0 || j == 0 || i == a.n0 - 1 || j == a.n1 - 1) { y[P] = 0.0; return; }
    const double v = mrow(a, x, i, j);
    y[P] = b ? b[P] - v : v;
}

// ------------------------------------------------------------------------
// mixed path: FP64 iterative refinement around the FP32 solve, and the p' sum

// ||q||, ||q - mean(q)|| (one pass: sum q^2 - sum(q) mean), p' = 0 and the
// refinement's start: tolerance, done flag, first inner tolerance
__global__ __launch_bounds__(NT) void k_q_norms(int n0, int n1, const double *__restrict__ q, double *__restrict__ pc,
                                                RedArgs ra) {
    ra = rz(ra);
    if (slot_skip(ra.scal, -1)) return;
    {
        const size_t o = zslot() * n0 * n1;
        q += o; pc += o;
    }
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int m = 0; m < FLAT_NPT; m++) {
        const FlatNode fn = flat_node(n0, n1, m);
        if (!fn.in) continue;
        const int i = fn.i, j = fn.j;
        const size_t P = fn.P;
        pc[P] = 0.0;
        if (inner_node(i, j, n0, n1)) {
            v[0] += q[P];
            v[1] += q[P] * q[P];
        }
    }
    block_partials<2>(v, ra.part, ra.pstride);
}

// FP64 defect r = q - L (p' + e) (faces stored in FP64), its norm with the mean
// removed, the refinement test and the next inner tolerance
__global__ __launch_bounds__(NT) void k_defect(FArgs<double> a, const double *__restrict__ q,
                                               const double *__restrict__ pc, const float *__restrict__ ef,
                                               double *__restrict__ r, RedArgs ra) {
    ra = rz(ra);
    if (slot_skip(ra.scal, S_RDONE)) return;
    {
        const size_t o = zslot() * a.n0 * a.n1;
        a = fz(a, o); q += o; pc += o; ef += o; r += o;
    }
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int m = 0; m < FLAT_NPT; m++) {
        const FlatNode fn = flat_node(a.n0, a.n1, m);
        if (!fn.in) continue;
        const int i = fn.i, j = fn.j;
        const size_t P = fn.P;
        double y = 0.0;
        if (inner_node(i, j, a.n0, a.n1)) {
            y = q[P] - frow_f<double>(a, i, j, [&](int ii, int jj) {
                    const size_t Q = idx(ii, jj, a.n0);
                    return pc[Q] + (double)ef[Q];
                });
            v[0] += y;
            v[1] += y * y;
        }
        r[P] = y;
    }
    block_partials<2>(v, ra.part, ra.pstride);
}

// end of the p' solve: (mixed) p' += e, e = 0; interior sum of p' -> S_PSUM
template <bool ACC>
__global__ __launch_bounds__(NT) void k_pc_final(int n0, int n1, double *__restrict__ pc, float *__restrict__ ef,
                                                 RedArgs ra) {
    ra = rz(ra);
    if (slot_skip(ra.scal, -1)) return;
    {
        const size_t o = zslot() * n0 * n1;
        pc += o;
        if (ACC) ef += o;
    }
    double v[1] = {0.0};
#pragma unroll
    for (int m = 0; m < FLAT_NPT; m++) {
        const FlatNode fn = flat_node(n0, n1, m);
        if (!fn.in) continue;
        const int i = fn.i, j = fn.j;
        const size_t P = fn.P;
        double x = pc[P];
        if (ACC) {
            x += (double)ef[P];
            pc[P] = x;
            ef[P] = 0.0f;
        }
        if (inner_node(i, j, n0, n1)) v[0] += x;
    }
    block_partials<1>(v, ra.part, ra.pstride);
}

// End of a mixed p' solve with one refinement step: p' was zeroed before it, so e
// is the whole correction.  One pass computes the FP64 defect q - L e (partials
// of its sum and square, the defect norm with the mean removed), sets p' = e and
// adds up the interior p' (third partial).  e is left as it is (the next solve
// starts from zero).
__global__ __launch_bounds__(NT) void k_defect_final(FArgs<double> a, const double *__restrict__ q,
                                                     double *__restrict__ pc, const float *__restrict__ ef,
                                                     RedArgs ra) {
    ra = rz(ra);
    if (slot_skip(ra.scal, -1)) return;
    {
        const size_t o = zslot() * a.n0 * a.n1;
        a = fz(a, o); q += o; pc += o; ef += o;
    }
    double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < FLAT_NPT; m++) {
        const FlatNode fn = flat_node(a.n0, a.n1, m);
        if (!fn.in) continue;
        const int i = fn.i, j = fn.j;
        const size_t P = fn.P;
        const double x = (double)ef[P];
        if (inner_node(i, j, a.n0, a.n1)) {
            const double y = q[P] - frow_f<double>(a, i, j, [&](int ii, int jj) { return (double)ef[idx(ii, jj, a.n0)]; });
            v[0] += y;
            v[1] += y * y;
            v[2] += x;
        }
        pc[P] = x;
    }
    block_partials<3>(v, ra.part, ra.pstride);
}

// totals of the case-level reductions and what follows from them (one block)
enum { CF_Q, CF_DEFECT, CF_PSUM, CF_DEFECT_PSUM };
// margin: fraction of the remaining FP64 tolerance the first FP32 solve aims at
// (1 when no further refinement step follows it)
template <int OP>
__global__ __launch_bounds__(FIN_NT) void k_cav_fin(int gs, RedArgs ra, int nparts, int n0, int n1, double rtol,
                                                    double margin) {
    ra = rz(ra);
    const bool skip = slot_skip(ra.scal, gs);  // read while the partials load
    constexpr int NV = OP == CF_PSUM ? 1 : (OP == CF_DEFECT_PSUM ? 3 : 2);
    double tot[NV];
    sum_partials<NV, FIN_NT>(ra.part, ra.pstride, nparts, tot);
    if (skip || tlin() != 0) return;
    double *s = ra.scal;
    const double nin = (double)(n0 - 2) * (double)(n1 - 2);
    if constexpr (OP == CF_DEFECT_PSUM) s[S_PSUM] = tot[NV - 1];
    if (OP == CF_PSUM) {
        s[S_PSUM] = tot[0];
    } else if (OP == CF_Q) {
        // ||q|| and ||q - mean(q)|| = sqrt(sum q^2 - sum(q) mean); refinement start
        // (a one-V-cycle slot has no tolerance)
        if (s[S_PMAX] < 0.0) rtol = 0.0;
        const double rr = fmax(tot[NV - 1] - tot[0] * (tot[0] / nin), 0.0);
        const double bn = sqrt(tot[NV - 1]), rn = sqrt(rr);
        s[S_BB64] = tot[NV - 1];
        s[S_RR64] = rr;
        s[S_BN64] = bn;
        s[S_TOL64] = rtol * bn;
        s[S_RN64] = rn;
        s[S_RELRES64] = bn > 0.0 ? rn / bn : 0.0;
        s[S_ITTOT] = 0.0;
        s[S_RDONE] = (bn == 0.0 || rn <= rtol * bn) ? 1.0 : 0.0;
        s[S_RTOL] = rn > 0.0 ? fmax(margin * rtol * bn / rn, 1e-5) : 1e-5;
        // the start of the first inner solve (its right-hand side is q)
        s[S_SUM] = tot[0];
        s[S_BB] = tot[NV - 1];
        s[S_DONE] = 0.0;
        s[S_IT] = 0.0;
        s[S_SCONV] = 0.0;
    } else {
        // refinement test on the FP64 defect (mean removed) and the next inner tolerance
        // (after a solve that the start test skipped there is nothing to add)
        if (OP == CF_DEFECT_PSUM && s[S_RDONE] != 0.0) return;
        s[S_ITTOT] += fmax(s[S_IT], 1.0);
        const double rr = fmax(tot[NV - 1] - tot[0] * (tot[0] / nin), 0.0);
        const double rn = sqrt(rr);
        s[S_RR64] = rr;
        s[S_RN64] = rn;
        s[S_RELRES64] = s[S_BN64] > 0.0 ? rn / s[S_BN64] : 0.0;
        if (rn <= s[S_TOL64]) s[S_RDONE] = 1.0;
        else s[S_RTOL] = fmax(0.5 * s[S_TOL64] / rn, 1e-5);
    }
}

}  // namespace cfd

// ==========================================================================
// host side

using namespace cfd;

namespace {

template <typename T>
int dalloc(Case *c, T **p, size_t n) {
    if (c->dry) {
        *p = nullptr;
        c->bytes += n * sizeof(T);
        return CFD_OK;
    }
    void *ptr = nullptr;
    if (hipMalloc(&ptr, n * sizeof(T)) != hipSuccess) return CFD_E_OOM;
    if (hipMemsetAsync(ptr, 0, n * sizeof(T), c->s) != hipSuccess) return CFD_E_HIP;
    c->allocs.push_back(ptr);
    c->bytes += n * sizeof(T);
    *p = static_cast<T *>(ptr);
    return CFD_OK;
}

#define RET(x) CFD_RET(x)

// ---- live kernel timing
// kinds: 0 p' operator apply (fine level), 1 predictor Jacobi sweep,
// 2 residual/RHS, 3 face divergence, 4 multigrid down (fine level), 5 multigrid up
// The pair of events that the next launch of `kind` is to be bound to (both null
// when profiling is off, the stream is being captured or the events are used up).
Timed prof_timed(Case *c, int kind) {
    if (!c->prof || c->capturing || c->ev_used + 2 > (int)c->ev.size()) return Timed{};
    const Timed t{c->ev[c->ev_used], c->ev[c->ev_used + 1]};
    c->ev_used += 2;
    c->ev_kind.push_back(kind);
    return t;
}
void prof_drain(Case *c) {
    if (!c->prof) return;
    // the caller may have seen the device's flags before the runtime marks the
    // events complete
    if (c->ev_used > 0) (void)hipEventSynchronize(c->ev[c->ev_used - 1]);
    for (int k = 0; k + 1 < c->ev_used; k += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]) == hipSuccess) {
            const int kind = c->ev_kind[k / 2];
            c->prof_ms[kind] += ms;
            c->prof_n[kind] += 1;
        }
    }
    (void)hipGetLastError();  // a failed query must not surface at the next launch check
    c->ev_used = 0;
    c->ev_kind.clear();
}
Timed cb_timed(void *o, int kind) { return prof_timed(static_cast<Case *>(o), kind); }
void cb_drain(void *o) { prof_drain(static_cast<Case *>(o)); }

size_t psz(const Case *c) { return (size_t)PS_SLOTS * c->pstride; }
PsCtx ctx(Case *c) {
    PsCtx cx{c->s, c->B, c->part, c->pstride, psz(c), c->scal, c->hscal, cb_timed, cb_drain, c,
             &c->allocs, &c->bytes, c->dry};
    if (c->dflag) {
        cx.flag_d = c->dflag;
        cx.flag_h = c->hflag;
        cx.seq = &c->flag_seq;
    }
    return cx;
}
RedArgs red(const Case *c, int slot = 0) {
    return RedArgs{c->part + (size_t)slot * c->pstride, c->pstride, psz(c), c->scal};
}
CGrid cg(const Case *c, const SlotP *sp = nullptr) {
    return CGrid{c->n0, c->n1, c->h1, c->h2, c->N, sp ? sp : c->spd, make_recip(c->h1, c->h2), c->prm.hyb_blend};
}
// the level's FAS sources when it uses them
Src srcs(const Case *c) {
    if (!c->use_src) return Src{nullptr, nullptr, nullptr, nullptr};
    return Src{c->src[0], c->src[1], c->src[2], c->src[3]};
}
dim3 grid_b(const Case *c) {
    dim3 g = grid2d(c->n0, c->n1);
    g.z = c->B;
    return g;
}
// 64 x 4 tiles of the face divergence and the correction, which write interior
// nodes only, and the number of their block partials
dim3 grid_in(const Case *c) { return dim3(tiles_inner(c->n0, BX), tiles_inner(c->n1, BY), c->B); }
int nblk_in(int n0, int n1) { return tiles_inner(n0, BX) * tiles_inner(n1, BY); }

// one-block-per-slot finish of a case-level reduction over the grid's block partials
template <int OP>
int cav_fin(Case *c, int gs = -1, double rtol = 0.0, double margin = 0.5, int nparts = -1) {
    if (nparts < 0) nparts = nblk(c->n0, c->n1);
    k_cav_fin<OP><<<dim3(1, 1, c->B), FIN_NT, 0, c->s>>>(gs, red(c), nparts, c->n0, c->n1, rtol, margin);
    CFD_LAUNCH_CHECK();
    return CFD_OK;
}

int sync_scal(Case *c) {
    PsCtx cx = ctx(c);
    return ps_sync(cx);
}

MArgs margs(const Case *c) {
    return MArgs{c->n0, c->n1, c->h1, c->h2, c->xix, c->xiy, c->etx, c->ety, c->ajac, c->d};
}

// coarse-level geometry arrays and the solver of the batch's precision
int build_levels(Case *c) {
    const int want = c->prm.mg_levels < 1 ? 1 : c->prm.mg_levels;
    Geom &G0 = c->geom[0];
    G0.n0 = c->n0; G0.n1 = c->n1; G0.h1 = c->h1; G0.h2 = c->h2;
    G0.xix = c->xix; G0.xiy = c->xiy; G0.etx = c->etx; G0.ety = c->ety; G0.jac = c->ajac;
    int l = 1;
    while (l < want && l < MAXLEV) {
        const Geom &F = c->geom[l - 1];
        if ((F.n0 & 1) == 0 || (F.n1 & 1) == 0 || F.n0 < 9 || F.n1 < 9) break;
        Geom &L = c->geom[l];
        L.n0 = (F.n0 + 1) / 2; L.n1 = (F.n1 + 1) / 2;
        L.h1 = 2.0 * F.h1; L.h2 = 2.0 * F.h2;
        const size_t M = (size_t)c->B * L.n0 * L.n1;
        RET(dalloc(c, &L.xix, M)); RET(dalloc(c, &L.xiy, M)); RET(dalloc(c, &L.etx, M));
        RET(dalloc(c, &L.ety, M)); RET(dalloc(c, &L.jac, M));
        l++;
    }
    c->nlev = l;
    PsCtx cx = ctx(c);
    if (c->mixed) {
        c->ps32 = new PSolver<float>();
        RET(c->ps32->init(cx, c->geom, c->nlev, c->prm));
    } else {
        c->ps64 = new PSolver<double>();
        RET(c->ps64->init(cx, c->geom, c->nlev, c->prm));
    }
    return CFD_OK;
}

// coarse metrics of slot z by repeated injection
int inject_slot(Case *c, int z) {
    const dim3 bl(BX, BY);
    for (int l = 1; l < c->nlev; l++) {
        const Geom &F = c->geom[l - 1];
        const Geom &L = c->geom[l];
        const size_t of = (size_t)z * F.n0 * F.n1, oc = (size_t)z * L.n0 * L.n1;
        const dim3 gr = grid2d(L.n0, L.n1);
        const double *src[5] = {F.xix, F.xiy, F.etx, F.ety, F.jac};
        double *dst[5] = {L.xix, L.xiy, L.etx, L.ety, L.jac};
        for (int k = 0; k < 5; k++) k_inject<<<gr, bl, 0, c->s>>>(F.n0, F.n1, src[k] + of, L.n0, L.n1, dst[k] + oc);
        CFD_LAUNCH_CHECK();
    }
    return CFD_OK;
}

// per outer iteration: faces and diagonals of every level from the current d
int mg_setup(Case *c) {
    PsCtx cx = ctx(c);
    if (c->mixed) return c->ps32->setup(cx, c->d, c->f64);
    return c->ps64->setup(cx, c->d, nullptr);
}

// p' solve into c->pc (cleared first), ending with the interior sum of p' in S_PSUM.
// FP64: BiCGStab + multigrid in FP64.  Mixed: FP64 iterative refinement around
// FP32 BiCGStab + multigrid solves, the FP64 defect (mean removed) deciding
// convergence; one refinement step when rtol >= 1e-4 (the FP32 solve alone can
// reach it, so it aims at rtol itself), up to six otherwise (each FP32 solve
// aiming at half the remaining tolerance).  graph: capturable (no host
// synchronisation).
// Iterations and relative residual: S_IT / S_RELRES (FP64), S_ITTOT / S_RELRES64.
// Times the multigrid block sizes once (PSolver::tune) unless switched off or the
// stream is being captured.  The timing launches overwrite the solver's scratch
// (level arrays, BiCGStab t, partial rows PS_APPLY_ROW and PS_APPLY_ROW + 1), so this runs only where none of
// them holds a value still needed: before an outer iteration or a p' solve starts.
int tune_mg(Case *c) {
    if (!c->prm.mg_tune || c->capturing) return CFD_OK;
    PsCtx cx = ctx(c);
    if (c->ps32) RET(c->ps32->tune(cx));
    if (c->ps64) RET(c->ps64->tune(cx));
    return CFD_OK;
}

int launch_residual(Case *c);

// Autotuning of the residual tile rows: k_residual is timed with 4- and 8-row
// tiles on the level's own buffers (one untimed launch of each, then the smaller
// minimum over five rounds of alternating timed launches) and the faster is kept.
// The launches write only what every SIMPLE iteration rewrites before reading it
// (R, U, V, d and the residual partial rows), so this runs before the first outer
// iteration (or capture).  Once per level; not while the stream is being captured.
int tune_res(Case *c) {
    if (!c->prm.mg_tune || c->capturing || c->res_tuned || c->dry) return CFD_OK;
    c->res_tuned = true;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CFD_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return CFD_E_HIP;
    }
    const int tys[2] = {4, 8};
    float best[2] = {3.0e38f, 3.0e38f};
    int rc = CFD_OK;
    auto launch = [&](int k) {
        c->res_ty = tys[k];
        if (launch_residual(c) != CFD_OK) rc = CFD_E_HIP;
    };
    launch(0);
    launch(1);
    for (int rep = 0; rep < 5 && rc == CFD_OK; rep++)
        for (int k = 0; k < 2 && rc == CFD_OK; k++) {
            float ms = 0.f;
            if (hipEventRecord(e0, c->s) != hipSuccess) rc = CFD_E_HIP;
            launch(k);
            if (rc == CFD_OK && (hipEventRecord(e1, c->s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                 hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = CFD_E_HIP;
            best[k] = std::min(best[k], ms);
        }
    c->res_ty = (rc == CFD_OK && best[0] < best[1]) ? 4 : 8;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// maxit < 0: p' is one V-cycle on the mean-free q (no Krylov iteration, no
// tolerance), as the oracle's vcycle_solve()
int psolve(Case *c, double rtol, int maxit, bool graph, bool q_norms_ready = false) {
    PsCtx cx = ctx(c);
    const int n0 = c->n0, n1 = c->n1;
    // the p' kernels here run one thread per node in memory order
    const dim3 gf(nflat(n0, n1), 1, c->B), bl(BX, BY);
    const int nf = nflat(n0, n1);
    const bool vc = maxit < 0;
    if (vc) rtol = 0.0;
    if (!c->mixed) {
        // with the face divergence's sums of q finished, the solve starts in one pass
        if (q_norms_ready) RET(cav_fin<CF_Q>(c, -1, rtol, 1.0, nblk_in(n0, n1)));
        if (vc) RET(c->ps64->vsolve(cx, c->q, c->pc, -1, q_norms_ready));
        else RET(c->ps64->solve(cx, c->q, c->pc, nullptr, rtol, -1, maxit, -1, graph, q_norms_ready));
        k_pc_final<false><<<gf, bl, 0, c->s>>>(n0, n1, c->pc, nullptr, red(c));
        CFD_LAUNCH_CHECK();
        return cav_fin<CF_PSUM>(c, -1, 0.0, 0.5, nf);
    }
    const int R = (vc || rtol >= 1e-4) ? 1 : 6;
    const FArgs<double> a64{n0, n1, c->f64[0], c->f64[1], c->f64[2], c->f64[3]};
    if (!q_norms_ready) {
        k_q_norms<<<gf, bl, 0, c->s>>>(n0, n1, c->q, c->pc, red(c));
        CFD_LAUNCH_CHECK();
    } else if (R > 1) {  // the refinement accumulates into p' from zero
        k_fill<double><<<(unsigned)((c->B * c->N + 255) / 256), 256, 0, c->s>>>(c->B * c->N, c->pc, 0.0);
        CFD_LAUNCH_CHECK();
    }
    RET(cav_fin<CF_Q>(c, -1, rtol, R == 1 ? 1.0 : 0.5, q_norms_ready ? nblk_in(n0, n1) : nf));
    for (int k = 0; k < R; k++) {
        if (!graph && k > 0) {
            RET(ps_sync(cx));
            bool all = true;
            for (int z = 0; z < c->B; z++) {
                const double *h = c->hscal + (size_t)z * SS;
                if (h[S_ACTIVE] != 0.0 && h[S_RDONE] == 0.0) all = false;
            }
            if (all) break;
        }
        if (vc) RET(c->ps32->vsolve(cx, c->q, c->ef, S_RDONE, true));
        else
            RET(c->ps32->solve(cx, k == 0 ? c->q : c->r64, c->ef, k == 0 ? nullptr : c->pc, 0.0, S_RTOL, maxit,
                               S_RDONE, graph, k == 0));
        if (R == 1) {  // defect, p' = e and its sum in one pass
            k_defect_final<<<gf, bl, 0, c->s>>>(a64, c->q, c->pc, c->ef, red(c));
            CFD_LAUNCH_CHECK();
            return cav_fin<CF_DEFECT_PSUM>(c, -1, 0.0, 0.5, nf);
        }
        k_defect<<<gf, bl, 0, c->s>>>(a64, c->q, c->pc, c->ef, c->r64, red(c));
        CFD_LAUNCH_CHECK();
        RET(cav_fin<CF_DEFECT>(c, S_RDONE, 0.0, 0.5, nf));
    }
    k_pc_final<true><<<gf, bl, 0, c->s>>>(n0, n1, c->pc, c->ef, red(c));
    CFD_LAUNCH_CHECK();
    return cav_fin<CF_PSUM>(c, -1, 0.0, 0.5, nf);
}

void psolve_result(const Case *c, int z, int *its, double *relres) {
    const double *h = c->hscal + (size_t)z * SS;
    *its = c->mixed ? (int)h[S_ITTOT] : (int)h[S_IT];
    *relres = c->mixed ? h[S_RELRES64] : h[S_RELRES];
}

GridPtrs gptrs(const Case *c) {
    return GridPtrs{c->xix, c->xiy, c->etx, c->ety, c->alph, c->beta, c->gamma, c->p1, c->q1};
}
WallIn wall_in(const Case *c) {
    return WallIn{c->u, c->v, c->T, c->p, c->xix, c->xiy, c->etx, c->ety, c->alph, c->beta, c->gamma,
                  c->p1, c->q1, c->ajac};
}

template <typename TR, int TY>
void launch_residual_ty(Case *c, const Timed &tm) {
    const dim3 gr = grid_res(c->n0, c->n1, c->B, TY);
    launch_timed(tm, k_residual<TR, TY>, gr, dim3(BX, BY), c->s, cg(c), (const double *)c->u, (const double *)c->v,
                 (const double *)c->T, (const double *)c->p, gptrs(c), (TR *)c->Ru, (TR *)c->Rv, (TR *)c->RT, c->d,
                 (TR *)c->U, (TR *)c->V, red(c, PS_RES), srcs(c));
}
template <typename TR>
int launch_residual_t(Case *c) {
    const Timed tm = prof_timed(c, 2);
    if (c->res_ty == 4) launch_residual_ty<TR, 4>(c, tm);
    else launch_residual_ty<TR, 8>(c, tm);
    CFD_LAUNCH_CHECK();
    return CFD_OK;
}
int launch_residual(Case *c) { return c->mixed ? launch_residual_t<float>(c) : launch_residual_t<double>(c); }

// predictor sweeps; *du.. receive the buffers holding the final corrections
template <typename TR>
int predictor_t(Case *c, void **du, void **dv, void **dT) {
    TR *a[3] = {(TR *)c->jb[0], (TR *)c->jb[1], (TR *)c->jb[2]};
    TR *b[3] = {(TR *)c->jb[3], (TR *)c->jb[4], (TR *)c->jb[5]};
    const TR *co[5];
    if (c->mixed) {
        for (int k = 0; k < 5; k++) co[k] = (const TR *)c->a32[k];
    } else {
        const double *m[5] = {c->alph, c->beta, c->gamma, c->p1, c->q1};
        for (int k = 0; k < 5; k++) co[k] = (const TR *)m[k];
    }
    const TR *Ru = (const TR *)c->Ru, *Rv = (const TR *)c->Rv, *RT = (const TR *)c->RT;
    const TR *U = (const TR *)c->U, *V = (const TR *)c->V;
    const dim3 gr = grid_b(c), bl(BX, BY);
    const dim3 g2(tiles_inner(c->n0, JT_X), tiles_inner(c->n1, JT_Y), c->B);
    const int m = c->prm.mom_sweeps;
    for (int s = 0; s < m;) {
        if (m - s >= 2) {  // four or two sweeps in one launch
            const JArgs<TR> ja{Ru, Rv, RT, U, V, co[0], co[1], co[2], co[3], co[4], a[0], a[1], a[2], b[0], b[1], b[2]};
            const int ns = m - s >= 4 ? 4 : 2;
            const Timed tm = s > 0 ? prof_timed(c, 1) : Timed{};
            if (ns == 4) {
                if (s == 0) launch_timed(tm, k_jacobin<TR, true, 4>, g2, dim3(NT), c->s, cg(c), ja);
                else launch_timed(tm, k_jacobin<TR, false, 4>, g2, dim3(NT), c->s, cg(c), ja);
            } else {
                if (s == 0) launch_timed(tm, k_jacobin<TR, true, 2>, g2, dim3(NT), c->s, cg(c), ja);
                else launch_timed(tm, k_jacobin<TR, false, 2>, g2, dim3(NT), c->s, cg(c), ja);
            }
            s += ns;
        } else if (s == 0) {
            k_jacobi<TR, true><<<gr, bl, 0, c->s>>>(cg(c), Ru, Rv, RT, U, V, co[0], co[1], co[2], co[3], co[4], a[0],
                                                    a[1], a[2], b[0], b[1], b[2]);
            s += 1;
        } else {
            k_jacobi<TR, false><<<gr, bl, 0, c->s>>>(cg(c), Ru, Rv, RT, U, V, co[0], co[1], co[2], co[3], co[4], a[0],
                                                     a[1], a[2], b[0], b[1], b[2]);
            s += 1;
        }
        CFD_LAUNCH_CHECK();
        for (int k = 0; k < 3; k++) std::swap(a[k], b[k]);
    }
    if (c->prm.mom_sweeps <= 0) {
        const size_t n = (size_t)c->B * c->N;
        const unsigned g = (unsigned)((n + 255) / 256);
        for (int k = 0; k < 3; k++) k_fill<TR><<<g, 256, 0, c->s>>>(n, a[k], TR(0));
        CFD_LAUNCH_CHECK();
    }
    *du = a[0]; *dv = a[1]; *dT = a[2];
    return CFD_OK;
}

// Wall conditions of the slots [z0, z0 + nz): velocity/temperature values
// (uvt), adiabatic sides, then the wall pressure (which reads the new wall T).
// Launched on pointers offset to slot z0, so other slots are untouched.
// sp (may be null; then z0 = 0 and nz = all slots): the slot table to launch with.
int wall_bcs(Case *c, int z0, int nz, bool uvt, const SlotP *sp = nullptr) {
    const int m = std::max(c->n0, c->n1);
    const size_t o = (size_t)z0 * c->N;
    const Src sr = srcs(c);
    const double *sT = sr.T ? sr.T + o : nullptr, *sm = sr.m ? sr.m + o : nullptr;
    CGrid g = cg(c, sp);
    g.sp += z0;
    WallIn w = wall_in(c);
    w.u += o; w.v += o; w.T += o; w.p += o; w.xix += o; w.xiy += o; w.etx += o; w.ety += o;
    w.alph += o; w.beta += o; w.gamma += o; w.p1 += o; w.q1 += o; w.ajac += o;
    const dim3 g1((m + 255) / 256, 1, nz);
    if (uvt) k_uvt_walls<<<g1, 256, 0, c->s>>>(g, c->u + o, c->v + o, c->T + o);
    const dim3 gs((c->n1 + WB - 1) / WB, 1, nz);
    k_t_sides<<<gs, WB, 0, c->s>>>(g, c->T + o, c->alph + o, c->beta + o, c->scratch + z0 * c->scr, c->scr, sT);
    k_p_walls<<<dim3((2 * m + WB - 1) / WB, 5, nz), WB, 0, c->s>>>(g, w, c->p + o, c->T + o, c->scratch + z0 * c->scr,
                                                               c->scr, sm, nullptr);
    CFD_LAUNCH_CHECK();
    return CFD_OK;
}

// û and the face divergence q of one outer iteration with its norms (mass
// partials in PS_MASS, Σq and Σq² for the p' solve in slots 0-1); du, dv: the
// predictor's final corrections
int launch_face_div(Case *c, const void *du, const void *dv) {
    const Timed tm = prof_timed(c, 3);
    const double *sm = srcs(c).m;
    if (c->mixed)
        launch_timed(tm, k_face_div<float>, grid_in(c), dim3(BX, BY), c->s, cg(c), c->u, c->v, (const float *)du,
                     (const float *)dv, c->p, c->d, c->xix, c->xiy, c->etx, c->ety, c->ajac, c->q,
                     red(c, PS_MASS), red(c), sm);
    else
        launch_timed(tm, k_face_div<double>, grid_in(c), dim3(BX, BY), c->s, cg(c), c->u, c->v,
                     (const double *)du, (const double *)dv, c->p, c->d, c->xix, c->xiy, c->etx, c->ety, c->ajac,
                     c->q, red(c, PS_MASS), red(c), sm);
    CFD_LAUNCH_CHECK();
    return CFD_OK;
}

// Everything of one SIMPLE iteration that runs on the stream.  graph: capturable
// (the p' solve is controlled on the device); otherwise the p' solve reads its
// convergence flags on the host between half-iterations.  stats: finish the
// iteration's statistics (the top level of a FAS hierarchy; a slot's first
// statistics also become its residual normalisers).
int simple_body(Case *c, bool graph, bool stats = true) {
    const dim3 bl(BX, BY);
    // 1-2 residuals and d (FP64 arithmetic; stored in the predictor precision)
    RET(launch_residual(c));
    // 3 predictor
    void *du, *dv, *dT;
    if (c->mixed) RET(predictor_t<float>(c, &du, &dv, &dT));
    else RET(predictor_t<double>(c, &du, &dv, &dT));
    // 4 u* = u + du, u-hat, face divergence and its norms (FP64)
    RET(launch_face_div(c, du, dv));
    // 5 p' solve (its right-hand side's norms come from step 4)
    RET(mg_setup(c));
    RET(psolve(c, c->prm.p_rtol, c->prm.p_maxit, graph, true));
    // 6-7 correction (u*, T += dT, p' mean removed) and BCs (FP64 accumulation)
    if (c->mixed)
        k_correct<float><<<grid_in(c), bl, 0, c->s>>>(cg(c), (const float *)du, (const float *)dv, (const float *)dT, c->pc,
                                              c->d, c->xix, c->xiy, c->etx, c->ety, c->scal, c->u, c->v, c->p,
                                              c->T, red(c));
    else
        k_correct<double><<<grid_in(c), bl, 0, c->s>>>(cg(c), (const double *)du, (const double *)dv, (const double *)dT,
                                               c->pc, c->d, c->xix, c->xiy, c->etx, c->ety, c->scal, c->u, c->v,
                                               c->p, c->T, red(c));
    CFD_LAUNCH_CHECK();
    RET(wall_bcs(c, 0, c->B, false));
    // 8 dmax (partials from the correction) and the wall Nusselt numbers (FP64)
    if (stats) {
        const dim3 grr = grid_res(c->n0, c->n1, 1, c->res_ty);
        k_dmax_nu_fin<<<dim3(1, 1, c->B), FIN_NT, 0, c->s>>>(cg(c), red(c), nblk_in(c->n0, c->n1), (int)(grr.x * grr.y),
                                                             c->T, c->ajac, c->etx, c->ety);
        CFD_LAUNCH_CHECK();
    }
    return CFD_OK;
}

// ---- FAS multigrid outer cycle (oracle: fas_cycle)

unsigned nblk256(size_t n) { return (unsigned)((n + 255) / 256); }

// R/dt's numerators (FP64) of u, v, T at the current state into fr[0..2], d, and the
// mass residual q of the state's own face velocities (no predictor correction),
// with the level's sources when with_src; sp: the slot table to launch with
int state_residual(Case *c, bool with_src, const SlotP *sp) {
    const bool keep = c->use_src;
    c->use_src = with_src && keep;
    const dim3 gr = grid_res(c->n0, c->n1, c->B, c->res_ty);
    if (c->res_ty == 4)
        k_residual<double, 4><<<gr, dim3(BX, BY), 0, c->s>>>(cg(c, sp), c->u, c->v, c->T, c->p, gptrs(c), c->fr[0],
                                                            c->fr[1], c->fr[2], c->d, c->uh, c->vh, red(c, PS_RES),
                                                            srcs(c));
    else
        k_residual<double, 8><<<gr, dim3(BX, BY), 0, c->s>>>(cg(c, sp), c->u, c->v, c->T, c->p, gptrs(c), c->fr[0],
                                                            c->fr[1], c->fr[2], c->d, c->uh, c->vh, red(c, PS_RES),
                                                            srcs(c));
    CFD_LAUNCH_CHECK();
    k_face_div<double><<<grid_in(c), dim3(BX, BY), 0, c->s>>>(cg(c, sp), c->u, c->v, (const double *)nullptr,
                                                             (const double *)nullptr, c->p, c->d, c->xix, c->xiy,
                                                             c->etx, c->ety, c->ajac, c->q, red(c, PS_MASS), red(c),
                                                             srcs(c).m);
    CFD_LAUNCH_CHECK();
    c->use_src = keep;
    return CFD_OK;
}

// the side-wall temperature update into scratch (with_src: the level's sources)
int side_walls(Case *c, bool with_src, const SlotP *sp) {
    const dim3 gs((c->n1 + WB - 1) / WB, 1, c->B);
    k_t_sides<<<gs, WB, 0, c->s>>>(cg(c, sp), c->T, c->alph, c->beta, c->scratch, c->scr,
                                   with_src ? srcs(c).T : nullptr);
    CFD_LAUNCH_CHECK();
    return CFD_OK;
}
// the wall pressure the wall update would set (after the side-wall update in
// scratch) into fr[3]
int wall_p_out(Case *c, bool with_src, const SlotP *sp) {
    const int m = std::max(c->n0, c->n1);
    k_p_walls<<<dim3((2 * m + WB - 1) / WB, 5, c->B), WB, 0, c->s>>>(cg(c, sp), wall_in(c), c->p, c->T, c->scratch,
                                                                 c->scr, with_src ? srcs(c).m : nullptr, c->fr[3]);
    CFD_LAUNCH_CHECK();
    return CFD_OK;
}

// A coarse FAS level that k_simple_small takes: at most SL_NODES nodes and SL_NT
// interior ones, at most CW_LEVS p' levels whose V-cycle arrays fit its LDS
bool small_fits(const Case *c) {
    if (!c->prm.fas_small || c->N > (size_t)SL_NODES || (c->n0 - 2) * (c->n1 - 2) > SL_NT) return false;
    if (c->nlev > CW_LEVS || c->prm.mg_pre > 2 || c->prm.mg_post > 2) return false;
    int tot = 0;
    for (int l = 0; l < c->nlev; l++) tot += c->geom[l].n0 * c->geom[l].n1;
    return tot <= CN_MAX;
}

// The face tables of a level that k_simple_small takes, and its arguments in device
// memory (written once here: everything in them is fixed when the level is created;
// the level's FAS sources are always applied, as on every coarse level)
int alloc_small(Case *c) {
    if (!small_fits(c)) return CFD_OK;
    for (int l = 0; l < c->nlev; l++)
        for (int k = 0; k < 5; k++) RET(dalloc(c, &c->sl_tab[l][k], (size_t)c->B * c->geom[l].n0 * c->geom[l].n1));
    RET(dalloc(c, &c->sl_args, 1));
    if (c->dry) return CFD_OK;
    SmallArgs a{};
    a.g = cg(c);
    a.gp = gptrs(c);
    a.jac = c->ajac;
    a.u = c->u; a.v = c->v; a.T = c->T; a.p = c->p;
    a.src = Src{c->src[0], c->src[1], c->src[2], c->src[3]};
    a.sweeps = c->prm.mom_sweeps;
    a.p_maxit = c->prm.p_maxit;
    a.p_rtol = c->prm.p_rtol;
    a.mg.nlev = c->nlev;
    a.mg.pre = c->prm.mg_pre;
    a.mg.post = c->prm.mg_post;
    a.mg.coarse = c->prm.mg_coarse;
    a.mg.omega = c->prm.mg_omega;
    int off = 0;
    for (int l = 0; l < c->nlev; l++) {
        double *const *t = c->sl_tab[l];
        a.mg.L[l] = CLev<double>{c->geom[l].n0, c->geom[l].n1, t[0], t[1], t[2], t[3], t[4]};
        a.mg.off[l] = off;
        off += CW_ARR * c->geom[l].n0 * c->geom[l].n1;
    }
    for (int l = 0; l < CW_LEVS; l++) a.lev[l] = c->geom[std::min(l, c->nlev - 1)];
    CFD_HIP(hipMemcpyAsync(c->sl_args, &a, sizeof(a), hipMemcpyHostToDevice, c->s));
    CFD_HIP(hipStreamSynchronize(c->s));
    return CFD_OK;
}

// SIMPLE iterations of every slot of a small coarse level in one launch: `iters`,
// or `iters_deep` on a slot that is its case's coarsest level
int simple_small(Case *c, int iters, int iters_deep) {
    if (iters <= 0 && iters_deep <= 0) return CFD_OK;
    if (!c->sl_args || !c->use_src) return CFD_E_ARG;
    // a level of at most 256 interior nodes (and its p' levels' face factors within
    // SL_FF_MAX) in four wavefronts (each with up to 512 registers per lane), a
    // larger one in sixteen
    size_t pnodes = 0;
    for (int l = 0; l < c->nlev; l++) pnodes += (size_t)c->geom[l].n0 * c->geom[l].n1;
    if ((c->n0 - 2) * (c->n1 - 2) <= 256 && 2 * (c->n0 + c->n1) <= 256 && pnodes <= (size_t)SL_FF_MAX)
        k_simple_small<256><<<dim3(1, 1, c->B), 256, 0, c->s>>>(c->sl_args, iters, iters_deep);
    else
        k_simple_small<SL_NT><<<dim3(1, 1, c->B), SL_NT, 0, c->s>>>(c->sl_args, iters, iters_deep);
    CFD_LAUNCH_CHECK();
    return CFD_OK;
}

// `iters` SIMPLE iterations of a level below the top of a FAS hierarchy
int coarse_iters(Case *c, bool graph, int iters) {
    if (small_fits(c)) return simple_small(c, iters, iters);
    for (int k = 0; k < iters; k++) RET(simple_body(c, graph, false));
    return CFD_OK;
}

// Restriction of level c's state to the coarse level cc, and cc's sources: spf is
// c's slot table with only the slots that have a coarse level active, spc cc's with
// only its coarse-level slots active, cmap (null: the same slot) c's slot -> cc's.
int fas_restrict(Case *c, Case *cc, const SlotP *spf, const SlotP *spc, const int *cmap) {
    // this level's residuals at its state, with its own sources
    RET(state_residual(c, true, spf));
    RET(side_walls(c, true, spf));
    RET(wall_p_out(c, true, spf));
    // restricted state and the coarse residuals of it without sources
    const int m0 = cc->n0, m1 = cc->n1;
    const size_t Mc = (size_t)m0 * m1;
    {
        FasInj a{c->n0, m0, m1, c->N, cc->N, spf, cmap, {c->u, c->v, c->T, c->p}, {cc->u, cc->v, cc->T, cc->p},
                 {cc->x0[0], cc->x0[1], cc->x0[2], cc->x0[3]}};
        k_fas_inject<<<dim3(nblk256(Mc), 1, c->B), 256, 0, c->s>>>(a);
        CFD_LAUNCH_CHECK();
    }
    RET(state_residual(cc, false, spc));
    RET(side_walls(cc, false, spc));
    {
        FasSrc a{c->n0, c->n1, m0, m1, c->N, cc->N, c->scr, cc->scr, spf, cmap,
                 {c->fr[0], c->fr[1], c->fr[2]}, c->q, c->T, c->scratch,
                 {cc->fr[0], cc->fr[1], cc->fr[2]}, cc->q, cc->T, cc->scratch,
                 {cc->src[0], cc->src[1], cc->src[2], cc->src[3]}};
        k_fas_src<<<dim3(nblk256(Mc), 1, c->B), 256, 0, c->s>>>(a);
        CFD_LAUNCH_CHECK();
    }
    cc->use_src = true;
    RET(side_walls(cc, true, spc));
    RET(wall_p_out(cc, false, spc));
    {
        FasSrcP a{c->n0, m0, m1, c->N, cc->N, spf, cmap, c->fr[3], c->p, cc->fr[3], cc->p, cc->src[3]};
        k_fas_src_p<<<dim3(nblk256((size_t)2 * (m0 + m1)), 1, c->B), 256, 0, c->s>>>(a);
        CFD_LAUNCH_CHECK();
    }
    return CFD_OK;
}

// Coarse-grid correction of level c from cc (the slots of spf, as fas_restrict's),
// then c's wall conditions on those slots
int fas_prolong(Case *c, Case *cc, const SlotP *spf, const int *cmap) {
    FasPro a{c->n0, c->n1, cc->n0, c->N, cc->N, spf, cmap, {cc->u, cc->v, cc->T, cc->p},
             {cc->x0[0], cc->x0[1], cc->x0[2], cc->x0[3]}, {c->u, c->v, c->T, c->p}};
    k_fas_prolong<<<dim3(nblk256(c->N), 1, c->B), 256, 0, c->s>>>(a);
    CFD_LAUNCH_CHECK();
    return wall_bcs(c, 0, c->B, true, spf);
}

int fas_cycle(Case *c, bool graph, bool top) {
    const cfd_params &pr = c->prm;
    Case *cc = c->coarse;
    // coarse levels control their p' solves on the device (no host round trips)
    const bool g = graph || !top;
    if (!cc) {
        if (!top) return coarse_iters(c, g, pr.fas_coarse);
        for (int k = 0; k < pr.fas_coarse; k++) RET(simple_body(c, g, top));
        return CFD_OK;
    }
    if (!top) RET(coarse_iters(c, g, pr.fas_pre));
    else
        for (int k = 0; k < pr.fas_pre; k++) RET(simple_body(c, g, top));
    RET(fas_restrict(c, cc, c->spd, cc->spd, nullptr));
    RET(fas_cycle(cc, graph, false));
    RET(fas_prolong(c, cc, c->spd, nullptr));
    if (!top) return coarse_iters(c, g, pr.fas_post);
    for (int k = 0; k < pr.fas_post; k++) RET(simple_body(c, g, top));
    return CFD_OK;
}

// one outer iteration: a FAS cycle when the case has coarse levels, else one SIMPLE iteration
int cycle_body(Case *c, bool graph) {
    if (c->coarse) return fas_cycle(c, graph, true);
    return simple_body(c, graph, true);
}

void drop_graph(Case *c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    (void)hipGetLastError();
    c->gexec = nullptr;
    c->graph = nullptr;
}

// capture the outer iteration once
int capture(Case *c) {
    drop_graph(c);
    CFD_HIP(hipStreamSynchronize(c->s));
    c->ev_used = 0;
    c->ev_kind.clear();
    CFD_HIP(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    const int rc = cycle_body(c, true);
    c->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->s, &g);
    if (rc != CFD_OK || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        fprintf(stderr, "cfd_gpu: outer-iteration capture failed (rc %d, %s)\n", rc, hipGetErrorString(e));
        c->ev_used = 0;
        c->ev_kind.clear();
        return rc != CFD_OK ? rc : CFD_E_HIP;
    }
    c->graph = g;
    CFD_HIP(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    return CFD_OK;
}

// launch one outer iteration of every active slot (no synchronisation)
int outer_launch(Case *c) {
    for (Case *l = c; l; l = l->coarse) {  // once, eagerly, before any capture
        RET(tune_mg(l));
        RET(tune_res(l));
    }
    // live kernel timing brackets individual launches, so a profiled iteration runs eagerly
    if (c->prm.graph && !c->prof) {
        if (!c->gexec) RET(capture(c));
        CFD_HIP(hipGraphLaunch(c->gexec, c->s));
        return CFD_OK;
    }
    return cycle_body(c, false);
}

// after an outer iteration: per-slot statistics from the scalars (host mirror current)
int outer_stats(Case *c, cfd_stats *st) {
    int rc = CFD_OK;
    for (int z = 0; z < c->B; z++) {
        const double *h = c->hscal + (size_t)z * SS;
        if (h[S_ACTIVE] == 0.0) {
            if (st) st[z] = cfd_stats{};
            continue;
        }
        c->iter[z]++;
        if (c->iter[z] == 1) {
            c->mass0[z] = h[S_MASS0];
            c->energy0[z] = h[S_ENERGY0];
        }
        if (st) {
            cfd_stats &s = st[z];
            psolve_result(c, z, &s.p_iters, &s.p_relres);
            s.mass_l1 = h[S_MASS];
            s.energy_l1 = h[S_ENERGY];
            s.mom_l1 = h[S_MOM];
            s.dmax = h[S_DMAX];
            s.nu_hot = h[S_NUH];
            s.nu_cold = h[S_NUC];
            s.mass_res = c->mass0[z] > 0 ? h[S_MASS] / c->mass0[z] : 0.0;
            s.energy_res = c->energy0[z] > 0 ? h[S_ENERGY] / c->energy0[z] : 0.0;
            s.iter = c->iter[z];
        }
        if (!std::isfinite(h[S_MASS]) || !std::isfinite(h[S_ENERGY])) rc = CFD_E_NAN;
    }
    return rc;
}

int outer_once(Case *c, cfd_stats *st) {
    RET(outer_launch(c));
    RET(sync_scal(c));
    return outer_stats(c, st);
}

// one launch of a fine-level solver kernel on the solver's own buffers (timing only)
template <typename T>
int time_solver_kernel(Case *c, PSolver<T> *ps, int kernel) {
    OpLevel<T> &L = ps->lev[0];
    dim3 gr = grid2d(L.n0, L.n1);
    gr.z = c->B;
    const dim3 bl(BX, BY);
    if (kernel == 0) {
        k_fapply<T><<<gr, bl, 0, c->s>>>(fargs(L), L.x, L.t);
    } else if (ps->lc == 0) {
        return CFD_E_ARG;  // no fine multigrid level on this grid
    } else if (kernel == 4) {
        OpLevel<T> &C = ps->lev[1];
        DownArgs<T> a{fargs(L), L.invd, L.b, nullptr, 0, nullptr, L.x, C.n0, C.n1, C.t, ps->lpart, (T)ps->omega,
                      ps->lps, c->scal, -1};
        dim3 g2 = ps->down_grid(0);
        g2.z = c->B;
        PSolver<T>::launch_down(g2, c->s, a, 2, ps->wdn[0], false);
    } else {
        OpLevel<T> &C = ps->lev[1];
        UpArgs<T> a{fargs(L), L.invd, L.b, L.x, C.n0, C.n1, C.x, L.t, (T)ps->omega, c->scal, -1};
        const dim3 g2 = ps->up_grid(0);
        PSolver<T>::template launch_up_p<2>(g2, c->s, a, ps->wup[0], Timed{});
    }
    CFD_LAUNCH_CHECK();
    return CFD_OK;
}

// device-side slot state: SlotP and the S_ACTIVE scalar
// fresh: the slot's next statistics become its residual normalisers.  Every level
// of a FAS hierarchy holds the slot alike.
int put_slot(Case *c, int z, bool fresh = false) {
    for (Case *l = c; l; l = l->coarse) {
        if (l != c) l->sph[z].active = c->sph[z].active;
        CFD_HIP(hipMemcpyAsync(l->spd + z, &l->sph[z], sizeof(SlotP), hipMemcpyHostToDevice, l->s));
        double *h = l->hscal + (size_t)z * SS;
        h[S_ACTIVE] = l->sph[z].active;
        h[S_FRESH] = fresh ? 1.0 : 0.0;
        h[S_PMAX] = l->sph[z].pmaxit;
        CFD_HIP(hipMemcpyAsync(l->scal + (size_t)z * SS + S_ACTIVE, h + S_ACTIVE, sizeof(double),
                               hipMemcpyHostToDevice, l->s));
        CFD_HIP(hipMemcpyAsync(l->scal + (size_t)z * SS + S_PMAX, h + S_PMAX, sizeof(double),
                               hipMemcpyHostToDevice, l->s));
        if (fresh)
            CFD_HIP(hipMemcpyAsync(l->scal + (size_t)z * SS + S_FRESH, h + S_FRESH, sizeof(double),
                                   hipMemcpyHostToDevice, l->s));
    }
    CFD_HIP(hipStreamSynchronize(c->s));
    return CFD_OK;
}

void batch_destroy(Case *c) {
    if (!c) return;
    if (c->s) (void)hipStreamSynchronize(c->s);
    batch_destroy(c->coarse);
    drop_graph(c);
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->hscal) (void)hipHostFree(c->hscal);
    if (c->hflag) (void)hipHostFree(c->hflag);
    if (c->s && c->own_stream) (void)hipStreamDestroy(c->s);
    delete c->ps64;
    delete c->ps32;
    delete c;
    (void)hipGetLastError();
}

// Smallest level of a FAS hierarchy (nodes per side): fas_min_n, or with
// fas_min_n <= 0 the smallest 2^k + 1 >= 17 whose cell Reynolds number Re / 2^k is at
// most 32, and >= 33 when Pr > 1 or the Rayleigh number Ra = |Ri| Re^2 Pr exceeds
// 2e5.  A 17 x 17 level misleads the cycle (it stalls away from the solution) for
// thin thermal layers and strong buoyancy: over 40 drawn mixed-convection cases at
// 129 x 129 (Re <= 512), every one with Pr 7 or Ra > 2e5 that converged with a
// 33 x 33 coarsest level stalled with 17 x 17, and every other one converged with
// 17 x 17 in fewer cycles.
int fas_min_eff(const cfd_params &p) {
    if (p.fas_min_n > 0) return p.fas_min_n < 5 ? 5 : p.fas_min_n;
    const double ra = std::fabs(p.Ri) * p.Re * p.Re * p.Pr;
    int m = (p.Pr <= 1.0 && ra <= 2.0e5) ? 16 : 32;
    while (32.0 * m < p.Re && m < (1 << 20)) m *= 2;
    return m + 1;
}

// the coarse level of a FAS hierarchy below an n0 x n1 level, if there is one
bool fas_coarsens(int n0, int n1, const cfd_params &p) {
    const int lo = fas_min_eff(p);
    return p.fas_levels > 1 && (n0 & 1) && (n1 & 1) && (n0 + 1) / 2 >= lo && (n1 + 1) / 2 >= lo;
}

// dry: only count the bytes the batch would hold (c->bytes), touching no device.
// coarse_level: a coarse level of a FAS hierarchy, on the finer level's stream share.
int batch_create(int n0, int n1, double dxi, double deta, int slots, const cfd_params *params, Case **out,
                 bool dry = false, hipStream_t share = nullptr, bool coarse_level = false) {
    if (!out || !params) return CFD_E_ARG;
    if (n0 < 5 || n1 < 5 || !(dxi > 0) || !(deta > 0) || slots < 1 || slots > 65535) return CFD_E_ARG;
    if ((long long)n0 * n1 > MAX_LEVEL_NODES) return CFD_E_ARG;
    if (params->precision != CFD_FP64 && params->precision != CFD_MIXED) return CFD_E_ARG;
    if (!coarse_level && params->p_maxit < 1) return CFD_E_ARG;  // (-1, one V-cycle: coarse levels)
    *out = nullptr;
    Case *c = new Case();
    c->dry = dry;
    c->n0 = n0; c->n1 = n1; c->B = slots; c->N = (size_t)n0 * n1; c->h1 = dxi; c->h2 = deta; c->prm = *params;
    c->mixed = params->precision == CFD_MIXED;
    c->sph.assign(slots, SlotP{});
    c->mass0.assign(slots, 0.0);
    c->energy0.assign(slots, 0.0);
    c->iter.assign(slots, 0);
    if (!dry && hipGetDevice(&c->device) != hipSuccess) { delete c; return CFD_E_HIP; }
    if (share) {
        c->s = share;
        c->own_stream = false;
    } else if (!dry && hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return CFD_E_HIP;
    }
    const bool is_coarse = coarse_level;
    const size_t BN = (size_t)slots * c->N;
    int rc = CFD_OK;
    double **fields[] = {&c->xg, &c->yg, &c->xix, &c->xiy, &c->etx, &c->ety, &c->alph, &c->beta, &c->gamma,
                         &c->ajac, &c->p1, &c->q1, &c->u, &c->v, &c->T, &c->p, &c->d,
                         &c->uh, &c->vh, &c->q, &c->pc};
    for (double **f : fields)
        if (rc == CFD_OK) rc = dalloc(c, f, BN);
    void **pred[] = {&c->Ru, &c->Rv, &c->RT, &c->U, &c->V, &c->jb[0], &c->jb[1], &c->jb[2], &c->jb[3],
                     &c->jb[4], &c->jb[5]};
    for (void **f : pred) {
        if (rc != CFD_OK) break;
        if (c->mixed) rc = dalloc(c, (float **)f, BN);
        else rc = dalloc(c, (double **)f, BN);
    }
    if (c->mixed) {
        for (int k = 0; k < 5 && rc == CFD_OK; k++) rc = dalloc(c, &c->a32[k], BN);
        if (rc == CFD_OK) rc = dalloc(c, &c->ef, BN);
        if (rc == CFD_OK) rc = dalloc(c, &c->r64, BN);
        for (int k = 0; k < 4 && rc == CFD_OK; k++) rc = dalloc(c, &c->f64[k], BN);
    }
    c->scr = (size_t)2 * (n0 + n1) + 8;
    if (rc == CFD_OK) rc = dalloc(c, &c->scratch, (size_t)slots * c->scr);
    // FAS: residual arrays on every level of a hierarchy, sources and the
    // restricted state on the coarse ones
    const bool coarsens = fas_coarsens(n0, n1, *params);
    if (coarsens || is_coarse)
        for (int k = 0; k < 4 && rc == CFD_OK; k++) rc = dalloc(c, &c->fr[k], BN);
    if (is_coarse)
        for (int k = 0; k < 4 && rc == CFD_OK; k++) {
            rc = dalloc(c, &c->src[k], BN);
            if (rc == CFD_OK) rc = dalloc(c, &c->x0[k], BN);
        }
    c->pstride = (nblk(n0, n1) + 255) / 256 * 256;
    if (rc == CFD_OK) rc = dalloc(c, &c->part, (size_t)slots * psz(c));
    if (rc == CFD_OK) rc = dalloc(c, &c->scal, (size_t)slots * SS);
    if (rc == CFD_OK) rc = dalloc(c, &c->spd, (size_t)slots);
    c->spd_nd = c->spd_cr = c->spd;
    cfd_params cp = *params;
    cp.fas_levels = params->fas_levels - 1;
    if (cp.fas_coarse_scheme > cp.wall_upwind) cp.wall_upwind = cp.fas_coarse_scheme;
    if (cp.fas_coarse_pmaxit != 0) cp.p_maxit = cp.fas_coarse_pmaxit < 0 ? -1 : cp.fas_coarse_pmaxit;
    if (dry) {
        if (rc == CFD_OK) rc = build_levels(c);
        if (rc == CFD_OK && is_coarse) rc = alloc_small(c);
        if (rc == CFD_OK && coarsens) {
            rc = batch_create((n0 + 1) / 2, (n1 + 1) / 2, 2 * dxi, 2 * deta, slots, &cp, &c->coarse, true, nullptr,
                              true);
            if (c->coarse) c->bytes += c->coarse->bytes;
        }
        *out = c;
        return rc;
    }
    if (rc == CFD_OK && hipHostMalloc((void **)&c->hscal, (size_t)slots * SS * sizeof(double)) != hipSuccess)
        rc = CFD_E_OOM;
    if (rc == CFD_OK) std::memset(c->hscal, 0, (size_t)slots * SS * sizeof(double));
    if (rc == CFD_OK &&
        (hipHostMalloc((void **)&c->hflag, (size_t)slots * sizeof(unsigned long long),
                       hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
         hipHostGetDevicePointer((void **)&c->dflag, c->hflag, 0) != hipSuccess))
        rc = CFD_E_OOM;
    if (rc == CFD_OK) std::memset(c->hflag, 0, (size_t)slots * sizeof(unsigned long long));
    if (rc == CFD_OK) rc = build_levels(c);
    if (rc == CFD_OK && is_coarse) rc = alloc_small(c);
    if (rc == CFD_OK && coarsens) {
        rc = batch_create((n0 + 1) / 2, (n1 + 1) / 2, 2 * dxi, 2 * deta, slots, &cp, &c->coarse, false, c->s, true);
        if (c->coarse) c->bytes += c->coarse->bytes;
    }
    if (rc == CFD_OK && hipStreamSynchronize(c->s) != hipSuccess) rc = CFD_E_HIP;
    if (rc != CFD_OK) {
        batch_destroy(c);
        return rc;
    }
    *out = c;
    return CFD_OK;
}

// free a dry-run hierarchy
void dry_free(Case *c) {
    if (!c) return;
    dry_free(c->coarse);
    delete c->ps64;
    delete c->ps32;
    delete c;
}

// Slot zl of level l from slot zf of the finer level f: grid arrays injected, the
// state at rest, the FP32 copies of the mixed path, the multigrid geometry.
int inject_level(const Case *f, int zf, Case *l, int zl) {
    const size_t of = (size_t)zf * f->N, ol = (size_t)zl * l->N;
    const double *src[12] = {f->xg, f->yg, f->xix, f->xiy, f->etx, f->ety, f->alph, f->beta, f->gamma,
                             f->ajac, f->p1, f->q1};
    double *dst[12] = {l->xg, l->yg, l->xix, l->xiy, l->etx, l->ety, l->alph, l->beta, l->gamma, l->ajac,
                       l->p1, l->q1};
    const dim3 gr = grid2d(l->n0, l->n1);
    for (int k = 0; k < 12; k++)
        k_inject<<<gr, dim3(BX, BY), 0, l->s>>>(f->n0, f->n1, src[k] + of, l->n0, l->n1, dst[k] + ol);
    CFD_LAUNCH_CHECK();
    double *state[] = {l->u, l->v, l->T, l->p, l->pc};
    for (double *a : state) CFD_HIP(hipMemsetAsync(a + ol, 0, l->N * sizeof(double), l->s));
    if (l->mixed) {
        const double *m[5] = {l->alph, l->beta, l->gamma, l->p1, l->q1};
        const int lb = 256, lg = (int)((l->N + lb - 1) / lb);
        for (int k = 0; k < 5; k++) k_cvt<double, float><<<lg, lb, 0, l->s>>>(l->N, m[k] + ol, l->a32[k] + ol);
        CFD_HIP(hipMemsetAsync(l->ef + ol, 0, l->N * sizeof(float), l->s));
        CFD_LAUNCH_CHECK();
    }
    return inject_slot(l, zl);
}

// slot z of the coarse levels below c: grid arrays injected from the finer level,
// the FP32 copies of the mixed path, the multigrid geometry, the slot's physics
int set_coarse_slots(Case *c, int z) {
    for (Case *f = c, *l = c->coarse; l; f = l, l = l->coarse) {
        RET(inject_level(f, z, l, z));
        l->sph[z] = f->sph[z];
        l->sph[z].scheme = l->prm.wall_upwind;
        l->sph[z].fas_src = 1.0;
        l->sph[z].deep = l->coarse ? 0.0 : 1.0;
        l->iter[z] = 0;
    }
    return CFD_OK;
}

// Slot z of level c from host grid arrays: copies, the FP32 copies of the mixed
// path, the multigrid geometry, the state at rest.
int load_grid(Case *c, int z, const double *x, const double *y, const double *xix, const double *xiy,
              const double *etx, const double *ety, const double *alph, const double *beta, const double *gamma,
              const double *ajac, const double *p1, const double *q1) {
    const size_t N = c->N, B = N * sizeof(double), o = (size_t)z * N;
    const double *src[] = {x, y, xix, xiy, etx, ety, alph, beta, gamma, ajac};
    double *dst[] = {c->xg, c->yg, c->xix, c->xiy, c->etx, c->ety, c->alph, c->beta, c->gamma, c->ajac};
    for (int k = 0; k < 10; k++) CFD_HIP(hipMemcpyAsync(dst[k] + o, src[k], B, hipMemcpyHostToDevice, c->s));
    const double *pq[2] = {p1, q1};
    double *pqd[2] = {c->p1, c->q1};
    for (int k = 0; k < 2; k++) {
        if (pq[k]) CFD_HIP(hipMemcpyAsync(pqd[k] + o, pq[k], B, hipMemcpyHostToDevice, c->s));
        else CFD_HIP(hipMemsetAsync(pqd[k] + o, 0, B, c->s));
    }
    double *state[] = {c->u, c->v, c->T, c->p, c->pc};
    for (double *f : state) CFD_HIP(hipMemsetAsync(f + o, 0, B, c->s));
    if (c->mixed) {
        const double *m[5] = {c->alph, c->beta, c->gamma, c->p1, c->q1};
        const int lb = 256, lg = (int)((N + lb - 1) / lb);
        for (int k = 0; k < 5; k++) k_cvt<double, float><<<lg, lb, 0, c->s>>>(N, m[k] + o, c->a32[k] + o);
        CFD_HIP(hipMemsetAsync(c->ef + o, 0, N * sizeof(float), c->s));
        CFD_LAUNCH_CHECK();
    }
    return inject_slot(c, z);
}

// a case's physics and wall lengths as the slot entry of its top level (inactive)
SlotP slot_physics(int n0, int n1, const double *x, const double *y, const cfd_params *phys, int scheme) {
    SlotP sp{};
    sp.Re = phys->Re; sp.Ri = phys->Ri; sp.Pr = phys->Pr; sp.dt = phys->dt;
    sp.alpha_u = phys->alpha_u; sp.alpha_t = phys->alpha_t; sp.alpha_p = phys->alpha_p;
    sp.lid = phys->lid_u; sp.t_hot = phys->t_hot; sp.t_cold = phys->t_cold;
    sp.dnu = phys->dt / phys->Re;
    sp.dnt = phys->dt / (phys->Re * phys->Pr);
    sp.scheme = scheme;
    sp.fas_src = 0.0;
    sp.deep = 0.0;
    sp.active = 0.0;
    sp.len_hot = sp.len_cold = 0.0;
    for (int i = 0; i < n0 - 1; i++) {
        const size_t a = (size_t)i, b = (size_t)i + 1;
        const size_t a1 = (size_t)(n1 - 1) * n0 + i, b1 = a1 + 1;
        sp.len_hot += std::hypot(x[b] - x[a], y[b] - y[a]);
        sp.len_cold += std::hypot(x[b1] - x[a1], y[b1] - y[a1]);
    }
    return sp;
}

bool grid_args_ok(const double *x, const double *y, const double *xix, const double *xiy, const double *etx,
                  const double *ety, const double *alph, const double *beta, const double *gamma,
                  const double *ajac) {
    return x && y && xix && xiy && etx && ety && alph && beta && gamma && ajac;
}

// Load a case into slot z: grid arrays (host), coarse metrics, physics; fields at
// rest with the wall conditions; the slot becomes active.
int batch_set(Case *c, int z, const double *x, const double *y, const double *xix, const double *xiy,
              const double *etx, const double *ety, const double *alph, const double *beta, const double *gamma,
              const double *ajac, const double *p1, const double *q1, const cfd_params *phys) {
    if (!c || z < 0 || z >= c->B || !grid_args_ok(x, y, xix, xiy, etx, ety, alph, beta, gamma, ajac) || !phys)
        return CFD_E_ARG;
    // a hierarchy built for other physics may have a level too coarse for this case
    if (c->coarse && c->prm.fas_min_n <= 0) {
        cfd_params q = c->prm;
        q.Re = phys->Re;
        q.Ri = phys->Ri;
        q.Pr = phys->Pr;
        const Case *l = c;
        while (l->coarse) l = l->coarse;
        if (std::min(l->n0, l->n1) < fas_min_eff(q)) return CFD_E_ARG;
    }
    // the slot's kernels must not run while its arrays are replaced
    c->sph[z].active = 0.0;
    RET(put_slot(c, z));
    RET(load_grid(c, z, x, y, xix, xiy, etx, ety, alph, beta, gamma, ajac, p1, q1));
    SlotP &sp = c->sph[z];
    sp = slot_physics(c->n0, c->n1, x, y, phys, c->prm.wall_upwind);
    sp.deep = c->coarse ? 0.0 : 1.0;
    RET(set_coarse_slots(c, z));
    sp.active = 1.0;
    RET(put_slot(c, z, true));
    c->iter[z] = 0;
    c->mass0[z] = c->energy0[z] = 0.0;
    RET(wall_bcs(c, z, 1, true));
    CFD_HIP(hipStreamSynchronize(c->s));
    return CFD_OK;
}

int batch_check_slot(const Case *c, int z) { return (!c || z < 0 || z >= c->B) ? CFD_E_ARG : CFD_OK; }

// ---- packs: cases of uneven size sharing every launch (cfd_pack_*)
//
// A pack is a chain of slabs, slab k a batch of slots of ((n0 - 1) >> k) + 1 x
// ((n1 - 1) >> k) + 1 nodes with spacing 2^k (dxi, deta).  A case whose finest grid
// is slab k's size takes one slot of slab k (its top level) and one slot of every
// slab below it down to its coarsest FAS level, so one slab holds the top levels of
// the cases of its size beside the coarse levels of all larger ones, and one FAS
// cycle of the whole pack is one launch sequence: each SIMPLE stage of each slab is
// one launch over all of its slots, whatever case they belong to.  The slots' roles
// are per-slot data (SlotP::scheme / fas_src / deep), the coarse slot of a fine one
// a per-slab map (cmap), and transfers run on the slot tables with only the slots
// taking part active (spd_nd, spd_cr), so every case computes what it computes
// alone in a batch of its own.  Requirements: a case's top level is not a small
// level (k_simple_small), its coarsest level is one.
struct Pack {
    int ns = 0;
    std::vector<Case *> slab;
    std::vector<int> off;                 // case ids of slab k's slots start at off[k]
    hipStream_t s = nullptr;
    cfd_params prm{};                     // the slabs' (single levels)
    cfd_params uprm{};                    // the caller's (a case's own FAS depth)
    std::vector<std::vector<int>> owner;  // case id owning each slab slot, or -1
    std::vector<std::vector<int>> cmap;   // slot of the next slab of the same case, or -1
    std::vector<int *> cmap_d;
    std::vector<int> last;                // per case id: its coarsest slab (-1: no case)
    std::vector<SlotP *> mask_alloc;
    bool tuned = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    size_t bytes = 0;
    bool dry = false;
};

void pack_destroy(Pack *P) {
    if (!P) return;
    if (P->s) (void)hipStreamSynchronize(P->s);
    if (P->gexec) (void)hipGraphExecDestroy(P->gexec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    for (int *m : P->cmap_d)
        if (m) (void)hipFree(m);
    for (SlotP *m : P->mask_alloc)
        if (m) (void)hipFree(m);
    for (Case *c : P->slab) {
        if (!c) continue;
        if (P->dry) dry_free(c);
        else batch_destroy(c);
    }
    if (P->s) (void)hipStreamDestroy(P->s);
    (void)hipGetLastError();
    delete P;
}

int pack_create(int n0, int n1, double dxi, double deta, int nslabs, const int *slots, const cfd_params *params,
                Pack **out, bool dry) {
    if (!out || !params || !slots || nslabs < 2 || nslabs > 16 || n0 < 5 || n1 < 5 || !(dxi > 0) || !(deta > 0))
        return CFD_E_ARG;
    if ((long long)n0 * n1 > MAX_LEVEL_NODES) return CFD_E_ARG;
    *out = nullptr;
    // the chain's shapes first (no device work for a bad one)
    for (int k = 0, a = n0, b = n1; k < nslabs; k++) {
        if (k > 0) {
            if (!(a & 1) || !(b & 1)) return CFD_E_ARG;
            a = (a + 1) / 2;
            b = (b + 1) / 2;
        }
        if (a < 5 || b < 5 || slots[k] < 0 || slots[k] > 65535) return CFD_E_ARG;
    }
    Pack *P = new Pack();
    P->dry = dry;
    P->ns = nslabs;
    P->prm = *params;
    P->uprm = *params;
    P->prm.fas_levels = 1;  // slabs are single levels; the pack links them
    // p' iterations launched per SIMPLE iteration: the larger of the top and coarse
    // levels' caps (slots with a smaller one stop early, pack_add)
    if (params->fas_coarse_pmaxit > params->p_maxit) P->prm.p_maxit = params->fas_coarse_pmaxit;
    P->slab.assign(nslabs, nullptr);
    P->off.assign(nslabs + 1, 0);
    P->owner.resize(nslabs);
    P->cmap.resize(nslabs);
    P->cmap_d.assign(nslabs, nullptr);
    int rc = CFD_OK;
    if (!dry && hipStreamCreateWithFlags(&P->s, hipStreamNonBlocking) != hipSuccess) rc = CFD_E_HIP;
    int a = n0, b = n1;
    double h1 = dxi, h2 = deta;
    for (int k = 0; k < nslabs && rc == CFD_OK; k++) {
        if (k > 0) {
            if (!(a & 1) || !(b & 1)) { rc = CFD_E_ARG; break; }
            a = (a + 1) / 2; b = (b + 1) / 2; h1 *= 2.0; h2 *= 2.0;
        }
        if (a < 5 || b < 5 || slots[k] < 0 || slots[k] > 65535) { rc = CFD_E_ARG; break; }
        P->off[k + 1] = P->off[k] + slots[k];
        P->owner[k].assign(slots[k], -1);
        P->cmap[k].assign(slots[k], -1);
        if (slots[k] == 0) continue;
        rc = batch_create(a, b, h1, h2, slots[k], &P->prm, &P->slab[k], dry, dry ? nullptr : P->s, true);
        if (rc != CFD_OK) break;
        Case *c = P->slab[k];
        if (dry) {
            c->bytes += 2 * (size_t)slots[k] * sizeof(SlotP) + (size_t)slots[k] * sizeof(int);
        } else {
            for (int m = 0; m < 2 && rc == CFD_OK; m++) {
                SlotP *d = nullptr;
                if (hipMalloc(&d, (size_t)slots[k] * sizeof(SlotP)) != hipSuccess ||
                    hipMemsetAsync(d, 0, (size_t)slots[k] * sizeof(SlotP), P->s) != hipSuccess)
                    rc = CFD_E_OOM;
                P->mask_alloc.push_back(d);
                (m == 0 ? c->spd_nd : c->spd_cr) = d;
            }
            if (rc == CFD_OK && hipMalloc(&P->cmap_d[k], (size_t)slots[k] * sizeof(int)) != hipSuccess) rc = CFD_E_OOM;
            if (rc == CFD_OK)
                rc = hipMemcpyAsync(P->cmap_d[k], P->cmap[k].data(), (size_t)slots[k] * sizeof(int),
                                    hipMemcpyHostToDevice, P->s) == hipSuccess ? CFD_OK : CFD_E_HIP;
            c->bytes += 2 * (size_t)slots[k] * sizeof(SlotP) + (size_t)slots[k] * sizeof(int);
        }
        P->bytes += c->bytes;
    }
    P->last.assign(P->off[nslabs], -1);
    if (rc == CFD_OK && !dry && hipStreamSynchronize(P->s) != hipSuccess) rc = CFD_E_HIP;
    if (rc != CFD_OK) {
        pack_destroy(P);
        return rc;
    }
    *out = P;
    return CFD_OK;
}

// slot z of slab k on the device: its entry, its masked entries (with a coarser
// level / as a coarse level), S_ACTIVE (and S_FRESH), its coarse-slot map entry
int pack_put(Pack *P, int k, int z, bool fresh) {
    Case *c = P->slab[k];
    RET(put_slot(c, z, fresh));
    SlotP nd = c->sph[z], cr = c->sph[z];
    nd.active = (nd.active != 0.0 && nd.deep == 0.0) ? 1.0 : 0.0;
    cr.active = (cr.active != 0.0 && cr.fas_src != 0.0) ? 1.0 : 0.0;
    CFD_HIP(hipMemcpyAsync(c->spd_nd + z, &nd, sizeof(SlotP), hipMemcpyHostToDevice, P->s));
    CFD_HIP(hipMemcpyAsync(c->spd_cr + z, &cr, sizeof(SlotP), hipMemcpyHostToDevice, P->s));
    CFD_HIP(hipMemcpyAsync(P->cmap_d[k] + z, &P->cmap[k][z], sizeof(int), hipMemcpyHostToDevice, P->s));
    CFD_HIP(hipStreamSynchronize(P->s));
    return CFD_OK;
}

int pack_release(Pack *P, int id) {
    if (!P || id < 0 || id >= P->off[P->ns] || P->last[id] < 0) return CFD_E_ARG;
    int k = 0;
    while (P->off[k + 1] <= id) k++;
    int z = id - P->off[k];
    const int kd = P->last[id];
    for (int l = k; l <= kd; l++) {
        const int zn = P->cmap[l][z];
        P->slab[l]->sph[z].active = 0.0;
        P->owner[l][z] = -1;
        P->cmap[l][z] = -1;
        RET(pack_put(P, l, z, false));
        z = zn;
    }
    P->last[id] = -1;
    return CFD_OK;
}

// A case of n0 x n1 nodes into the first free slots of its slabs; *id: the case
// id (the top slot's).  CFD_E_FULL when a slab of its chain has no free slot.
int pack_add(Pack *P, int n0, int n1, const double *x, const double *y, const double *xix, const double *xiy,
             const double *etx, const double *ety, const double *alph, const double *beta, const double *gamma,
             const double *ajac, const double *p1, const double *q1, const cfd_params *phys, int *id) {
    if (!P || !id || !phys || !grid_args_ok(x, y, xix, xiy, etx, ety, alph, beta, gamma, ajac)) return CFD_E_ARG;
    int k = -1;
    for (int l = 0; l < P->ns; l++)
        if (P->slab[l] && P->slab[l]->n0 == n0 && P->slab[l]->n1 == n1) k = l;
    if (k < 0) return CFD_E_ARG;
    // the case's own hierarchy (as a batch of its own would have it): its coarsest level
    cfd_params q = P->uprm;
    q.Re = phys->Re; q.Ri = phys->Ri; q.Pr = phys->Pr;
    const int coarsest = cfd_fas_coarsest(n0, n1, &q);
    int kd = -1;
    for (int l = k + 1; l < P->ns; l++)
        if (P->slab[l] && std::min(P->slab[l]->n0, P->slab[l]->n1) == coarsest) kd = l;
    if (kd < 0) return CFD_E_ARG;
    for (int l = k; l <= kd; l++)
        if (!P->slab[l]) return CFD_E_ARG;
    if (small_fits(P->slab[k]) || !small_fits(P->slab[kd])) return CFD_E_ARG;
    std::vector<int> zs(kd + 1, -1);
    for (int l = k; l <= kd; l++) {
        for (int z = 0; z < P->slab[l]->B; z++)
            if (P->owner[l][z] < 0) { zs[l] = z; break; }
        if (zs[l] < 0) return CFD_E_FULL;
    }
    const int cid = P->off[k] + zs[k];
    Case *top = P->slab[k];
    RET(load_grid(top, zs[k], x, y, xix, xiy, etx, ety, alph, beta, gamma, ajac, p1, q1));
    const SlotP base = slot_physics(n0, n1, x, y, phys, P->prm.wall_upwind);
    const int cscheme = std::max(P->prm.wall_upwind, P->prm.fas_coarse_scheme);
    // a case's levels cap their p' solves as a batch of its own does
    const int launched = P->prm.p_maxit, tpmax = P->uprm.p_maxit;
    const int cfp = P->uprm.fas_coarse_pmaxit;
    const int cpmax = cfp != 0 ? (cfp < 0 ? -1 : cfp) : P->uprm.p_maxit;  // -1: one V-cycle
    for (int l = k; l <= kd; l++) {
        Case *c = P->slab[l];
        const int z = zs[l];
        if (l > k) RET(inject_level(P->slab[l - 1], zs[l - 1], c, z));
        SlotP &sp = c->sph[z];
        sp = base;
        sp.pmaxit = tpmax < launched ? tpmax : 0.0;
        if (l > k) {
            sp.scheme = cscheme;
            sp.fas_src = 1.0;
            sp.pmaxit = cpmax < launched ? cpmax : 0.0;
        }
        sp.deep = l == kd ? 1.0 : 0.0;
        sp.active = 1.0;
        P->owner[l][z] = cid;
        P->cmap[l][z] = l < kd ? zs[l + 1] : -1;
        c->iter[z] = 0;
        c->mass0[z] = c->energy0[z] = 0.0;
    }
    for (int l = kd; l >= k; l--) RET(pack_put(P, l, zs[l], l == k));
    RET(wall_bcs(top, zs[k], 1, true));
    CFD_HIP(hipStreamSynchronize(P->s));
    P->last[cid] = kd;
    *id = cid;
    return CFD_OK;
}

// one FAS cycle of slab k and the slabs below it: the pre-smoothing (on a small
// slab, a slot at its coarsest level runs its fas_coarse iterations there), the
// restriction of the slots that have a coarser level, the slabs below, the
// correction, the post-smoothing
int pack_cycle(Pack *P, int k) {
    Case *c = P->slab[k];
    Case *cc = k + 1 < P->ns ? P->slab[k + 1] : nullptr;
    const cfd_params &pr = P->prm;
    const bool small = small_fits(c);
    if (small) RET(simple_small(c, pr.fas_pre, pr.fas_coarse));
    else
        for (int it = 0; it < pr.fas_pre; it++) RET(simple_body(c, true, true));
    if (cc) {
        RET(fas_restrict(c, cc, c->spd_nd, cc->spd_cr, P->cmap_d[k]));
        RET(pack_cycle(P, k + 1));
        RET(fas_prolong(c, cc, c->spd_nd, P->cmap_d[k]));
    }
    if (small) return simple_small(c, pr.fas_post, 0);
    for (int it = 0; it < pr.fas_post; it++) RET(simple_body(c, true, true));
    return CFD_OK;
}

int pack_launch(Pack *P) {
    if (!P || P->dry) return CFD_E_ARG;
    int k0 = 0;
    while (k0 < P->ns && !P->slab[k0]) k0++;
    if (k0 == P->ns) return CFD_OK;
    for (int k = k0; k < P->ns; k++)
        if (!P->slab[k]) return CFD_E_ARG;  // the chain below the largest slab is complete
    if (!P->tuned) {
        for (int k = k0; k < P->ns; k++) {
            RET(tune_mg(P->slab[k]));
            RET(tune_res(P->slab[k]));
        }
        P->tuned = true;
    }
    if (!P->prm.graph) return pack_cycle(P, k0);
    if (!P->gexec) {
        CFD_HIP(hipStreamSynchronize(P->s));
        CFD_HIP(hipStreamBeginCapture(P->s, hipStreamCaptureModeThreadLocal));
        for (int k = k0; k < P->ns; k++) P->slab[k]->capturing = true;
        const int rc = pack_cycle(P, k0);
        for (int k = k0; k < P->ns; k++) P->slab[k]->capturing = false;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(P->s, &g);
        if (rc != CFD_OK || e != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            fprintf(stderr, "cfd_gpu: pack cycle capture failed (rc %d, %s)\n", rc, hipGetErrorString(e));
            return rc != CFD_OK ? rc : CFD_E_HIP;
        }
        P->graph = g;
        CFD_HIP(hipGraphInstantiate(&P->gexec, P->graph, nullptr, nullptr, 0));
    }
    CFD_HIP(hipGraphLaunch(P->gexec, P->s));
    return CFD_OK;
}

// statistics of every case (st: one entry per case id, zeros where there is none)
int pack_finish(Pack *P, cfd_stats *st) {
    if (!P || P->dry) return CFD_E_ARG;
    CFD_HIP(hipStreamSynchronize(P->s));
    for (int k = 0; k < P->ns; k++) {
        Case *c = P->slab[k];
        if (!c) continue;
        bool tops = false;
        for (int z = 0; z < c->B; z++) tops |= P->owner[k][z] == P->off[k] + z;
        if (tops) RET(sync_scal(c));
    }
    for (int k = 0; k < P->ns; k++) {
        Case *c = P->slab[k];
        for (int z = 0; c && z < c->B; z++) {
            const int id = P->off[k] + z;
            if (P->owner[k][z] != id) {
                if (st) st[id] = cfd_stats{};
                continue;
            }
            const double *h = c->hscal + (size_t)z * SS;
            c->iter[z]++;
            if (c->iter[z] == 1) {
                c->mass0[z] = h[S_MASS0];
                c->energy0[z] = h[S_ENERGY0];
            }
            if (!st) continue;
            cfd_stats &o = st[id];
            psolve_result(c, z, &o.p_iters, &o.p_relres);
            o.mass_l1 = h[S_MASS];
            o.energy_l1 = h[S_ENERGY];
            o.mom_l1 = h[S_MOM];
            o.dmax = h[S_DMAX];
            o.nu_hot = h[S_NUH];
            o.nu_cold = h[S_NUC];
            o.mass_res = c->mass0[z] > 0 ? h[S_MASS] / c->mass0[z] : 0.0;
            o.energy_res = c->energy0[z] > 0 ? h[S_ENERGY] / c->energy0[z] : 0.0;
            o.iter = c->iter[z];
        }
    }
    return CFD_OK;
}

int pack_case(const Pack *P, int id, Case **c, int *z) {
    if (!P || id < 0 || id >= P->off[P->ns] || P->last[id] < 0) return CFD_E_ARG;
    int k = 0;
    while (P->off[k + 1] <= id) k++;
    *c = P->slab[k];
    *z = id - P->off[k];
    return CFD_OK;
}

}  // namespace

// ==========================================================================
// C ABI

extern "C" {

void cfd_params_default(cfd_params *p) {
    if (!p) return;
    p->Re = 100.0; p->Ri = 1.0; p->Pr = 0.71; p->dt = 1.0e4;
    p->alpha_u = 0.7; p->alpha_p = 0.3; p->alpha_t = 0.85;
    p->lid_u = 1.0; p->t_hot = 1.0; p->t_cold = 0.0;
    p->mom_sweeps = 8; p->p_rtol = 1.0e-2; p->p_maxit = 2;
    p->mg_levels = 12; p->mg_pre = 2; p->mg_post = 2; p->mg_coarse = 4; p->mg_omega = 0.8;
    p->precision = CFD_FP64;
    p->graph = 0;
    p->mg_tune = 1;
    p->wall_upwind = 1; p->hyb_blend = 1.0;
    p->fas_levels = 8; p->fas_pre = 2; p->fas_post = 2; p->fas_coarse = 40; p->fas_min_n = 0;
    p->fas_coarse_scheme = 2;
    p->fas_small = 1;
    p->fas_coarse_pmaxit = -1;
}

int cfd_fas_coarsest(int n0, int n1, const cfd_params *p) {
    if (!p || n0 < 5 || n1 < 5) return CFD_E_ARG;
    int a = n0, b = n1;
    for (int l = 1; l < p->fas_levels && fas_coarsens(a, b, *p); l++) {
        a = (a + 1) / 2;
        b = (b + 1) / 2;
    }
    return std::min(a, b);
}

// ---- batches

int cfd_batch_create(int n0, int n1, double dxi, double deta, int slots, const cfd_params *params,
                     cfd_batch *out) {
    if (!out) return CFD_E_ARG;
    Case *c = nullptr;
    RET(batch_create(n0, n1, dxi, deta, slots, params, &c));
    *out = reinterpret_cast<cfd_batch>(c);
    return CFD_OK;
}

int cfd_batch_destroy(cfd_batch h) {
    batch_destroy(reinterpret_cast<Case *>(h));
    return CFD_OK;
}

int cfd_batch_set_case(cfd_batch h, int slot, const double *x, const double *y, const double *xix,
                       const double *xiy, const double *etx, const double *ety, const double *alph,
                       const double *beta, const double *gamma, const double *ajac, const double *p1,
                       const double *q1, const cfd_params *params) {
    return batch_set(reinterpret_cast<Case *>(h), slot, x, y, xix, xiy, etx, ety, alph, beta, gamma, ajac, p1, q1,
                     params);
}

int cfd_batch_release(cfd_batch h, int slot) {
    Case *c = reinterpret_cast<Case *>(h);
    RET(batch_check_slot(c, slot));
    c->sph[slot].active = 0.0;
    return put_slot(c, slot);
}

// a slot whose residuals went non-finite shows it in its statistics; the batch call succeeds
int cfd_batch_outer(cfd_batch h, int n, cfd_stats *st) {
    Case *c = reinterpret_cast<Case *>(h);
    if (!c || n < 0) return CFD_E_ARG;
    for (int k = 0; k < n; k++) {
        const int rc = outer_once(c, st);
        if (rc != CFD_OK && rc != CFD_E_NAN) return rc;
    }
    return CFD_OK;
}

int cfd_batch_launch(cfd_batch h) {
    Case *c = reinterpret_cast<Case *>(h);
    if (!c) return CFD_E_ARG;
    return outer_launch(c);
}

int cfd_batch_finish(cfd_batch h, cfd_stats *st) {
    Case *c = reinterpret_cast<Case *>(h);
    if (!c) return CFD_E_ARG;
    RET(sync_scal(c));
    const int rc = outer_stats(c, st);
    return rc == CFD_E_NAN ? CFD_OK : rc;
}

int cfd_batch_get_fields(cfd_batch h, int slot, double *u, double *v, double *T, double *p) {
    Case *c = reinterpret_cast<Case *>(h);
    RET(batch_check_slot(c, slot));
    const size_t B = c->N * sizeof(double), o = (size_t)slot * c->N;
    if (u) CFD_HIP(hipMemcpyAsync(u, c->u + o, B, hipMemcpyDeviceToHost, c->s));
    if (v) CFD_HIP(hipMemcpyAsync(v, c->v + o, B, hipMemcpyDeviceToHost, c->s));
    if (T) CFD_HIP(hipMemcpyAsync(T, c->T + o, B, hipMemcpyDeviceToHost, c->s));
    if (p) CFD_HIP(hipMemcpyAsync(p, c->p + o, B, hipMemcpyDeviceToHost, c->s));
    CFD_HIP(hipStreamSynchronize(c->s));
    return CFD_OK;
}

int cfd_batch_set_fields(cfd_batch h, int slot, const double *u, const double *v, const double *T,
                         const double *p) {
    Case *c = reinterpret_cast<Case *>(h);
    RET(batch_check_slot(c, slot));
    if (!u || !v || !T || !p) return CFD_E_ARG;
    const size_t B = c->N * sizeof(double), o = (size_t)slot * c->N;
    CFD_HIP(hipMemcpyAsync(c->u + o, u, B, hipMemcpyHostToDevice, c->s));
    CFD_HIP(hipMemcpyAsync(c->v + o, v, B, hipMemcpyHostToDevice, c->s));
    CFD_HIP(hipMemcpyAsync(c->T + o, T, B, hipMemcpyHostToDevice, c->s));
    CFD_HIP(hipMemcpyAsync(c->p + o, p, B, hipMemcpyHostToDevice, c->s));
    CFD_HIP(hipStreamSynchronize(c->s));
    c->iter[slot] = 0;
    c->mass0[slot] = c->energy0[slot] = 0.0;
    return put_slot(c, slot, true);
}

int cfd_batch_bytes(cfd_batch h, size_t *bytes) {
    Case *c = reinterpret_cast<Case *>(h);
    if (!c || !bytes) return CFD_E_ARG;
    *bytes = c->bytes;
    return CFD_OK;
}

int cfd_batch_bytes_needed(int n0, int n1, double dxi, double deta, int slots, const cfd_params *params,
                           size_t *bytes) {
    if (!bytes) return CFD_E_ARG;
    Case *c = nullptr;
    const int rc = batch_create(n0, n1, dxi, deta, slots, params, &c, true);
    if (c) {
        *bytes = c->bytes;
        dry_free(c);
    }
    return rc;
}

// ---- single cases: a batch of one slot

// ---- packs

int cfd_pack_create(int n0, int n1, double dxi, double deta, int nslabs, const int *slots, const cfd_params *params,
                    cfd_pack *out) {
    if (!out) return CFD_E_ARG;
    Pack *P = nullptr;
    RET(pack_create(n0, n1, dxi, deta, nslabs, slots, params, &P, false));
    *out = reinterpret_cast<cfd_pack>(P);
    return CFD_OK;
}

int cfd_pack_destroy(cfd_pack h) {
    pack_destroy(reinterpret_cast<Pack *>(h));
    return CFD_OK;
}

int cfd_pack_bytes_needed(int n0, int n1, double dxi, double deta, int nslabs, const int *slots,
                          const cfd_params *params, size_t *bytes) {
    if (!bytes) return CFD_E_ARG;
    Pack *P = nullptr;
    const int rc = pack_create(n0, n1, dxi, deta, nslabs, slots, params, &P, true);
    if (rc == CFD_OK) *bytes = P->bytes;
    pack_destroy(P);
    return rc;
}

int cfd_pack_bytes(cfd_pack h, size_t *bytes) {
    const Pack *P = reinterpret_cast<const Pack *>(h);
    if (!P || !bytes) return CFD_E_ARG;
    *bytes = P->bytes;
    return CFD_OK;
}

int cfd_pack_capacity(cfd_pack h, int *ids) {
    const Pack *P = reinterpret_cast<const Pack *>(h);
    if (!P || !ids) return CFD_E_ARG;
    *ids = P->off[P->ns];
    return CFD_OK;
}

int cfd_pack_add_case(cfd_pack h, int n0, int n1, const double *x, const double *y, const double *xix,
                      const double *xiy, const double *etx, const double *ety, const double *alph,
                      const double *beta, const double *gamma, const double *ajac, const double *p1,
                      const double *q1, const cfd_params *params, int *id) {
    return pack_add(reinterpret_cast<Pack *>(h), n0, n1, x, y, xix, xiy, etx, ety, alph, beta, gamma, ajac, p1, q1,
                    params, id);
}

int cfd_pack_release(cfd_pack h, int id) { return pack_release(reinterpret_cast<Pack *>(h), id); }

int cfd_pack_launch(cfd_pack h) { return pack_launch(reinterpret_cast<Pack *>(h)); }

int cfd_pack_finish(cfd_pack h, cfd_stats *st) { return pack_finish(reinterpret_cast<Pack *>(h), st); }

int cfd_pack_outer(cfd_pack h, int n, cfd_stats *st) {
    Pack *P = reinterpret_cast<Pack *>(h);
    if (!P || n < 0) return CFD_E_ARG;
    for (int k = 0; k < n; k++) {
        RET(pack_launch(P));
        RET(pack_finish(P, st));
    }
    return CFD_OK;
}

int cfd_pack_get_fields(cfd_pack h, int id, double *u, double *v, double *T, double *p) {
    Case *c = nullptr;
    int z = 0;
    RET(pack_case(reinterpret_cast<const Pack *>(h), id, &c, &z));
    return cfd_batch_get_fields(reinterpret_cast<cfd_batch>(c), z, u, v, T, p);
}

int cfd_case_create(int n0, int n1, double dxi, double deta, const double *x, const double *y,
                    const double *xix, const double *xiy, const double *etx, const double *ety,
                    const double *alph, const double *beta, const double *gamma, const double *ajac,
                    const double *p1, const double *q1, const cfd_params *params, cfd_case *out) {
    if (!out || !params || !x || !y || !xix || !xiy || !etx || !ety || !alph || !beta || !gamma || !ajac)
        return CFD_E_ARG;
    *out = nullptr;
    Case *c = nullptr;
    RET(batch_create(n0, n1, dxi, deta, 1, params, &c));
    const int rc = batch_set(c, 0, x, y, xix, xiy, etx, ety, alph, beta, gamma, ajac, p1, q1, params);
    if (rc != CFD_OK) {
        batch_destroy(c);
        return rc;
    }
    *out = reinterpret_cast<cfd_case>(c);
    return CFD_OK;
}

int cfd_case_destroy(cfd_case h) {
    batch_destroy(reinterpret_cast<Case *>(h));
    return CFD_OK;
}

int cfd_case_set_fields(cfd_case h, const double *u, const double *v, const double *T, const double *p) {
    return cfd_batch_set_fields(reinterpret_cast<cfd_batch>(h), 0, u, v, T, p);
}

int cfd_case_get_fields(cfd_case h, double *u, double *v, double *T, double *p) {
    return cfd_batch_get_fields(reinterpret_cast<cfd_batch>(h), 0, u, v, T, p);
}

int cfd_case_apply_bcs(cfd_case h) {
    Case *c = reinterpret_cast<Case *>(h);
    if (!c) return CFD_E_ARG;
    RET(wall_bcs(c, 0, c->B, true));
    CFD_HIP(hipStreamSynchronize(c->s));
    return CFD_OK;
}

// n outer iterations.  The scalars are read back after the last one and after a
// case's first (which sets the residual normalisers); the iterations in between
// are enqueued back to back (the p' solve's flag waits still pace the host), so
// the next iteration's launches queue behind the previous one's tail.  A residual
// that turns non-finite stays so, and is reported after the last iteration.
int cfd_case_outer(cfd_case h, int n, cfd_stats *st) {
    Case *c = reinterpret_cast<Case *>(h);
    if (!c || n < 0) return CFD_E_ARG;
    cfd_stats tmp{};
    for (int k = 0; k < n; k++) {
        bool first = false;
        for (int z = 0; z < c->B; z++) first |= c->sph[z].active && c->iter[z] == 0;
        if (k == n - 1 || first || c->prof) {
            RET(outer_once(c, st ? st : &tmp));
            continue;
        }
        RET(outer_launch(c));
        for (int z = 0; z < c->B; z++)
            if (c->sph[z].active) c->iter[z]++;
    }
    return CFD_OK;
}

int cfd_case_solve(cfd_case h, double tol, int max_outer, cfd_stats *st) {
    Case *c = reinterpret_cast<Case *>(h);
    if (!c || max_outer < 1) return CFD_E_ARG;
    cfd_stats s{};
    for (int k = 0; k < max_outer; k++) {
        RET(outer_once(c, &s));
        if (s.mass_res <= tol && s.energy_res <= tol) {
            if (st) *st = s;
            return CFD_OK;
        }
    }
    if (st) *st = s;
    return CFD_E_NOTCONV;
}

int cfd_case_bytes(cfd_case h, size_t *bytes) { return cfd_batch_bytes(reinterpret_cast<cfd_batch>(h), bytes); }

// residual arrays are returned in FP64 whatever the predictor precision
int cfd_case_residuals(cfd_case h, double *Ru, double *Rv, double *RT, double *d) {
    Case *c = reinterpret_cast<Case *>(h);
    if (!c) return CFD_E_ARG;
    RET(launch_residual(c));
    {
        const int m = std::max(c->n0, c->n1);
        k_d_border<<<dim3((m + 255) / 256, 1, c->B), 256, 0, c->s>>>(c->n0, c->n1, c->d);
        CFD_LAUNCH_CHECK();
    }
    const size_t N = c->N, B = N * sizeof(double);
    void *src[3] = {c->Ru, c->Rv, c->RT};
    double *dst[3] = {Ru, Rv, RT};
    for (int k = 0; k < 3; k++) {
        if (!dst[k]) continue;
        if (c->mixed) {
            const int lb = 256, lg = (int)((N + lb - 1) / lb);
            k_cvt<float, double><<<lg, lb, 0, c->s>>>(N, (const float *)src[k], c->uh);
            CFD_LAUNCH_CHECK();
            CFD_HIP(hipMemcpyAsync(dst[k], c->uh, B, hipMemcpyDeviceToHost, c->s));
            CFD_HIP(hipStreamSynchronize(c->s));
        } else {
            CFD_HIP(hipMemcpyAsync(dst[k], src[k], B, hipMemcpyDeviceToHost, c->s));
        }
    }
    if (d) CFD_HIP(hipMemcpyAsync(d, c->d, B, hipMemcpyDeviceToHost, c->s));
    CFD_HIP(hipStreamSynchronize(c->s));
    return CFD_OK;
}

int cfd_case_p_apply(cfd_case h, const double *d, const double *x, double *y) {
    Case *c = reinterpret_cast<Case *>(h);
    if (!c || !d || !x || !y) return CFD_E_ARG;
    const size_t B = c->N * sizeof(double);
    CFD_HIP(hipMemcpyAsync(c->d, d, B, hipMemcpyHostToDevice, c->s));
    CFD_HIP(hipMemcpyAsync(c->uh, x, B, hipMemcpyHostToDevice, c->s));
    k_mapply<<<grid2d(c->n0, c->n1), dim3(BX, BY), 0, c->s>>>(margs(c), c->uh, nullptr, c->vh);
    CFD_LAUNCH_CHECK();
    CFD_HIP(hipMemcpyAsync(y, c->vh, B, hipMemcpyDeviceToHost, c->s));
    CFD_HIP(hipStreamSynchronize(c->s));
    return CFD_OK;
}

int cfd_case_p_solve(cfd_case h, const double *d, const double *q, double *x, double rtol, int maxit,
                     int *iters, double *relres) {
    Case *c = reinterpret_cast<Case *>(h);
    if (!c) return CFD_E_ARG;
    const size_t B = c->N * sizeof(double);
    if (d) CFD_HIP(hipMemcpyAsync(c->d, d, B, hipMemcpyHostToDevice, c->s));
    if (q) CFD_HIP(hipMemcpyAsync(c->q, q, B, hipMemcpyHostToDevice, c->s));
    RET(mg_setup(c));
    RET(tune_mg(c));
    RET(psolve(c, rtol, maxit, false));
    RET(sync_scal(c));
    int its = 0;
    double rr = 0.0;
    psolve_result(c, 0, &its, &rr);
    if (x) CFD_HIP(hipMemcpyAsync(x, c->pc, B, hipMemcpyDeviceToHost, c->s));
    CFD_HIP(hipStreamSynchronize(c->s));
    if (iters) *iters = its;
    if (relres) *relres = rr;
    return CFD_OK;
}

int cfd_case_profile(cfd_case h, int enable) {
    Case *c = reinterpret_cast<Case *>(h);
    if (!c) return CFD_E_ARG;
    CFD_HIP(hipStreamSynchronize(c->s));
    if (enable && c->ev.empty()) {
        c->ev.resize(PROF_EVENTS);
        for (auto &e : c->ev) CFD_HIP(hipEventCreate(&e));
    }
    if (enable != 2) c->prof = enable != 0;
    c->ev_used = 0;
    c->ev_kind.clear();
    for (int k = 0; k < CFD_PROF_KINDS; k++) {
        c->prof_ms[k] = 0.0;
        c->prof_n[k] = 0;
    }
    return CFD_OK;
}

int cfd_case_profile_read(cfd_case h, double *ms_total, long *launches, int nkinds) {
    Case *c = reinterpret_cast<Case *>(h);
    if (!c || !ms_total || !launches || nkinds < 1) return CFD_E_ARG;
    CFD_HIP(hipStreamSynchronize(c->s));
    prof_drain(c);
    for (int k = 0; k < nkinds; k++) {
        ms_total[k] = k < CFD_PROF_KINDS ? c->prof_ms[k] : 0.0;
        launches[k] = k < CFD_PROF_KINDS ? c->prof_n[k] : 0;
    }
    return CFD_OK;
}

int cfd_case_time_kernel(cfd_case h, int kernel, int reps, double *ms_per_launch) {
    Case *c = reinterpret_cast<Case *>(h);
    if (!c || reps < 1 || !ms_per_launch || kernel < 0 || kernel > 5) return CFD_E_ARG;
    hipEvent_t e0, e1;
    CFD_HIP(hipEventCreate(&e0));
    CFD_HIP(hipEventCreate(&e1));
    CFD_HIP(hipEventRecord(e0, c->s));
    for (int k = 0; k < reps; k++) {
        switch (kernel) {
        case 0:
        case 4:
        case 5:
            RET(c->mixed ? time_solver_kernel<float>(c, c->ps32, kernel)
                         : time_solver_kernel<double>(c, c->ps64, kernel));
            break;
        case 1:
            if (c->mixed) {
                void *a, *b, *t;
                (void)predictor_t<float>(c, &a, &b, &t);
            } else {
                void *a, *b, *t;
                (void)predictor_t<double>(c, &a, &b, &t);
            }
            break;
        case 2:
            (void)launch_residual(c);
            break;
        case 3:
            (void)launch_face_div(c, c->jb[0], c->jb[1]);
            break;
        }
    }
    CFD_LAUNCH_CHECK();
    CFD_HIP(hipEventRecord(e1, c->s));
    CFD_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CFD_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_launch = (double)ms / reps;
    return CFD_OK;
}

}  // extern "C"
