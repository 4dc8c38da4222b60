// Shared device helpers for the gfx950 kernels of the SIMPLE / pressure-correction path.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstddef>
#include <cstdio>

#include "../../include/cfd_gpu.h"

#define CFD_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            fprintf(stderr, "cfd_gpu: %s failed at %s:%d: %s\n", #call, __FILE__, __LINE__, \
                    hipGetErrorString(e_));                                                \
            return CFD_E_HIP;                                                              \
        }                                                                                  \
    } while (0)

#define CFD_LAUNCH_CHECK()                                                                 \
    do {                                                                                   \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess) {                                                            \
            fprintf(stderr, "cfd_gpu: launch failed at %s:%d: %s\n", __FILE__, __LINE__,   \
                    hipGetErrorString(e_));                                                \
            return CFD_E_HIP;                                                              \
        }                                                                                  \
    } while (0)

#define CFD_RET(x)                   \
    do {                             \
        int r_ = (x);                \
        if (r_ != CFD_OK) return r_; \
    } while (0)

namespace cfd {

// A launch that a pair of HIP events may time.  With events, hipExtLaunchKernelGGL
// binds both to the kernel's own dispatch, so their elapsed time is the kernel's
// duration without the gap to the launches around it; without events it is a plain
// launch (the form a graph capture records).
struct Timed {
    hipEvent_t start = nullptr, stop = nullptr;
};
template <typename... P, typename... A>
inline void launch_timed(const Timed &t, void (*k)(P...), dim3 g, dim3 b, hipStream_t s, A... a) {
    if (t.start) hipExtLaunchKernelGGL(k, g, b, 0, s, t.start, t.stop, 0, a...);
    else k<<<g, b, 0, s>>>(a...);
}

// Tile (x, y) of this block, numbered so that each XCD works on one contiguous
// band of tiles: the dispatcher hands the blocks of a launch to the 8 XCDs in
// turn (block b to XCD b % 8), so renumbering b to xcd * (its share) + b / 8
// keeps neighbouring tiles, which read each other's halo rows, in one L2.  A
// bijection on the tiles of one slot (blockIdx.z).
__device__ __forceinline__ int xcd_order(int b, int nb) {
    constexpr int NX = 8;
    const int xcd = b % NX, q = nb / NX, rm = nb % NX;
    return xcd * q + min(xcd, rm) + b / NX;
}
__device__ __forceinline__ int2 xcd_tile() {
    const int gx = (int)gridDim.x;
    const int t = xcd_order((int)blockIdx.x + gx * (int)blockIdx.y, gx * (int)gridDim.y);
    return make_int2(t % gx, t / gx);
}

// 2-D launch shape: 64 lanes along xi (one wavefront per row segment), 4 rows per block.
constexpr int BX = 64;
constexpr int BY = 4;

inline dim3 grid2d(int n0, int n1) { return dim3((n0 + BX - 1) / BX, (n1 + BY - 1) / BY); }

// Tiles of width w covering nodes 0 .. n - 2 of one direction, for kernels that
// write interior nodes only (or whose wall nodes the tiles next to them set).  On
// a 2^k + 1 grid this is (n - 1) / w: the tile column (row) that ceil(n / w) adds
// would hold only the east (north) wall, and would add a partial round of blocks
// to the launch.
inline int tiles_inner(int n, int w) { return (n - 1 + w - 1) / w; }

// v clamped into [lo, hi] (lo <= hi): one min and one max
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return max(lo, min(v, hi)); }

// Wavefront (64 lanes) reductions of a double with DPP lane moves instead of LDS
// permutes (a dependent ds_bpermute costs an LDS round trip; a DPP move is a VALU
// operand modifier).  The double moves as two 32-bit halves.  A lane whose source
// is outside its row keeps `old`: zero for the sum, its own value for the maximum.
// Steps: swap neighbours and pairs within each quad (every lane then holds its
// quad's total), row_shr 4 and 8 (lane 15 of each 16-lane row holds the row's
// total), then the gfx9 row broadcasts of lane 15 into the next row and of lane 31
// into rows 2 and 3, which leave the wavefront's total in lane 63.  The order of the
// additions is fixed, so results are deterministic; the total is returned in every
// lane (read from lane 63).  Every lane of the wavefront must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v, double old) {
    const long long b = __double_as_longlong(v), o = __double_as_longlong(old);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double lane63(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
enum : int {
    DPP_QUAD_1032 = 0xB1,  // quad_perm [1, 0, 3, 2]
    DPP_QUAD_2301 = 0x4E,  // quad_perm [2, 3, 0, 1]
    DPP_ROW_SHR4 = 0x114,
    DPP_ROW_SHR8 = 0x118,
    DPP_ROW_BCAST15 = 0x142,
    DPP_ROW_BCAST31 = 0x143,
};
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move<DPP_QUAD_1032>(v, 0.0);
    v += dpp_move<DPP_QUAD_2301>(v, 0.0);
    v += dpp_move<DPP_ROW_SHR4>(v, 0.0);
    v += dpp_move<DPP_ROW_SHR8>(v, 0.0);
    v += dpp_move<DPP_ROW_BCAST15>(v, 0.0);
    v += dpp_move<DPP_ROW_BCAST31>(v, 0.0);
    return lane63(v);
}
__device__ __forceinline__ double wave_max(double v) {
    v = fmax(v, dpp_move<DPP_QUAD_1032>(v, v));
    v = fmax(v, dpp_move<DPP_QUAD_2301>(v, v));
    v = fmax(v, dpp_move<DPP_ROW_SHR4>(v, v));
    v = fmax(v, dpp_move<DPP_ROW_SHR8>(v, v));
    v = fmax(v, dpp_move<DPP_ROW_BCAST15>(v, v));
    v = fmax(v, dpp_move<DPP_ROW_BCAST31>(v, v));
    return lane63(v);
}

// Block reduction (sum) for up to 1024 threads; result valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double s_part[NT / 64];
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    v = wave_sum(v);
    if ((lane & 63) == 0) s_part[lane >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (lane < 64) {
        r = (lane < NT / 64) ? s_part[lane] : 0.0;
        r = wave_sum(r);
    }
    __syncthreads();
    return r;
}
// Block sums of NV values at once for up to 1024 threads, valid in thread 0: per
// value the additions of block_sum in the same order (the same bits), with one
// barrier and one LDS round trip for all NV values instead of one per value.
template <int NV, int NT>
__device__ __forceinline__ void block_sums(double (&v)[NV]) {
    __shared__ double s_part[NV][NT / 64];
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
    if ((lane & 63) == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) s_part[k][lane >> 6] = v[k];
    }
    __syncthreads();
    if (lane < 64) {
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] = wave_sum((lane < NT / 64) ? s_part[k][lane] : 0.0);
    }
    __syncthreads();
}
template <int NT>
__device__ __forceinline__ double block_max(double v) {
    __shared__ double s_part[NT / 64];
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    v = wave_max(v);
    if ((lane & 63) == 0) s_part[lane >> 6] = v;
    __syncthreads();
    double r = -1.0e300;
    if (lane < 64) {
        r = (lane < NT / 64) ? s_part[lane] : -1.0e300;
        r = wave_max(r);
    }
    __syncthreads();
    return r;
}

}  // namespace cfd
