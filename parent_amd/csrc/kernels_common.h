// Device-side helpers shared by the engine's HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace parent {

constexpr int kModfitBins = 100;   // PARENT.cpp MODFITNBINS
// Bin indices are stored as uint8 while every count of a matrix is <= 255
// (0xFF marks padding frames), as uint16 up to kMaxBins (0xFFFF marks them).
constexpr int kMaxBins8 = 255;
constexpr int kMaxBins = 4096;     // bins per dof the library accepts
constexpr uint8_t kPadBin = 0xFF;
constexpr uint16_t kPadBin16 = 0xFFFF;
constexpr int kFrameAlign = 4096;  // frame stride of the dof-major bin-index rows

// error bits raised by kernels (engine flags word)
enum : int { kFlagNegBond = 1, kFlagRange = 2, kFlagMist = 4 };

struct Bins3 {
    int b, a, d;
    __host__ __device__ int of(int t) const { return t == 0 ? b : (t == 1 ? a : d); }
};

__host__ __device__ __forceinline__ int dof_type(int nD, int g) {
    return g < nD + 2 ? 0 : (g < 2 * nD + 3 ? 1 : 2);
}
__host__ __device__ __forceinline__ int dof_first(int nD, int t) {
    return t == 0 ? 0 : (t == 1 ? nD + 2 : 2 * nD + 3);
}

// .par body position of pair gi < gj (io_binary.cpp:606-631, PARENT.cpp:1107-1112)
__host__ __device__ __forceinline__ long long par_offset(int nD, int gi, int gj) {
    const long long nB = nD + 2, nA = nD + 1, nd = nD;
    const int t1 = dof_type(nD, gi), t2 = dof_type(nD, gj);
    const long long i = gi - dof_first(nD, t1), j = gj - dof_first(nD, t2);
    const long long oBA = nB * (nB - 1) / 2, oBD = oBA + nB * nA, oAA = oBD + nB * nd,
                    oAD = oAA + nA * (nA - 1) / 2, oDD = oAD + nA * nd;
    if (t1 == 0) {
        if (t2 == 0) return (nB - i) * (nB - i - 1) / 2 + i - j;
        if (t2 == 1) return oBA + i * nA + j;
        return oBD + i * nd + j;
    }
    if (t1 == 1) {
        if (t2 == 1) return oAA + (nA - i) * (nA - i - 1) / 2 + i - j;
        return oAD + i * nd + j;
    }
    return oDD + (nd - i) * (nd - i - 1) / 2 + i - j;
}

// Row-major upper-triangle pair index p -> (gi < gj).
__host__ __device__ __forceinline__ void pair_of_index(long long p, int D, int* gi, int* gj) {
    const double b = 2.0 * D - 1.0;
    long long i = (long long)((b - sqrt(b * b - 8.0 * (double)p)) * 0.5);
    if (i < 0) i = 0;
    if (i > D - 2) i = D - 2;
    auto start = [&](long long r) { return r * (2LL * D - r - 1) / 2; };
    while (i > 0 && start(i) > p) i--;
    while (i < D - 2 && start(i + 1) <= p) i++;
    *gi = (int)i;
    *gj = (int)(p - start(i) + i + 1);
}

// Wave-wide sums without LDS traffic: DPP lane moves inside each 16-lane row
// (pair swap, quad swap, half-row mirror, row mirror: every lane then holds
// its row's sum), then the four row sums are read from lanes 0, 16, 32, 48
// and added in that order.  The result is wave-uniform and its summation
// order fixed, so it is bit-reproducible.  Every lane of the wave must be
// active.  (The LDS-based shuffles this replaces compete with the LDS atomics
// of other workgroups on the CU.)
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int v) {
    // every lane reads a lane of its own row, so no lane keeps its old value
    // and none needs initialising
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_mov<CTRL>((int)(b & 0xFFFFFFFFLL)), hi = dpp_mov<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ int lane_value(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ double lane_value(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xFFFFFFFFLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <class T>
__device__ __forceinline__ T wave_total(T v) {
    v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_mov<0x141>(v);  // row_half_mirror
    v += dpp_mov<0x140>(v);  // row_mirror
    return ((lane_value(v, 0) + lane_value(v, 16)) + lane_value(v, 32)) + lane_value(v, 48);
}

}  // namespace parent
