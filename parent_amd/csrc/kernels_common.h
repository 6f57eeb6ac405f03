// Device-side helpers shared by the engine's HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace parent {

constexpr int kModfitBins = 100;   // PARENT.cpp MODFITNBINS
constexpr int kMaxBins = 255;      // bin indices are stored as uint8; 0xFF marks padding frames
constexpr uint8_t kPadBin = 0xFF;
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

}  // namespace parent
