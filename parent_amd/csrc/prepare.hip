// Pre-binning kernels: per-dof extrema, dihedral modfit, bin indices and the
// per-dof 1D statistics (PARENT.cpp:503-772).
//
// Compiled with -ffp-contract=off: the range rules, the modfit shift and the
// bin-index divisions must round exactly as the reference's FP64 code does, so
// that every index is bit-identical (no FMA contraction, IEEE division).
//
// Data layout in HBM
//   values   frame-major [N][D] float or double (the trajectory as handed in)
//   bins2    dof-major  [D][Np] uint8, Np = N rounded up to 4096; frames >= N
//            hold 0xFF.  One row per dof so a pair kernel streams two rows.
//   bins1    same layout for the 1D bin counts, only when they differ from the
//            2D ones for some dof type.
#include "kernels.h"

#include <algorithm>

namespace parent {

__device__ __forceinline__ unsigned long long ord_enc(double x) {
    unsigned long long u = (unsigned long long)__double_as_longlong(x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double ord_dec(unsigned long long e) {
    unsigned long long u = (e >> 63) ? (e & 0x7FFFFFFFFFFFFFFFull) : ~e;
    return __longlong_as_double((long long)u);
}

// x + modFit - 2*pi*int((x+modFit)/(2*pi))   (PARENT.cpp:610-612)
__device__ __forceinline__ double modfit_shift(double x, double mf, double twopi) {
    const double s = x + mf;
    const int k = (int)(s / twopi);
    return s - twopi * (double)k;
}

// Block order of the frame-major scans (1D grid of 8 * ceil(nbx * nby / 8)
// workgroups).  Workgroups go to the 8 XCDs round-robin by id and each XCD
// has its own L2.  A frame row is D values long, not a multiple of 128 bytes,
// so neighbouring 64-dof blocks of one frame range share the cache lines at
// their boundary: id b runs logical block L = (b % 8) * ceil(T / 8) + b / 8
// (T = nbx * nby, dof block fastest), which gives each XCD a contiguous run
// of logical blocks and one L2 both sides of each boundary.  The grid holds
// 8 * ceil(T / 8) ids, so L covers [0, 8 ceil(T / 8)) once; ids with L >= T
// have no block and return false.
__device__ __forceinline__ bool xcd_block(int nbx, int nby, int* bx, int* by) {
    const int T = nbx * nby, per = (T + 7) / 8;
    const int b = (int)blockIdx.x;
    const int L = (b % 8) * per + b / 8;
    if (L >= T) return false;
    *bx = L % nbx;
    *by = L / nbx;
    return true;
}

inline unsigned xcd_grid(int nbx, int nby) { return 8u * (unsigned)((nbx * nby + 7) / 8); }

// Visits dof g's values at frames f = f_begin, f_begin + 4, ... < f_end of
// the frame-major array v (D values a frame), loading KB of them before using
// any, so that KB loads per thread are in flight (one at a time, each load's
// latency was exposed).  use(f, x) gets the value widened to double.
template <int KB, class T, class F>
__device__ __forceinline__ void for_frames(const T* __restrict__ v, int D, int g, int f_begin, int f_end, F&& use) {
    int f = f_begin;
    for (; f + 4 * (KB - 1) < f_end; f += 4 * KB) {
        T xv[KB];
#pragma unroll
        for (int u = 0; u < KB; u++) xv[u] = v[(size_t)(f + 4 * u) * D + g];
#pragma unroll
        for (int u = 0; u < KB; u++) use(f + 4 * u, (double)xv[u]);
    }
    for (; f < f_end; f += 4) use(f, (double)v[(size_t)f * D + g]);
}

// Pass 1: min/max of every dof over a slice of frames (exact in the input
// precision), plus the 100-bin modfit histogram of each dihedral
// (PARENT.cpp:513-521, 537-545, 568-572).  Block = 64 dofs x 4 frame lanes;
// one wave reads 64 consecutive values of a frame row.
template <class T>
__global__ __launch_bounds__(256) void k_scan_raw(const T* __restrict__ v, int N, int D, int nD,
                                                  int fpb, double mf_bin, int nbx, int nby,
                                                  unsigned long long* __restrict__ emin,
                                                  unsigned long long* __restrict__ emax,
                                                  unsigned* __restrict__ mfh, int* flags) {
    __shared__ unsigned hist[64 * kModfitBins];
    __shared__ double smin[256], smax[256];
    int bx, by;
    if (!xcd_block(nbx, nby, &bx, &by)) return;
    const int dl = threadIdx.x & 63, fl = threadIdx.x >> 6;
    const int g = bx * 64 + dl;
    const int dih0 = 2 * nD + 3;
    const bool blk_dih = bx * 64 + 63 >= dih0;
    if (blk_dih)
        for (int i = threadIdx.x; i < 64 * kModfitBins; i += 256) hist[i] = 0;
    __syncthreads();
    double mn = 1.0e300, mx = -1.0e300;
    const int f0 = by * fpb, f1 = min(N, f0 + fpb);
    if (g < D) {
        const bool isd = g >= dih0;
        bool bad = false;
        // (min, max and the integer histogram do not depend on the order)
        for_frames<8>(v, D, g, f0 + fl, f1, [&](int, double x) {
            mn = x < mn ? x : mn;
            mx = x > mx ? x : mx;
            if (isd) {
                const int k = (int)(x / mf_bin);
                if (k < 0 || k >= kModfitBins) bad = true;
                else atomicAdd(&hist[dl * kModfitBins + k], 1u);
            }
        });
        if (bad) atomicOr(flags, kFlagRange);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    // dihedral extrema come from the shifted values (k_scan_dihedral), so only
    // bonds and angles publish their raw extrema here
    if (fl == 0 && g < dih0) {
        for (int q = 1; q < 4; q++) {
            const double a = smin[dl + 64 * q], b = smax[dl + 64 * q];
            mn = a < mn ? a : mn;
            mx = b > mx ? b : mx;
        }
        atomicMin(&emin[g], ord_enc(mn));
        atomicMax(&emax[g], ord_enc(mx));
    }
    if (blk_dih) {
        for (int i = threadIdx.x; i < 64 * kModfitBins; i += 256) {
            const int gg = bx * 64 + i / kModfitBins;
            if (gg >= dih0 && gg < D && hist[i]) atomicAdd(&mfh[(size_t)(gg - dih0) * kModfitBins + i % kModfitBins], hist[i]);
        }
    }
}

// The empty-bin scans of one dihedral's 100-bin histogram (PARENT.cpp:573-607),
// done by one wave over ballot masks instead of two serial passes per thread.
// Returns the modfit bin: with an empty bin, the start of the longest run of
// empty bins; otherwise the first of the least populated bins.
//
// The reference walks k = 0 ... 199 over bin k % 100 and keeps a run only
// when it is strictly longer than the longest so far, so among the longest
// runs the one that ends first wins.  Every run but one ends in the order of
// its start bin; the run that holds bins 99 and 0 ends last, in the second
// pass, and the walk's first run, which begins at bin 0, is then a strict
// part of it and never wins.  So: the first start bin (a run start is an
// empty bin after an occupied one, circularly) of the longest runs, lengths
// counted circularly.  No occupied bin at all leaves the reference's -1.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const unsigned lo = __shfl_xor((unsigned)v, m), hi = __shfl_xor((unsigned)(v >> 32), m);
        const unsigned long long o = ((unsigned long long)hi << 32) | lo;
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ int trailing_ones(unsigned long long x) {
    return x == ~0ull ? 64 : __builtin_ctzll(~x);
}

__device__ int modfit_bin(const unsigned* __restrict__ histo, int lane) {
    static_assert(kModfitBins > 64 && kModfitBins <= 128, "two ballot words per histogram");
    constexpr int kHigh = kModfitBins - 64;
    const unsigned c0 = histo[lane], c1 = lane < kHigh ? histo[64 + lane] : 1u;
    // bit k of z0 (bins 0-63) and of z1 (bins 64-99): bin empty
    const unsigned long long z0 = __ballot(c0 == 0u), z1 = __ballot(lane < kHigh && c1 == 0u);
    if (z0 == 0 && z1 == 0) {
        // key = ~count : ~bin, so the largest key is the lowest count's first bin
        unsigned long long key = ((unsigned long long)~c0 << 32) | (unsigned)~lane;
        if (lane < kHigh) {
            const unsigned long long k1 = ((unsigned long long)~c1 << 32) | (unsigned)~(64 + lane);
            key = k1 > key ? k1 : key;
        }
        return (int)~(unsigned)(wave_max_u64(key) & 0xFFFFFFFFull);
    }
    if (z0 == ~0ull && z1 == (1ull << kHigh) - 1) return -1;
    auto empty = [&](int k) { return (int)((k < 64 ? z0 >> k : z1 >> (k - 64)) & 1u); };
    // empty bins from bin k on, up to bin 99 (z1 holds no bits past it)
    auto empties_from = [&](int k) {
        if (k >= 64) return trailing_ones(z1 >> (k - 64));
        const int t = trailing_ones(z0 >> k);
        return t == 64 - k ? t + trailing_ones(z1) : t;
    };
    // key = run length : ~start bin, so the largest key is the longest run's first start
    unsigned long long key = 0;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int k = lane + 64 * h;
        if (k < kModfitBins && empty(k) && !empty(k == 0 ? kModfitBins - 1 : k - 1)) {
            int len = empties_from(k);
            if (k + len == kModfitBins) len += empties_from(0);
            const unsigned long long kk = ((unsigned long long)len << 32) | (unsigned)~k;
            key = kk > key ? kk : key;
        }
    }
    return (int)~(unsigned)(wave_max_u64(key) & 0xFFFFFFFFull);
}

// Pass 2: range rules of bonds and angles, and the modfit shift of each
// dihedral from its 100-bin histogram (PARENT.cpp:522-533, 546-562, 573-609).
// kFinalizeBlock dofs a block: all threads stage the block's histograms in
// LDS, each wave then takes every fourth dihedral of the block, and one
// thread per bond or angle applies its rules.
constexpr int kFinalizeBlock = 16, kFinalizeThreads = 256;
__global__ __launch_bounds__(kFinalizeThreads) void k_finalize_ranges(int D, int nD, double pi, double mf_bin,
                                  const unsigned long long* __restrict__ emin,
                                  const unsigned long long* __restrict__ emax,
                                  const unsigned* __restrict__ mfh, double* vmin, double* vmax,
                                  double* modfit, int* flags) {
    // coalesced staging, all of a thread's loads issued before the first store
    constexpr int kWords = kFinalizeBlock * kModfitBins, kPerThread = (kWords + kFinalizeThreads - 1) / kFinalizeThreads;
    __shared__ unsigned sh[kWords];
    const int g0 = blockIdx.x * kFinalizeBlock, dih0 = 2 * nD + 3;
    const bool blk_dih = g0 + kFinalizeBlock > dih0;
    if (blk_dih) {
        unsigned w[kPerThread];
#pragma unroll
        for (int u = 0; u < kPerThread; u++) {
            const int i = threadIdx.x + u * kFinalizeThreads, gg = g0 + i / kModfitBins;
            w[u] = i < kWords && gg >= dih0 && gg < D ? mfh[(size_t)(gg - dih0) * kModfitBins + i % kModfitBins] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kPerThread; u++) {
            const int i = threadIdx.x + u * kFinalizeThreads;
            if (i < kWords) sh[i] = w[u];
        }
    }
    __syncthreads();
    if (blk_dih) {
        const int lane = threadIdx.x & 63;
        for (int j = threadIdx.x >> 6; j < kFinalizeBlock; j += kFinalizeThreads / 64) {
            const int g = g0 + j;
            if (g < dih0 || g >= D) continue;  // the same for the whole wave
            const int pos = modfit_bin(sh + j * kModfitBins, lane);
            if (lane == 0) modfit[g] = 2 * pi - (pos + 0.5) * mf_bin;
        }
    }
    if (threadIdx.x >= kFinalizeBlock) return;
    const int g = g0 + threadIdx.x;
    if (g >= D || g >= dih0) return;
    const int t = dof_type(nD, g);
    double tmpMin = ord_dec(emin[g]), tmpMax = ord_dec(emax[g]);
    if (t == 0) {
        if ((tmpMin < 0.0) || (tmpMax < 0.0)) atomicOr(flags, kFlagNegBond);
        tmpMin -= 0.000000005;
        tmpMax += 0.000000005;
        if (tmpMin < 0.0) tmpMin = 0.0;
        if ((tmpMax - tmpMin) < 1.0e-4) tmpMax += 0.05;
    } else {
        tmpMin -= 0.000000005;
        tmpMax += 0.000000005;
        if (tmpMin < 0) tmpMin = 0;
        if (tmpMax > pi) tmpMax = pi;
        if ((tmpMax - tmpMin) < 1.0e-4) tmpMax += 0.05;
    }
    vmin[g] = tmpMin;
    vmax[g] = tmpMax;
    modfit[g] = 0.0;
}

// Pass 3: extrema of the shifted dihedrals (PARENT.cpp:614-623), FP64.
template <class T>
__global__ __launch_bounds__(256) void k_scan_dihedral(const T* __restrict__ v, int N, int D, int nD,
                                                       int fpb, double twopi, int nbx, int nby,
                                                       const double* __restrict__ modfit,
                                                       unsigned long long* __restrict__ emin,
                                                       unsigned long long* __restrict__ emax) {
    __shared__ double smin[256], smax[256];
    int bx, by;
    if (!xcd_block(nbx, nby, &bx, &by)) return;
    const int dl = threadIdx.x & 63, fl = threadIdx.x >> 6;
    const int dih0 = 2 * nD + 3;
    const int g = dih0 + bx * 64 + dl;
    double mn = 1.0e300, mx = -1.0e300;
    const int f0 = by * fpb, f1 = min(N, f0 + fpb);
    if (g < D) {
        const double mf = modfit[g];
        for_frames<8>(v, D, g, f0 + fl, f1, [&](int, double raw) {
            const double x = modfit_shift(raw, mf, twopi);
            mn = x < mn ? x : mn;
            mx = x > mx ? x : mx;
        });
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    if (fl == 0 && g < D) {
        for (int q = 1; q < 4; q++) {
            const double a = smin[dl + 64 * q], b = smax[dl + 64 * q];
            mn = a < mn ? a : mn;
            mx = b > mx ? b : mx;
        }
        atomicMin(&emin[g], ord_enc(mn));
        atomicMax(&emax[g], ord_enc(mx));
    }
}

// Pass 4: dihedral range rules (PARENT.cpp:624-636).
__global__ void k_finalize_dihedral(int D, int nD, double twopi,
                                    const unsigned long long* __restrict__ emin,
                                    const unsigned long long* __restrict__ emax, double* vmin,
                                    double* vmax) {
    const int g = 2 * nD + 3 + blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= D) return;
    double tmpMin = ord_dec(emin[g]), tmpMax = ord_dec(emax[g]);
    tmpMin -= 0.000000005;
    tmpMax += 0.000000005;
    if (tmpMin < 0.0) tmpMin = 0.0;
    if (tmpMax > twopi) tmpMax = twopi;
    if ((tmpMax - tmpMin) < 1.0e-4) tmpMax += 0.05;
    vmin[g] = tmpMin;
    vmax[g] = tmpMax;
}

// Pass 5: bin indices int((x-min)/((max-min)/B)) for every (frame, dof)
// (PARENT.cpp:654-656 and the 2D fills, 802-804 ...), written transposed to
// the dof-major rows through an LDS tile of 64 dofs x 256 frames so both the
// frame-major reads and the dof-major writes are coalesced.  I2 / I1: the
// index types (uint8_t, or uint16_t for bin counts above 255) of the 2D and
// 1D matrices; the tiles are dynamic LDS of 64 x kPrebinRow indices each.
constexpr int kPrebinRow = 260;  // 256 frames + 4: rows start in different LDS banks

template <class I2, class I1, bool NEED1>
constexpr size_t prebin_lds() {
    return (size_t)64 * kPrebinRow * (sizeof(I2) + (NEED1 ? sizeof(I1) : 0));
}

template <class T, bool NEED1, class I2, class I1>
__global__ __launch_bounds__(256) void k_prebin(const T* __restrict__ v, int N, int Np, int D, int nD,
                                                const double* __restrict__ vmin,
                                                const double* __restrict__ vmax,
                                                const double* __restrict__ modfit, double twopi,
                                                Bins3 B2, Bins3 B1, int nbx, int nby,
                                                void* __restrict__ bins2, void* __restrict__ bins1,
                                                int* flags) {
    extern __shared__ __attribute__((aligned(16))) unsigned char prebin_smem[];
    I2* tile2 = reinterpret_cast<I2*>(prebin_smem);
    I1* tile1 = reinterpret_cast<I1*>(prebin_smem + (size_t)64 * kPrebinRow * sizeof(I2));
    constexpr I2 pad2 = (I2)~(I2)0;
    constexpr I1 pad1 = (I1)~(I1)0;
    int bx, by;
    if (!xcd_block(nbx, nby, &bx, &by)) return;
    const int dl = threadIdx.x & 63, fl = threadIdx.x >> 6;
    const int d0 = bx * 64, f0 = by * 256;
    const int g = d0 + dl;
    const int dih0 = 2 * nD + 3;
    double mn = 0, bs2 = 1, bs1 = 1, mf = 0;
    int nb2 = 1, nb1 = 1;
    bool isd = false, sep1 = false;
    if (g < D) {
        const int t = dof_type(nD, g);
        mn = vmin[g];
        nb2 = B2.of(t);
        nb1 = B1.of(t);
        bs2 = (vmax[g] - vmin[g]) / nb2;
        bs1 = (vmax[g] - vmin[g]) / nb1;
        isd = g >= dih0;
        mf = isd ? modfit[g] : 0.0;
        sep1 = NEED1 && nb1 != nb2;
    }
    bool bad = false;
    // padding: frames past N, and every frame of slots past the last dof
    const int fend = g < D ? min(N, f0 + 256) : f0;
    for (int r = fl; r < 256; r += 4)
        if (f0 + r >= fend) {
            tile2[dl * kPrebinRow + r] = pad2;
            if constexpr (NEED1) tile1[dl * kPrebinRow + r] = pad1;
        }
    for_frames<16>(v, D, g, f0 + fl, fend, [&](int f, double x) {
        const int r = f - f0;
        if (isd) x = modfit_shift(x, mf, twopi);
        int k = (int)((x - mn) / bs2);
        if (k < 0 || k >= nb2) {
            bad = true;
            k = k < 0 ? 0 : nb2 - 1;
        }
        tile2[dl * kPrebinRow + r] = (I2)k;
        if constexpr (NEED1) {
            int k1 = k;
            if (sep1) {
                k1 = (int)((x - mn) / bs1);
                if (k1 < 0 || k1 >= nb1) {
                    bad = true;
                    k1 = k1 < 0 ? 0 : nb1 - 1;
                }
            }
            tile1[dl * kPrebinRow + r] = (I1)k1;
        }
    });
    if (bad) atomicOr(flags, kFlagRange);
    __syncthreads();
    // rows of 256 frames leave as 32-bit words: 64 (uint8) or 128 (uint16) a row
    constexpr int w2 = 256 * (int)sizeof(I2) / 4, w1 = 256 * (int)sizeof(I1) / 4;
    for (int w = threadIdx.x; w < 64 * w2; w += 256) {
        const int dof = w / w2, wr = w % w2;
        const int gg = d0 + dof;
        if (gg >= D) continue;
        const uint32_t word = *reinterpret_cast<const uint32_t*>(tile2 + dof * kPrebinRow + wr * (4 / sizeof(I2)));
        *reinterpret_cast<uint32_t*>(static_cast<I2*>(bins2) + (size_t)gg * Np + f0 + wr * (4 / sizeof(I2))) = word;
    }
    if constexpr (NEED1) {
        for (int w = threadIdx.x; w < 64 * w1; w += 256) {
            const int dof = w / w1, wr = w % w1;
            const int gg = d0 + dof;
            if (gg >= D) continue;
            const uint32_t word = *reinterpret_cast<const uint32_t*>(tile1 + dof * kPrebinRow + wr * (4 / sizeof(I1)));
            *reinterpret_cast<uint32_t*>(static_cast<I1*>(bins1) + (size_t)gg * Np + f0 + wr * (4 / sizeof(I1))) = word;
        }
    }
}

// Histogram of one dof's index row (Np indices of type I, padding skipped)
// into the LDS counters h.
template <class I>
__device__ __forceinline__ void row_histogram(const I* __restrict__ row, int Np, unsigned* h) {
    constexpr int per = 16 / (int)sizeof(I);  // indices per 16-byte load
    constexpr unsigned pad = (unsigned)(I)~(I)0;
    const uint4* r = reinterpret_cast<const uint4*>(row);
    for (int i = threadIdx.x; i < Np / per; i += 256) {
        const uint4 w = r[i];
        const unsigned ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int q = 0; q < per; q++) {
            const unsigned b = sizeof(I) == 1 ? (ws[q >> 2] >> (8 * (q & 3))) & 0xFFu
                                              : (ws[q >> 1] >> (16 * (q & 1))) & 0xFFFFu;
            if (b != pad) atomicAdd(&h[b], 1u);
        }
    }
}

// The same into 16 lane-sliced copies, h[b * 16 + lane % 16], for bin counts
// up to 64: the 16 lanes the LDS serves together never share a word, where a
// single histogram of a few dozen bins serialises many of a wave's atomics
// (the 1D pass of 4,494 dofs x 100k frames: 0.98 ms at 5 bins, 0.39 ms at 20
// and 0.23 ms at 50 with one histogram; 0.12 ms sliced).
template <class I>
__device__ __forceinline__ void row_histogram_sliced(const I* __restrict__ row, int Np, unsigned* h) {
    constexpr int per = 16 / (int)sizeof(I);
    constexpr unsigned pad = (unsigned)(I)~(I)0;
    const uint4* r = reinterpret_cast<const uint4*>(row);
    unsigned* hl = h + (threadIdx.x & 15);
    for (int i = threadIdx.x; i < Np / per; i += 256) {
        const uint4 w = r[i];
        const unsigned ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int q = 0; q < per; q++) {
            const unsigned b = sizeof(I) == 1 ? (ws[q >> 2] >> (8 * (q & 3))) & 0xFFu
                                              : (ws[q >> 1] >> (16 * (q & 1))) & 0xFFFFu;
            if (b != pad) atomicAdd(&hl[b * 16], 1u);
        }
    }
}

// Pass 6: per-dof histograms, the 1D entropy with the reference's formula
// (PARENT.cpp:650-671 bonds, 694-715 angles, 738-757 dihedrals), and the
// per-dof constants of the 2D epilogue:
//   jtab[g][k] = Jacobian of 2D bin k (blen^2, sin(theta) or 1, with the
//                reference's bin-centre recurrences, PARENT.cpp:808-819)
//   c2[g]      = ln(binsize2) + (1/N) sum_k h2[k] ln J_k
// With these, S_ij = ln N + c2[i] + c2[j] - (1/N) sum n ln n + (occ-1)/(2N),
// which is the reference epilogue regrouped (the marginal of a 2D histogram
// is the dof's 1D histogram at the 2D bin count).
//
// One workgroup per dof.  The bin-centre recurrences run in one thread (each
// centre is the previous one plus the bin size, as in the reference); the
// per-bin terms (divisions, log, sin) then run one bin per thread, and one
// thread adds them in bin order, so every sum has the reference's order and
// bits.  Dynamic LDS (dof_stats_lds): counters h2 [nb2] and h1 [nh1] (the
// separate 1D histogram, nh1 = 0 without a 1D matrix), then doubles t2
// [nb2] and t1 [nb1]: the bin centres, overwritten by the terms.
template <class I2, class I1>
__global__ __launch_bounds__(256) void k_dof_stats(const void* __restrict__ idx2, const void* __restrict__ idx1,
                                                   int N, int Np, int nD, Bins3 B1, Bins3 B2, int nb2, int nh1,
                                                   int nb1, const double* __restrict__ vmin,
                                                   const double* __restrict__ vmax, double* s1d, double* c2,
                                                   double* jtab, int jstride) {
    extern __shared__ __attribute__((aligned(16))) unsigned char stats_smem[];
    unsigned* h2 = reinterpret_cast<unsigned*>(stats_smem);
    unsigned* h1 = h2 + nb2;
    double* t2 = reinterpret_cast<double*>(stats_smem + (size_t)((nb2 + nh1 + 1) / 2) * 8);
    double* t1 = t2 + nb2;
    const int g = blockIdx.x;
    const int t = dof_type(nD, g);
    const bool sep1 = idx1 != nullptr && B1.of(t) != B2.of(t);
    // counts of up to kSliceBins bins go through 16 lane-sliced copies (a
    // bin's 16 copies in consecutive words), then add up in h2 / h1
    constexpr int kSliceBins = 64;
    __shared__ unsigned hs[2][16 * kSliceBins];
    const bool sl2 = nb2 <= kSliceBins, sl1 = sep1 && nh1 <= kSliceBins;
    for (int i = threadIdx.x; i < nb2 + nh1; i += 256) h2[i] = 0;
    for (int i = threadIdx.x; i < 2 * 16 * kSliceBins; i += 256) hs[0][i] = 0;
    __syncthreads();
    if (sl2) row_histogram_sliced(static_cast<const I2*>(idx2) + (size_t)g * Np, Np, hs[0]);
    else row_histogram(static_cast<const I2*>(idx2) + (size_t)g * Np, Np, h2);
    if (sep1) {
        if (sl1) row_histogram_sliced(static_cast<const I1*>(idx1) + (size_t)g * Np, Np, hs[1]);
        else row_histogram(static_cast<const I1*>(idx1) + (size_t)g * Np, Np, h1);
    }
    if (sl2 || sl1) {
        __syncthreads();
        // threads 0-127: (bin, histogram); copy (l + k) % 16 first, so that
        // the 64 lanes of a read touch 16 different banks, not one
        const int k = threadIdx.x % kSliceBins, side = threadIdx.x / kSliceBins;
        if (side < 2 && (side ? sl1 && k < nh1 : sl2 && k < nb2)) {
            unsigned n = 0;
#pragma unroll
            for (int l = 0; l < 16; l++) n += hs[side][k * 16 + ((l + k) & 15)];
            (side ? h1 : h2)[k] = n;
        }
    }
    const unsigned* histo = sep1 ? h1 : h2;
    const int numFrames = N;
    const int bins = B1.of(t), bins2 = B2.of(t);
    const double binsize = (vmax[g] - vmin[g]) / (bins);
    const double binsize2 = (vmax[g] - vmin[g]) / bins2;
    // bin centres (PARENT.cpp:660-667, 704-712; 2D: 808-819)
    if (threadIdx.x == 0 && t != 2) {
        double c = vmin[g] + (binsize / 2.0);
        for (int k = 0; k < bins; k++) {
            t1[k] = c;
            c += binsize;
        }
        c = vmin[g] + binsize2 / 2.0;
        for (int k = 0; k < bins2; k++) {
            t2[k] = c;
            c = c + binsize2;
        }
    }
    __syncthreads();
    // per-bin terms: t1[k] = J p ln p (the 1D sum's term when p > 0), t2[k] =
    // h2[k] ln J; the Jacobian table
    double* jt = jtab + (size_t)g * jstride;
    for (int k = threadIdx.x; k < bins; k += 256) {
        double term = 0.0;
        if (t == 0) {
            const double blen = t1[k];
            const double probDens = histo[k] / (numFrames * binsize * blen * blen);
            if (probDens > 0) term = blen * blen * probDens * log(probDens);
        } else if (t == 1) {
            const double theta = t1[k];
            const double probDens = histo[k] / (numFrames * binsize * sin(theta));
            if (probDens > 0) term = sin(theta) * probDens * log(probDens);
        } else {
            const double probDens = histo[k] / (numFrames * binsize);
            if (probDens > 0) term = probDens * log(probDens);
        }
        t1[k] = term;
    }
    for (int k = threadIdx.x; k < bins2; k += 256) {
        double J = 1.0;
        if (t == 0) J = t2[k] * t2[k];
        else if (t == 1) J = sin(t2[k]);
        jt[k] = J;
        t2[k] = (t != 2 && h2[k]) ? (double)h2[k] * log(J) : 0.0;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    // the sums in bin order (PARENT.cpp:650-671 ...): a bin counts when its
    // density is > 0, i.e. when its count is
    int occupbins = 0;
    double plnpsum = 0;
    for (int k = 0; k < bins; k++)
        if (histo[k]) {
            plnpsum = plnpsum + t1[k];
            occupbins = occupbins + 1;
        }
    plnpsum = -plnpsum * binsize;
    s1d[g] = plnpsum + (occupbins - 1.0) / (2.0 * numFrames);
    double acc = 0.0;
    if (t != 2)
        for (int k = 0; k < bins2; k++)
            if (h2[k]) acc += t2[k];
    c2[g] = log(binsize2) + acc / numFrames;
}

size_t dof_stats_lds(int nb2, int nh1, int nb1) { return (size_t)((nb2 + nh1 + 1) / 2) * 8 + (size_t)(nb2 + nb1) * 8; }

// n ln n - 1/2 for n = 1..N and 0 for n = 0, in FP64.  The 2D epilogue sums
// table entries over the cells: the occupancy term (occ - 1) / (2N) of
// S_ij then comes out of the same sum, S_ij = ln N + c_i + c_j - 1/(2N)
// - (1/N) sum (n ln n - 1/2 [n > 0]), with no separate count of occupied
// cells.
__device__ __forceinline__ double nlogn_entry(int n) { return n == 0 ? 0.0 : (double)n * log((double)n) - 0.5; }

__global__ void k_nlogn(int N, double* tab) {
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n <= N; n += gridDim.x * blockDim.x) tab[n] = nlogn_entry(n);
}

// The set-up of a prepare in one launch (each was its own small kernel, a
// few microseconds apiece on short trajectories): the flags word, and with
// the ranges the extrema accumulators (min at all ones, max at 0) and the
// modfit histograms, and the n ln n table when nlogn is given.
__global__ void k_prepare_init(int* flags, unsigned long long* emin, unsigned long long* emax, int D,
                               unsigned* mfh, size_t nmf, double* nlogn, int N) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    if (tid == 0) *flags = 0;
    if (emin)
        for (size_t i = tid; i < (size_t)D; i += stride) {
            emin[i] = ~0ull;
            emax[i] = 0ull;
        }
    for (size_t i = tid; i < nmf; i += stride) mfh[i] = 0u;
    if (nlogn)
        for (size_t n = tid; n <= (size_t)N; n += stride) nlogn[n] = nlogn_entry((int)n);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

template <class T, bool NEED1, class I2, class I1>
hipError_t prebin_launch(const PrepareArgs& a, unsigned g5, int nbx5, int nby5, double twopi, hipStream_t s) {
    constexpr size_t lds = prebin_lds<I2, I1, NEED1>();
    if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_prebin<T, NEED1, I2, I1>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    k_prebin<T, NEED1, I2, I1><<<g5, 256, lds, s>>>(static_cast<const T*>(a.values), a.N, a.Np, a.D, a.nD, a.vmin,
                                                     a.vmax, a.modfit, twopi, a.B2, a.B1, nbx5, nby5, a.bins2,
                                                     a.bins1, a.flags);
    return hipSuccess;
}

// k_prebin for the matrices' index types
template <class T>
hipError_t prebin_index_types(const PrepareArgs& a, unsigned g5, int nbx5, int nby5, double twopi, hipStream_t s) {
    if (!a.bins1) {
        return a.idx2 == 1 ? prebin_launch<T, false, uint8_t, uint8_t>(a, g5, nbx5, nby5, twopi, s)
                           : prebin_launch<T, false, uint16_t, uint16_t>(a, g5, nbx5, nby5, twopi, s);
    }
    if (a.idx2 == 1)
        return a.idx1 == 1 ? prebin_launch<T, true, uint8_t, uint8_t>(a, g5, nbx5, nby5, twopi, s)
                           : prebin_launch<T, true, uint8_t, uint16_t>(a, g5, nbx5, nby5, twopi, s);
    return a.idx1 == 1 ? prebin_launch<T, true, uint16_t, uint8_t>(a, g5, nbx5, nby5, twopi, s)
                       : prebin_launch<T, true, uint16_t, uint16_t>(a, g5, nbx5, nby5, twopi, s);
}

// Frames per block of the scans: 2,048 (4 frame lanes of 512), halved down to
// 128 while the grid would have fewer than 1,024 blocks, so that a short or
// narrow trajectory still spreads over the CUs instead of leaving a few
// blocks to walk long frame ranges one load latency after another.
static int scan_frames_per_block(int nbx, int N) {
    int fpb = 2048;
    while (fpb > 128 && (long long)nbx * ((N + fpb - 1) / fpb) < 1024) fpb /= 2;
    return fpb;
}

hipError_t launch_prepare(const PrepareArgs& a, hipStream_t s) {
    const int D = a.D, N = a.N, nD = a.nD;
    const int nDih = nD;
    const double pi = a.pi, twopi = 2 * pi;
    const double mf_bin = (2 * pi + 0.000000005) / kModfitBins;
    {
        const bool ranges = (a.stages & kPrepareRanges) != 0;
        const size_t nmf = ranges ? (size_t)nDih * kModfitBins : 0;
        double* tab = (a.stages & kPrepareNlogn) ? a.nlogn : nullptr;
        const size_t work = std::max(std::max(ranges ? (size_t)D : 1, nmf), tab ? (size_t)N + 1 : 1);
        const unsigned blocks = (unsigned)std::min<size_t>(256, (work + 255) / 256);
        k_prepare_init<<<blocks, 256, 0, s>>>(a.flags, ranges ? a.emin : nullptr, a.emax, D, a.mfhist, nmf, tab, N);
    }
    if (a.stages & kPrepareRanges) {
        const int nbx1 = (D + 63) / 64, fpb = scan_frames_per_block(nbx1, N);
        const int nby1 = (N + fpb - 1) / fpb;
        const unsigned g1 = xcd_grid(nbx1, nby1);
        if (a.double_prec)
            k_scan_raw<double><<<g1, 256, 0, s>>>((const double*)a.values, N, D, nD, fpb, mf_bin, nbx1, nby1, a.emin, a.emax, a.mfhist, a.flags);
        else
            k_scan_raw<float><<<g1, 256, 0, s>>>((const float*)a.values, N, D, nD, fpb, mf_bin, nbx1, nby1, a.emin, a.emax, a.mfhist, a.flags);
        k_finalize_ranges<<<(D + kFinalizeBlock - 1) / kFinalizeBlock, kFinalizeThreads, 0, s>>>(D, nD, pi, mf_bin, a.emin, a.emax, a.mfhist, a.vmin, a.vmax, a.modfit, a.flags);
        if (nDih > 0) {
            const int nbx3 = (nDih + 63) / 64, fpb3 = scan_frames_per_block(nbx3, N);
            const int nby3 = (N + fpb3 - 1) / fpb3;
            const unsigned g3 = xcd_grid(nbx3, nby3);
            if (a.double_prec)
                k_scan_dihedral<double><<<g3, 256, 0, s>>>((const double*)a.values, N, D, nD, fpb3, twopi, nbx3, nby3, a.modfit, a.emin, a.emax);
            else
                k_scan_dihedral<float><<<g3, 256, 0, s>>>((const float*)a.values, N, D, nD, fpb3, twopi, nbx3, nby3, a.modfit, a.emin, a.emax);
            k_finalize_dihedral<<<(nDih + 255) / 256, 256, 0, s>>>(D, nD, twopi, a.emin, a.emax, a.vmin, a.vmax);
        }
    }
    if (a.stages & kPrepareBins) {
        const int nbx5 = (D + 63) / 64, nby5 = a.Np / 256;
        const unsigned g5 = xcd_grid(nbx5, nby5);
        hipError_t e = a.double_prec ? prebin_index_types<double>(a, g5, nbx5, nby5, twopi, s)
                                     : prebin_index_types<float>(a, g5, nbx5, nby5, twopi, s);
        if (e != hipSuccess) return e;
        const int nb2 = max(a.B2.b, max(a.B2.a, a.B2.d)), nb1 = max(a.B1.b, max(a.B1.a, a.B1.d));
        const int nh1 = a.bins1 ? nb1 : 0;
        const size_t lds = dof_stats_lds(nb2, nh1, nb1);
        const void* fn = a.idx2 == 1 ? (a.idx1 == 1 ? (const void*)&k_dof_stats<uint8_t, uint8_t> : (const void*)&k_dof_stats<uint8_t, uint16_t>)
                                     : (a.idx1 == 1 ? (const void*)&k_dof_stats<uint16_t, uint8_t> : (const void*)&k_dof_stats<uint16_t, uint16_t>);
        if (lds > 65536 && (e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)) != hipSuccess)
            return e;
#define PARENT_STATS(I2, I1)                                                                                   \
    k_dof_stats<I2, I1><<<D, 256, lds, s>>>(a.bins2, a.bins1, N, a.Np, nD, a.B1, a.B2, nb2, nh1, nb1, a.vmin, \
                                            a.vmax, a.s1d, a.c2, a.jtab, a.jstride)
        if (a.idx2 == 1 && a.idx1 == 1) PARENT_STATS(uint8_t, uint8_t);
        else if (a.idx2 == 1) PARENT_STATS(uint8_t, uint16_t);
        else if (a.idx1 == 1) PARENT_STATS(uint16_t, uint8_t);
        else PARENT_STATS(uint16_t, uint16_t);
#undef PARENT_STATS
    }
    return hipGetLastError();
}

// .bat frame records -> frame-major values.  A record (io_binary.cpp:281-368)
// is 11 floats (time, xtcPrec, box) then, in the file's precision, 6 external
// coordinates, bond0, bond1, angle0 and nD triples (bond[i+2], angle[i+1],
// dihedral[i]).  Value k of a record sits at byte 44 + k * sizeof(T); output
// dof g takes
//   bonds      g = 0, 1 -> k = 6, 7;   g = 2 + i      -> k = 9 + 3i
//   angles     a = 0    -> k = 8;      a = 1 + i      -> k = 10 + 3i
//   dihedrals  d = i                   -> k = 11 + 3i
// Records are 4-byte aligned only (44 + 8k for doubles), so doubles are
// assembled from two 32-bit words.  One workgroup per frame, threads over
// the frame's D values: coalesced writes, stride-3 reads.
template <bool DP>
__global__ __launch_bounds__(256) void k_bat_records(const uint32_t* __restrict__ rec, long long rec_words,
                                                     int nD, void* __restrict__ out) {
    const int f = blockIdx.x;
    const int nB = nD + 2, nA = nD + 1, D = 3 * nD + 3;
    const uint32_t* r = rec + (size_t)f * rec_words + 11;
    for (int g = threadIdx.x; g < D; g += 256) {
        int k;
        if (g < nB) k = g < 2 ? 6 + g : 9 + 3 * (g - 2);
        else if (g < nB + nA) k = g == nB ? 8 : 10 + 3 * (g - nB - 1);
        else k = 11 + 3 * (g - nB - nA);
        if (DP) {
            const unsigned long long lo = r[2 * k], hi = r[2 * k + 1];
            reinterpret_cast<unsigned long long*>(out)[(size_t)f * D + g] = lo | (hi << 32);
        } else {
            reinterpret_cast<uint32_t*>(out)[(size_t)f * D + g] = r[k];
        }
    }
}

hipError_t launch_bat_records(const void* records, long long rec_bytes, int n_frames, int nD, int dp,
                              void* out, hipStream_t s) {
    if (n_frames <= 0) return hipSuccess;
    const uint32_t* r = reinterpret_cast<const uint32_t*>(records);
    if (dp) k_bat_records<true><<<n_frames, 256, 0, s>>>(r, rec_bytes / 4, nD, out);
    else k_bat_records<false><<<n_frames, 256, 0, s>>>(r, rec_bytes / 4, nD, out);
    return hipGetLastError();
}

hipError_t launch_nlogn(int N, double* nlogn, hipStream_t s) {
    k_nlogn<<<(N + 256) / 256, 256, 0, s>>>(N, nlogn);
    return hipGetLastError();
}

}  // namespace parent
