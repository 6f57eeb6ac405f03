// 2D histograms larger than the LDS, and bin indices wider than 8 bits
// (PARENT.cpp:774-1479 at any bin count).
//
// The reference takes any positive bin count: its per-thread histogram is a
// long long histo[B_i][biggest] on the stack (PARENT.cpp:790).  The LDS
// kernels keep a whole B_i x B_j histogram in LDS, which stops at about 200
// bins a side (160 KB), and read uint8 indices (up to 255 bins).  Here one
// workgroup takes one dof pair and builds its histogram in row slabs: slab s
// holds rows [s R, s R + R) of the histogram (R rows of B_j u32 counters fill
// the LDS given), the workgroup streams both index rows once per slab and
// counts only the frames whose row bin falls in the slab, then adds the
// slab's n ln n table terms and clears it for the next one.  Every frame is
// counted in exactly one slab, so the counts are exact and the epilogue sum
// is the same as for the whole histogram:
//     S_ij = ln N + c_i + c_j - 1/(2N) - (1/N) sum (n ln n - 1/2 [n > 0]).
// Padding frames carry the index type's all-ones value, which is at least
// B_i and so in no slab.
#include "kernels.h"

namespace parent {

template <class I>
__global__ __launch_bounds__(512) void k_pairs_slabs(PairInputs in, long long p0, long long p1, int lds_words,
                                                     int max_rows, const int* __restrict__ plist,
                                                     const long long* __restrict__ dump_off,
                                                     int* __restrict__ dump_out) {
    extern __shared__ unsigned hist[];
    __shared__ double wsum[8];
    constexpr int per = 16 / (int)sizeof(I);  // indices per 16-byte load
    const long long p = p0 + blockIdx.x;
    if (p >= p1) return;
    int gi, gj;
    if (plist) {
        gi = plist[2 * p];
        gj = plist[2 * p + 1];
    } else {
        pair_of_index(p, in.D, &gi, &gj);
    }
    const int Bi = in.B.of(dof_type(in.nD, gi)), Bj = in.B.of(dof_type(in.nD, gj));
    int R = min(Bi, lds_words / Bj);
    if (max_rows > 0) R = min(R, max_rows);
    const uint4* xi = reinterpret_cast<const uint4*>(reinterpret_cast<const I*>(in.bins) + (size_t)gi * in.Np);
    const uint4* xj = reinterpret_cast<const uint4*>(reinterpret_cast<const I*>(in.bins) + (size_t)gj * in.Np);
    const int nvec = in.Np / per;
    const int tid = threadIdx.x;
    double T = 0.0;
    for (int k0 = 0; k0 < Bi; k0 += R) {
        const unsigned kr = (unsigned)min(R, Bi - k0);
        const int cells = (int)kr * Bj;
        for (int c = tid; c < cells; c += 512) hist[c] = 0;
        __syncthreads();
        for (int v = tid; v < nvec; v += 512) {
            const uint4 a4 = xi[v], b4 = xj[v];
            const unsigned aw[4] = {a4.x, a4.y, a4.z, a4.w}, bw[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
            for (int q = 0; q < per; q++) {
                unsigned a, b;
                if constexpr (sizeof(I) == 1) {
                    a = (aw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                    b = (bw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                } else {
                    a = (aw[q >> 1] >> (16 * (q & 1))) & 0xFFFFu;
                    b = (bw[q >> 1] >> (16 * (q & 1))) & 0xFFFFu;
                }
                const unsigned r = a - (unsigned)k0;
                if (r < kr) atomicAdd(&hist[r * Bj + b], 1u);
            }
        }
        __syncthreads();
        if (dump_out) {
            int* out = dump_out + dump_off[p] + (size_t)k0 * Bj;
            for (int c = tid; c < cells; c += 512) out[c] = (int)hist[c];
        } else {
            for (int c = tid; c < cells; c += 512) T += in.nlogn[hist[c]];
        }
        __syncthreads();  // the next slab clears the counters
    }
    if (dump_out) return;
    T = wave_total(T);
    if ((tid & 63) == 0) wsum[tid >> 6] = T;
    __syncthreads();
    if (tid == 0) {
        double tot = 0.0;
        for (int w = 0; w < 8; w++) tot += wsum[w];
        const int N = in.N;
        in.s2d[par_offset(in.nD, gi, gj)] = ((in.logN + in.c2[gi]) + in.c2[gj] - 0.5 / N) - tot / N;
    }
}

hipError_t launch_pairs_slabs(const PairInputs& in, size_t lds_bytes, int max_rows, const CountsDump* dump,
                              hipStream_t s) {
    const int bmax = max(in.B.b, max(in.B.a, in.B.d));
    const int lds_words = (int)(lds_bytes / 4);
    if (bmax < 1 || lds_words < bmax || (in.idx_bytes != 1 && in.idx_bytes != 2)) return hipErrorInvalidValue;
    const void* fn = in.idx_bytes == 2 ? reinterpret_cast<const void*>(&k_pairs_slabs<uint16_t>)
                                       : reinterpret_cast<const void*>(&k_pairs_slabs<uint8_t>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    const long long total = dump ? dump->n : (long long)in.D * (in.D - 1) / 2;
    const long long chunk = 1LL << 23;  // grid x block stays below 2^32 work-items
    for (long long p0 = 0; p0 < total; p0 += chunk) {
        const long long p1 = p0 + chunk < total ? p0 + chunk : total;
        const int* pl = dump ? dump->pairs : nullptr;
        const long long* off = dump ? dump->offsets : nullptr;
        int* out = dump ? dump->out : nullptr;
        if (in.idx_bytes == 2)
            k_pairs_slabs<uint16_t><<<(unsigned)(p1 - p0), 512, lds_bytes, s>>>(in, p0, p1, lds_words, max_rows, pl, off, out);
        else
            k_pairs_slabs<uint8_t><<<(unsigned)(p1 - p0), 512, lds_bytes, s>>>(in, p0, p1, lds_words, max_rows, pl, off, out);
    }
    return hipGetLastError();
}

}  // namespace parent
