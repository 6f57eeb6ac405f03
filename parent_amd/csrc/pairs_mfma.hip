// One-hot int8 MFMA contraction for the 2D histograms (PARENT.cpp:774-1479).
//
// For dofs i, j with bin indices a_f, b_f over frames f, the histogram is
// counts_ij = A_i^T A_j with A_i[f][k] = (a_f == k): an int8 x int8 -> int32
// matrix product over the frame axis, exact below 2^31 frames.  A workgroup
// owns a 128 x 128 output tile: TD = floor(128 / B) row dofs x TD column dofs,
// rows/columns (dof, bin) packed dof by dof, the 128 - TD*B leftover rows zero.
//
// The one-hot operands are never compared lane by lane: for each chunk of 128
// frames every thread scatters the ones of its (dof, 16-frame group) units
// into the LDS operand tiles (one byte store per dof-frame), the four waves
// run 4 k-steps of v_mfma_i32_32x32x32_i8 on 2 x 2 blocks each, and the same
// thread clears its bytes again.  Padding frames carry bin 255 and set
// nothing.  The accumulators then go to an LDS count tile and each wave
// reduces whole pairs with the fused entropy epilogue of the atomic kernels.
//
// Lane layouts used (gfx950 32x32x32 int8): A operand lane l = row l % 32,
// 16 bytes = k slots 16*(l/32) .. +15; B operand lane l = column l % 32, the
// same k slots; result value v of lane l = row (v/4)*8 + (l/32)*4 + v%4,
// column l % 32.  Any bijection of k slots onto frames gives the same sums as
// long as A and B share it, which they do here: frame 16g + e of a chunk sits
// in k slot 32(e/4) + 4g + e%4, so the 8 frame groups of one dof write 8
// consecutive LDS words in each scatter instruction (bank-spread stores).
#include "kernels.h"

namespace parent {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kTile = 128;        // output rows / columns per workgroup
constexpr int kChunk = 128;       // frames per LDS operand chunk
constexpr int kStride = kChunk + 16;  // operand row stride in bytes (bank spread for 16-byte reads)
constexpr int kBuf = 2 * kTile * kStride;   // one operand buffer: the A tile then the B tile
constexpr int kOperandBytes = 2 * kBuf;     // double-buffered: chunk c+1 is scattered while chunk c multiplies
constexpr int kCountBytes = kTile * kTile * 4;
constexpr int kLds = kCountBytes > kOperandBytes ? kCountBytes : kOperandBytes;
constexpr int kMaxUnits = 4;      // (dof, frame group) units per thread: 2 * TD * 8 <= 1024

// Workgroup barrier that waits for the wave's LDS operations only (global
// loads in flight are not waited for, unlike __syncthreads()).  The memory
// clobber keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

}  // namespace

template <bool BATCH>
__global__ __launch_bounds__(256, 2) void k_pairs_mfma(PairInputs in0, int B, int TD, TileQueue q,
                                                       int t_begin, int t_end, const int* __restrict__ dump_begin,
                                                       const int* __restrict__ dump_entries, int* __restrict__ dump_out) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    int* cnt = reinterpret_cast<int*>(smem);

    const int chunk_t = (t_end - t_begin + 7) / 8;
    const int t = t_begin + (int)(blockIdx.x % 8) * chunk_t + (int)(blockIdx.x / 8);
    if (t >= t_end || (int)(blockIdx.x / 8) >= chunk_t) return;
    const int* tp = q.tiles + (size_t)(BATCH ? 3 : 2) * t;
    const int r0 = tp[0], c0 = tp[1];
    const PairInputs in = BATCH ? q.descs[tp[2]] : in0;
    const int D = in.D, N = in.N, nD = in.nD;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wr = wid >> 1, wc = wid & 1;

    for (int i = tid; i < kOperandBytes / 16; i += 256) reinterpret_cast<uint4*>(smem)[i] = make_uint4(0, 0, 0, 0);

    // Static unit assignment: unit u = (side, dof slot, frame group of 16).
    // Units of dof slots past the last dof read dof D-1 and fill operand rows
    // that no epilogue or dump reads; units past the table (u >= nunits, only
    // in the last partial wave) are masked.
    const int nunits = 2 * TD * 8;
    int ubase[kMaxUnits];            // byte offset of the unit's (row block, frame group) origin
    const uint8_t* usrc[kMaxUnits];  // the dof's bin-index row at the unit's frame group
    bool uact[kMaxUnits];
#pragma unroll
    for (int k = 0; k < kMaxUnits; k++) {
        const int u = tid + k * 256;
        const int side = u / (TD * 8), rem = u % (TD * 8), d = rem / 8, g = rem % 8;
        uact[k] = u < nunits;
        const int dof = uact[k] ? min((side ? c0 : r0) + d, D - 1) : r0;
        usrc[k] = in.bins + (size_t)dof * in.Np + g * 16;
        ubase[k] = uact[k] ? (side ? kTile * kStride : 0) + d * B * kStride + g * 4 : 0;
    }
    const int nch = (N + kChunk - 1) / kChunk;  // chunks holding real frames
    auto load = [&](uint4* dst, int c) {
        const int f = min(c, nch - 1) * kChunk;
#pragma unroll
        for (int k = 0; k < kMaxUnits; k++)
            if (k * 256 < nunits) dst[k] = *reinterpret_cast<const uint4*>(usrc[k] + f);
    };
    // Set (val 1) or clear (val 0) the ones of the thread's units of chunk c
    // in buffer `boff`.  Only the last chunk can hold padding frames (bin
    // index 255), so only it tests the index.
    auto scatter = [&](int boff, const uint4* src, int c, unsigned char val) {
        const bool check = (c + 1) * kChunk > N;
#pragma unroll
        for (int k = 0; k < kMaxUnits; k++) {
            if (k * 256 >= nunits || !uact[k]) continue;
            const unsigned words[4] = {src[k].x, src[k].y, src[k].z, src[k].w};
            unsigned char* base = smem + boff + ubase[k];
            if (!check) {
#pragma unroll
                for (int e = 0; e < 16; e++) {
                    const unsigned b = (words[e >> 2] >> (8 * (e & 3))) & 0xFFu;
                    base[__umul24(b, kStride) + (e >> 2) * 32 + (e & 3)] = val;
                }
            } else {
#pragma unroll
                for (int e = 0; e < 16; e++) {
                    const unsigned b = (words[e >> 2] >> (8 * (e & 3))) & 0xFFu;
                    if (b < (unsigned)B) base[__umul24(b, kStride) + (e >> 2) * 32 + (e & 3)] = val;
                }
            }
        }
    };
    uint4 w[kMaxUnits], wn[kMaxUnits], wn2[kMaxUnits];
    load(w, 0);
    load(wn, 1);
    lds_barrier();
    scatter(0, w, 0, 1);
    lds_barrier();

    v16i acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[i][j][v] = 0;

    // Chunk c multiplies out of buffer c % 2 while chunk c + 1 is scattered
    // into the other buffer; after the barrier each thread clears its own
    // bytes of chunk c (no other thread writes them), so every buffer is
    // clean again before its next scatter.  The barriers wait for LDS traffic
    // only: the global loads of chunk c + 2 stay in flight across them.
    for (int c = 0; c < nch; c++) {
        const int p = (c & 1) * kBuf, q = kBuf - p;
        const unsigned char* opA = smem + p;
        const unsigned char* opB = smem + p + kTile * kStride;
#pragma unroll
        for (int ks = 0; ks < kChunk / 32; ks++) {
            v4i a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; i++)
                a[i] = *reinterpret_cast<const v4i*>(opA + (wr * 64 + i * 32 + (lane & 31)) * kStride + ks * 32 + (lane >> 5) * 16);
#pragma unroll
            for (int j = 0; j < 2; j++)
                b[j] = *reinterpret_cast<const v4i*>(opB + (wc * 64 + j * 32 + (lane & 31)) * kStride + ks * 32 + (lane >> 5) * 16);
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (c + 1 < nch) scatter(q, wn, c + 1, 1);
        load(wn2, c + 2);
        lds_barrier();
        scatter(p, w, c, 0);
#pragma unroll
        for (int k = 0; k < kMaxUnits; k++) {
            w[k] = wn[k];
            wn[k] = wn2[k];
        }
    }
    __syncthreads();

    // counts to LDS: count tile [128][128] int32
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int row = wr * 64 + i * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                const int col = wc * 64 + j * 32 + (lane & 31);
                cnt[row * kTile + col] = acc[i][j][v];
            }
    __syncthreads();

    if (dump_out) {
        for (int e = dump_begin[t]; e < dump_begin[t + 1]; e++) {
            const int sl = dump_entries[3 * e];
            const long long off = ((long long)dump_entries[3 * e + 1] << 31) | (long long)dump_entries[3 * e + 2];
            const int di = sl / TD, dj = sl % TD;
            const int Bi = in.B.of(dof_type(nD, r0 + di)), Bj = in.B.of(dof_type(nD, c0 + dj));
            for (int c = tid; c < Bi * Bj; c += 256) dump_out[off + c] = cnt[(di * B + c / Bj) * kTile + dj * B + c % Bj];
        }
        return;
    }
    for (int p = wid; p < TD * TD; p += 4) {
        const int di = p / TD, dj = p % TD;
        const int gi = r0 + di, gj = c0 + dj;
        if (!(gi < gj && gj < D)) continue;
        double T = 0.0;
        // branch-free (the table is 0 at n = 0): the unrolled table loads overlap
#pragma unroll 4
        for (int c = lane; c < B * B; c += 64) T += in.nlogn[cnt[(di * B + c / B) * kTile + dj * B + c % B]];
        T = wave_total(T);
        if (lane == 0) in.s2d[par_offset(nD, gi, gj)] = ((in.logN + in.c2[gi]) + in.c2[gj] - 0.5 / N) - T / N;
    }
}

// At least 2 dofs a side (B <= 64: above that the one-hot product spends more
// MFMA work per update than the LDS-atomic kernels), at most 64 (the unit
// table of 4 units per thread).
int mfma_tile_dofs(int bins) {
    const int td = bins >= 1 ? kTile / bins : 0;
    return td >= 2 && 2 * td * 8 <= kMaxUnits * 256 ? td : 0;
}

size_t mfma_lds_bytes() { return kLds; }

int mfma_tiles_per_launch() { return tiles_per_launch(256); }

hipError_t launch_pairs_mfma(const PairInputs& in, int bins, const TileQueue& q, const int* dump_begin,
                             const int* dump_entries, int* dump_out, hipStream_t s) {
    const int ntiles = q.ntiles;
    const int TD = mfma_tile_dofs(bins);
    if (TD < 1) return hipErrorInvalidValue;
    const void* fn = q.descs ? reinterpret_cast<const void*>(&k_pairs_mfma<true>)
                             : reinterpret_cast<const void*>(&k_pairs_mfma<false>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kLds);
    if (e != hipSuccess) return e;
    const int per_launch = mfma_tiles_per_launch();
    for (int t0 = 0; t0 < ntiles; t0 += per_launch) {
        const int t1 = t0 + per_launch < ntiles ? t0 + per_launch : ntiles;
        const int chunk = (t1 - t0 + 7) / 8;
        if (q.descs)
            k_pairs_mfma<true><<<chunk * 8, 256, kLds, s>>>(in, bins, TD, q, t0, t1, dump_begin, dump_entries, dump_out);
        else
            k_pairs_mfma<false><<<chunk * 8, 256, kLds, s>>>(in, bins, TD, q, t0, t1, dump_begin, dump_entries, dump_out);
    }
    return hipGetLastError();
}

}  // namespace parent
