// One-hot int8 MFMA contraction for the 2D histograms (PARENT.cpp:774-1479).
//
// For dofs i, j with bin indices a_f, b_f over frames f, the histogram is
// counts_ij = A_i^T A_j with A_i[f][k] = (a_f == k): an int8 x int8 -> int32
// matrix product over the frame axis, exact below 2^31 frames.  A workgroup
// owns a 128 x 128 output tile: TD = floor(128 / B) row dofs x TD column dofs,
// rows/columns (dof, bin) packed dof by dof, the 128 - TD*B leftover rows zero.
//
// Per chunk of 128 frames the four waves run 4 k-steps of
// v_mfma_i32_32x32x32_i8 on 2 x 2 blocks each.  The one-hot operands come
// from one of two builders:
//   * up to 12 bins (PERM 1, 2): the chunk's bin indices are staged in LDS
//     (one 16-byte store per dof and 16 frames) and each lane turns the 16
//     index bytes of its operand row's dof into the row's one-hot bytes with
//     one v_perm_b32 table lookup per 4 frames; frame f of the chunk is k
//     slot f (from 9 bins, the rows of bins 8-11 read a second staged copy
//     with the indices remapped, so that no lane masks its bytes);
//   * 13 to 64 bins (PERM 0): every thread scatters the ones of its (dof,
//     16-frame group) units into LDS operand tiles (one byte store per
//     dof-frame) and clears its bytes again after the multiplies; frame
//     16g + e sits in k slot 32(e/4) + 4g + e%4, so the 8 frame groups of one
//     dof write 8 consecutive LDS words in each scatter instruction.
// Padding frames carry bin 255 and set nothing.  The accumulators then go to
// an LDS count tile, and one thread (up to 12 bins) or one wave reduces each
// pair with the fused entropy epilogue of the atomic kernels.
//
// Variant 1 (8 to 12 bins, fewer than 2^24 frames) builds FP4 operands for
// v_mfma_scale_f32_32x32x64_f8f6f4 instead: twice the k slots per
// instruction, one-hot nibbles (1.0 = 0b0010, unit E8M0 scales) and FP32
// accumulators, whose sums of 0/1 products are exact integers below 2^24.
//
// Lane layouts used (gfx950 32x32x32 int8): A operand lane l = row l % 32,
// 16 bytes = k slots 16*(l/32) .. +15; B operand lane l = column l % 32, the
// same k slots; result value v of lane l = row (v/4)*8 + (l/32)*4 + v%4,
// column l % 32.  The FP4 form: 32 k slots a lane, 32*(l/32) .. +31, in its
// first four operand registers; the same result layout.  Any bijection of k
// slots onto frames gives the same sums as long as A and B share it.
#include "kernels.h"

#include <type_traits>

namespace parent {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

// FP4 operands (e2m1) of v_mfma_scale_f32_32x32x64_f8f6f4: a one-hot value
// is 1.0 = 0b0010 in its nibble; the E8M0 scale 127 = 2^0 leaves products
// at 0 or 1, and FP32 sums of them are exact integers below 2^24.
constexpr unsigned kUnitScales = 0x7F7F7F7Fu;
constexpr int kFp4Format = 4;  // cbsz / blgp operand format code of FP4 (e2m1)

constexpr int kTile = 128;        // output rows / columns per workgroup
constexpr int kChunk = 128;       // frames per LDS operand chunk
constexpr int kStride = kChunk + 16;  // operand row stride in bytes (bank spread for 16-byte reads)
constexpr int kBuf = 2 * kTile * kStride;   // one operand buffer: the A tile then the B tile
constexpr int kOperandBytes = 2 * kBuf;     // double-buffered: chunk c+1 is scattered while chunk c multiplies
constexpr int kCountBytes = kTile * kTile * 4;
constexpr int kLds = kCountBytes > kOperandBytes ? kCountBytes : kOperandBytes;
constexpr int kMaxUnits = 4;      // (dof, frame group) units per thread: 2 * TD * 8 <= 1024
constexpr int kPermMaxBins = 12;  // the register-built operands (PERM) serve B <= 12
constexpr int kIdxStride = kChunk + 16;  // staged index row stride in bytes (PERM)
constexpr unsigned kPadIndex = 12;       // staged value of padding frames: v_perm selector 12 gives 0x00
// PERM: two staging buffers of both sides' index rows (at most 64 dofs a
// side; PERM 2, at most 14 dofs a side, adds a remapped copy of both), then
// the count tile's row phases in the same memory: at most max over B <= 12
// of ceil(64 / B) * B = 72 rows (B = 9 and 12)
constexpr int kPermStage = 2 * 2 * 64 * kIdxStride;
static_assert(2 * (2 * 2 * 14 * kIdxStride) + 64 <= kPermStage, "PERM 2 staging with its remapped copy");
constexpr int kPermCountRows = 72;
constexpr int kPermCountBytes = kPermCountRows * kTile * 4;
constexpr int kPermLds = kPermCountBytes > kPermStage ? kPermCountBytes : kPermStage;

// The 64-bit v_perm_b32 table {hi, lo} whose byte lookup turns a word of 4
// bin indices into the 4 one-hot bytes (index == k).  v_perm_b32 selector
// bytes 0-7 pick bytes of {hi, lo} (0-3 from lo), 8-11 replicate the top bit
// of byte 1, 3, 5 or 7, and 12 gives 0x00.  Bins 0-7 set byte k to 1; the
// other staged values (bins 8-11, padding 12) select 0x00: the table has no
// top bit set.  Rows of bins 8-11 read a second staged copy of the indices,
// remapped so that bins 8-11 become selectors 0-3 and everything else 4-7
// (remap_high_bins), with byte k - 8 of the table set.  A row outside the
// tile (ok false) gets the zero table.
// Returned as {lo, hi} by value: through pointers the tables went to scratch
// memory and every chunk then waited for all its global loads.
__device__ __forceinline__ uint2 onehot_table(bool ok, int k) {
    if (!ok) return make_uint2(0u, 0u);
    if (k < 4) return make_uint2(1u << (8 * k), 0u);
    if (k < 8) return make_uint2(0u, 1u << (8 * (k - 4)));
    return make_uint2(1u << (8 * (k - 8)), 0u);
}

// Staged index bytes (bins 0-11, padding 12) -> selectors of the bin 8-11
// rows: b - 8 for b in 8..11, (b & 7) | 4 (a zero byte of their tables) for
// the rest.
__device__ __forceinline__ unsigned remap_high_bins(unsigned x) {
    return (x & 0x07070707u) | ((~x & 0x08080808u) >> 1);
}

// Workgroup barrier that waits for the wave's LDS operations only (global
// loads in flight are not waited for, unlike __syncthreads()).  The memory
// clobber keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

}  // namespace

// PERM 1 (B <= 8) and 2 (B <= 12): each chunk's bin indices are staged in
// LDS once, one 16-byte store per (dof, 16 frames), and every lane builds its
// own MFMA operand row from them with one v_perm_b32 per 4 frames (PERM 2
// stages a remapped second copy for the rows of bins 8-11).  PERM 0: the
// one-hot bytes are scattered into LDS operand
// tiles and cleared again (every B <= 64).  UPT (PERM only): staging units
// per thread, ceil(2 * TD * 8 / 256); every thread issues all UPT loads, so
// the waits for a staged chunk count the loads in flight exactly.  PERM
// kernels with UPT < 4 fit 128 VGPRs and 36 KB of LDS: four workgroups per
// CU (UPT 4, at 2 and 3 bins, three: it would spill at 128).  FP4: the
// operands of v_mfma_scale_f32_32x32x64_f8f6f4 (64 frames a k-step, one-hot
// nibbles 0b0010 = 1.0, FP32 sums), built with two v_perm_b32 and a
// v_lshl_or_b32 per 8 frames; these kernels need 162 VGPRs and run three
// workgroups per CU (at 128 the accumulators went to scratch), for one
// staging unit per thread (8 to 12 bins) only.
template <bool BATCH, int PERM, int UPT, bool FP4 = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PERM ? (UPT < 4 && !FP4 ? 4 : 3) : 2))) void k_pairs_mfma(PairInputs in0, int B, int TD, TileQueue q,
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

    static_assert(!FP4 || PERM != 0, "FP4 operands are register-built");
    typedef typename std::conditional<FP4, v16f, v16i>::type acc_t;
    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[i][j][v] = 0;
    if constexpr (PERM != 0) {
        // Staging unit u = (side, dof slot, frame group of 16): one uint4 of
        // the dof's bin-index row per chunk, stored at its slot's row of the
        // side's staging block.  Dof slots past the last dof read dof D-1
        // (rows no epilogue or dump reads); units past 2 * TD * 8 load the
        // row dof's first group and store nothing.  Every dof is clamped to
        // D-1, so the empty padding tiles of a batch queue (r0 = c0 = D)
        // read inside the index matrix too.
        const int side_bytes = TD * kIdxStride, buf_bytes = 2 * side_bytes;
        // PERM 2: the remapped copy for the rows of bins 8-11, after both
        // buffers and 64 bytes further on, so that the copy of a dof sits 16
        // banks away from the dof's own row: a 32-row operand block reads both
        // for the same dofs (at an offset that is a multiple of the bank count
        // the two reads conflicted, 35% of LDS-active cycles at 10 bins).
        const int alt = PERM == 2 ? 2 * buf_bytes + 64 : 0;
        const int nunits = 2 * TD * 8;
        int ubase[UPT];
        const uint8_t* usrc[UPT];
#pragma unroll
        for (int k = 0; k < UPT; k++) {
            const int u = tid + k * 256;
            const int side = u / (TD * 8), rem = u % (TD * 8), d = rem / 8, g = rem % 8;
            const bool act = u < nunits;
            const int dof = min(act ? (side ? c0 : r0) + d : r0, D - 1);
            usrc[k] = in.bins + (size_t)dof * in.Np + (act ? g * 16 : 0);
            ubase[k] = act ? side * side_bytes + d * kIdxStride + g * 16 : -1;
        }
        const int nch = (N + kChunk - 1) / kChunk;
        auto load = [&](uint4* dst, int c) {
            const int f = min(c, nch - 1) * kChunk;
#pragma unroll
            for (int k = 0; k < UPT; k++) dst[k] = *reinterpret_cast<const uint4*>(usrc[k] + f);
        };
        // Only the last chunk holds padding frames (index 255, the only
        // index with its top bit set): they are staged as kPadIndex.
        auto pad = [](unsigned x) {
            const unsigned m = ((x >> 7) & 0x01010101u) * 0xFFu;
            return (x & ~m) | (kPadIndex * 0x01010101u & m);
        };
        auto stage = [&](int buf, const uint4* src, int c) {
            const bool check = (c + 1) * kChunk > N;
#pragma unroll
            for (int k = 0; k < UPT; k++) {
                if (ubase[k] < 0) continue;
                uint4 x = src[k];
                if (check) x = make_uint4(pad(x.x), pad(x.y), pad(x.z), pad(x.w));
                *reinterpret_cast<uint4*>(smem + buf * buf_bytes + ubase[k]) = x;
                if constexpr (PERM == 2)
                    *reinterpret_cast<uint4*>(smem + alt + buf * buf_bytes + ubase[k]) =
                        make_uint4(remap_high_bins(x.x), remap_high_bins(x.y), remap_high_bins(x.z),
                                   remap_high_bins(x.w));
            }
        };
        // The lane's operand rows: A rows wr*64 + i*32 + lane%32 of the row
        // side, B rows wc*64 + j*32 + lane%32 of the column side; row d*B + k
        // is the one-hot of bin k of dof slot d.  Lanes 32-63 take k slots
        // 16-31 of each 32-frame step (frame = k slot within the chunk).
        unsigned thi[2][2], tlo[2][2];
        int toff[2][2];
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++)
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int row = (s2 ? wc : wr) * 64 + i * 32 + (lane & 31);
                const bool ok = row < TD * B;
                const uint2 tab = onehot_table(ok, row % B);
                // (FP4: table bytes 0x02, the e2m1 1.0 in a low nibble)
                tlo[s2][i] = FP4 ? tab.x << 1 : tab.x;
                thi[s2][i] = FP4 ? tab.y << 1 : tab.y;
                toff[s2][i] = s2 * side_bytes + (ok ? row / B : 0) * kIdxStride + (lane >> 5) * (FP4 ? 32 : 16) +
                              (ok && row % B >= 8 ? alt : 0);
            }
        auto onehot = [&](int s2, int i, const uint4& x) {
            const unsigned words[4] = {x.x, x.y, x.z, x.w};
            v4i r;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                r[e] = (int)__builtin_amdgcn_perm(thi[s2][i], tlo[s2][i], words[e]);
            }
            return r;
        };
        // FP4: the lane's 32 frames of a 64-frame k-step, two words of 4
        // frames per operand register: byte b of register v holds frame
        // 8v + b in its low nibble and frame 8v + 4 + b in its high one (a
        // bijection of k slots onto frames, the same for A and B).
        auto onehot4 = [&](int s2, int i, const uint4& x0, const uint4& x1) {
            const unsigned w[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
            v8i r;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const unsigned lo = __builtin_amdgcn_perm(thi[s2][i], tlo[s2][i], w[2 * v]);
                const unsigned hi = __builtin_amdgcn_perm(thi[s2][i], tlo[s2][i], w[2 * v + 1]);
                r[v] = (int)(lo | (hi << 4));
                r[v + 4] = 0;
            }
            return r;
        };
        // Chunk c multiplies out of staging buffer c % 2; chunk c + 1 is
        // staged into the other one, which every wave finished reading before
        // the previous barrier.  The barrier waits for LDS traffic only.  The
        // loop is unrolled by two so that the register sets swap roles
        // without copies: x holds chunk c + 1 at even c, y at odd c, and each
        // is reloaded (chunk c + 3) right after it is staged, two chunks of
        // multiplies before it is needed.
        uint4 x[UPT], y[UPT];
        load(y, 0);
        stage(0, y, 0);
        load(x, 1);
        load(y, 2);
        lds_barrier();
        // Every step issues its loads, also the one past the last chunk when
        // nch is odd, so that all paths into the loop head have the same
        // loads in flight and each wait counts them exactly.
        auto step = [&](int c, uint4* next) {
            if (c < nch) {
                const unsigned char* st = smem + (c & 1) * buf_bytes;
                // (Issuing all 16 operand reads of the chunk first, pinned by a
                // scheduling barrier, needs 64 more VGPRs: at three waves per
                // SIMD it was 4% slower than this at four.)
                if constexpr (FP4) {
                    // B operands first, then one A operand at a time, each
                    // group fenced off from the next
                    auto build = [&](int s2, int i, int ks) {
                        const int off = toff[s2][i] + ks * 64;
                        const uint4* p = reinterpret_cast<const uint4*>(st + off);
                        return onehot4(s2, i, p[0], p[1]);
                    };
#pragma unroll
                    for (int ks = 0; ks < kChunk / 64; ks++) {
                        const v8i b0 = build(1, 0, ks), b1 = build(1, 1, ks);
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int i = 0; i < 2; i++) {
                            const v8i a = build(0, i, ks);
                            acc[i][0] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                                a, b0, acc[i][0], kFp4Format, kFp4Format, 0, kUnitScales, 0, kUnitScales);
                            acc[i][1] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                                a, b1, acc[i][1], kFp4Format, kFp4Format, 0, kUnitScales, 0, kUnitScales);
                            __builtin_amdgcn_sched_barrier(0);
                        }
                    }
                } else {
#pragma unroll
                    for (int ks = 0; ks < kChunk / 32; ks++) {
                        v4i a[2], b[2];
#pragma unroll
                        for (int i = 0; i < 2; i++) a[i] = onehot(0, i, *reinterpret_cast<const uint4*>(st + toff[0][i] + ks * 32));
#pragma unroll
                        for (int j = 0; j < 2; j++) b[j] = onehot(1, j, *reinterpret_cast<const uint4*>(st + toff[1][j] + ks * 32));
#pragma unroll
                        for (int i = 0; i < 2; i++)
#pragma unroll
                            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[i], b[j], acc[i][j], 0, 0, 0);
                    }
                }
                if (c + 1 < nch) stage((c + 1) & 1, next, c + 1);
            }
            load(next, c + 3);
            lds_barrier();
        };
        for (int c = 0; c < nch; c += 2) {
            step(c, x);
            step(c + 1, y);
        }
    } else {
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
            const int dof = min(uact[k] ? (side ? c0 : r0) + d : r0, D - 1);
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
    }
    // Counts to LDS, then the epilogue.  The register-built kernels (PERM)
    // hold the count tile in two row phases, dof slots [0, D0) then [D0, TD)
    // with D0 = ceil(64 / B), each at most kPermCountRows rows, so that their
    // LDS is 36 KB and three workgroups share a CU.  PERM 0 writes the whole
    // 128 x 128 tile in one phase.
    constexpr int kPhases = PERM ? 2 : 1;
    const int D0 = PERM ? (64 + B - 1) / B : TD;
    for (int h = 0; h < kPhases; h++) {
        const int dlo = h ? D0 : 0, dhi = h ? TD : D0;
        const int rlo = dlo * B, rhi = h == kPhases - 1 ? kTile : D0 * B;
        __syncthreads();  // the staging buffers or the previous phase are no longer read
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int row = wr * 64 + i * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                    const int col = wc * 64 + j * 32 + (lane & 31);
                    if (row >= rlo && row < rhi) cnt[(row - rlo) * kTile + col] = (int)acc[i][j][v];
                }
        __syncthreads();

        if (dump_out) {
            for (int e = dump_begin[t]; e < dump_begin[t + 1]; e++) {
                const int sl = dump_entries[3 * e];
                const long long off = ((long long)dump_entries[3 * e + 1] << 31) | (long long)dump_entries[3 * e + 2];
                const int di = sl / TD, dj = sl % TD;
                if (di < dlo || di >= dhi) continue;
                const int Bi = in.B.of(dof_type(nD, r0 + di)), Bj = in.B.of(dof_type(nD, c0 + dj));
                for (int c = tid; c < Bi * Bj; c += 256)
                    dump_out[off + c] = cnt[((di - dlo) * B + c / Bj) * kTile + dj * B + c % Bj];
            }
            continue;
        }
        if constexpr (PERM != 0) {
            // Up to 12 bins (at most 144 cells a pair, 144 to 4,096 pairs a
            // tile): one thread per pair, cells in row order, its table loads
            // in flight together.  A wave per pair left most lanes idle and
            // waited out one LDS read and one table load per pair.
            for (int p = tid; p < (dhi - dlo) * TD; p += 256) {
                const int di = dlo + p / TD, dj = p % TD;
                const int gi = r0 + di, gj = c0 + dj;
                if (!(gi < gj && gj < D)) continue;
                const int* cell = cnt + (di - dlo) * B * kTile + dj * B;
                double T = 0.0;
                for (int r = 0; r < B; r++, cell += kTile) {
#pragma unroll 4
                    for (int c = 0; c < B; c++) T += in.nlogn[cell[c]];
                }
                in.s2d[par_offset(nD, gi, gj)] = ((in.logN + in.c2[gi]) + in.c2[gj] - 0.5 / N) - T / N;
            }
        } else {
            for (int p = wid; p < (dhi - dlo) * TD; p += 4) {
                const int di = dlo + p / TD, dj = p % TD;
                const int gi = r0 + di, gj = c0 + dj;
                if (!(gi < gj && gj < D)) continue;
                double T = 0.0;
                // branch-free (the table is 0 at n = 0): the unrolled table loads overlap
#pragma unroll 4
                for (int c = lane; c < B * B; c += 64) T += in.nlogn[cnt[((di - dlo) * B + c / B) * kTile + dj * B + c % B]];
                T = wave_total(T);
                if (lane == 0) in.s2d[par_offset(nD, gi, gj)] = ((in.logN + in.c2[gi]) + in.c2[gj] - 0.5 / N) - T / N;
            }
        }
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

// FP4 sums are FP32: exact while every count stays below 2^24.  The FP4
// kernels exist for one staging unit per thread (8 to 12 bins): with two,
// their operand builds spill even at three waves per SIMD.
bool mfma_variant_fits(int variant, int bmax, int n_frames) {
    if (mfma_tile_dofs(bmax) < 1) return false;
    if (variant == 0) return true;
    return variant == 1 && bmax >= 8 && bmax <= kPermMaxBins && n_frames < (1 << 24);
}

int mfma_tiles_per_launch() { return tiles_per_launch(256); }

hipError_t launch_pairs_mfma(const PairInputs& in, int bins, int variant, const TileQueue& q, const int* dump_begin,
                             const int* dump_entries, int* dump_out, hipStream_t s) {
    const int ntiles = q.ntiles;
    const int TD = mfma_tile_dofs(bins);
    if (TD < 1) return hipErrorInvalidValue;
    const int perm = bins <= 8 ? 1 : bins <= kPermMaxBins ? 2 : 0;
    const int upt = (2 * TD * 8 + 255) / 256;  // 1, 2, 3 or 4; 3 runs as 4
    const bool batch = q.descs != nullptr;
    // variant 1: FP4 operands, register-built kernels of one staging unit per thread
    const bool fp4 = variant == 1;
    if (variant < 0 || variant > 1 || (fp4 && (perm == 0 || upt != 1))) return hipErrorInvalidValue;
    // kernel index: 0 scatter, 1-3 PERM 1 with 1, 2, 4 units per thread, 4 PERM 2 (TD <= 14: 1 unit)
    const int kid = perm == 0 ? 0 : perm == 2 ? 4 : upt == 1 ? 1 : upt == 2 ? 2 : 3;
#define PARENT_MFMA_FN(BT, PM, U, F) reinterpret_cast<const void*>(&k_pairs_mfma<BT, PM, U, F>)
    const void* fns[2][2][5] = {
        {{PARENT_MFMA_FN(false, 0, 4, false), PARENT_MFMA_FN(false, 1, 1, false), PARENT_MFMA_FN(false, 1, 2, false),
          PARENT_MFMA_FN(false, 1, 4, false), PARENT_MFMA_FN(false, 2, 1, false)},
         {nullptr, PARENT_MFMA_FN(false, 1, 1, true), nullptr, nullptr, PARENT_MFMA_FN(false, 2, 1, true)}},
        {{PARENT_MFMA_FN(true, 0, 4, false), PARENT_MFMA_FN(true, 1, 1, false), PARENT_MFMA_FN(true, 1, 2, false),
          PARENT_MFMA_FN(true, 1, 4, false), PARENT_MFMA_FN(true, 2, 1, false)},
         {nullptr, PARENT_MFMA_FN(true, 1, 1, true), nullptr, nullptr, PARENT_MFMA_FN(true, 2, 1, true)}}};
#undef PARENT_MFMA_FN
    if (perm == 2 && upt != 1) return hipErrorInvalidValue;
    if (perm && (64 + bins - 1) / bins * bins > kPermCountRows) return hipErrorInvalidValue;
    const int lds = perm ? kPermLds : kLds;
    hipError_t e = hipFuncSetAttribute(fns[batch][fp4][kid], hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
    const int per_launch = mfma_tiles_per_launch();
    for (int t0 = 0; t0 < ntiles; t0 += per_launch) {
        const int t1 = t0 + per_launch < ntiles ? t0 + per_launch : ntiles;
        const dim3 grid((t1 - t0 + 7) / 8 * 8);
#define PARENT_MFMA_LAUNCH(BT, PM, U, F) \
    k_pairs_mfma<BT, PM, U, F><<<grid, 256, lds, s>>>(in, bins, TD, q, t0, t1, dump_begin, dump_entries, dump_out)
#define PARENT_MFMA_SWITCH(BT)                                   \
    if (fp4) {                                                   \
        if (kid == 4) PARENT_MFMA_LAUNCH(BT, 2, 1, true);        \
        else PARENT_MFMA_LAUNCH(BT, 1, 1, true);                 \
    } else {                                                     \
        switch (kid) {                                           \
            case 1: PARENT_MFMA_LAUNCH(BT, 1, 1, false); break;  \
            case 2: PARENT_MFMA_LAUNCH(BT, 1, 2, false); break;  \
            case 3: PARENT_MFMA_LAUNCH(BT, 1, 4, false); break;  \
            case 4: PARENT_MFMA_LAUNCH(BT, 2, 1, false); break;  \
            default: PARENT_MFMA_LAUNCH(BT, 0, 4, false); break; \
        }                                                        \
    }
        if (batch) {
            PARENT_MFMA_SWITCH(true)
        } else {
            PARENT_MFMA_SWITCH(false)
        }
#undef PARENT_MFMA_SWITCH
#undef PARENT_MFMA_LAUNCH
    }
    return hipGetLastError();
}

}  // namespace parent
