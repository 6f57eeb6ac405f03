// 2D histogram + entropy kernels built on LDS atomics (PARENT.cpp:774-1479).
//
// Both kernels read the dof-major uint8 bin indices written by the pre-bin
// pass, so the reference's two FP64 divisions per pair-frame (PARENT.cpp:803)
// happen once per dof-frame instead.  Counts are exact uint32 (N < 2^32).
//
//  k_pairs_plain   one dof pair per workgroup; per-cell FP64 epilogue in the
//                  reference's form p = n/(N d1 d2 J1 J2), S = -sum J1 J2 p ln p
//                  d1 d2 + (occ-1)/(2N).  The baseline GPU kernel.
//  k_pairs_atomic  a TR x TC tile of dof pairs per workgroup: the tile's row
//                  and column index rows are streamed once and feed TR*TC LDS
//                  histograms; fused epilogue S = ln N + c_i + c_j - 1/(2N)
//                  - (1/N) sum (n ln n - 1/2 [n > 0]) over an FP64 table of
//                  that summand (the occupancy term (occ-1)/(2N) folded in).
#include "kernels.h"

#include <algorithm>
#include <type_traits>

namespace parent {

namespace {

// Workgroup barrier that waits for this wave's LDS operations only: global
// loads in flight stay in flight (__syncthreads() would wait for them too).
// The memory clobber keeps LDS accesses from moving across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

__device__ __forceinline__ unsigned byte_of(const uint4& w, int e) {
    const unsigned v = e < 4 ? w.x : (e < 8 ? w.y : (e < 12 ? w.z : w.w));
    return (v >> (8 * (e & 3))) & 0xFFu;
}

// (byte E of w) << amount, and 1 << (byte E of s): one v_lshlrev_b32 with a
// byte-select (SDWA) operand each.  Written out because the compiler emits a
// shift + mask pair for some bytes.  Both operands must be VGPRs.
#define PARENT_SDWA_SHL(E)                                                                             \
    if constexpr (SEL == E) {                                                                          \
        asm("v_lshlrev_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD "        \
            "src1_sel:BYTE_" #E                                                                        \
            : "=v"(r)                                                                                  \
            : "v"(amount), "v"(w));                                                                    \
    }
template <int SEL>
__device__ __forceinline__ unsigned byte_shl(unsigned w, unsigned amount) {
    unsigned r;
    PARENT_SDWA_SHL(0) PARENT_SDWA_SHL(1) PARENT_SDWA_SHL(2) PARENT_SDWA_SHL(3)
    return r;
}
#undef PARENT_SDWA_SHL
#define PARENT_SDWA_ONE_SHL(E)                                                                         \
    if constexpr (SEL == E) {                                                                          \
        asm("v_lshlrev_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_" #E      \
            " src1_sel:DWORD"                                                                          \
            : "=v"(r)                                                                                  \
            : "v"(s), "v"(one));                                                                       \
    }
template <int SEL>
__device__ __forceinline__ unsigned one_shl_byte(unsigned s, unsigned one) {
    unsigned r;
    PARENT_SDWA_ONE_SHL(0) PARENT_SDWA_ONE_SHL(1) PARENT_SDWA_ONE_SHL(2) PARENT_SDWA_ONE_SHL(3)
    return r;
}
#undef PARENT_SDWA_ONE_SHL

// low halves a.lo + b.lo, high halves a.hi + b.hi: one v_add_u32 with
// halfword-select (SDWA) operands each
__device__ __forceinline__ unsigned add_lo16(unsigned a, unsigned b) {
    unsigned r;
    asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:WORD_0 src1_sel:WORD_0"
        : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ unsigned add_hi16(unsigned a, unsigned b) {
    unsigned r;
    asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:WORD_1 src1_sel:WORD_1"
        : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// operand type of v_dot2_u32_u16 (__builtin_amdgcn_udot2)
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// halves added separately, modulo 2^16 each: one v_pk_add_u16.  Written
// out: as 16-bit vector arithmetic the compiler splits the 16-byte LDS reads
// that feed it into 12 + 4 bytes.
__device__ __forceinline__ unsigned pk_add16(unsigned a, unsigned b) {
    unsigned r;
    asm("v_pk_add_u16 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// Largest N at which a sum of 8 (4) 16-bit sub-counters of the sliced
// kernels stays below 2^16: each holds at most N/16 + 16 counts.
constexpr int kPk8Frames = (65535 / 8 - 16) * 16;
constexpr int kPk4Frames = (65535 / 4 - 16) * 16;
// an LDS word addressed by its 32-bit LDS byte address
typedef __attribute__((address_space(3))) unsigned lds_u32;

// Raw buffer over one dof's index row (Np bytes): 16-byte loads at a
// per-thread byte offset plus a uniform (SGPR) offset need no VALU address
// math, and the hardware returns 0 past the row's end.  Word 3 is the raw
// 32-bit data format of gfx9 buffer resources.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t index_row(const uint8_t* row, int Np) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(row), (short)0, Np, 0x00020000);
}
__device__ __forceinline__ uint4 load16(__amdgpu_buffer_rsrc_t rs, int voff, int soff) {
    return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0));
}

}  // namespace

// ---------------------------------------------------------------------------
// plain: one pair per workgroup
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pairs_plain(PairInputs in, long long p0, long long p1,
                                                     const int* __restrict__ plist,
                                                     const long long* __restrict__ dump_off,
                                                     int* __restrict__ dump_out) {
    extern __shared__ unsigned hist[];
    __shared__ double rs[256];
    __shared__ int ro[256];
    const long long p = p0 + blockIdx.x;
    if (p >= p1) return;
    int gi, gj;
    if (plist) {
        gi = plist[2 * p];
        gj = plist[2 * p + 1];
    } else {
        pair_of_index(p, in.D, &gi, &gj);
    }
    const int ti = dof_type(in.nD, gi), tj = dof_type(in.nD, gj);
    const int Bi = in.B.of(ti), Bj = in.B.of(tj);
    const int cells = Bi * Bj;
    for (int c = threadIdx.x; c < cells; c += 256) hist[c] = 0;
    __syncthreads();
    const uint32_t* xi = reinterpret_cast<const uint32_t*>(in.bins + (size_t)gi * in.Np);
    const uint32_t* xj = reinterpret_cast<const uint32_t*>(in.bins + (size_t)gj * in.Np);
    const int N = in.N;
    const int nwords = (N + 3) / 4;
    for (int w = threadIdx.x; w < nwords; w += 256) {
        const uint32_t a4 = xi[w], b4 = xj[w];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            if (4 * w + e < N)
                atomicAdd(&hist[((a4 >> (8 * e)) & 0xFFu) * Bj + ((b4 >> (8 * e)) & 0xFFu)], 1u);
        }
    }
    __syncthreads();
    if (dump_out) {
        int* out = dump_out + dump_off[p];
        for (int c = threadIdx.x; c < cells; c += 256) out[c] = (int)hist[c];
        return;
    }
    const double b1 = (in.vmax[gi] - in.vmin[gi]) / Bi;
    const double b2 = (in.vmax[gj] - in.vmin[gj]) / Bj;
    const double* J1 = in.jtab + (size_t)gi * in.jstride;
    const double* J2 = in.jtab + (size_t)gj * in.jstride;
    const int numFrames = N;
    double sum = 0.0;
    int occ = 0;
    for (int c = threadIdx.x; c < cells; c += 256) {
        const unsigned n = hist[c];
        if (n) {
            const int k = c / Bj, l = c - k * Bj;
            const double j1 = J1[k], j2 = J2[l];
            const double probDens = n / (numFrames * b1 * b2 * j1 * j2);
            if (probDens > 0) {
                sum = sum + j1 * j2 * probDens * log(probDens) * b1 * b2;
                occ = occ + 1;
            }
        }
    }
    rs[threadIdx.x] = sum;
    ro[threadIdx.x] = occ;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            rs[threadIdx.x] += rs[threadIdx.x + s];
            ro[threadIdx.x] += ro[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        in.s2d[par_offset(in.nD, gi, gj)] = -rs[0] + (ro[0] - 1.0) / (2.0 * numFrames);
}

// dynamic histogram + the static reduction arrays (rs, ro)
size_t plain_lds_bytes(int bmax) { return (size_t)bmax * bmax * sizeof(unsigned) + 256 * (sizeof(double) + sizeof(int)); }

hipError_t launch_pairs_plain(const PairInputs& in, const CountsDump* dump, hipStream_t s) {
    const int bmax = max(in.B.b, max(in.B.a, in.B.d));
    const size_t lds = (size_t)bmax * bmax * sizeof(unsigned);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pairs_plain),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const long long total = dump ? dump->n : (long long)in.D * (in.D - 1) / 2;
    const long long chunk = 1LL << 23;
    for (long long p0 = 0; p0 < total; p0 += chunk) {
        const long long p1 = p0 + chunk < total ? p0 + chunk : total;
        k_pairs_plain<<<(unsigned)(p1 - p0), 256, lds, s>>>(in, p0, p1, dump ? dump->pairs : nullptr,
                                                            dump ? dump->offsets : nullptr,
                                                            dump ? dump->out : nullptr);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// tiled: TR x TC pairs per workgroup
// ---------------------------------------------------------------------------
// tiles: [ntiles][2] = (first row dof, first column dof).  Block b of a launch
// takes tile (b % 8) * chunk + b / 8 of its range, so the workgroups that the
// dispatcher places on one XCD (round robin over 8 XCDs) walk one contiguous
// run of tiles and share its row-dof index rows in that XCD's L2.
//
// SLOTW is the compile-time word stride between the tile's histograms
// (>= bins_i * bins_j), so a histogram's base folds into the ds_add_u32
// immediate offset.  Tiles whose TR*TC pairs are all valid and whose columns
// share one bin count (every tile off the diagonal band when the 2D bin counts
// are equal) take a branch-free fill: per frame one 24-bit multiply per row
// and one add + one LDS atomic per pair.
template <int TR, int TC, int BLOCK, int SLOTW, bool BATCH>
__global__ __launch_bounds__(BLOCK) void k_pairs_atomic(PairInputs in0, TileQueue q,
                                                        int t_begin, int t_end,
                                                        const int* __restrict__ dump_begin,
                                                        const int* __restrict__ dump_entries,
                                                        int* __restrict__ dump_out) {
    extern __shared__ unsigned hist[];
    const int chunk = (t_end - t_begin + 7) / 8;
    const int t = t_begin + (int)(blockIdx.x % 8) * chunk + (int)(blockIdx.x / 8);
    if (t >= t_end || (int)(blockIdx.x / 8) >= chunk) return;
    const int* tp = q.tiles + (size_t)(BATCH ? 3 : 2) * t;
    const int r0 = tp[0], c0 = tp[1];
    const PairInputs in = BATCH ? q.descs[tp[2]] : in0;
    const int D = in.D, N = in.N, nD = in.nD;

    int gr[TR], gc[TC], Bc[TC];
#pragma unroll
    for (int r = 0; r < TR; r++) gr[r] = r0 + r;
#pragma unroll
    for (int c = 0; c < TC; c++) {
        gc[c] = c0 + c;
        Bc[c] = gc[c] < D ? in.B.of(dof_type(nD, gc[c])) : 1;
    }
    unsigned valid = 0;  // bit r*TC+c
    bool uniform = true;
#pragma unroll
    for (int r = 0; r < TR; r++)
#pragma unroll
        for (int c = 0; c < TC; c++)
            if (gr[r] < D && gc[c] < D && gr[r] < gc[c]) valid |= 1u << (r * TC + c);
#pragma unroll
    for (int c = 1; c < TC; c++) uniform = uniform && Bc[c] == Bc[0];
    const bool fast = uniform && valid == (1u << (TR * TC)) - 1u;

    const int per_iter = BLOCK * 16;
    const int niter = (N + per_iter - 1) / per_iter;
    const int nfull = N / per_iter;
    const uint8_t* rows[TR];
    const uint8_t* cols[TC];
#pragma unroll
    for (int r = 0; r < TR; r++) rows[r] = in.bins + (size_t)(gr[r] < D ? gr[r] : 0) * in.Np;
#pragma unroll
    for (int c = 0; c < TC; c++) cols[c] = in.bins + (size_t)(gc[c] < D ? gc[c] : 0) * in.Np;
    // index words of the fast fill's first iteration, requested before the
    // clear; the barrier after it waits for LDS operations only
    auto load = [&](uint4* R, uint4* C, int it) {
        const int f = it * per_iter + threadIdx.x * 16;
        if (f < in.Np) {
#pragma unroll
            for (int r = 0; r < TR; r++) R[r] = *reinterpret_cast<const uint4*>(rows[r] + f);
#pragma unroll
            for (int c = 0; c < TC; c++) C[c] = *reinterpret_cast<const uint4*>(cols[c] + f);
        }
    };
    uint4 R0[TR], C0[TC];
    if (fast && nfull > 0) load(R0, C0, 0);

    for (int i = threadIdx.x; i < TR * TC * SLOTW; i += BLOCK) hist[i] = 0;
    lds_barrier();

    int it0 = 0;
    bool tail_fast = false;
    if (fast) {
        const unsigned B = (unsigned)Bc[0];
        auto process = [&](const uint4* R, const uint4* C) {
#pragma unroll
            for (int e = 0; e < 16; e++) {
                unsigned ro[TR];
#pragma unroll
                for (int r = 0; r < TR; r++) ro[r] = __umul24(byte_of(R[r], e), B);
#pragma unroll
                for (int r = 0; r < TR; r++)
#pragma unroll
                    for (int c = 0; c < TC; c++) atomicAdd(&hist[(r * TC + c) * SLOTW + ro[r] + byte_of(C[c], e)], 1u);
            }
        };
        // two register sets: the next iteration's words load while this
        // iteration's atomics issue
        uint4 R1[TR], C1[TC];
        int it = 0;
        for (; it + 2 <= nfull; it += 2) {
            load(R1, C1, it + 1);
            process(R0, C0);
            load(R0, C0, it + 2);
            process(R1, C1);
        }
        if (it < nfull) {
            process(R0, C0);
            it++;
        }
        it0 = it;
        // the last, partial iteration: a thread whose 16 frames are all
        // valid stays on the fast path; only the thread with the final
        // partial chunk goes through the general loop below
        if (it0 < niter && it0 * per_iter + (int)threadIdx.x * 16 + 16 <= N) {
            load(R1, C1, it0);
            process(R1, C1);
            tail_fast = true;
        }
    }
    // general fill: diagonal-band / edge tiles, mixed bin counts, and the
    // final partial iteration
    for (int it = it0; it < niter; it++) {
        if (tail_fast) break;  // this thread's last iteration ran on the fast path
        const int f = it * per_iter + threadIdx.x * 16;
        uint4 R[TR], C[TC];
#pragma unroll
        for (int r = 0; r < TR; r++)
            R[r] = f < in.Np ? *reinterpret_cast<const uint4*>(rows[r] + f) : make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int c = 0; c < TC; c++)
            C[c] = f < in.Np ? *reinterpret_cast<const uint4*>(cols[c] + f) : make_uint4(0, 0, 0, 0);
        const int nvalid = min(16, max(0, N - f));
        for (int e = 0; e < nvalid; e++) {
#pragma unroll
            for (int r = 0; r < TR; r++) {
                const unsigned a = byte_of(R[r], e);
#pragma unroll
                for (int c = 0; c < TC; c++) {
                    if (valid & (1u << (r * TC + c)))
                        atomicAdd(&hist[(r * TC + c) * SLOTW + __umul24(a, (unsigned)Bc[c]) + byte_of(C[c], e)], 1u);
                }
            }
        }
    }
    __syncthreads();

    if (dump_out) {
        // entries of this tile: (slot, offset) with offset split in two ints
        const int e0 = dump_begin[t], e1 = dump_begin[t + 1];
        for (int e = e0; e < e1; e++) {
            const int sl = dump_entries[3 * e];
            const long long off = ((long long)dump_entries[3 * e + 1] << 31) | (long long)dump_entries[3 * e + 2];
            const int r = sl / TC, c = sl % TC;
            const int cells = in.B.of(dof_type(nD, gr[r])) * Bc[c];
            for (int i = threadIdx.x; i < cells; i += BLOCK) dump_out[off + i] = (int)hist[sl * SLOTW + i];
        }
        return;
    }

    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    constexpr int kWaves = BLOCK / 64;
    for (int sl = wid; sl < TR * TC; sl += kWaves) {
        if (!(valid & (1u << sl))) continue;
        const int r = sl / TC, c = sl % TC;
        const int gi = gr[r], gj = gc[c];
        const int cells = in.B.of(dof_type(nD, gi)) * Bc[c];
        const unsigned* h = hist + sl * SLOTW;
        double T = 0.0;
        // branch-free (the table is 0 at n = 0) so the unrolled table loads
        // are in flight together
#pragma unroll 4
        for (int i = lane; i < cells; i += 64) T += in.nlogn[h[i]];
        T = wave_total(T);
        if (lane == 0) in.s2d[par_offset(nD, gi, gj)] = ((in.logN + in.c2[gi]) + in.c2[gj] - 0.5 / N) - T / N;
    }
}

// ---------------------------------------------------------------------------
// sliced: bank-conflict-free LDS histograms
// ---------------------------------------------------------------------------
// The LDS serves ds_add_u32 16 lanes per clock over 32 banks; random cells
// from 16 lanes collide in banks and cost ~half the peak rate.  Here every
// cell owns 16 u32 sub-counters, one per lane position within a 16-lane
// group: word(cell, s) = cell * 16 + s with s = lane % 16, whose bank is
// s + 16 * (cell & 1) -- distinct for the 16 lanes served together whatever
// their cells.  Lanes of different groups share words, which the hardware
// serves in different cycles.  The count of a cell is the sum of its 16
// sub-counters (epilogue).  64 bytes per cell: one 50 x 50 histogram fills
// the 160 KB of LDS.
//
// PACK = 2 halves that: the cells (2k, b) and (2k+1, b) share their 16
// words as 16-bit halves (word = ((a/2) * Bj + b) * 16 + s; a lane adds 1
// or 65536 by the row's parity, so the increment is shared by every column
// of a tile row).  The bank pattern is unchanged, so the fill stays
// conflict-free, and a 50 x 50 histogram takes 80 KB, two workgroups per CU.
// A half counts at most the frames of one lane position, N/16 + 16, so
// PACK = 2 needs N <= kPackedMaxFrames.
//
// The fill computes byte offsets straight from the index bytes (a 24-bit
// multiply with a byte operand for the row, a shift for the column, one
// three-way add), four frames per index word.
//
// DOT (1 x 1 tiles) forms each offset in two instructions: v_perm_b32 puts
// the frame's row byte a and column byte b into the two 16-bit lanes of one
// register, and v_dot2_u32_u16 returns a * rowmul + b * 64 + 4 s.  With
// PACK = 2 it also interleaves the halves instead of pairing rows: cell
// (a, b) is at byte a * S * 32 + b * 64 + 4 s (S = Bj rounded up to even),
// in the low half for even a and the high half for odd a.  For each parity
// the map (a, b, s) -> word is one-to-one, the word's bank is still
// s + 16 * (word / 16 & 1), and the row byte needs no mask.  A slot holds
// (Bi - 1) * S / 2 + Bj groups of 16 words; every half that no cell owns
// stays 0.
template <int TR, int TC, int BLOCK, int SLOTC, int PACK, bool DOT, bool BATCH>
__global__ __launch_bounds__(BLOCK, 1) void k_pairs_sliced(PairInputs in0, TileQueue q,
                                                        int t_begin, int t_end,
                                                        const int* __restrict__ dump_begin,
                                                        const int* __restrict__ dump_entries,
                                                        int* __restrict__ dump_out) {
    static_assert(!DOT || (TR == 1 && TC == 1), "DOT fill is for one-pair tiles");
    extern __shared__ unsigned hist[];
    // slot size: SLOTC cells, or SLOTC groups of 16 words for the interleaved halves
    constexpr int kSlotW = (DOT && PACK == 2) ? SLOTC * 16 : SLOTC / PACK * 16;
    const int chunk = (t_end - t_begin + 7) / 8;
    const int t = t_begin + (int)(blockIdx.x % 8) * chunk + (int)(blockIdx.x / 8);
    if (t >= t_end || (int)(blockIdx.x / 8) >= chunk) return;
    const int* tp = q.tiles + (size_t)(BATCH ? 3 : 2) * t;
    const int r0 = tp[0], c0 = tp[1];
    const PairInputs in = BATCH ? q.descs[tp[2]] : in0;
    const int D = in.D, N = in.N, nD = in.nD;
    const unsigned s16 = threadIdx.x & 15;

    int gr[TR], gc[TC], Bc[TC];
#pragma unroll
    for (int r = 0; r < TR; r++) gr[r] = r0 + r;
#pragma unroll
    for (int c = 0; c < TC; c++) {
        gc[c] = c0 + c;
        Bc[c] = gc[c] < D ? in.B.of(dof_type(nD, gc[c])) : 1;
    }
    unsigned valid = 0;
    bool uniform = true;
#pragma unroll
    for (int r = 0; r < TR; r++)
#pragma unroll
        for (int c = 0; c < TC; c++)
            if (gr[r] < D && gc[c] < D && gr[r] < gc[c]) valid |= 1u << (r * TC + c);
#pragma unroll
    for (int c = 1; c < TC; c++) uniform = uniform && Bc[c] == Bc[0];
    const bool fast = uniform && valid == (1u << (TR * TC)) - 1u;

    const int per_iter = BLOCK * 16;
    const int niter = (N + per_iter - 1) / per_iter;
    const int nfull = N / per_iter;
    const uint8_t* rows[TR];
    const uint8_t* cols[TC];
#pragma unroll
    for (int r = 0; r < TR; r++) rows[r] = in.bins + (size_t)(gr[r] < D ? gr[r] : 0) * in.Np;
#pragma unroll
    for (int c = 0; c < TC; c++) cols[c] = in.bins + (size_t)(gc[c] < D ? gc[c] : 0) * in.Np;
    // The fast fill's first index words are requested before the histogram
    // clear, and the barrier after the clear waits for LDS operations only,
    // so the load latency of a new tile overlaps the clear.
    __amdgpu_buffer_rsrc_t rrs[TR], crs[TC];
#pragma unroll
    for (int r = 0; r < TR; r++) rrs[r] = index_row(rows[r], in.Np);
#pragma unroll
    for (int c = 0; c < TC; c++) crs[c] = index_row(cols[c], in.Np);
    const int toff = (int)threadIdx.x * 16;
    uint4 R0[TR], C0[TC];
    if (fast && nfull > 0) {
#pragma unroll
        for (int r = 0; r < TR; r++) R0[r] = load16(rrs[r], toff, 0);
#pragma unroll
        for (int c = 0; c < TC; c++) C0[c] = load16(crs[c], toff, 0);
    }

    {
        uint4* h4 = reinterpret_cast<uint4*>(hist);
        for (int i = threadIdx.x; i < TR * TC * kSlotW / 4; i += BLOCK) h4[i] = make_uint4(0, 0, 0, 0);
    }
    lds_barrier();
    // c2 of the tile's dofs for the epilogue: scalar loads (the constant
    // address space; c2 is not written while pairs run), issued after the
    // barrier so that it does not wait for them, and in flight during the fill
    typedef __attribute__((address_space(4))) const double cdouble;
    double c2r[TR], c2c[TC];
#pragma unroll
    for (int r = 0; r < TR; r++) c2r[r] = ((cdouble*)in.c2)[gr[r] < D ? gr[r] : 0];
#pragma unroll
    for (int c = 0; c < TC; c++) c2c[c] = ((cdouble*)in.c2)[gc[c] < D ? gc[c] : 0];

    // word of cell (a, b) in a slot with Bj columns, and the lane's increment
    auto cell_word = [&](unsigned a, unsigned b, unsigned Bj) {
        if constexpr (PACK == 1) return ((__umul24(a, Bj) + b) << 4) + s16;
        else if constexpr (DOT) return (__umul24(a, (Bj + 1u) & ~1u) << 3) + (b << 4) + s16;
        else return ((__umul24(a >> 1, Bj) + b) << 4) + s16;
    };
    auto cell_inc = [&](unsigned a) {
        if constexpr (PACK == 1) return 1u;
        else return 1u << ((a & 1u) << 4);
    };
    unsigned char* hb = reinterpret_cast<unsigned char*>(hist);
    constexpr int kSlotB = kSlotW * 4;
    auto comp = [](const uint4& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); };
    int it0 = 0;
    bool tail_fast = false;
    if (fast) {
        // byte offset of row byte a: (a or a & ~1) * rowmul; of column byte b: b * 64
        const unsigned rowmul = (unsigned)Bc[0] * (PACK == 1 ? 64u : 32u);
        const unsigned s4 = s16 * 4u;
        // shift amount and increment base for the byte-select shifts (VGPRs)
        unsigned six = 6u, one = 1u;
        asm("" : "+v"(six), "+v"(one));
        // DOT: 16-bit lanes (row multiplier, 64) of the dot product, whose
        // accumulator is the LDS byte address of the lane's sub-counter 0
        const unsigned rowmul_dot = PACK == 1 ? (unsigned)Bc[0] * 64u : (((unsigned)Bc[0] + 1u) & ~1u) * 32u;
        const unsigned kdot = rowmul_dot | (64u << 16);
        const unsigned acc = (unsigned)(uintptr_t)(lds_u32*)hist + s4;
        auto process = [&](const uint4* R, const uint4* C) {
            if constexpr (DOT) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const unsigned rw = comp(R[0], k), cw = comp(C[0], k);
                    unsigned rsh = 0;
                    if constexpr (PACK == 2) {
                        rsh = (rw & 0x01010101u) << 4;
                        asm("" : "+v"(rsh));
                    }
                    // the word's four addresses first: a dot product's result
                    // is read by its atomic a few instructions later
                    unsigned addr[4], inc[4];
                    auto frame = [&](auto ec) {
                        constexpr int e = decltype(ec)::value;
                        // bytes: [row byte e, 0, column byte e, 0]
                        constexpr unsigned sel = 0x0C000C00u | ((unsigned)e << 16) | (4u + (unsigned)e);
                        const unsigned ab = __builtin_amdgcn_perm(rw, cw, sel);
                        addr[e] = __builtin_amdgcn_udot2(__builtin_bit_cast(u16x2, ab),
                                                         __builtin_bit_cast(u16x2, kdot), acc, false);
                        inc[e] = PACK == 1 ? 1u : one_shl_byte<e>(rsh, one);
                    };
                    frame(std::integral_constant<int, 0>{});
                    frame(std::integral_constant<int, 1>{});
                    frame(std::integral_constant<int, 2>{});
                    frame(std::integral_constant<int, 3>{});
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        __hip_atomic_fetch_add((lds_u32*)(uintptr_t)addr[e], inc[e], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                return;
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                unsigned rw[TR], rsh[TR], cw[TC];
#pragma unroll
                for (int r = 0; r < TR; r++) {
                    const unsigned w = comp(R[r], k);
                    rw[r] = PACK == 1 ? w : (w & 0xFEFEFEFEu);
                    rsh[r] = (w & 0x01010101u) << 4;
                    // Keep the masked words opaque: otherwise the masks are
                    // folded back into per-byte shift/and pairs (~7 VALU per
                    // update) instead of byte-select (SDWA) operands of the
                    // multiply and the increment shift (~4.5 VALU).
                    if constexpr (PACK == 2) asm("" : "+v"(rw[r]), "+v"(rsh[r]));
                }
#pragma unroll
                for (int c = 0; c < TC; c++) cw[c] = comp(C[c], k);
                auto frame = [&](auto ec) {
                    constexpr int e = decltype(ec)::value;
                    unsigned ro[TR], inc[TR], co[TC];
#pragma unroll
                    for (int r = 0; r < TR; r++) {
                        ro[r] = __umul24((rw[r] >> (8 * e)) & 0xFFu, rowmul) + s4;
                        if constexpr (PACK == 1) inc[r] = 1u;
                        else inc[r] = one_shl_byte<e>(rsh[r], one);
                    }
#pragma unroll
                    for (int c = 0; c < TC; c++) co[c] = byte_shl<e>(cw[c], six);
#pragma unroll
                    for (int r = 0; r < TR; r++)
#pragma unroll
                        for (int c = 0; c < TC; c++)
                            atomicAdd(reinterpret_cast<unsigned*>(hb + (r * TC + c) * kSlotB + ro[r] + co[c]), inc[r]);
                };
                frame(std::integral_constant<int, 0>{});
                frame(std::integral_constant<int, 1>{});
                frame(std::integral_constant<int, 2>{});
                frame(std::integral_constant<int, 3>{});
            }
        };
        // two register sets: the next iteration's index words load while
        // this iteration's atomics issue, without copies between the sets
        // full iterations only: the partial last one goes by 4-frame words
        auto load = [&](uint4* R, uint4* C, int it) {
            if (it < nfull) {
                const int fb = it * per_iter;
#pragma unroll
                for (int r = 0; r < TR; r++) R[r] = load16(rrs[r], toff, fb);
#pragma unroll
                for (int c = 0; c < TC; c++) C[c] = load16(crs[c], toff, fb);
            }
        };
        uint4 R1[TR], C1[TC];  // R0 / C0 hold iteration 0 already
        int it = 0;
        for (; it + 2 <= nfull; it += 2) {
            load(R1, C1, it + 1);
            process(R0, C0);
            load(R0, C0, it + 2);
            process(R1, C1);
        }
        if (it < nfull) {
            process(R0, C0);
            it++;
        }
        // the last, partial iteration, one 4-frame index word per thread:
        // the workgroup waits for its slowest thread, and 16-frame chunks
        // would leave most threads idle here.  A lane position still counts
        // at most N / 16 + 4 frames.
        const int nw = (N + 3) / 4;
        for (int w = nfull * (per_iter / 4) + (int)threadIdx.x; w < nw; w += BLOCK) {
            unsigned rwv[TR], cwv[TC];
#pragma unroll
            for (int r = 0; r < TR; r++) rwv[r] = reinterpret_cast<const unsigned*>(rows[r])[w];
#pragma unroll
            for (int c = 0; c < TC; c++) cwv[c] = reinterpret_cast<const unsigned*>(cols[c])[w];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                if (4 * w + e >= N) break;
#pragma unroll
                for (int r = 0; r < TR; r++) {
                    const unsigned a = (rwv[r] >> (8 * e)) & 0xFFu;
#pragma unroll
                    for (int c = 0; c < TC; c++)
                        atomicAdd(&hist[(r * TC + c) * kSlotW + cell_word(a, (cwv[c] >> (8 * e)) & 0xFFu,
                                                                            (unsigned)Bc[c])],
                                  cell_inc(a));
                }
            }
        }
        tail_fast = true;
    }
    for (int it = it0; it < niter; it++) {
        if (tail_fast) break;  // the fast path above covered every frame
        const int f = it * per_iter + threadIdx.x * 16;
        uint4 R[TR], C[TC];
#pragma unroll
        for (int r = 0; r < TR; r++)
            R[r] = f < in.Np ? *reinterpret_cast<const uint4*>(rows[r] + f) : make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int c = 0; c < TC; c++)
            C[c] = f < in.Np ? *reinterpret_cast<const uint4*>(cols[c] + f) : make_uint4(0, 0, 0, 0);
        const int nvalid = min(16, max(0, N - f));
        for (int e = 0; e < nvalid; e++) {
#pragma unroll
            for (int r = 0; r < TR; r++) {
                const unsigned a = byte_of(R[r], e);
#pragma unroll
                for (int c = 0; c < TC; c++) {
                    if (valid & (1u << (r * TC + c))) {
                        const unsigned b = byte_of(C[c], e);
                        atomicAdd(&hist[(r * TC + c) * kSlotW + cell_word(a, b, (unsigned)Bc[c])], cell_inc(a));
                    }
                }
            }
        }
    }
    __syncthreads();

    // count of cell i (row-major over Bj columns) of slot sl: sum of its 16 sub-counters
    auto cell_count = [&](int sl, int i, int Bj) {
        const unsigned a = (unsigned)(i / Bj), b = (unsigned)(i - (int)a * Bj);
        const uint4* q = reinterpret_cast<const uint4*>(hist + sl * kSlotW + (cell_word(a, b, (unsigned)Bj) - s16));
        unsigned n = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint4 v = q[k];
            if constexpr (PACK == 1) {
                n += v.x + v.y + v.z + v.w;
            } else {
                const unsigned sh = (a & 1u) * 16u;
                n += ((v.x >> sh) & 0xFFFFu) + ((v.y >> sh) & 0xFFFFu) + ((v.z >> sh) & 0xFFFFu) + ((v.w >> sh) & 0xFFFFu);
            }
        }
        return n;
    };
    if (dump_out) {
        const int e0 = dump_begin[t], e1 = dump_begin[t + 1];
        for (int e = e0; e < e1; e++) {
            const int sl = dump_entries[3 * e];
            const long long off = ((long long)dump_entries[3 * e + 1] << 31) | (long long)dump_entries[3 * e + 2];
            const int r = sl / TC, c = sl % TC;
            const int cells = in.B.of(dof_type(nD, gr[r])) * Bc[c];
            for (int i = threadIdx.x; i < cells; i += BLOCK) dump_out[off + i] = (int)cell_count(sl, i, Bc[c]);
        }
        return;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    constexpr int kWaves = BLOCK / 64;
    for (int sl = 0; sl < TR * TC; sl++) {
        if (!(valid & (1u << sl))) continue;
        const int r = sl / TC, c = sl % TC;
        const int gi = gr[r], gj = gc[c];
        const int Bi = in.B.of(dof_type(nD, gi)), Bj = Bc[c];
        // all waves share a slot: per-wave partial sums, then wave 0 combines
        __shared__ double sT[kWaves];
        double T = 0.0;
        // Word groups (the 16 sub-counter words of a cell, or of a cell pair
        // when packed), one per thread: thread t reads word (s + t) % 16 of
        // its group at step s, so the 16 lanes served together hit 16
        // distinct banks.
        const int ngroups = (DOT && PACK == 2) ? (Bi - 1) * (((Bj + 1) & ~1) / 2) + Bj
                                               : (PACK == 1 ? Bi : (Bi + 1) / 2) * Bj;
        const uint4* hs4 = reinterpret_cast<const uint4*>(hist + sl * kSlotW);
        // a group is 4 uint4; thread t reads them in the order rotated by
        // (t / 2) % 4, so the 8 lanes a 128-byte read cycle serves (4 banks
        // each, groups of alternating parity) cover all 32 banks
        int quad[4];
#pragma unroll
        for (int k = 0; k < 4; k++) quad[k] = (k + (threadIdx.x >> 1)) & 3;
        // The thread's groups (at most kMaxG) are counted first and their
        // table loads issued together, then summed in group order, low half
        // then high half; a group past the end counts 0, whose entry is 0.
        constexpr int kMaxG = (kSlotW / 16 + BLOCK - 1) / BLOCK;
        unsigned cnt[2 * kMaxG];
#pragma unroll
        for (int k = 0; k < kMaxG; k++) {
            const int g = (int)threadIdx.x + k * BLOCK;
            unsigned lo = 0, hi = 0;
            if (g < ngroups) {
#pragma unroll
                for (int kk = 0; kk < 4; kk++) {
                    const uint4 v = hs4[g * 4 + quad[kk]];
                    if constexpr (PACK == 1) {
                        lo += (v.x + v.y) + (v.z + v.w);
                    } else {
                        lo += add_lo16(v.x, v.y) + add_lo16(v.z, v.w);
                        hi += add_hi16(v.x, v.y) + add_hi16(v.z, v.w);
                    }
                }
            }
            cnt[2 * k] = lo;
            // the odd row of a packed pair exists unless Bi is odd and this
            // is the last pair row (whose high halves stay 0); interleaved
            // halves that no cell owns are 0 already
            cnt[2 * k + 1] = PACK == 1 ? 0u : (DOT || 2 * (g / Bj) + 1 < Bi ? hi : 0u);
        }
        double tab[2 * kMaxG];
#pragma unroll
        for (int k = 0; k < 2 * kMaxG; k++) tab[k] = (PACK == 1 && (k & 1)) ? 0.0 : in.nlogn[cnt[k]];
#pragma unroll
        for (int k = 0; k < 2 * kMaxG; k++) T += tab[k];
        T = wave_total(T);
        if (lane == 0) sT[wid] = T;
        __syncthreads();
        if (wid == 0) {
            // the waves' partial sums in one cross-lane reduction (fixed
            // order): the workgroup, and its LDS, ends sooner than with a
            // serial loop
            const double tt = wave_total(lane < kWaves ? sT[lane] : 0.0);
            if (lane == 0) in.s2d[par_offset(nD, gi, gj)] = ((in.logN + c2r[r]) + c2c[c] - 0.5 / N) - tt / N;
        }
        if (sl + 1 < TR * TC) __syncthreads();  // sT is reused by the next slot
    }
}

// ---------------------------------------------------------------------------
// multi-tile sliced kernel: one-pair tiles, perm + dot2 fill, TPW tiles per workgroup
// ---------------------------------------------------------------------------
// A workgroup takes TPW consecutive tiles of its XCD's range, with the
// interleaved (PACK = 2) or 32-bit (PACK = 1) layout of the DOT sliced
// kernel, with SL sub-counters per cell: 16 (conflict-free) or 4.  With 4,
// the 4 lanes of a 16-lane LDS cycle that share a sub-counter index meet in
// a bank only when their cells fall on the same 8-group residue, so a fill
// costs about 1.7 cycles per 16 lanes instead of about 3 for unsliced
// histograms, at a quarter of the LDS: 100 x 100 bins in 80,800 bytes.  Per tile: fill; barrier; an epilogue that zeroes each group it
// reads (the next tile needs no clear); the next tile's queue entry and first
// index words requested while this tile's table loads are in flight;
// barrier; the tile's cross-wave sum on wave 0 while the other waves start
// the next fill.  Workgroups still start when slots free, at the
// dispatcher's times, so the two workgroups of a CU do not fall into step.
// Register use decides the speed here: a version that spilled to scratch
// (the tile loop unrolled, the kernel-argument inputs copied per tile) ran
// at half the rate, so the loop stays rolled and one trajectory's inputs are
// read from the argument in place.
template <int BLOCK, int SLOTC, int PACK, int TPW, bool BATCH, int SL = 16>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(BLOCK / 128))) void k_pairs_multi(
    PairInputs in0, TileQueue q, int t_begin, int t_end, const int* __restrict__ dump_begin,
    const int* __restrict__ dump_entries, int* __restrict__ dump_out) {
    extern __shared__ unsigned hist[];
    constexpr int kWaves = BLOCK / 64;
    static_assert(SL == 16 || SL == 4, "sub-counters per cell");
    constexpr int kSlotW = SLOTC * SL;                   // SL-word groups (PACK 2) or cells (PACK 1)
    constexpr int NQ = SL / 4;                           // uint4 per group
    constexpr int kMaxG = (SLOTC + BLOCK - 1) / BLOCK;   // groups per thread
    constexpr int per_iter = BLOCK * 16;
    __shared__ double sT[2][kWaves];
    const int chunk = (t_end - t_begin + 7) / 8;        // tiles of this XCD
    const int tbase = t_begin + (int)(blockIdx.x % 8) * chunk;
    const int l0 = (int)(blockIdx.x / 8) * TPW;
    auto has = [&](int local) { return local < chunk && tbase + local < t_end; };
    if (!has(l0)) return;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const unsigned s16 = threadIdx.x & (SL - 1);  // the lane's sub-counter
    const int toff = (int)threadIdx.x * 16;
    const unsigned acc = (unsigned)(uintptr_t)(lds_u32*)hist + s16 * 4u;
    uint4* h4 = reinterpret_cast<uint4*>(hist);
    unsigned one = 1u;
    asm("" : "+v"(one));
    int quad[NQ];
#pragma unroll
    for (int k = 0; k < NQ; k++) quad[k] = (k + (threadIdx.x >> 1)) & (NQ - 1);

    // the queue and the batch descriptors are not written while the kernel
    // runs: read through the constant address space they are scalar loads
    typedef __attribute__((address_space(4))) const int qint;
    typedef __attribute__((address_space(4))) const unsigned qword;
    struct TileIdx {
        int t, gi, gj, k;
    };
    auto fetch = [&](int local) {
        TileIdx x;
        x.t = tbase + local;
        qint* tp = (qint*)(q.tiles + (size_t)(BATCH ? 3 : 2) * x.t);
        x.gi = tp[0];
        x.gj = tp[1];
        x.k = BATCH ? tp[2] : 0;
        return x;
    };
    auto inputs = [&](const TileIdx& x) -> PairInputs {
        if constexpr (BATCH) {
            static_assert(sizeof(PairInputs) % 4 == 0, "descriptor copied as 32-bit words");
            qword* src = (qword*)(q.descs + x.k);
            PairInputs r;
            unsigned* dst = reinterpret_cast<unsigned*>(&r);
#pragma unroll
            for (int i = 0; i < (int)(sizeof(PairInputs) / 4); i++) dst[i] = src[i];
            return r;
        } else {
            return in0;
        }
    };
    auto cell_word = [&](unsigned a, unsigned b, unsigned Bj) {
        if constexpr (PACK == 1) return (__umul24(a, Bj) + b) * SL + s16;
        else return __umul24(a, (Bj + 1u) & ~1u) * (SL / 2) + b * SL + s16;
    };
    auto cell_inc = [&](unsigned a) {
        if constexpr (PACK == 1) return 1u;
        else return 1u << ((a & 1u) << 4);
    };
    auto first_words = [&](const PairInputs& in, const TileIdx& x, uint4& R, uint4& C) {
        if (x.gi < x.gj && x.gj < in.D && in.N > 0) {
            R = load16(index_row(in.bins + (size_t)x.gi * in.Np, in.Np), toff, 0);
            C = load16(index_row(in.bins + (size_t)x.gj * in.Np, in.Np), toff, 0);
        }
    };

    int local = l0;
    TileIdx x = fetch(local);
    PairInputs in = inputs(x);
    uint4 R0, C0;
    first_words(in, x, R0, C0);
    {
        // the clear unrolled: fixed LDS offsets instead of a counted loop
        constexpr int kZ = kSlotW / 4;
#pragma unroll
        for (int r = 0; r < kZ / BLOCK; r++) h4[threadIdx.x + r * BLOCK] = make_uint4(0, 0, 0, 0);
        if constexpr (kZ % BLOCK != 0) {
            if ((int)threadIdx.x < kZ % BLOCK) h4[threadIdx.x + (kZ / BLOCK) * BLOCK] = make_uint4(0, 0, 0, 0);
        }
    }
    lds_barrier();
    int par = 0;
    // not unrolled: one tile's state at a time stays in registers
#pragma unroll 1
    for (int k = 0; k < TPW; k++) {
        const int nlocal = local + 1;
        const bool more = k + 1 < TPW && has(nlocal);
        const TileIdx nx = more ? fetch(nlocal) : x;
        const bool valid = x.gi < x.gj && x.gj < in.D;
        const int Bi = valid ? in.B.of(dof_type(in.nD, x.gi)) : 1;
        const int Bj = valid ? in.B.of(dof_type(in.nD, x.gj)) : 1;
        // ---- fill
        if (valid) {
            const int N = in.N, nfull = N / per_iter;
            const uint8_t* rowp = in.bins + (size_t)x.gi * in.Np;
            const uint8_t* colp = in.bins + (size_t)x.gj * in.Np;
            const __amdgpu_buffer_rsrc_t rr = index_row(rowp, in.Np), cr = index_row(colp, in.Np);
            // byte offsets: row a * rowmul, column b * 4 SL
            const unsigned rowmul = PACK == 1 ? (unsigned)Bj * (4u * SL) : (((unsigned)Bj + 1u) & ~1u) * (2u * SL);
            const unsigned kdot = rowmul | ((4u * SL) << 16);
            auto process = [&](const uint4& R, const uint4& C) {
#pragma unroll
                for (int kk = 0; kk < 4; kk++) {
                    const unsigned rw = kk == 0 ? R.x : (kk == 1 ? R.y : (kk == 2 ? R.z : R.w));
                    const unsigned cw = kk == 0 ? C.x : (kk == 1 ? C.y : (kk == 2 ? C.z : C.w));
                    unsigned rsh = 0;
                    if constexpr (PACK == 2) {
                        rsh = (rw & 0x01010101u) << 4;
                        asm("" : "+v"(rsh));
                    }
                    unsigned addr[4], inc[4];
                    auto frame = [&](auto ec) {
                        constexpr int e = decltype(ec)::value;
                        constexpr unsigned sel = 0x0C000C00u | ((unsigned)e << 16) | (4u + (unsigned)e);
                        const unsigned ab = __builtin_amdgcn_perm(rw, cw, sel);
                        addr[e] = __builtin_amdgcn_udot2(__builtin_bit_cast(u16x2, ab),
                                                         __builtin_bit_cast(u16x2, kdot), acc, false);
                        inc[e] = PACK == 1 ? 1u : one_shl_byte<e>(rsh, one);
                    };
                    frame(std::integral_constant<int, 0>{});
                    frame(std::integral_constant<int, 1>{});
                    frame(std::integral_constant<int, 2>{});
                    frame(std::integral_constant<int, 3>{});
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        __hip_atomic_fetch_add((lds_u32*)(uintptr_t)addr[e], inc[e], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            };
            // Iterations of per_iter frames; the last may be partial.  Its
            // whole 16-frame chunks take the same path as the rest, and its
            // one ragged chunk (N not a multiple of 16) goes frame by frame.
            // Each iteration's words are requested one iteration ahead.
            const int nit = (N + per_iter - 1) / per_iter;
            uint4 R1 = make_uint4(0, 0, 0, 0), C1 = R1;
            int it = 0;
            for (; it + 2 <= nfull; it += 2) {
                R1 = load16(rr, toff, (it + 1) * per_iter);
                C1 = load16(cr, toff, (it + 1) * per_iter);
                process(R0, C0);
                // unconditional (past the row a buffer load reads zeros):
                // without a branch the wait before process(R1) leaves these
                // two loads in flight
                R0 = load16(rr, toff, (it + 2) * per_iter);
                C0 = load16(cr, toff, (it + 2) * per_iter);
                process(R1, C1);
            }
            if (it < nfull) {
                if (it + 1 < nit) {
                    R1 = load16(rr, toff, (it + 1) * per_iter);
                    C1 = load16(cr, toff, (it + 1) * per_iter);
                }
                process(R0, C0);
                R0 = R1;
                C0 = C1;
                it++;
            }
            if (it < nit) {
                const int left = N - it * per_iter - toff;
                if (left >= 16) {
                    process(R0, C0);
                } else if (left > 0) {
                    for (int e = 0; e < left; e++) {
                        const unsigned a = byte_of(R0, e), b = byte_of(C0, e);
                        atomicAdd(&hist[cell_word(a, b, (unsigned)Bj)], cell_inc(a));
                    }
                }
            }
        }
        lds_barrier();
        // ---- counts of the tile's cells (parity checks), before the zeroing epilogue
        if (dump_out && valid) {
            for (int e = dump_begin[x.t]; e < dump_begin[x.t + 1]; e++) {
                const long long off = ((long long)dump_entries[3 * e + 1] << 31) | (long long)dump_entries[3 * e + 2];
                for (int i = threadIdx.x; i < Bi * Bj; i += BLOCK) {
                    const unsigned a = (unsigned)(i / Bj), b = (unsigned)(i - (int)a * Bj);
                    const uint4* g4 = reinterpret_cast<const uint4*>(hist + (cell_word(a, b, (unsigned)Bj) - s16));
                    const unsigned sh = PACK == 2 ? (a & 1u) * 16u : 0u;
                    const unsigned m = PACK == 2 ? 0xFFFFu : 0xFFFFFFFFu;
                    unsigned n = 0;
#pragma unroll
                    for (int kk = 0; kk < NQ; kk++) {
                        const uint4 v = g4[kk];
                        n += ((v.x >> sh) & m) + ((v.y >> sh) & m) + ((v.z >> sh) & m) + ((v.w >> sh) & m);
                    }
                    dump_out[off + i] = (int)n;
                }
            }
            __syncthreads();
        }
        // ---- epilogue: read and zero this thread's groups
        const int ngroups = !valid ? 0 : PACK == 2 ? (Bi - 1) * (((Bj + 1) & ~1) / 2) + Bj : Bi * Bj;
        // A 16-bit half holds at most N/16 + 16 counts (one sub-counter
        // index), so halves of 8 (or 4) words add as packed 16-bit integers
        // without carrying out when N is at most kPk8Frames (kPk4Frames).
        const int pk = PACK == 2 && SL == 16 && valid ? (in.N <= kPk8Frames ? 8 : in.N <= kPk4Frames ? 4 : 1) : 1;
        unsigned cnt[2 * kMaxG];
#pragma unroll
        for (int kk = 0; kk < kMaxG; kk++) {
            const int g = (int)threadIdx.x + kk * BLOCK;
            unsigned lo = 0, hi = 0;
            if (g < ngroups) {
                uint4 v[NQ];
#pragma unroll
                for (int q4 = 0; q4 < NQ; q4++) {
                    v[q4] = h4[g * NQ + quad[q4]];
                    if (more) h4[g * NQ + quad[q4]] = make_uint4(0, 0, 0, 0);
                }
                if constexpr (PACK == 1) {
#pragma unroll
                    for (int q4 = 0; q4 < NQ; q4++) lo += (v[q4].x + v[q4].y) + (v[q4].z + v[q4].w);
                } else if constexpr (SL == 16) {
                    if (pk == 8) {
                        const unsigned y0 = pk_add16(pk_add16(pk_add16(v[0].x, v[0].y), pk_add16(v[0].z, v[0].w)),
                                                     pk_add16(pk_add16(v[1].x, v[1].y), pk_add16(v[1].z, v[1].w)));
                        const unsigned y1 = pk_add16(pk_add16(pk_add16(v[2].x, v[2].y), pk_add16(v[2].z, v[2].w)),
                                                     pk_add16(pk_add16(v[3].x, v[3].y), pk_add16(v[3].z, v[3].w)));
                        lo = add_lo16(y0, y1);
                        hi = add_hi16(y0, y1);
                    } else if (pk == 4) {
                        unsigned y[4];
#pragma unroll
                        for (int q4 = 0; q4 < 4; q4++)
                            y[q4] = pk_add16(pk_add16(v[q4].x, v[q4].y), pk_add16(v[q4].z, v[q4].w));
                        lo = add_lo16(y[0], y[1]) + add_lo16(y[2], y[3]);
                        hi = add_hi16(y[0], y[1]) + add_hi16(y[2], y[3]);
                    } else {
#pragma unroll
                        for (int q4 = 0; q4 < NQ; q4++) {
                            lo += add_lo16(v[q4].x, v[q4].y) + add_lo16(v[q4].z, v[q4].w);
                            hi += add_hi16(v[q4].x, v[q4].y) + add_hi16(v[q4].z, v[q4].w);
                        }
                    }
                } else {
#pragma unroll
                    for (int q4 = 0; q4 < NQ; q4++) {
                        lo += add_lo16(v[q4].x, v[q4].y) + add_lo16(v[q4].z, v[q4].w);
                        hi += add_hi16(v[q4].x, v[q4].y) + add_hi16(v[q4].z, v[q4].w);
                    }
                }
            }
            cnt[2 * kk] = lo;
            cnt[2 * kk + 1] = hi;
        }
        // this tile's table loads, then the next tile's first index words
        double tab[2 * kMaxG];
#pragma unroll
        for (int kk = 0; kk < 2 * kMaxG; kk++) tab[kk] = (PACK == 1 && (kk & 1)) ? 0.0 : in.nlogn[cnt[kk]];
        // (a batch tile's inputs are its trajectory's descriptor; one
        // trajectory's are the kernel argument, never copied)
        PairInputs nin = in;
        if constexpr (BATCH) {
            if (more) nin = inputs(nx);
        }
        if (more) first_words(nin, nx, R0, C0);
        // groups in order, low half then high half (unused halves add 0)
        double T = 0.0;
#pragma unroll
        for (int kk = 0; kk < 2 * kMaxG; kk++) T += tab[kk];
        T = wave_total(T);
        if (lane == 0) sT[par][wid] = T;
        lds_barrier();
        if (wid == 0 && valid) {
            const double tt = wave_total(lane < kWaves ? sT[par][lane] : 0.0);
            if (lane == 0)
                in.s2d[par_offset(in.nD, x.gi, x.gj)] =
                    ((in.logN + in.c2[x.gi]) + in.c2[x.gj] - 0.5 / in.N) - tt / in.N;
        }
        if (!more) break;
        x = nx;
        if constexpr (BATCH) in = nin;
        local = nlocal;
        par ^= 1;
    }
}

namespace {

// (kind, tile rows, tile columns, threads, slot size).  kind 0: plain LDS
// histograms, slot size in words (2560 -> B <= 50, 5632 -> B <= 75,
// 10240 -> B <= 101, 40960 -> B <= 202); kind 1: sliced histograms, slot size
// in cells (64 bytes each; 2500 -> B <= 50, 1225 -> B <= 35, 900 -> B <= 30,
// 400 -> B <= 20, 100 -> B <= 10); kind 2: packed sliced histograms, slot
// size in cells of B x (B rounded up to even) (32 bytes each; 2500 -> B <= 50,
// 3600 -> B <= 60, 1296 -> B <= 36, 900 -> B <= 30); kind 3: sliced
// histograms filled through v_perm + v_dot2, slot size in cells as kind 1;
// kind 4: interleaved packed histograms filled the same way, slot size in
// groups of 16 words, (B - 1) * S / 2 + B for B x B cells with S = B rounded
// up to even (1275 -> B <= 50, 820 -> B <= 40, 465 -> B <= 30); kinds 5
// and 6: the layout of kind 4 in k_pairs_multi, 2 consecutive tiles per
// workgroup (5) or 1 (6); kind 7: k_pairs_multi with 4 sub-counters per
// cell, slot size in groups of 4 words (5050 -> B <= 100, 2887 -> B <= 75).
// A CU holds 32 waves: 1,024-thread workgroups of the 820-group layout
// (52.5 KB) run two to a CU, 512-thread ones three (variants 59, 60: 40 bins
// 150.7 -> 135.3 ms on 4,494 dofs x 100k frames).  Block sizes stay
// multiples of 256, one wave per SIMD each: 384 and 640 threads left SIMDs
// unevenly loaded and ran slower (128.7 against 126.1 ms at 30 bins, 152.5
// against 135.5 ms at 40).  Variants 61-64: 4 sub-counters per cell sized
// for 50, 30, 40 and 20 bins (20.4, 7.4, 13.1 and 3.4 KB): for short
// trajectories, whose tiles spend much of their time clearing and reading a
// 16-copy histogram (the 25 trajectories of under 18,000 frames of a
// 64-trajectory batch at 30 bins: 119.6 against 152.4 ms).
#define PARENT_ATOMIC_VARIANTS(X)     \
    X(0, 0, 1, 1, 256, 2560)          \
    X(1, 0, 2, 2, 256, 2560)          \
    X(2, 0, 2, 2, 512, 2560)          \
    X(3, 0, 2, 4, 512, 2560)          \
    X(4, 0, 1, 4, 256, 2560)          \
    X(5, 0, 1, 8, 512, 2560)          \
    X(6, 0, 3, 3, 512, 2560)          \
    X(7, 0, 4, 4, 1024, 2560)         \
    X(8, 0, 2, 2, 1024, 2560)         \
    X(9, 0, 1, 2, 256, 2560)          \
    X(10, 0, 1, 1, 256, 5632)         \
    X(11, 0, 2, 2, 512, 5632)         \
    X(12, 0, 1, 4, 256, 5632)         \
    X(13, 0, 2, 3, 512, 5632)         \
    X(14, 0, 1, 1, 256, 10240)        \
    X(15, 0, 2, 2, 512, 10240)        \
    X(16, 0, 1, 2, 256, 10240)        \
    X(17, 0, 1, 3, 512, 10240)        \
    X(18, 0, 1, 1, 256, 40960)        \
    X(19, 1, 1, 1, 1024, 2500)        \
    X(20, 1, 1, 1, 512, 2500)         \
    X(21, 1, 1, 1, 1024, 1225)        \
    X(22, 1, 1, 2, 1024, 1225)        \
    X(23, 1, 1, 1, 512, 1225)         \
    X(24, 1, 1, 2, 1024, 900)         \
    X(25, 1, 1, 1, 512, 900)          \
    X(26, 1, 2, 2, 1024, 400)         \
    X(27, 1, 1, 4, 1024, 400)         \
    X(28, 1, 1, 1, 256, 400)          \
    X(29, 1, 2, 2, 512, 100)          \
    X(30, 1, 4, 4, 1024, 100)         \
    X(31, 2, 1, 1, 512, 2500)         \
    X(32, 2, 1, 1, 1024, 2500)        \
    X(33, 2, 1, 1, 256, 2500)         \
    X(34, 2, 1, 2, 1024, 2500)        \
    X(35, 2, 1, 1, 512, 3600)         \
    X(36, 2, 1, 1, 512, 1296)         \
    X(37, 2, 1, 2, 512, 1296)         \
    X(38, 2, 2, 2, 1024, 900)         \
    X(39, 2, 1, 1, 256, 900)          \
    X(40, 3, 1, 1, 1024, 2500)        \
    X(41, 3, 1, 1, 512, 2500)         \
    X(42, 3, 1, 1, 1024, 900)         \
    X(43, 3, 1, 1, 256, 900)          \
    X(44, 3, 1, 1, 256, 400)          \
    X(45, 4, 1, 1, 1024, 1275)        \
    X(46, 4, 1, 1, 512, 1275)         \
    X(47, 4, 1, 1, 256, 1275)         \
    X(48, 4, 1, 1, 1024, 820)         \
    X(49, 4, 1, 1, 256, 465)          \
    X(50, 4, 1, 1, 1024, 465)         \
    X(51, 5, 1, 1, 256, 465)          \
    X(52, 5, 1, 1, 256, 1275)         \
    X(53, 6, 1, 1, 1024, 1275)        \
    X(54, 6, 1, 1, 1024, 820)         \
    X(55, 6, 1, 1, 256, 465)          \
    X(56, 7, 1, 1, 1024, 5050)        \
    X(57, 7, 1, 1, 1024, 2887)        \
    X(58, 7, 1, 1, 512, 2887)         \
    X(59, 6, 1, 1, 512, 820)          \
    X(60, 5, 1, 1, 512, 820)          \
    X(61, 7, 1, 1, 256, 1275)         \
    X(62, 7, 1, 1, 256, 465)          \
    X(63, 7, 1, 1, 256, 820)          \
    X(64, 7, 1, 1, 256, 210)

struct AtomicVariant {
    AtomicConfig cfg;
    int kind;
    int slot;  // words (kind 0) or cells (kind 1)
};

// tiles per workgroup of the k_pairs_multi kinds (1 for the others)
constexpr int multi_tiles_per_workgroup(int kind) {
    return kind == 5 ? 2 : 1;
}

// One launch of a configuration; batch queues (q.descs set) run the BATCH
// instantiation, which reads each tile's inputs from q.descs.
template <int KIND, int TR, int TC, int BLK, int SL>
hipError_t launch_variant(dim3 grid, size_t lds, hipStream_t s, const PairInputs& in, const TileQueue& q, int t0,
                          int t1, const int* db, const int* de, int* dout) {
    if constexpr (KIND >= 5) {
        constexpr int TPW = multi_tiles_per_workgroup(KIND);
        constexpr int NS = KIND == 7 ? 4 : 16;
        const void* mf = q.descs ? reinterpret_cast<const void*>(&k_pairs_multi<BLK, SL, 2, TPW, true, NS>)
                                 : reinterpret_cast<const void*>(&k_pairs_multi<BLK, SL, 2, TPW, false, NS>);
        hipError_t e = hipFuncSetAttribute(mf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        // TPW consecutive tiles of an XCD's range per workgroup
        const int chunk = (t1 - t0 + 7) / 8;
        const dim3 g((unsigned)((chunk + TPW - 1) / TPW) * 8);
        if (q.descs) k_pairs_multi<BLK, SL, 2, TPW, true, NS><<<g, BLK, lds, s>>>(in, q, t0, t1, db, de, dout);
        else k_pairs_multi<BLK, SL, 2, TPW, false, NS><<<g, BLK, lds, s>>>(in, q, t0, t1, db, de, dout);
        return hipGetLastError();
    }
    constexpr int PACK = KIND == 2 || KIND == 4 ? 2 : 1;
    constexpr bool DOT = KIND >= 3;
    const void* fn;
    if constexpr (KIND == 0)
        fn = q.descs ? reinterpret_cast<const void*>(&k_pairs_atomic<TR, TC, BLK, SL, true>)
                     : reinterpret_cast<const void*>(&k_pairs_atomic<TR, TC, BLK, SL, false>);
    else
        fn = q.descs ? reinterpret_cast<const void*>(&k_pairs_sliced<TR, TC, BLK, SL, PACK, DOT, true>)
                     : reinterpret_cast<const void*>(&k_pairs_sliced<TR, TC, BLK, SL, PACK, DOT, false>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    if constexpr (KIND == 0) {
        if (q.descs) k_pairs_atomic<TR, TC, BLK, SL, true><<<grid, BLK, lds, s>>>(in, q, t0, t1, db, de, dout);
        else k_pairs_atomic<TR, TC, BLK, SL, false><<<grid, BLK, lds, s>>>(in, q, t0, t1, db, de, dout);
    } else {
        if (q.descs)
            k_pairs_sliced<TR, TC, BLK, SL, PACK, DOT, true><<<grid, BLK, lds, s>>>(in, q, t0, t1, db, de, dout);
        else
            k_pairs_sliced<TR, TC, BLK, SL, PACK, DOT, false><<<grid, BLK, lds, s>>>(in, q, t0, t1, db, de, dout);
    }
    return hipGetLastError();
}

#define PARENT_ATOMIC_ENTRY(ID, K, TR, TC, BLK, SL) AtomicVariant{AtomicConfig{TR, TC, BLK}, K, SL},
const AtomicVariant kAtomicVariants[] = {PARENT_ATOMIC_VARIANTS(PARENT_ATOMIC_ENTRY)};
#undef PARENT_ATOMIC_ENTRY

}  // namespace

int atomic_config_count() { return (int)(sizeof(kAtomicVariants) / sizeof(kAtomicVariants[0])); }
// Variant 45 (k_pairs_sliced, 1,024 threads, interleaved halves up to 50
// bins) stays out of the autotuners' candidates: on their samples it times
// up to 1% ahead of variant 53 (k_pairs_multi, the same layout and fill),
// and over whole passes it runs 1.3% behind it (163.3 against 161.0 ms at
// 4,494 dofs x 100k frames x 50 bins, on samples up to 3.4M tiles as well).
// It remains selectable with parent_engine_set_variant.
bool atomic_tuned(int id) { return id != 45; }
AtomicConfig atomic_config(int id) { return kAtomicVariants[id].cfg; }
int atomic_kind(int id) { return kAtomicVariants[id].kind; }
size_t atomic_lds_bytes(int id) {
    const AtomicVariant& v = kAtomicVariants[id];
    const size_t per_slot = v.kind == 0   ? (size_t)v.slot * 4
                            : v.kind == 2 ? (size_t)v.slot / 2 * 64
                            : v.kind == 7 ? (size_t)v.slot * 16    // 4-word groups
                                          : (size_t)v.slot * 64;   // cells (1, 3) or 16-word groups (4-6)
    return (size_t)v.cfg.tr * v.cfg.tc * per_slot;
}
bool atomic_fits(int id, int bmax, size_t lds_max, int n_frames) {
    const AtomicVariant& v = kAtomicVariants[id];
    const long long even = (bmax + 1) & ~1;
    const bool halves = v.kind == 2 || v.kind >= 4;
    const long long cells = v.kind == 2  ? (long long)bmax * even
                            : v.kind >= 4 ? (long long)(bmax - 1) * (even / 2) + bmax
                                                         : (long long)bmax * bmax;
    // a 16-bit half counts the frames of one sub-counter index: N/16 + 16
    // with 16 of them, N/4 + 16 with 4
    if (halves && n_frames > (v.kind == 7 ? (65535 - 16) * 4 : kPackedMaxFrames)) return false;
    return cells <= v.slot && atomic_lds_bytes(id) <= lds_max;
}

// the k_pairs_multi kinds with several tiles per workgroup cover as many
// times the tiles of a launch
int atomic_tiles_per_launch(int id) {
    return tiles_per_launch(kAtomicVariants[id].cfg.block) * multi_tiles_per_workgroup(kAtomicVariants[id].kind);
}

hipError_t launch_pairs_atomic(const PairInputs& in, int id, const TileQueue& q,
                               const int* dump_begin, const int* dump_entries, int* dump_out,
                               hipStream_t s) {
    if (id < 0 || id >= atomic_config_count()) return hipErrorInvalidValue;
    const size_t lds = atomic_lds_bytes(id);
    const int ntiles = q.ntiles;
    const int per_launch = atomic_tiles_per_launch(id);
    for (int t0 = 0; t0 < ntiles; t0 += per_launch) {
        const int t1 = t0 + per_launch < ntiles ? t0 + per_launch : ntiles;
        const int chunk = (t1 - t0 + 7) / 8;
        dim3 grid((unsigned)chunk * 8);
        hipError_t e = hipErrorInvalidValue;
        switch (id) {
#define PARENT_ATOMIC_CASE(ID, K, TR, TC, BLK, SL)                                                           \
    case ID:                                                                                                  \
        e = launch_variant<K, TR, TC, BLK, SL>(grid, lds, s, in, q, t0, t1, dump_begin, dump_entries, dump_out); \
        break;
            PARENT_ATOMIC_VARIANTS(PARENT_ATOMIC_CASE)
#undef PARENT_ATOMIC_CASE
            default: break;
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace parent
