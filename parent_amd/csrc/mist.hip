// MIST: maximum-information spanning tree over all dofs (get_PAR_MIST.cpp:396-766).
//
// The reference rescans every (In, Out) pair at each of the V-1 steps, O(V^3).
// Here every vertex v still Out keeps key[v] = max MI(u, v) over the In
// vertices u and the u achieving it, so a step is one argmax over Out plus
// one key update against the vertex just added: O(V^2) in total.
//
// Same result as the reference's single-thread scan.  That scan visits the
// candidates (u in In, v in Out) in the order
//     (type(u), type(v), position of u in In[type(u)], index of v)
// (category loops BB BA BD AB AA AD DB DA DD, get_PAR_MIST.cpp:480-606, then
// In position, then Out position), keeps a candidate only if its MI is
// strictly greater than the best so far, and starts from -1e200
// (get_PAR_MIST.cpp:448, 477).  So among equal MI the first candidate in
// that order wins, and an MI that is NaN or <= -1e200 never does.  The keys
// keep, per v, the u that comes first for that v (lower type, then earlier
// insertion), and the argmax breaks ties by the same 4-tuple, packed into a
// 64-bit order word (order_word below).
// In positions are insertion steps, which order each type's In vector the
// same way as the reference's push_back.
//
// MI(u, v) = (S1D[lo] + S1D[hi]) - S2D[lo, hi] with lo < hi the global ids:
// the expression of get_mutual (util.cpp:47-81) that fills the reference's
// symmetric MI matrices (get_PAR_MIST.cpp:289-326), so every MI here has the
// reference's bits.
//
// k_mist_reg runs all V-1 steps in one workgroup of 1,024 threads.  Vertex v
// belongs to thread v % 1024, slot v / 1024; a thread keeps its slots' keys
// and In positions in registers (K <= 16 slots: V <= 16,384; larger V runs
// k_mist_mem, the same steps over keys in global memory).  A step is
//   1. each thread's best candidate (kept up to date by the updates), then a
//      wave argmax (DPP lane moves),
//   2. one barrier: the 16 wave winners in LDS (double-buffered by step),
//   3. lanes 0-15 of every row read them and reduce them by DPP: every lane
//      then knows the new vertex w,
//   4. thread 0 records the edge; each thread updates its slots with MI(w, v)
//      from row w of a dense MI matrix (coalesced) or, without that matrix,
//      gathered from S1D / S2D.
// One barrier per step; the only per-step memory traffic is row w.
#include "kernels.h"

#include <map>
#include <mutex>

namespace parent {

namespace {

__device__ __forceinline__ double mutual(int nD, const double* s1d, const double* s2d, int u, int v) {
    const int lo = u < v ? u : v, hi = u < v ? v : u;
    return s1d[lo] + s1d[hi] - s2d[par_offset(nD, lo, hi)];
}

// mutual() that may also be asked for u == v (then it reads S2D[0] and the
// caller discards the value): lets a thread issue its slots' loads without
// branches
__device__ __forceinline__ double mutual_any(int nD, const double* s1d, const double* s2d, int u, int v) {
    const int lo = u < v ? u : v, hi = u < v ? v : u;
    return s1d[lo] + s1d[hi] - s2d[lo == hi ? 0 : par_offset(nD, lo, hi)];
}

// Candidate keys are compared as (MI, then the smaller order word).  An MI is
// a candidate only when it is > -1e200 (so not NaN); candidates map to an
// order-preserving unsigned 64-bit image of the double (-0 and +0 to the
// same image: the reference's '>' treats them as equal), non-candidates to
// 0.  The MI recorded for an edge is the winning key's own bits, written by
// the thread that holds it.
constexpr double kNoMi = -1e200;  // get_PAR_MIST.cpp:448, 477
__device__ __forceinline__ unsigned long long ordered(double k) {
    if (!(k > kNoMi)) return 0ull;
    const unsigned long long b = k == 0.0 ? 0ull : (unsigned long long)__double_as_longlong(k);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// (k1, o1) beats (k2, o2): larger key, then smaller order word
__device__ __forceinline__ bool wins(unsigned long long k1, unsigned long long o1, unsigned long long k2,
                                     unsigned long long o2) {
    return k1 > k2 || (k1 == k2 && o1 < o2);
}

// The order word of candidate (u, v): (type(u) << 62) | (type(v) << 60) |
// (In position of u << 30) | v.  Within a type the global id v orders like
// the index of v, so comparing these words compares the reference's scan
// order (type(u), type(v), In position of u, index of v).
__device__ __forceinline__ unsigned long long order_word(int nD, unsigned upos, int v) {
    return ((unsigned long long)(upos >> 30) << 62) | ((unsigned long long)dof_type(nD, v) << 60) |
           ((unsigned long long)(upos & 0x3FFFFFFFu) << 30) | (unsigned long long)v;
}
__device__ __forceinline__ int vertex_of(unsigned long long o) { return (int)(o & 0x3FFFFFFFull); }
__device__ __forceinline__ int in_position_of(unsigned long long o) { return (int)((o >> 30) & 0x3FFFFFFFull); }
__device__ __forceinline__ int u_type_of(unsigned long long o) { return (int)(o >> 62); }
__device__ __forceinline__ int v_type_of(unsigned long long o) { return (int)((o >> 60) & 3); }

// (type(u) << 30) | In position of u, as a slot keeps it
__device__ __forceinline__ unsigned slot_upos(unsigned tu, unsigned pos) { return (tu << 30) | pos; }

// Argmax of (key, order word) pairs over lanes, without LDS traffic: DPP lane
// moves inside each 16-lane row (pair swap, quad swap, half-row mirror, row
// mirror) leave every lane with its row's best; the four rows' bests are then
// read from lanes 0, 16, 32, 48.  Every lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp64(unsigned long long x) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)x, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), CTRL, 0xF, 0xF, false);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
template <int CTRL>
__device__ __forceinline__ void dpp_argmax_step(unsigned long long& bk, unsigned long long& bo) {
    const unsigned long long ok = dpp64<CTRL>(bk), oo = dpp64<CTRL>(bo);
    if (wins(ok, oo, bk, bo)) {
        bk = ok;
        bo = oo;
    }
}
__device__ __forceinline__ void row_argmax(unsigned long long& bk, unsigned long long& bo) {
    dpp_argmax_step<0xB1>(bk, bo);   // quad_perm [1,0,3,2]
    dpp_argmax_step<0x4E>(bk, bo);   // quad_perm [2,3,0,1]
    dpp_argmax_step<0x141>(bk, bo);  // row_half_mirror
    dpp_argmax_step<0x140>(bk, bo);  // row_mirror
}
__device__ __forceinline__ unsigned long long lane64(unsigned long long x, int l) {
    const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)x, l);
    const unsigned hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
    return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ void wave_argmax(unsigned long long& bk, unsigned long long& bo) {
    row_argmax(bk, bo);
    unsigned long long k0 = lane64(bk, 0), o0 = lane64(bo, 0);
#pragma unroll
    for (int r = 1; r < 4; r++) {
        const unsigned long long k1 = lane64(bk, 16 * r), o1 = lane64(bo, 16 * r);
        if (wins(k1, o1, k0, o0)) {
            k0 = k1;
            o0 = o1;
        }
    }
    bk = k0;
    bo = o0;
}

}  // namespace

// Step structure shared by the two MIST kernels, around the winner of the
// workgroup (wave argmax, one barrier, row argmax of the 16 wave winners,
// which lanes 0-15 of every row read: each lane then holds the winner).
__device__ __forceinline__ void block_argmax(unsigned long long (*part_k)[16], unsigned long long (*part_o)[16],
                                             int p, unsigned long long& bk, unsigned long long& bo) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    wave_argmax(bk, bo);
    if (lane == 0) {
        part_k[p][wid] = bk;
        part_o[p][wid] = bo;
    }
    __syncthreads();
    bk = part_k[p][lane & 15];
    bo = part_o[p][lane & 15];
    row_argmax(bk, bo);
}

template <int K, bool DENSE>
__global__ __launch_bounds__(1024) void k_mist_reg(MistArgs a) {
    __shared__ unsigned long long part_k[2][16], part_o[2][16];
    __shared__ int order[K * 1024];  // the vertex inserted at each step (thread 0 only)
    if (a.skip && *a.skip) return;   // the tree path has written the result
    const int V = a.V, nD = a.nD, tid = threadIdx.x;
    double key[K];
    unsigned upos[K];  // slot_upos of the slot's best u
    unsigned live = 0; // bit k: slot k is a vertex still Out
#pragma unroll
    for (int k = 0; k < K; k++) {
        key[k] = kNoMi;
        upos[k] = 0;
        const int v = tid + k * 1024;
        if (v < V && v != 0) live |= 1u << k;
    }
    if (tid == 0) {
        order[0] = 0;  // the root: bond 0 (get_PAR_MIST.cpp:419-422)
        a.flags[0] = 0;
    }
    // The thread's best candidate over its live slots, kept up to date: a
    // slot's candidate only improves (its key grows, or stays and gets an
    // earlier u), so an update compares the improved slot with the kept
    // best; only the thread that loses its slot to the tree rescans.
    unsigned long long cbk = 0, cbo = ~0ull;
    // Key update against the vertex w inserted at step `pos`.  Every slot's
    // MI is loaded, live or not (slots past V read vertex V-1), so that the
    // loads are issued together (loads under a per-slot branch would each
    // wait for the one before), in groups of G slots (with more loads in
    // flight the kernel spills).  `t` is the thread id made opaque per step, so that the
    // slots' vertex ids and offsets are recomputed rather than held in
    // registers across the steps.
    constexpr int G = K <= 8 || DENSE ? K : 4;
    auto update = [&](int w, unsigned pos, int t) {
        const unsigned tw = (unsigned)dof_type(nD, w);
        const double* row = DENSE ? a.mi_dense + (size_t)w * V : nullptr;
#pragma unroll
        for (int k0 = 0; k0 < K; k0 += G) {
            double m[G];
#pragma unroll
            for (int j = 0; j < G; j++) {
                const int v = min(t + (k0 + j) * 1024, V - 1);
                m[j] = DENSE ? row[v] : mutual_any(nD, a.s1d, a.s2d, w, v);
            }
#pragma unroll
            for (int j = 0; j < G; j++) {
                const int k = k0 + j;
                // equal MI: the u of the lower type comes first; of equal types
                // the one inserted earlier, i.e. the one kept
                if (((live >> k) & 1u) && (m[j] > key[k] || (m[j] == key[k] && tw < (upos[k] >> 30)))) {
                    key[k] = m[j];
                    upos[k] = slot_upos(tw, pos);
                    const unsigned long long kk = ordered(m[j]), o = order_word(nD, upos[k], t + k * 1024);
                    if (wins(kk, o, cbk, cbo)) {
                        cbk = kk;
                        cbo = o;
                    }
                }
            }
        }
    };
    update(0, 0, tid);
    for (int step = 0; step < V - 1; step++) {
        int t = tid;
        asm volatile("" : "+v"(t));
        unsigned long long bk = cbk, bo = cbo;
        block_argmax(part_k, part_o, step & 1, bk, bo);
        if (bk == 0) {  // no candidate left (every remaining MI NaN or <= -1e200)
            if (tid == 0) a.flags[0] = kFlagMist;
            return;
        }
        const int w = vertex_of(bo);
        if (tid == 0) {
            const int tu = u_type_of(bo), tw = v_type_of(bo);
            const int u = order[in_position_of(bo)];
            a.edges[4 * step + 0] = tu;
            a.edges[4 * step + 1] = tw;
            a.edges[4 * step + 2] = u - dof_first(nD, tu);
            a.edges[4 * step + 3] = w - dof_first(nD, tw);
            a.left[step] = V - 1 - step;
            order[step + 1] = w;
        }
        const bool owner = (w & 1023) == tid;
        if (owner) {
            const int kw = w >> 10;
#pragma unroll
            for (int k = 0; k < K; k++)
                if (k == kw) a.mi[step] = key[k];
            live &= ~(1u << kw);
        }
        update(w, (unsigned)(step + 1), t);
        if (owner) {
            // w was this thread's best: rescan its remaining slots
            cbk = 0;
            cbo = ~0ull;
#pragma unroll
            for (int k = 0; k < K; k++) {
                if (!((live >> k) & 1u)) continue;
                const unsigned long long kk = ordered(key[k]), o = order_word(nD, upos[k], t + k * 1024);
                if (wins(kk, o, cbk, cbo)) {
                    cbk = kk;
                    cbo = o;
                }
            }
        }
    }
}

// The same steps for larger V, with the keys, In positions and Out flags in
// global memory (key, upos, live: [V] each).  One pass per step updates the
// thread's slots against the new vertex and finds their best candidate, in
// groups of 8 slots whose loads are issued together (slots past V read
// vertex V-1 and are discarded).  A slot belongs to one thread for the whole
// run, so no other thread reads what it writes.
__global__ __launch_bounds__(1024) void k_mist_mem(MistArgs a) {
    __shared__ unsigned long long part_k[2][16], part_o[2][16];
    if (a.skip && *a.skip) return;
    const int V = a.V, nD = a.nD, tid = threadIdx.x;
    for (int v = tid; v < V; v += 1024) {
        a.key[v] = kNoMi;
        a.upos[v] = 0;
        a.live[v] = v != 0;
    }
    if (tid == 0) {
        a.order[0] = 0;
        a.flags[0] = 0;
    }
    auto pass = [&](int w, unsigned pos, unsigned long long& bk, unsigned long long& bo) {
        const unsigned tw = (unsigned)dof_type(nD, w);
        bk = 0;
        bo = ~0ull;
        for (int v0 = tid; v0 < V; v0 += 8 * 1024) {
            double m[8], k[8];
            unsigned up[8];
            int lv[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int v = min(v0 + j * 1024, V - 1);
                m[j] = a.mi_dense ? a.mi_dense[(size_t)w * V + v] : mutual_any(nD, a.s1d, a.s2d, w, v);
                k[j] = a.key[v];
                up[j] = a.upos[v];
                lv[j] = a.live[v];
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int v = v0 + j * 1024;
                if (v >= V || !lv[j]) continue;
                if (m[j] > k[j] || (m[j] == k[j] && tw < (up[j] >> 30))) {
                    k[j] = m[j];
                    up[j] = slot_upos(tw, pos);
                    a.key[v] = k[j];
                    a.upos[v] = up[j];
                }
                const unsigned long long kk = ordered(k[j]), o = order_word(nD, up[j], v);
                if (wins(kk, o, bk, bo)) {
                    bk = kk;
                    bo = o;
                }
            }
        }
    };
    unsigned long long bk, bo;
    pass(0, 0, bk, bo);
    for (int step = 0; step < V - 1; step++) {
        block_argmax(part_k, part_o, step & 1, bk, bo);
        if (bk == 0) {
            if (tid == 0) a.flags[0] = kFlagMist;
            return;
        }
        const int w = vertex_of(bo);
        if (tid == 0) {
            const int tu = u_type_of(bo), tw = v_type_of(bo);
            const int u = a.order[in_position_of(bo)];
            a.edges[4 * step + 0] = tu;
            a.edges[4 * step + 1] = tw;
            a.edges[4 * step + 2] = u - dof_first(nD, tu);
            a.edges[4 * step + 3] = w - dof_first(nD, tw);
            a.left[step] = V - 1 - step;
            a.order[step + 1] = w;
        }
        if ((w & 1023) == tid) {
            a.mi[step] = a.key[w];
            a.live[w] = 0;
        }
        pass(w, (unsigned)(step + 1), bk, bo);
    }
}

// Dense MI matrix [V][V] (diagonal 0): 64 x 64 tiles of the upper triangle,
// each read once from S1D / S2D and written as the tile and, through LDS,
// as its transpose, so both writes are coalesced.
__global__ __launch_bounds__(256) void k_mi_dense(int nD, int V, const double* __restrict__ s1d,
                                                  const double* __restrict__ s2d, double* __restrict__ out,
                                                  int nt) {
    __shared__ double t[64][65];
    // block -> tile (R, C) with R <= C of nt x nt tiles, row-major upper triangle
    long long b = blockIdx.x;
    int R = 0;
    while (b >= nt - R) {
        b -= nt - R;
        R++;
    }
    const int C = R + (int)b;
    const int r0 = R * 64, c0 = C * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int i = ty; i < 64; i += 4) {
        const int r = r0 + i, c = c0 + tx;
        double m = 0.0;
        if (r < V && c < V && r != c) m = mutual(nD, s1d, s2d, r, c);
        t[i][tx] = m;
        if (r < V && c < V) out[(size_t)r * V + c] = m;
    }
    if (R == C) return;
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
        const int r = c0 + i, c = r0 + tx;
        if (r < V && c < V) out[(size_t)r * V + c] = t[tx][i];
    }
}

// MIST totals in the reference's order (get_PAR_MIST.cpp:276-284, 663-741,
// 754-765): per-type S1D sums, per-category tree MI sums in step order, and
// total = sum S1D - sum tree MI.  One workgroup; the values are staged in LDS
// 1,024 at a time and thread 0 adds them in sequence.  status = 0, or
// PARENT_E_DATA (4) when the tree could not be completed.
__global__ __launch_bounds__(1024) void k_mist_totals(int nD, const double* __restrict__ s1d,
                                                      const int* __restrict__ edges, const double* __restrict__ mi,
                                                      const int* __restrict__ flags, double* totals, int* status) {
    __shared__ double buf[1024];
    __shared__ int cat[1024];
    const int V = 3 * nD + 3, tid = threadIdx.x;
    double t1[3] = {0, 0, 0}, c6[6] = {0, 0, 0, 0, 0, 0}, tmi = 0;
    for (int t = 0; t < 3; t++) {
        const int f0 = dof_first(nD, t), n = t == 0 ? nD + 2 : (t == 1 ? nD + 1 : nD);
        for (int b0 = 0; b0 < n; b0 += 1024) {
            __syncthreads();
            if (b0 + tid < n) buf[tid] = s1d[f0 + b0 + tid];
            __syncthreads();
            if (tid == 0)
                for (int i = 0; i < min(1024, n - b0); i++) t1[t] += buf[i];
        }
    }
    const bool ok = flags[0] == 0;
    for (int b0 = 0; ok && b0 < V - 1; b0 += 1024) {
        __syncthreads();
        if (b0 + tid < V - 1) {
            const int ta = edges[4 * (b0 + tid)], tb = edges[4 * (b0 + tid) + 1];
            const int lo = min(ta, tb), hi = max(ta, tb);
            buf[tid] = mi[b0 + tid];
            cat[tid] = lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
        }
        __syncthreads();
        if (tid == 0)
            for (int i = 0; i < min(1024, V - 1 - b0); i++) {
                tmi += buf[i];
                c6[cat[i]] += buf[i];
            }
    }
    if (tid != 0) return;
    for (int t = 0; t < 3; t++) totals[t] = t1[t];
    for (int c = 0; c < 6; c++) totals[3 + c] = c6[c];
    totals[9] = t1[0] + t1[1] + t1[2] - tmi;
    *status = ok ? 0 : 4;
}

// K = 16 slots take up to 116 VGPRs (1,024 threads allow 128); K = 32 spills
// ---------------------------------------------------------------------------
// Tree path: the same result from the maximum spanning tree
// ---------------------------------------------------------------------------
// If every edge Borůvka adds is the strictly largest candidate leaving its
// component (cut property with '>'), every such edge is in every maximum
// spanning tree, so that tree T is unique.  Then each non-tree pair (x, y)
// has an MI strictly below every tree edge on the path from x to y (an equal
// one could replace the path's smallest edge in T), so at any Prim step a
// non-tree candidate is strictly below the tree edge where its path crosses
// the cut.  The reference's winner at every step is therefore a tree edge
// from In to Out, and among tree edges of equal MI its scan order decides,
// which the order words reproduce.  Prim over the tree edges alone (each Out
// vertex has at most one tree neighbour in In, its parent with the tree
// rooted at vertex 0) then gives the reference's edges, MI, order and left
// counts.  Equal candidates in any component's choice, an MI that is NaN or
// <= -1e200, or V above the register form's limit: the path gives up and
// launch_mist runs the O(V^2) Prim.
//
// Borůvka: per round one wave per vertex scans its dense MI row for the best
// and second-best candidate leaving its component; one workgroup then takes
// each component's best, hooks components along the chosen edges (a pair of
// components that chose each other chose the same edge, the only cycle
// possible), and relabels by pointer jumping.  At most ceil(log2 V) rounds.
// Rooting: the tree's Euler tour (arcs by source vertex), ranked by pointer
// jumping; an arc u->v that comes before v->u makes u the parent of v.

// (best, second best, argmax) of a wave: lane shuffles
__device__ __forceinline__ unsigned long long shfl_xor64(unsigned long long x, int m) {
    const int lo = __shfl_xor((int)(unsigned)x, m), hi = __shfl_xor((int)(x >> 32), m);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

__global__ __launch_bounds__(1024) void k_tree_init(int V, MistTreeArgs t) {
    for (int v = threadIdx.x; v < V; v += 1024) t.comp[v] = v;
    if (threadIdx.x < 8) t.state[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void k_tree_scan(int V, const double* __restrict__ mi, MistTreeArgs t) {
    if (t.state[0] || t.state[2]) return;
    const int v = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (v >= V) return;
    const int cv = t.comp[v];
    const double* row = mi + (size_t)v * V;
    unsigned long long b1 = 0, b2 = 0;
    int bi = -1;
    bool bad = false;
    for (int u = lane; u < V; u += 64) {
        if (u == v) continue;
        const unsigned long long o = ordered(row[u]);
        bad |= o == 0;
        if (t.comp[u] == cv) continue;
        if (o > b1) {
            b2 = b1;
            b1 = o;
            bi = u;
        } else if (o > b2) {
            b2 = o;  // an o equal to b1 lands here: b2 = b1 marks the tie
        }
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const unsigned long long o1 = shfl_xor64(b1, m), o2 = shfl_xor64(b2, m);
        const int oi = __shfl_xor(bi, m);
        if (o1 > b1) {
            b2 = b1 > o2 ? b1 : o2;
            b1 = o1;
            bi = oi;
        } else if (o1 == b1) {
            b2 = b1;  // the same best on two lanes: two candidates (or none, at 0)
            bi = min(bi, oi);
        } else if (o1 > b2) {
            b2 = o1;
        }
    }
    if (__any(bad) && lane == 0) atomicOr(&t.state[0], 1);
    if (lane == 0) {
        t.vbest[v] = b1;
        t.vsec[v] = b2;
        t.vnbr[v] = bi;
    }
}

template <class T>
__device__ __forceinline__ T agent_load(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(1024) void k_tree_merge(int V, MistTreeArgs t) {
    __shared__ int changed;
    if (t.state[0] || t.state[2]) return;
    const int tid = threadIdx.x;
    for (int c = tid; c < V; c += 1024) {
        t.cbest[c] = 0;
        t.ccount[c] = 0;
    }
    __syncthreads();
    for (int v = tid; v < V; v += 1024) atomicMax(&t.cbest[t.comp[v]], t.vbest[v]);
    __syncthreads();
    for (int v = tid; v < V; v += 1024) {
        const int c = t.comp[v];
        const unsigned long long b = t.vbest[v];
        if (b != 0 && b == agent_load(&t.cbest[c])) {
            atomicAdd(&t.ccount[c], 1);
            __hip_atomic_store(&t.csel[c], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (t.vsec[v] == b) atomicOr(&t.state[0], 1);  // a tie inside v's row
        }
    }
    __syncthreads();
    // each component's edge, and the component it joins
    for (int c = tid; c < V; c += 1024) {
        if (t.comp[c] != c) continue;
        const unsigned long long b = agent_load(&t.cbest[c]);
        int h = c;
        if (b != 0) {
            if (agent_load(&t.ccount[c]) != 1) atomicOr(&t.state[0], 1);  // two vertices share the best
            else h = t.comp[t.vnbr[agent_load(&t.csel[c])]];
        }
        t.hook[c] = h;
    }
    __syncthreads();
    if (agent_load(&t.state[0])) return;
    // a pair that chose each other keeps its edge once, at the smaller id,
    // which stays a root; every other chosen edge is added
    for (int c = tid; c < V; c += 1024) {
        if (t.comp[c] != c) continue;
        const int d = t.hook[c];
        int h = d;
        bool add = d != c;
        if (d != c && t.hook[d] == c) {
            if (c < d) h = c;
            else add = false;
        }
        t.hook2[c] = h;
        if (add) {
            const int x = agent_load(&t.csel[c]);
            const int k = atomicAdd(&t.state[1], 1);
            if (k < V - 1) {
                t.tree[2 * k] = x;
                t.tree[2 * k + 1] = t.vnbr[x];
            } else {
                atomicOr(&t.state[0], 1);
            }
        }
    }
    __syncthreads();
    // roots by pointer jumping (labels only move towards a root, so a racing
    // read sees an ancestor either way); 2^40 exceeds any chain, so a pass
    // count past 40 can only mean a cycle: give up rather than loop
    for (int pass = 0;; pass++) {
        if (pass == 40) {
            if (tid == 0) atomicOr(&t.state[0], 1);
            return;
        }
        if (tid == 0) changed = 0;
        __syncthreads();
        for (int c = tid; c < V; c += 1024) {
            if (t.comp[c] != c) continue;
            const int h = agent_load(&t.hook2[c]), hh = agent_load(&t.hook2[h]);
            if (hh != h) {
                __hip_atomic_store(&t.hook2[c], hh, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                changed = 1;
            }
        }
        __syncthreads();
        const bool again = changed;
        __syncthreads();
        if (!again) break;
    }
    for (int v = tid; v < V; v += 1024) t.comp[v] = agent_load(&t.hook2[t.comp[v]]);
    __syncthreads();
    if (tid == 0 && agent_load(&t.state[1]) == V - 1) t.state[2] = 1;
}

// inclusive scan of n ints in place by one workgroup of 1,024 threads
__device__ void block_scan(int* x, int n, int* lds) {
    int carry = 0;
    for (int b0 = 0; b0 < n; b0 += 1024) {
        const int i = b0 + (int)threadIdx.x;
        int v = i < n ? agent_load(&x[i]) : 0;
        lds[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const int o = threadIdx.x >= (unsigned)d ? lds[threadIdx.x - d] : 0;
            __syncthreads();
            lds[threadIdx.x] += o;
            __syncthreads();
        }
        if (i < n) x[i] = lds[threadIdx.x] + carry;
        carry += lds[1023];
        __syncthreads();
    }
}

__global__ __launch_bounds__(1024) void k_tree_root(int V, const double* __restrict__ mi, MistTreeArgs t) {
    __shared__ int lds[1024];
    __shared__ int changed;
    if (t.state[0] || !t.state[2]) return;
    const int tid = threadIdx.x, E = V - 1, A = 2 * E;
    // arcs: 2i = tree[2i] -> tree[2i+1] and 2i + 1 the reverse, so arc a
    // runs from tree[a] to tree[a ^ 1]
    for (int v = tid; v <= V; v += 1024) t.arc_off[v] = 0;
    for (int v = tid; v < V; v += 1024) t.hook[v] = 0;  // fill counters
    __syncthreads();
    for (int a = tid; a < A; a += 1024) atomicAdd(&t.arc_off[t.tree[a] + 1], 1);  // source of arc a is tree[a]
    __syncthreads();
    block_scan(t.arc_off, V + 1, lds);
    __syncthreads();
    for (int a = tid; a < A; a += 1024) {
        const int u = t.tree[a];
        const int k = atomicAdd(&t.hook[u], 1);
        t.arcs[agent_load(&t.arc_off[u]) + k] = a;
        t.arc_pos[a] = k;
    }
    __syncthreads();
    // Euler tour from the first arc out of vertex 0: after arc u->v comes the
    // arc after v->u in v's list (cyclically); the arc leading back to the
    // first one ends the list
    const int a0 = agent_load(&t.arcs[agent_load(&t.arc_off[0])]);
    int* succ0 = t.succ;
    int* dist0 = t.dist;
    for (int a = tid; a < A; a += 1024) {
        const int b = a ^ 1, v = t.tree[b];  // b = v -> u
        const int off = agent_load(&t.arc_off[v]), deg = agent_load(&t.arc_off[v + 1]) - off;
        int s = agent_load(&t.arcs[off + (agent_load(&t.arc_pos[b]) + 1) % deg]);
        if (s == a0) s = -1;
        succ0[a] = s;
        dist0[a] = s < 0 ? 0 : 1;
    }
    __syncthreads();
    // list ranking: dist = arcs after a (at most 18 passes for 2^17 arcs;
    // past 40 the links cannot be a list: give up)
    int cur = 0;
    for (int pass = 0;; pass++) {
        if (pass == 40) {
            if (tid == 0) atomicOr(&t.state[0], 1);
            return;
        }
        int* sc = t.succ + cur * A;
        int* dc = t.dist + cur * A;
        int* sn = t.succ + (cur ^ 1) * A;
        int* dn = t.dist + (cur ^ 1) * A;
        if (tid == 0) changed = 0;
        __syncthreads();
        for (int a = tid; a < A; a += 1024) {
            const int s = agent_load(&sc[a]);
            if (s >= 0) {
                dn[a] = agent_load(&dc[a]) + agent_load(&dc[s]);
                sn[a] = agent_load(&sc[s]);
                changed = 1;
            } else {
                dn[a] = agent_load(&dc[a]);
                sn[a] = -1;
            }
        }
        __syncthreads();
        cur ^= 1;
        const bool again = changed;
        __syncthreads();
        if (!again) break;
    }
    int* dist = t.dist + cur * A;
    if (tid == 0) {
        t.parent[0] = -1;
        t.pmi[0] = 0.0;
    }
    for (int a = tid; a < A; a += 1024) {
        // a = u -> v comes before v -> u: u is v's parent
        if (agent_load(&dist[a]) > agent_load(&dist[a ^ 1])) {
            const int u = t.tree[a], v = t.tree[a ^ 1];
            t.parent[v] = u;
            t.pmi[v] = mi[(size_t)u * V + v];
        }
    }
    __syncthreads();
    if (tid == 0) t.state[4] = 1;
}


// Rank of each tree edge's MI (the edge into v, v != 0) among the V - 1 tree
// edges: the number of smaller ones; state[5] marks two equal ones.
__global__ __launch_bounds__(256) void k_tree_rank(int V, MistTreeArgs t) {
    __shared__ unsigned long long ks[256];
    if (t.state[0] || !t.state[4]) return;
    const int v = blockIdx.x * 256 + threadIdx.x;
    const unsigned long long kv = v >= 1 && v < V ? ordered(t.pmi[v]) : 0ull;
    int less = 0, equal = 0;
    for (int u0 = 1; u0 < V; u0 += 256) {
        const int u = u0 + threadIdx.x;
        ks[threadIdx.x] = u < V ? ordered(t.pmi[u]) : ~0ull;
        __syncthreads();
        const int n = min(256, V - u0);
        for (int j = 0; j < n; j++) {
            less += ks[j] < kv;
            equal += ks[j] == kv;
        }
        __syncthreads();
    }
    if (v >= 1 && v < V) {
        t.rank[v] = less;
        if (equal > 1) atomicOr(&t.state[5], 1);
    }
}

template <int CTRL>
__device__ __forceinline__ unsigned dpp_max(unsigned x) {
    return max(x, (unsigned)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true));
}
__device__ __forceinline__ unsigned row_max(unsigned x) {
    x = dpp_max<0xB1>(x);   // quad_perm [1,0,3,2]
    x = dpp_max<0x4E>(x);   // quad_perm [2,3,0,1]
    x = dpp_max<0x141>(x);  // row_half_mirror
    return dpp_max<0x140>(x);  // row_mirror
}

// Prim over the tree with every tree edge's MI distinct: a candidate is the
// 32-bit word (rank << 14) | v, unique, so a step is a maximum of words (DPP
// max within rows, then across rows and waves) instead of comparisons of
// (MI, order word) pairs.  Same steps, edges, MI and left counts as
// k_mist_tree; V <= 16,384.
template <int K>
__global__ __launch_bounds__(1024) void k_mist_tree32(MistArgs a, MistTreeArgs t) {
    __shared__ unsigned part[2][16];
    __shared__ int parl[K * 1024];
    if (t.state[0] || !t.state[4] || t.state[5]) return;
    const int V = a.V, nD = a.nD, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    for (int v = tid; v < V; v += 1024) parl[v] = t.parent[v];
    int par[K];
    unsigned key[K];
    double pm[K];
    unsigned live = 0, keyed = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int v = tid + k * 1024;
        par[k] = v < V ? t.parent[v] : -1;
        key[k] = v >= 1 && v < V ? ((unsigned)t.rank[v] << 14) | (unsigned)v : 0u;
        pm[k] = v < V ? t.pmi[v] : 0.0;
        if (v >= 1 && v < V) live |= 1u << k;
    }
    __syncthreads();
    unsigned cand = 0;  // the thread's best live candidate word, 0 for none
    auto update = [&](int w) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (((live >> k) & 1u) && par[k] == w) {
                keyed |= 1u << k;
                cand = max(cand, key[k]);
            }
        }
    };
    update(0);
    for (int step = 0; step < V - 1; step++) {
        unsigned b = row_max(cand);
        b = max(max((unsigned)__builtin_amdgcn_readlane((int)b, 0), (unsigned)__builtin_amdgcn_readlane((int)b, 16)),
                max((unsigned)__builtin_amdgcn_readlane((int)b, 32), (unsigned)__builtin_amdgcn_readlane((int)b, 48)));
        if (lane == 0) part[step & 1][wid] = b;
        __syncthreads();
        b = row_max(part[step & 1][lane & 15]);
        b = (unsigned)__builtin_amdgcn_readfirstlane((int)b);
        if (b == 0) return;  // no candidate: cannot happen on a spanning tree
        const int w = (int)(b & 0x3FFFu);
        if (tid == 0) {
            const int u = parl[w];
            const int tu = dof_type(nD, u), tw = dof_type(nD, w);
            a.edges[4 * step + 0] = tu;
            a.edges[4 * step + 1] = tw;
            a.edges[4 * step + 2] = u - dof_first(nD, tu);
            a.edges[4 * step + 3] = w - dof_first(nD, tw);
            a.left[step] = V - 1 - step;
        }
        const bool owner = (w & 1023) == tid;
        if (owner) {
            const int kw = w >> 10;
#pragma unroll
            for (int k = 0; k < K; k++)
                if (k == kw) a.mi[step] = pm[k];
            live &= ~(1u << kw);
        }
        update(w);
        if (owner) {
            cand = 0;
#pragma unroll
            for (int k = 0; k < K; k++)
                if ((live >> k) & (keyed >> k) & 1u) cand = max(cand, key[k]);
        }
    }
    __syncthreads();
    if (tid == 0) {
        a.flags[0] = 0;
        t.state[3] = 1;
    }
}

// Prim over the tree: k_mist_reg's steps, where inserting w makes the Out
// vertices whose parent is w candidates with key MI(w, v)
template <int K>
__global__ __launch_bounds__(1024) void k_mist_tree(MistArgs a, MistTreeArgs t) {
    __shared__ unsigned long long part_k[2][16], part_o[2][16];
    __shared__ int order[K * 1024];
    if (t.state[0] || !t.state[4] || !t.state[5]) return;  // k_mist_tree32 serves distinct MI
    const int V = a.V, nD = a.nD, tid = threadIdx.x;
    int par[K];
    double pm[K];
    unsigned upos[K];
    unsigned live = 0, keyed = 0;  // keyed: bit k once slot k is a candidate
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int v = tid + k * 1024;
        par[k] = v < V ? t.parent[v] : -1;
        pm[k] = v < V ? t.pmi[v] : 0.0;
        upos[k] = 0;
        if (v < V && v != 0) live |= 1u << k;
    }
    if (tid == 0) order[0] = 0;
    unsigned long long cbk = 0, cbo = ~0ull;
    auto update = [&](int w, unsigned pos, int tt) {
        const unsigned tw = (unsigned)dof_type(nD, w);
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (((live >> k) & 1u) && par[k] == w) {
                keyed |= 1u << k;
                upos[k] = slot_upos(tw, pos);
                const unsigned long long kk = ordered(pm[k]), o = order_word(nD, upos[k], tt + k * 1024);
                if (wins(kk, o, cbk, cbo)) {
                    cbk = kk;
                    cbo = o;
                }
            }
        }
    };
    update(0, 0, tid);
    for (int step = 0; step < V - 1; step++) {
        int tt = tid;
        asm volatile("" : "+v"(tt));
        unsigned long long bk = cbk, bo = cbo;
        block_argmax(part_k, part_o, step & 1, bk, bo);
        if (bk == 0) return;  // cannot happen under the certificate; launch_mist then redoes it
        const int w = vertex_of(bo);
        if (tid == 0) {
            const int tu = u_type_of(bo), twt = v_type_of(bo);
            const int u = order[in_position_of(bo)];
            a.edges[4 * step + 0] = tu;
            a.edges[4 * step + 1] = twt;
            a.edges[4 * step + 2] = u - dof_first(nD, tu);
            a.edges[4 * step + 3] = w - dof_first(nD, twt);
            a.left[step] = V - 1 - step;
            order[step + 1] = w;
        }
        const bool owner = (w & 1023) == tid;
        if (owner) {
            const int kw = w >> 10;
#pragma unroll
            for (int k = 0; k < K; k++)
                if (k == kw) a.mi[step] = pm[k];
            live &= ~(1u << kw);
        }
        update(w, (unsigned)(step + 1), tt);
        if (owner) {
            cbk = 0;
            cbo = ~0ull;
#pragma unroll
            for (int k = 0; k < K; k++) {
                if (!((live >> k) & 1u) || !((keyed >> k) & 1u)) continue;
                const unsigned long long kk = ordered(pm[k]), o = order_word(nD, upos[k], tt + k * 1024);
                if (wins(kk, o, cbk, cbo)) {
                    cbk = kk;
                    cbo = o;
                }
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        a.flags[0] = 0;
        t.state[3] = 1;
    }
}

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// MistTreeArgs carved from one allocation of mist_tree_bytes(V) at base
// (base == nullptr: only the size)
MistTreeArgs mist_tree_layout(char* base, int V, size_t* bytes) {
    MistTreeArgs t{};
    size_t off = 0;
    auto take = [&](size_t n) {
        char* p = base ? base + off : nullptr;
        off += align256(n);
        return p;
    };
    const size_t E2 = 2 * (size_t)(V - 1);
    t.comp = (int*)take(4 * (size_t)V);
    t.hook = (int*)take(4 * (size_t)V);
    t.hook2 = (int*)take(4 * (size_t)V);
    t.vbest = (unsigned long long*)take(8 * (size_t)V);
    t.vsec = (unsigned long long*)take(8 * (size_t)V);
    t.vnbr = (int*)take(4 * (size_t)V);
    t.cbest = (unsigned long long*)take(8 * (size_t)V);
    t.ccount = (int*)take(4 * (size_t)V);
    t.csel = (int*)take(4 * (size_t)V);
    t.tree = (int*)take(4 * E2);
    t.arc_off = (int*)take(4 * ((size_t)V + 1));
    t.arcs = (int*)take(4 * E2);
    t.arc_pos = (int*)take(4 * E2);
    t.succ = (int*)take(8 * E2);
    t.dist = (int*)take(8 * E2);
    t.parent = (int*)take(4 * (size_t)V);
    t.pmi = (double*)take(8 * (size_t)V);
    t.rank = (int*)take(4 * (size_t)V);
    t.state = (int*)take(32);
    if (bytes) *bytes = off;
    return t;
}

int mist_reg_max_vertices() { return 16 * 1024; }

hipError_t launch_mi_dense(int nD, const double* s1d, const double* s2d, double* out, hipStream_t s) {
    const int V = 3 * nD + 3, nt = (V + 63) / 64;
    k_mi_dense<<<(unsigned)((long long)nt * (nt + 1) / 2), 256, 0, s>>>(nD, V, s1d, s2d, out, nt);
    return hipGetLastError();
}

bool mist_keys_in_memory(int V, bool) { return V > mist_reg_max_vertices(); }

size_t mist_tree_bytes(int V) {
    size_t n = 0;
    (void)mist_tree_layout(nullptr, V, &n);
    return n;
}

hipError_t launch_mist_tree(const MistArgs& a, const MistTreeArgs& t, hipStream_t s) {
    const int V = a.V;
    if (!a.mi_dense || V < 2 || V > mist_reg_max_vertices()) return hipErrorInvalidValue;
    k_tree_init<<<1, 1024, 0, s>>>(V, t);
    int rounds = 0;
    while ((1 << rounds) < V) rounds++;  // components at least halve each round
    for (int r = 0; r <= rounds; r++) {
        k_tree_scan<<<(unsigned)((V + 3) / 4), 256, 0, s>>>(V, a.mi_dense, t);
        k_tree_merge<<<1, 1024, 0, s>>>(V, t);
    }
    k_tree_root<<<1, 1024, 0, s>>>(V, a.mi_dense, t);
    k_tree_rank<<<(unsigned)((V + 255) / 256), 256, 0, s>>>(V, t);
    const int k = (V + 1023) / 1024;
    // one of the two runs (tree-edge MIs distinct or not), the other returns
#define PARENT_MIST_TREE(KK)                              \
    if (k <= KK) {                                        \
        k_mist_tree32<KK><<<1, 1024, 0, s>>>(a, t);       \
        k_mist_tree<KK><<<1, 1024, 0, s>>>(a, t);         \
        return hipGetLastError();                         \
    }
    PARENT_MIST_TREE(1) PARENT_MIST_TREE(2) PARENT_MIST_TREE(4) PARENT_MIST_TREE(8) PARENT_MIST_TREE(16)
#undef PARENT_MIST_TREE
    return hipErrorInvalidValue;
}

hipError_t launch_mist(const MistArgs& a, hipStream_t s) {
    if (a.V < 2 || a.V > (1 << 24)) return hipErrorInvalidValue;
    if (mist_keys_in_memory(a.V, a.mi_dense != nullptr)) {
        if (!a.key || !a.upos || !a.live) return hipErrorInvalidValue;
        k_mist_mem<<<1, 1024, 0, s>>>(a);
        return hipGetLastError();
    }
    const int k = (a.V + 1023) / 1024;
    const bool d = a.mi_dense != nullptr;
#define PARENT_MIST(KK)                                                   \
    if (k <= KK) {                                                        \
        if (d) k_mist_reg<KK, true><<<1, 1024, 0, s>>>(a);                \
        else k_mist_reg<KK, false><<<1, 1024, 0, s>>>(a);                 \
        return hipGetLastError();                                         \
    }
    PARENT_MIST(1) PARENT_MIST(2) PARENT_MIST(4) PARENT_MIST(8) PARENT_MIST(16)
#undef PARENT_MIST
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// MIST without device allocations: its scratch (order, flags and, above 8,192
// vertices, the keys) lives in one file-scope device array per GPU, whose
// device address each GPU's first user looks up with hipGetSymbolAddress and
// keeps.  Every call on a GPU shares that GPU's array: calls that overlap on
// one GPU race.  The kernels initialise every scratch word they read.
// ---------------------------------------------------------------------------
namespace {

constexpr int kNoAllocMaxV = kMistNoAllocMaxVertices;
// byte offsets inside the array
constexpr size_t kNaFlags = 0, kNaOrder = 256, kNaKey = kNaOrder + (size_t)kNoAllocMaxV * 4,
                 kNaUpos = kNaKey + (size_t)kNoAllocMaxV * 8, kNaLive = kNaUpos + (size_t)kNoAllocMaxV * 4,
                 kNaBytes = kNaLive + (size_t)kNoAllocMaxV * 4;

}  // namespace

// 1.25 MiB per GPU, held until the process ends.  In host code `scratch`
// names a host-side stand-in: kernels get the address from
// mist_noalloc_scratch.
static __device__ __align__(256) unsigned char scratch[kNaBytes];

size_t mist_noalloc_scratch_bytes() { return kNaBytes; }

// The device address of this GPU's array (the caller has made `device`
// current).  Looked up once per device and kept.
hipError_t mist_noalloc_scratch(int device, void** out) {
    static std::mutex mu;
    static std::map<int, void*> addr;
    std::lock_guard<std::mutex> g(mu);
    auto it = addr.find(device);
    if (it != addr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
    *out = p;
    return hipSuccess;
}

hipError_t launch_mist_noalloc(int nD, const double* s1d, const double* s2d, int* edges, double* mi, int* left,
                               double* totals, int* status, void* scratch_dev, hipStream_t s) {
    const int V = 3 * nD + 3;
    if (V < 2 || V > kNoAllocMaxV || !scratch_dev) return hipErrorInvalidValue;
    unsigned char* base = static_cast<unsigned char*>(scratch_dev);
    MistArgs a{};
    a.nD = nD;
    a.V = V;
    a.s1d = s1d;
    a.s2d = s2d;
    a.mi_dense = nullptr;
    a.flags = reinterpret_cast<int*>(base + kNaFlags);
    a.order = reinterpret_cast<int*>(base + kNaOrder);
    a.key = reinterpret_cast<double*>(base + kNaKey);
    a.upos = reinterpret_cast<unsigned*>(base + kNaUpos);
    a.live = reinterpret_cast<int*>(base + kNaLive);
    a.edges = edges;
    a.mi = mi;
    a.left = left;
    hipError_t e = launch_mist(a, s);
    if (e != hipSuccess) return e;
    return launch_mist_totals(nD, s1d, edges, mi, a.flags, totals, status, s);
}

hipError_t launch_mist_totals(int nD, const double* s1d, const int* edges, const double* mi, const int* flags,
                              double* totals, int* status, hipStream_t s) {
    k_mist_totals<<<1, 1024, 0, s>>>(nD, s1d, edges, mi, flags, totals, status);
    return hipGetLastError();
}

}  // namespace parent
