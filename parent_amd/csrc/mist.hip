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
int mist_reg_max_vertices() { return 16 * 1024; }

hipError_t launch_mi_dense(int nD, const double* s1d, const double* s2d, double* out, hipStream_t s) {
    const int V = 3 * nD + 3, nt = (V + 63) / 64;
    k_mi_dense<<<(unsigned)((long long)nt * (nt + 1) / 2), 256, 0, s>>>(nD, V, s1d, s2d, out, nt);
    return hipGetLastError();
}

bool mist_keys_in_memory(int V, bool) { return V > mist_reg_max_vertices(); }

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
