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
// 64-bit order word
//     (type(u) << 60) | (type(v) << 56) | (In position of u << 28) | index of v.
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
// and In positions in registers (K slots: V <= 1024 K).  A step is
//   1. each thread's best candidate, then a wave argmax (lane shuffles),
//   2. one barrier: the 16 wave winners in LDS (double-buffered by step),
//   3. every thread reduces the 16 winners itself: all know the new vertex w,
//   4. thread 0 records the edge; each thread updates its slots with MI(w, v)
//      from row w of a dense MI matrix (coalesced) or, without that matrix,
//      gathered from S1D / S2D.
// One barrier per step; the only per-step memory traffic is row w.
#include "kernels.h"

namespace parent {

namespace {

__device__ __forceinline__ double mutual(int nD, const double* s1d, const double* s2d, int u, int v) {
    const int lo = u < v ? u : v, hi = u < v ? v : u;
    return s1d[lo] + s1d[hi] - s2d[par_offset(nD, lo, hi)];
}

// Candidate keys are compared as (MI, then the smaller order word).  An MI is
// a candidate only when it is > -1e200 (so not NaN); candidates map to an
// order-preserving unsigned 64-bit image of the double, non-candidates to 0.
constexpr double kNoMi = -1e200;  // get_PAR_MIST.cpp:448, 477
__device__ __forceinline__ unsigned long long ordered(double k) {
    if (!(k > kNoMi)) return 0ull;
    const unsigned long long b = (unsigned long long)__double_as_longlong(k);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double unordered(unsigned long long o) {
    const unsigned long long b = (o >> 63) ? (o & 0x7FFFFFFFFFFFFFFFull) : ~o;
    return __longlong_as_double((long long)b);
}

// (k1, o1) beats (k2, o2): larger key, then smaller order word
__device__ __forceinline__ bool wins(unsigned long long k1, unsigned long long o1, unsigned long long k2,
                                     unsigned long long o2) {
    return k1 > k2 || (k1 == k2 && o1 < o2);
}

__device__ __forceinline__ unsigned long long shfl_xor64(unsigned long long x, int m) {
    const int lo = __shfl_xor((int)(unsigned)x, m), hi = __shfl_xor((int)(x >> 32), m);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

__device__ __forceinline__ int vertex_of(int nD, unsigned long long o) {
    return dof_first(nD, (int)((o >> 56) & 0xF)) + (int)(o & 0xFFFFFFFull);
}

// (type(v) << 56) | index of v within its type
__device__ __forceinline__ unsigned long long vertex_word(int nD, int v) {
    const int t = dof_type(nD, v);
    return ((unsigned long long)t << 56) | (unsigned long long)(v - dof_first(nD, t));
}

}  // namespace

template <int K, bool DENSE>
__global__ __launch_bounds__(1024) void k_mist_reg(MistArgs a) {
    __shared__ unsigned long long part_k[2][16], part_o[2][16];
    const int V = a.V, nD = a.nD, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    double key[K];
    unsigned upos[K];  // (type(u) << 28) | In position of u, for the slot's best u
    unsigned live = 0; // bit k: slot k is a vertex still Out
#pragma unroll
    for (int k = 0; k < K; k++) {
        key[k] = kNoMi;
        upos[k] = 0;
        const int v = tid + k * 1024;
        if (v < V && v != 0) live |= 1u << k;
    }
    if (tid == 0) {
        a.order[0] = 0;  // the root: bond 0 (get_PAR_MIST.cpp:419-422)
        a.flags[0] = 0;
    }
    // key update against the vertex w inserted at step `pos`
    auto update = [&](int w, unsigned pos) {
        const unsigned tw = (unsigned)dof_type(nD, w);
        const double* row = DENSE ? a.mi_dense + (size_t)w * V : nullptr;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int v = tid + k * 1024;
            if (!((live >> k) & 1u)) continue;
            const double m = DENSE ? row[v] : mutual(nD, a.s1d, a.s2d, w, v);
            // equal MI: the u of the lower type comes first; of equal types the
            // one inserted earlier, i.e. the one kept
            if (m > key[k] || (m == key[k] && tw < (upos[k] >> 28))) {
                key[k] = m;
                upos[k] = (tw << 28) | pos;
            }
        }
    };
    update(0, 0);
    for (int step = 0; step < V - 1; step++) {
        unsigned long long bk = 0, bo = ~0ull;
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!((live >> k) & 1u)) continue;
            const unsigned long long kk = ordered(key[k]);
            const unsigned long long o = ((unsigned long long)(upos[k] >> 28) << 60) |
                                         ((unsigned long long)(upos[k] & 0xFFFFFFFu) << 28) |
                                         vertex_word(nD, tid + k * 1024);
            if (wins(kk, o, bk, bo)) {
                bk = kk;
                bo = o;
            }
        }
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            const unsigned long long ok = shfl_xor64(bk, m), oo = shfl_xor64(bo, m);
            if (wins(ok, oo, bk, bo)) {
                bk = ok;
                bo = oo;
            }
        }
        const int p = step & 1;
        if (lane == 0) {
            part_k[p][wid] = bk;
            part_o[p][wid] = bo;
        }
        __syncthreads();
        bk = part_k[p][0];
        bo = part_o[p][0];
#pragma unroll
        for (int q = 1; q < 16; q++) {
            const unsigned long long ok = part_k[p][q], oo = part_o[p][q];
            if (wins(ok, oo, bk, bo)) {
                bk = ok;
                bo = oo;
            }
        }
        if (bk == 0) {  // no candidate left (every remaining MI NaN or <= -1e200)
            if (tid == 0) a.flags[0] = kFlagMist;
            return;
        }
        const int w = vertex_of(nD, bo);
        if (tid == 0) {
            const int tu = (int)(bo >> 60), tw = (int)((bo >> 56) & 0xF);
            const int u = a.order[(int)((bo >> 28) & 0xFFFFFFFull)];
            a.edges[4 * step + 0] = tu;
            a.edges[4 * step + 1] = tw;
            a.edges[4 * step + 2] = u - dof_first(nD, tu);
            a.edges[4 * step + 3] = w - dof_first(nD, tw);
            a.mi[step] = unordered(bk);
            a.left[step] = V - 1 - step;
            a.order[step + 1] = w;
        }
        if ((w & 1023) == tid) live &= ~(1u << (w >> 10));
        update(w, (unsigned)(step + 1));
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

int mist_max_vertices() { return 32 * 1024; }

hipError_t launch_mi_dense(int nD, const double* s1d, const double* s2d, double* out, hipStream_t s) {
    const int V = 3 * nD + 3, nt = (V + 63) / 64;
    k_mi_dense<<<(unsigned)((long long)nt * (nt + 1) / 2), 256, 0, s>>>(nD, V, s1d, s2d, out, nt);
    return hipGetLastError();
}

hipError_t launch_mist(const MistArgs& a, hipStream_t s) {
    if (a.V < 2 || a.V > mist_max_vertices()) return hipErrorInvalidValue;
    const int k = (a.V + 1023) / 1024;
    const bool d = a.mi_dense != nullptr;
#define PARENT_MIST(KK)                                                   \
    if (k <= KK) {                                                        \
        if (d) k_mist_reg<KK, true><<<1, 1024, 0, s>>>(a);                \
        else k_mist_reg<KK, false><<<1, 1024, 0, s>>>(a);                 \
        return hipGetLastError();                                         \
    }
    PARENT_MIST(1) PARENT_MIST(2) PARENT_MIST(4) PARENT_MIST(8) PARENT_MIST(16) PARENT_MIST(32)
#undef PARENT_MIST
    return hipErrorInvalidValue;
}

hipError_t launch_mist_totals(int nD, const double* s1d, const int* edges, const double* mi, const int* flags,
                              double* totals, int* status, hipStream_t s) {
    k_mist_totals<<<1, 1024, 0, s>>>(nD, s1d, edges, mi, flags, totals, status);
    return hipGetLastError();
}

}  // namespace parent
