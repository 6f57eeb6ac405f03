// MIST: maximum-information spanning tree over all dofs (get_PAR_MIST.cpp:396-766).
//
// The reference rescans every (In, Out) pair at each of the V-1 steps, O(V^3).
// This kernel keeps, for every vertex v still Out, key[v] = max MI(u, v) over
// In vertices u and upar[v] = the u achieving it, so a step is one argmax over
// Out plus one key update against the vertex just added: O(V^2) total.
//
// Same result as the reference's single-thread scan: its scan visits the
// candidate (u in In, v in Out) in the order
//     (type(u), type(v), position of u in In[type(u)], index of v)
// (category loops BB BA BD AB AA AD DB DA DD, get_PAR_MIST.cpp:480-606, then
// In position, then Out position, with a strict '>'), so among equal MI the
// first candidate in that order wins.  key/upar keep, per v, the u that comes
// first for that v (lower type, then earlier insertion), and the argmax breaks
// ties by the same 4-tuple.  In-positions are insertion steps, which order
// the In vector of each type the same way.
//
// MI(u, v) = (S1D[lo] + S1D[hi]) - S2D[lo, hi] with lo < hi global ids, the
// expression of get_mutual (util.cpp:47-81) that fills the reference's
// symmetric MI matrices (get_PAR_MIST.cpp:289-326).
//
// One workgroup of 1024 threads runs all steps; V <= 2^24.
#include "kernels.h"

namespace parent {

namespace {

__device__ __forceinline__ double mutual(int nD, const double* s1d, const double* s2d, int u, int v) {
    const int lo = u < v ? u : v, hi = u < v ? v : u;
    return s1d[lo] + s1d[hi] - s2d[par_offset(nD, lo, hi)];
}

// tie order of candidate (upar, v): smaller comes first in the reference scan
__device__ __forceinline__ unsigned long long tie_key(int nD, int u, int upos, int v) {
    const unsigned long long tu = dof_type(nD, u), tv = dof_type(nD, v);
    const unsigned long long iv = v - dof_first(nD, (int)tv);
    return (tu << 60) | (tv << 56) | ((unsigned long long)upos << 28) | iv;
}

__device__ __forceinline__ bool better(double k1, unsigned long long o1, double k2, unsigned long long o2) {
    return k1 > k2 || (k1 == k2 && o1 < o2);
}

}  // namespace

__global__ __launch_bounds__(1024) void k_mist(MistArgs a) {
    __shared__ double sk[1024];
    __shared__ unsigned long long so[1024];
    __shared__ int sv[1024];
    const int V = a.V, nD = a.nD, tid = threadIdx.x;
    for (int v = tid; v < V; v += 1024) {
        if (v == 0) {
            a.state[v] = 1;
            a.inpos[v] = 0;
        } else {
            a.state[v] = 0;
            a.key[v] = mutual(nD, a.s1d, a.s2d, 0, v);
            a.upar[v] = 0;
        }
    }
    __syncthreads();
    for (int step = 0; step < V - 1; step++) {
        double bk = -1.0e300;
        unsigned long long bo = ~0ull;
        int bv = -1;
        for (int v = tid; v < V; v += 1024) {
            if (a.state[v]) continue;
            const double k = a.key[v];
            const int u = a.upar[v];
            const unsigned long long o = tie_key(nD, u, a.inpos[u], v);
            if (bv < 0 || better(k, o, bk, bo)) {
                bk = k;
                bo = o;
                bv = v;
            }
        }
        sk[tid] = bk;
        so[tid] = bo;
        sv[tid] = bv;
        __syncthreads();
        for (int s = 512; s > 0; s >>= 1) {
            if (tid < s) {
                const int ov = sv[tid + s];
                if (ov >= 0 && (sv[tid] < 0 || better(sk[tid + s], so[tid + s], sk[tid], so[tid]))) {
                    sk[tid] = sk[tid + s];
                    so[tid] = so[tid + s];
                    sv[tid] = ov;
                }
            }
            __syncthreads();
        }
        const int w = sv[0];
        if (w < 0) {
            if (tid == 0) atomicOr(a.flags, kFlagMist);
            return;
        }
        if (tid == 0) {
            const int u = a.upar[w];
            const int tu = dof_type(nD, u), tw = dof_type(nD, w);
            a.edges[4 * step + 0] = tu;
            a.edges[4 * step + 1] = tw;
            a.edges[4 * step + 2] = u - dof_first(nD, tu);
            a.edges[4 * step + 3] = w - dof_first(nD, tw);
            a.mi[step] = a.key[w];
            a.left[step] = V - 1 - step;
            a.state[w] = 1;
            a.inpos[w] = step + 1;
        }
        __syncthreads();
        const int tw = dof_type(nD, w);
        for (int v = tid; v < V; v += 1024) {
            if (a.state[v]) continue;
            const double m = mutual(nD, a.s1d, a.s2d, w, v);
            const double k = a.key[v];
            if (m > k || (m == k && tw < dof_type(nD, a.upar[v]))) {
                a.key[v] = m;
                a.upar[v] = w;
            }
        }
        __syncthreads();
    }
}

hipError_t launch_mist(const MistArgs& a, hipStream_t s) {
    k_mist<<<1, 1024, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace parent
