// Kernel launch interfaces between the engine (host) and the HIP translation
// units.  Internal; not part of the C ABI.
#pragma once

#include "kernels_common.h"

namespace parent {

// stages of launch_prepare: the ranges and dihedral shifts (scans), the bin
// indices and 1D statistics (which read vmin/vmax/modfit), the n ln n table
enum : int { kPrepareRanges = 1, kPrepareBins = 2, kPrepareNlogn = 4 };

struct PrepareArgs {
    int stages;          // kPrepare* bits
    const void* values;  // frame-major [N][D]
    int double_prec, nD, D, N, Np;
    double pi;
    Bins3 B1, B2;
    unsigned long long *emin, *emax;  // [D] scratch
    unsigned* mfhist;                 // [nD][100] scratch
    int* flags;
    double *vmin, *vmax, *modfit;     // [D]
    void* bins2;                      // [D][Np] indices of idx2 bytes (uint8 or uint16)
    void* bins1;                      // [D][Np] of idx1 bytes, or nullptr when 1D bins == 2D bins for all types
    int idx2, idx1;                   // index bytes: 1 while the counts are <= 255, else 2
    double *s1d, *c2;                 // [D]
    double* jtab;                     // [D][jstride]: 2D-bin Jacobians
    int jstride;                      // >= every 2D bin count
    double* nlogn;                    // [N+1]: n ln n - 1/2, and 0 at n = 0
};
hipError_t launch_prepare(const PrepareArgs& a, hipStream_t s);
// n ln n for n = 0..N (FP64 table of the fused epilogues)
hipError_t launch_nlogn(int N, double* nlogn, hipStream_t s);
// raw .bat frame records (n_frames x rec_bytes, 4-byte aligned) -> frame-major
// values (float, or double when dp)
hipError_t launch_bat_records(const void* records, long long rec_bytes, int n_frames, int nD, int dp,
                              void* out, hipStream_t s);

// Common inputs of every pair kernel.
struct PairInputs {
    const uint8_t* bins;  // dof-major [D][Np] indices of idx_bytes each (uint8 for all kernels but the slab one)
    int nD, D, N, Np;
    Bins3 B;              // 2D bins per type
    const double *vmin, *vmax, *jtab, *c2, *nlogn;
    double logN;
    double* s2d;          // [D(D-1)/2] .par order
    int idx_bytes;        // 1 or 2
    int jstride;          // doubles per dof row of jtab
};

// A work list of pair tiles.  One trajectory: tiles[t] = (r0, c0), stride 2,
// descs == nullptr, the launch's PairInputs apply.  Batch of trajectories:
// tiles[t] = (r0, c0, k), stride 3, descs[k] = trajectory k's inputs (device
// memory, read once per workgroup).
struct TileQueue {
    const int* tiles;
    int ntiles;
    int stride;
    const PairInputs* descs;
};

// Optional histogram dump of selected pairs (parity checks).
struct CountsDump {
    int n;                      // number of requested pairs
    const int* pairs;           // [n][2]
    const long long* offsets;   // [n] offset of each pair's counts in `out`
    int* out;
};

// plain: one pair per workgroup
hipError_t launch_pairs_plain(const PairInputs& in, const CountsDump* dump, hipStream_t s);
// LDS bytes the plain kernel needs for 2D bin counts up to bmax (histogram and reduction arrays)
size_t plain_lds_bytes(int bmax);

// slabs: one pair per workgroup, the histogram in row slabs that fit
// lds_bytes of LDS (every bin count up to kMaxBins, uint8 or uint16
// indices); max_rows > 0 caps the rows of a slab
hipError_t launch_pairs_slabs(const PairInputs& in, size_t lds_bytes, int max_rows, const CountsDump* dump,
                              hipStream_t s);

// tiled LDS-atomic kernel
struct AtomicConfig {
    int tr, tc;    // row / column dofs per tile
    int block;     // threads per workgroup
};
int atomic_config_count();
// whether the autotuners time configuration id
bool atomic_tuned(int id);
AtomicConfig atomic_config(int id);
// 0 plain LDS histograms, 1 sliced (bank-conflict-free), 2 packed sliced,
// 3 sliced with the perm + dot2 fill, 4 interleaved packed with that fill,
// 5 and 6 the layout of 4 in k_pairs_multi (2 and 1 tiles per workgroup),
// 7 k_pairs_multi with 4 sub-counters per cell
int atomic_kind(int id);
size_t atomic_lds_bytes(int id);
// Packed sliced histograms keep 16-bit counters of the frames of one lane
// position (N / 16 + 16 of them at most).
constexpr int kPackedMaxFrames = (65535 - 16) * 16;
// histograms of bmax x bmax fit its slots and the LDS, and its counters hold n_frames
bool atomic_fits(int id, int bmax, size_t lds_max, int n_frames);
hipError_t launch_pairs_atomic(const PairInputs& in, int config_id, const TileQueue& q,
                               const int* dump_begin, const int* dump_entries, int* dump_out,
                               hipStream_t s);

// Tiles per launch of a tiled kernel with `block` threads: a multiple of 8
// (the XCD block-to-slot map) with grid x block < 2^32 (the dispatch packet
// counts work-items in 32 bits), at most 2^22.  Longer queues are issued as
// several launches of this many tiles.
inline int tiles_per_launch(int block) {
    const unsigned cap = (0xFFFFFFFFu / (unsigned)block) & ~7u;
    return cap < (1u << 22) ? (int)cap : (1 << 22);
}
int atomic_tiles_per_launch(int config_id);
int mfma_tiles_per_launch();

// one-hot int8 MFMA kernel: square tiles of mfma_tile_dofs(bmax) dofs a side
// (0 when the 2D bin count is outside the kernel's range)
int mfma_tile_dofs(int bmax);
size_t mfma_lds_bytes();
// variant 0: int8 operands (v_mfma_i32_32x32x32_i8); variant 1: FP4 one-hot
// operands (v_mfma_scale_f32_32x32x64_f8f6f4, FP32 sums), 8 to 12 bins and
// below 2^24 frames (mfma_variant_fits).  The autotuners offer variant 0
// only: variant 1 ran within -2% to +5% of it (faster at 8 bins only), and
// the tuning samples did not tell the two apart reliably.
bool mfma_variant_fits(int variant, int bmax, int n_frames);
hipError_t launch_pairs_mfma(const PairInputs& in, int bmax, int variant, const TileQueue& q, const int* dump_begin,
                             const int* dump_entries, int* dump_out, hipStream_t s);

// device Prim (MIST)
struct MistArgs {
    int nD, V;
    const double *s1d, *s2d;
    const double* mi_dense;  // [V][V] MI matrix (launch_mi_dense), or nullptr: MI gathered from s1d / s2d
    int* order;              // [V] scratch: the vertex inserted at each step (keys in memory only)
    int* flags;              // [1] scratch: kFlagMist when no candidate is left
    int* edges;              // [V-1][4] outputs, as parent_engine_mist
    double* mi;              // [V-1]
    int* left;               // [V-1]
    double* key;             // [V] scratch, only when mist_keys_in_memory(V, mi_dense != nullptr)
    unsigned* upos;          // [V] idem
    int* live;               // [V] idem
    const int* skip;         // or nullptr: nonzero when the tree path below has written the result
};

// Scratch of the tree path of MIST (launch_mist_tree), all device memory.
struct MistTreeArgs {
    int* comp;                  // [V] component of each vertex (a member's id)
    int* hook;                  // [V] per component: the component its edge joins
    int* hook2;                 // [V] idem after breaking mutual pairs, then the root
    unsigned long long* vbest;  // [V] per vertex: best outgoing candidate (order image)
    unsigned long long* vsec;   // [V] the second best
    int* vnbr;                  // [V] the neighbour of the best
    unsigned long long* cbest;  // [V] per component: the best of its vertices
    int* ccount;                // [V] its vertices holding that best
    int* csel;                  // [V] one such vertex
    int* tree;                  // [2(V-1)] tree edges (x, y)
    int* arc_off;               // [V+1] arcs by source vertex: offsets
    int* arcs;                  // [2(V-1)] arc ids by source vertex
    int* arc_pos;               // [2(V-1)] position of an arc in its source's list
    int* succ;                  // [2][2(V-1)] Euler tour links, double-buffered
    int* dist;                  // [2][2(V-1)] arcs to the end of the tour, double-buffered
    int* parent;                // [V] the tree rooted at vertex 0
    double* pmi;                // [V] MI(parent(v), v)
    int* rank;                  // [V] rank of pmi[v] among the tree edges
    int* state;                 // [8] 0 failed, 1 tree edges, 2 spanning, 3 result written, 4 rooted,
                                //     5 two tree edges of equal MI
};
// Tree path of MIST (needs a.mi_dense, V <= mist_reg_max_vertices()): the
// maximum spanning tree by Borůvka rounds, each component's edge required to
// be its strictly largest outgoing candidate, then Prim over the tree's edges
// only.  Writes a's outputs and sets t.state[3] when it succeeds; otherwise
// leaves them to launch_mist (a.skip = t.state + 3).
hipError_t launch_mist_tree(const MistArgs& a, const MistTreeArgs& t, hipStream_t s);
// device bytes of MistTreeArgs for V vertices, and its pointers carved from
// an allocation of that size at base
size_t mist_tree_bytes(int V);
MistTreeArgs mist_tree_layout(char* base, int V, size_t* bytes);
// vertices whose keys one workgroup keeps in registers
int mist_reg_max_vertices();
// whether launch_mist needs MistArgs key/upos/live (V above the register form's limit)
bool mist_keys_in_memory(int V, bool dense);
hipError_t launch_mist(const MistArgs& a, hipStream_t s);
// dense MI matrix out[V][V] of the .par entropies (diagonal 0)
hipError_t launch_mi_dense(int nD, const double* s1d, const double* s2d, double* out, hipStream_t s);
// MIST with its scratch in a static device array (no allocation), up to
// kMistNoAllocMaxVertices vertices; scratch_dev = mist_noalloc_scratch's
// address for the current device.  Outputs as launch_mist, plus the totals
// and status of launch_mist_totals.
constexpr int kMistNoAllocMaxVertices = 65536;
size_t mist_noalloc_scratch_bytes();
hipError_t mist_noalloc_scratch(int device, void** out);
hipError_t launch_mist_noalloc(int nD, const double* s1d, const double* s2d, int* edges, double* mi, int* left,
                               double* totals, int* status, void* scratch_dev, hipStream_t s);
// MIST totals (PARENT_TOT_* order) and status (0, or PARENT_E_DATA) on the device
hipError_t launch_mist_totals(int nD, const double* s1d, const int* edges, const double* mi, const int* flags,
                              double* totals, int* status, hipStream_t s);

}  // namespace parent
