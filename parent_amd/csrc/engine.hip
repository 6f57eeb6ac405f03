// Engine: device buffers, phase sequencing and the C ABI of parent_hip.h.
//
// One engine = one HIP stream on one GPU.  The trajectory lives in HBM as
// frame-major values; prepare() turns it into dof-major uint8 bin indices
// plus per-dof constants; pairs() runs one of the pair kernels over all
// D(D-1)/2 dof pairs and leaves S2D in .par body order on the device.
//
// State is split in two: a Traj (the trajectory on the device) and one
// BinSet per bin set (indices, per-dof constants, S1D/S2D and the kernel
// chosen for it).  prepare()/pairs() use one BinSet; a sweep fills several
// over the same trajectory and replays them as one captured hipGraph; a batch
// (parent_batch_*) holds one (Traj, BinSet) per trajectory and runs all their
// pair tiles through one launch.
#include "../../include/parent_hip.h"
#include "host_common.h"
#include "kernels.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

using namespace parent;

namespace {

// Owning device allocation (move-only).
struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    hipError_t ensure(size_t bytes) {
        if (bytes <= n && p) return hipSuccess;
        release();
        if (bytes == 0) return hipSuccess;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = bytes;
        else p = nullptr;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

struct TileList {
    int tr = -1, tc = -1, D = -1, Np = -1;  // the list's tile shape, dofs and band (by Np)
    int ntiles = 0;
    DevBuf dev;
};

// A trajectory resident on the device.
struct Traj {
    int nD = -1, N = 0, dp = 0, D = 0, Np = 0;
    const void* values = nullptr;
    DevBuf values_own;
};

// One bin set's results over a trajectory.
struct BinSet {
    int bins[6] = {0, 0, 0, 0, 0, 0};
    bool prepared = false;   // indices and 1D results valid
    bool s2d_valid = false;  // S2D valid
    int ent_nD = -1;         // nD of the entropies in s1d/s2d
    int method = PARENT_METHOD_PLAIN, variant = 0;  // kernel of the last pair pass
    int idx2 = 1, idx1 = 1;  // bytes per bin index of bins2 / bins1: 1 up to 255 bins, else 2
    int jstride = 256;       // doubles per dof row of jtab
    DevBuf flags, vmin, vmax, modfit, bins2, bins1, s1d, c2, jtab, s2d;
    TileList tiles;
    // the set whose vmin/vmax/modfit this one uses: itself (nullptr), or, in a
    // sweep, the first set over the same trajectory (the ranges and the
    // dihedral shift do not depend on the bin counts, PARENT.cpp:503-637)
    const BinSet* ranges_from = nullptr;
    const BinSet& ranges() const { return ranges_from ? *ranges_from : *this; }
};

// Scratch of the prepare and MIST kernels, shared by the bin sets of a stream
// (they run one after another on it).
struct Scratch {
    DevBuf emin, emax, mfhist, nlogn;
    DevBuf mdense, morder, mkey, mupos, mlive, medges, mmi, mleft, mflags, mtree;
};

}  // namespace

struct parent_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    std::string err;
    int lds_max = 65536;
    int cus = 256;  // compute units, sizes the autotuner's samples
    void* noalloc_scratch = nullptr;  // this device's static MIST scratch (device address)

    Traj traj;
    Scratch scratch;
    // sets[0] belongs to prepare()/pairs(); a sweep fills sets[0..n-1]; the
    // readers, pair_counts and mist act on sets[active]
    std::vector<std::unique_ptr<BinSet>> sets;
    int active = 0;
    parent_stats stats{};

    // autotuner: the fastest pair-kernel configuration per problem key
    struct Tuned {
        int b, a, d, N, D;
        int method, variant;
    };
    std::vector<Tuned> tuned;
    std::vector<parent_tune_entry> last_tune;
    int forced_atomic = -1;  // parent_engine_set_variant
    int forced_mfma = -1;    // parent_engine_set_variant(PARENT_METHOD_MFMA, ...)
    int forced_slab_rows = 0;
    int mist_path_mode = 0;    // parent_engine_set_mist_path  // idem for PARENT_METHOD_SLABS: rows per slab (0: as many as the LDS holds)

    // the captured sweep
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<int> graph_key;
    long long traj_serial = 0;  // bumped by every set_trajectory
};

struct parent_batch {
    parent_engine* eng = nullptr;
    int bins[6] = {0, 0, 0, 0, 0, 0};
    struct Item {
        Traj traj;
        BinSet set;
    };
    std::vector<std::unique_ptr<Item>> items;
    // One kernel launch of a run: some of the items (indices into items) and
    // the configuration they run with.
    struct Plan {
        std::vector<int> members;
        int method = -1, variant = -1;
    };
    // A launch's merged tile queue on the device, built for one Plan.
    struct Queue {
        Plan plan;
        DevBuf tiles;
        int ntiles = 0;
    };
    std::vector<std::unique_ptr<Queue>> queues;  // the last run's launches
    DevBuf descs;                                // PairInputs of every item
    int descs_items = -1;
    const void* descs_nlogn = nullptr;  // the n ln n table the uploaded descriptors point to
    int method = PARENT_METHOD_PLAIN, variant = 0;
    int classes = 1;                      // parent_batch_set_classes
    int tuned_items = -1;                 // items when tuned_plan was made
    std::vector<Plan> tuned_plan;         // the batch autotuner's launches
};

namespace {

int fail(parent_engine* e, int code, const std::string& msg) {
    e->err = msg;
    return code;
}

int hip_fail(parent_engine* e, hipError_t h, const char* what) {
    return fail(e, PARENT_E_DEVICE, std::string(what) + ": " + hipGetErrorString(h));
}

#define HIPCHK(eng, call)                                          \
    do {                                                           \
        hipError_t _h = (call);                                    \
        if (_h != hipSuccess) return hip_fail((eng), _h, #call);   \
    } while (0)

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

Bins3 bins2d(const BinSet& s) { return Bins3{s.bins[3], s.bins[4], s.bins[5]}; }
Bins3 bins1d(const BinSet& s) { return Bins3{s.bins[0], s.bins[1], s.bins[2]}; }
int bmax2d(const BinSet& s) { return std::max(s.bins[3], std::max(s.bins[4], s.bins[5])); }
int bmax1d(const BinSet& s) { return std::max(s.bins[0], std::max(s.bins[1], s.bins[2])); }

BinSet& cur(parent_engine* e) { return *e->sets[e->active]; }

int check_bins(parent_engine* e, const int* bins) {
    for (int k = 0; k < 6; k++) {
        if (bins[k] < 1) return fail(e, PARENT_E_INVALID, "number of bins < 1");
        if (bins[k] > kMaxBins) return fail(e, PARENT_E_UNSUPPORTED, "more than 4096 bins per dof is not supported");
    }
    return PARENT_OK;
}

PairInputs pair_inputs(const Traj& t, BinSet& s) {
    PairInputs in{};
    in.bins = s.bins2.as<uint8_t>();
    in.nD = t.nD;
    in.D = t.D;
    in.N = t.N;
    in.Np = t.Np;
    in.B = bins2d(s);
    in.vmin = s.ranges().vmin.as<double>();
    in.vmax = s.ranges().vmax.as<double>();
    in.jtab = s.jtab.as<double>();
    in.c2 = s.c2.as<double>();
    in.logN = log((double)t.N);
    in.s2d = s.s2d.as<double>();
    in.idx_bytes = s.idx2;
    in.jstride = s.jstride;
    return in;
}

// Device buffers of bin set s over trajectory t, and the shared scratch.
int alloc_set(parent_engine* e, const Traj& t, BinSet& s) {
    const int D = t.D, nD = t.nD;
    const bool need1 = s.bins[0] != s.bins[3] || s.bins[1] != s.bins[4] || s.bins[2] != s.bins[5];
    s.idx2 = bmax2d(s) > kMaxBins8 ? 2 : 1;
    s.idx1 = need1 ? (bmax1d(s) > kMaxBins8 ? 2 : 1) : s.idx2;
    s.jstride = std::max(bmax2d(s), 1);
    Scratch& x = e->scratch;
    HIPCHK(e, x.emin.ensure(D * 8));
    HIPCHK(e, x.emax.ensure(D * 8));
    HIPCHK(e, x.mfhist.ensure((size_t)nD * kModfitBins * 4));
    HIPCHK(e, x.nlogn.ensure(((size_t)t.N + 1) * 8));
    HIPCHK(e, s.flags.ensure(64));
    if (s.ranges_from) {
        s.vmin.release();
        s.vmax.release();
        s.modfit.release();
    } else {
        HIPCHK(e, s.vmin.ensure(D * 8));
        HIPCHK(e, s.vmax.ensure(D * 8));
        HIPCHK(e, s.modfit.ensure(D * 8));
    }
    HIPCHK(e, s.bins2.ensure((size_t)D * t.Np * s.idx2));
    if (need1) HIPCHK(e, s.bins1.ensure((size_t)D * t.Np * s.idx1));
    else s.bins1.release();
    HIPCHK(e, s.s1d.ensure(D * 8));
    HIPCHK(e, s.c2.ensure(D * 8));
    HIPCHK(e, s.jtab.ensure((size_t)D * s.jstride * 8));
    HIPCHK(e, s.s2d.ensure((size_t)std::max<long long>(n_pairs(nD), 1) * 8));
    return PARENT_OK;
}

// Ranges, modfit, pre-binning and 1D entropies of set s: kernels only, no
// host synchronisation (capturable).  The n ln n table of t.N goes to the
// shared scratch.  A set that takes its ranges from another one runs only
// the pre-binning and 1D kernels.
hipError_t enqueue_prepare(parent_engine* e, const Traj& t, BinSet& s) {
    const bool need1 = s.bins1.p != nullptr;
    PrepareArgs a{};
    a.stages = s.ranges_from ? kPrepareBins : (kPrepareRanges | kPrepareBins | kPrepareNlogn);
    a.values = t.values;
    a.double_prec = t.dp;
    a.nD = t.nD;
    a.D = t.D;
    a.N = t.N;
    a.Np = t.Np;
    a.pi = acos(-1.0);  // PARENT.cpp:112
    a.B1 = bins1d(s);
    a.B2 = bins2d(s);
    a.emin = e->scratch.emin.as<unsigned long long>();
    a.emax = e->scratch.emax.as<unsigned long long>();
    a.mfhist = e->scratch.mfhist.as<unsigned>();
    a.flags = s.flags.as<int>();
    a.vmin = const_cast<double*>(s.ranges().vmin.as<double>());
    a.vmax = const_cast<double*>(s.ranges().vmax.as<double>());
    a.modfit = const_cast<double*>(s.ranges().modfit.as<double>());
    a.bins2 = s.bins2.p;
    a.bins1 = need1 ? s.bins1.p : nullptr;
    a.idx2 = s.idx2;
    a.idx1 = s.idx1;
    a.jstride = s.jstride;
    a.s1d = s.s1d.as<double>();
    a.c2 = s.c2.as<double>();
    a.jtab = s.jtab.as<double>();
    a.nlogn = e->scratch.nlogn.as<double>();
    return launch_prepare(a, e->stream);
}

// Reads set s's data flags (synchronises) and maps them to the reference's errors.
int check_flags(parent_engine* e, BinSet& s) {
    int flags = 0;
    HIPCHK(e, hipMemcpyAsync(&flags, s.flags.p, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    if (flags & kFlagNegBond) return fail(e, PARENT_E_DATA, "ERROR: a bond length is smaller than 0.0");
    if (flags & kFlagRange)
        return fail(e, PARENT_E_RANGE, "ERROR: a value falls outside its histogram range (angle >= pi or dihedral outside [0, 2pi+5e-9))");
    return PARENT_OK;
}

// Eager prepare of set s: enqueue, time, check.
int run_prepare(parent_engine* e, const Traj& t, BinSet& s) {
    s.prepared = false;
    s.s2d_valid = false;
    if (int rc = alloc_set(e, t, s)) return rc;
    HIPCHK(e, hipEventRecord(e->ev[2], e->stream));
    HIPCHK(e, enqueue_prepare(e, t, s));
    HIPCHK(e, hipEventRecord(e->ev[3], e->stream));
    if (int rc = check_flags(e, s)) return rc;
    e->stats.ms_prepare = elapsed(e->ev[2], e->ev[3]);
    s.prepared = true;
    s.ent_nD = t.nD;
    return PARENT_OK;
}

// Tiles of a TR x TC configuration over the upper triangle: row blocks at
// multiples of TR, column blocks at multiples of TC, kept when some column of
// the block lies above the block's first row.  Order: bands of ~32 rows, and
// within a band column by column, so the workgroups running at the same time
// (one contiguous run of the list per XCD) read the same column-dof index rows
// from L2 while their band's row-dof rows stay resident there.  A taller band
// re-reads each column row from HBM fewer times, as long as the band's rows fit
// the XCD's 4 MB L2 beside the columns in flight: bands of about 3.2 MB of
// index rows, 8 to 32 rows (32 at 100k frames: half the HBM reads of 16 rows;
// 16 at 200k frames, where 32 rows of 6.4 MB ran 1.7% slower).
void push_band(std::vector<int>& t, int D, int tr, int tc, int b0, int b1) {
    const int cfirst = ((b0 + 1) / tc) * tc;
    for (int c0 = cfirst; c0 < D; c0 += tc)
        for (int r0 = b0; r0 < b1; r0 += tr)
            if (c0 + tc - 1 > r0) {
                t.push_back(r0);
                t.push_back(c0);
            }
}

int tile_band(int tr, int Np) {
    // MFMA tiles (8 dofs a side and more, B <= 16; the atomic kernels' tiles
    // have at most 4): 8 tile rows a band, so the ~128 workgroups of an XCD
    // run about 8 x 16 tiles at once and read each chunk of a row or column
    // dof's indices from L2 for all of them
    if (tr >= 8) return 8 * tr;
    const int rows = std::min(32, std::max(8, (int)(3276800LL / std::max(Np, 1))));
    return std::max(tr, rows / tr * tr);
}
std::vector<int> make_tiles(int D, int Np, int tr, int tc) {
    std::vector<int> t;
    const int band = tile_band(tr, Np);
    for (int b0 = 0; b0 < D - 1; b0 += band) push_band(t, D, tr, tc, b0, std::min(b0 + band, D - 1));
    return t;
}

// Number of tiles make_tiles(D, Np, tr, tc) lists, without building the list.
long long count_tiles(int D, int Np, int tr, int tc) {
    long long n = 0;
    const int band = tile_band(tr, Np);
    for (int b0 = 0; b0 < D - 1; b0 += band) {
        const int b1 = std::min(b0 + band, D - 1);
        const int cfirst = ((b0 + 1) / tc) * tc;
        for (int r0 = b0; r0 < b1; r0 += tr) {
            // columns c0 = cfirst + k tc < D with c0 + tc - 1 > r0
            const int lo = r0 - tc + 2;
            const int c_start = cfirst >= lo ? cfirst : cfirst + (lo - cfirst + tc - 1) / tc * tc;
            if (c_start < D) n += (D - c_start + tc - 1) / tc;
        }
    }
    return n;
}

// Tile shape (row dofs x column dofs) of a tiled method's variant.
void tile_shape(const BinSet& s, int method, int variant, int* tr, int* tc) {
    if (method == PARENT_METHOD_MFMA) {
        *tr = *tc = mfma_tile_dofs(bmax2d(s));
    } else {
        const AtomicConfig c = atomic_config(variant);
        *tr = c.tr;
        *tc = c.tc;
    }
}

int ensure_tiles(parent_engine* e, const Traj& t, BinSet& s, int method, int variant) {
    int tr, tc;
    tile_shape(s, method, variant, &tr, &tc);
    if (s.tiles.tr == tr && s.tiles.tc == tc && s.tiles.D == t.D && s.tiles.Np == t.Np) return PARENT_OK;
    std::vector<int> tl = make_tiles(t.D, t.Np, tr, tc);
    HIPCHK(e, s.tiles.dev.ensure(std::max<size_t>(tl.size(), 3) * sizeof(int)));
    if (!tl.empty())
        HIPCHK(e, hipMemcpyAsync(s.tiles.dev.p, tl.data(), tl.size() * sizeof(int), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    s.tiles.tr = tr;
    s.tiles.tc = tc;
    s.tiles.D = t.D;
    s.tiles.Np = t.Np;
    s.tiles.ntiles = (int)(tl.size() / 2);
    return PARENT_OK;
}

bool variant_fits(const parent_engine* e, const BinSet& s, int id, int n_frames) {
    return atomic_fits(id, bmax2d(s), (size_t)e->lds_max, n_frames);
}

bool mfma_fits(const parent_engine* e, const BinSet& s) {
    return mfma_tile_dofs(bmax2d(s)) > 0 && mfma_lds_bytes() <= (size_t)e->lds_max;
}

PairInputs with_nlogn(parent_engine* e, PairInputs in) {
    in.nlogn = e->scratch.nlogn.as<double>();
    return in;
}

// one launch of a tiled method's variant over a tile queue
hipError_t launch_tiled(parent_engine* e, const PairInputs& in, int bmax, int method, int id, const TileQueue& q,
                        const int* db, const int* de, int* dout) {
    if (method == PARENT_METHOD_MFMA) return launch_pairs_mfma(in, bmax, id, q, db, de, dout, e->stream);
    return launch_pairs_atomic(in, id, q, db, de, dout, e->stream);
}

// pairs (gi < gj < D) covered by a work list of tr x tc tiles
long long tile_pairs(int D, const std::vector<int>& t, int tr, int tc) {
    long long n = 0;
    for (size_t k = 0; k + 1 < t.size(); k += 2)
        for (int gi = t[k]; gi < std::min(t[k] + tr, D); gi++) {
            const int lo = std::max(t[k + 1], gi + 1), hi = std::min(t[k + 1] + tc, D);
            if (hi > lo) n += hi - lo;
        }
    return n;
}

int choose_atomic_config(const parent_engine* e, const BinSet& s, int n_frames) {
    if (e->forced_atomic >= 0) return variant_fits(e, s, e->forced_atomic, n_frames) ? e->forced_atomic : -1;
    // fallback order when not autotuning: the best of the autotuner's tables
    // at 50 bins first, then configurations with larger histogram slots
    static const int order[] = {2, 3, 1, 9, 0, 11, 13, 10, 15, 16, 14, 18};
    for (int id : order)
        if (variant_fits(e, s, id, n_frames)) return id;
    return -1;
}

// Rows sampled by the autotuner: `runs` runs of `height` rows (a multiple of
// 12) spread over the dof range, each starting on a multiple of 12.  Each
// configuration's rate is its own work list's pair count (tile heights that
// do not divide 12 cover a few more rows) over its time.
std::vector<std::pair<int, int>> tune_rows(int D, int runs = 8, int height = 12) {
    std::vector<std::pair<int, int>> rows;
    for (int k = 0; k < runs; k++) {
        int r = (int)((k + 0.5) * (D - 1) / runs) / 12 * 12;
        if (!rows.empty()) r = std::max(r, rows.back().second);
        int r1 = std::min(r + height, D - 1);
        if (r < r1) rows.push_back({r, r1});
    }
    return rows;
}

std::vector<int> tune_work(int D, int tr, int tc, const std::vector<std::pair<int, int>>& rows) {
    std::vector<int> t;
    for (auto& rr : rows) push_band(t, D, tr, tc, rr.first, rr.second);
    return t;
}

// A tuning work list of `runs` runs for tiles of tr x tc dofs, the runs made
// taller (12, 24, 48, ... rows) until the list holds at least min_tiles tiles
// or covers every row: a configuration with large tiles (the MFMA kernel at a
// few bins: 25 x 25 dofs) otherwise gets a fraction of a wave of workgroups
// and is timed far below its full-size rate.
std::vector<int> tune_sample(int D, int tr, int tc, int runs, size_t min_tiles) {
    for (int h = 12;; h *= 2) {
        std::vector<int> t = tune_work(D, tr, tc, tune_rows(D, runs, h));
        if (t.size() / 2 >= min_tiles || (size_t)h * runs >= (size_t)D) return t;
    }
}

// Times `launch` `reps` times with events on the engine stream and keeps the
// fastest run after the first (which warms the caches).
template <class F>
int time_launch(parent_engine* e, F launch, double* ms, int reps = 2) {
    *ms = 0;
    for (int rep = 0; rep < reps; rep++) {
        HIPCHK(e, hipEventRecord(e->ev[4], e->stream));
        HIPCHK(e, launch());
        HIPCHK(e, hipEventRecord(e->ev[5], e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
        const double t = elapsed(e->ev[4], e->ev[5]);
        if (rep == 1 || (rep > 1 && t < *ms)) *ms = t;
    }
    return PARENT_OK;
}

// Two stages: every fitting configuration is timed on one run in the middle
// of the dof range, then the kTuneFinalists fastest are timed again, best of
// kFinalReps, on kFinalRuns runs spread over it, and the fastest of those is
// kept.  Samples hold at least kProbeTilesPerCu / kFinalTilesPerCu tiles per
// compute unit.  The table holds one row per configuration, the finalists'
// rows from the second stage.
constexpr int kTuneFinalists = 3, kFinalRuns = 32, kFinalReps = 3;
constexpr int kProbeTilesPerCu = 16, kFinalTilesPerCu = 32;

int run_autotune(parent_engine* e, const Traj& tj, BinSet& s) {
    e->last_tune.clear();
    const int D = tj.D;
    const PairInputs in = with_nlogn(e, pair_inputs(tj, s));
    DevBuf dt;
    // candidates: every fitting atomic configuration, then the MFMA kernel
    std::vector<std::pair<int, int>> cand;
    for (int id = 0; id < atomic_config_count(); id++)
        if (atomic_tuned(id) && variant_fits(e, s, id, tj.N)) cand.push_back({PARENT_METHOD_ATOMIC, id});
    if (mfma_fits(e, s)) cand.push_back({PARENT_METHOD_MFMA, 0});
    auto measure = [&](int m, int id, int runs, int tiles_per_cu, int reps, parent_tune_entry* te) -> int {
        int tr, tc;
        tile_shape(s, m, id, &tr, &tc);
        const int block = m == PARENT_METHOD_MFMA ? 256 : atomic_config(id).block;
        std::vector<int> t = tune_sample(D, tr, tc, runs, (size_t)tiles_per_cu * e->cus);
        const double updates = (double)tile_pairs(D, t, tr, tc) * tj.N;
        HIPCHK(e, dt.ensure(std::max<size_t>(t.size(), 2) * sizeof(int)));
        HIPCHK(e, hipMemcpyAsync(dt.p, t.data(), t.size() * sizeof(int), hipMemcpyHostToDevice, e->stream));
        double ms = 0;
        const TileQueue q{dt.as<int>(), (int)(t.size() / 2), 2, nullptr};
        if (int rc = time_launch(e, [&] { return launch_tiled(e, in, bmax2d(s), m, id, q, nullptr, nullptr, nullptr); }, &ms, reps))
            return rc;
        *te = parent_tune_entry{m, id, tr, tc, block, ms, updates, updates / (ms * 1e-3)};
        return PARENT_OK;
    };
    const bool two_stage = (int)cand.size() > kTuneFinalists;
    for (auto& mv : cand) {
        parent_tune_entry te;
        const int rc = two_stage ? measure(mv.first, mv.second, 1, kProbeTilesPerCu, 2, &te)
                                 : measure(mv.first, mv.second, kFinalRuns, kFinalTilesPerCu, kFinalReps, &te);
        if (rc) return rc;
        e->last_tune.push_back(te);
    }
    // the contenders: every row, or the finalists re-timed on the larger sample
    std::vector<int> order(e->last_tune.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = (int)k;
    if (two_stage) {
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return e->last_tune[x].updates_per_s > e->last_tune[y].updates_per_s;
        });
        order.resize(kTuneFinalists);
        for (int k : order) {
            parent_tune_entry& te = e->last_tune[k];
            if (int rc = measure(te.method, te.variant, kFinalRuns, kFinalTilesPerCu, kFinalReps, &te)) return rc;
        }
    }
    int best_m = PARENT_METHOD_PLAIN, best_v = 0;
    double best_rate = -1;
    for (int k : order) {
        const parent_tune_entry& te = e->last_tune[k];
        if (te.updates_per_s > best_rate) {
            best_rate = te.updates_per_s;
            best_m = te.method;
            best_v = te.variant;
        }
    }
    e->tuned.push_back({s.bins[3], s.bins[4], s.bins[5], tj.N, D, best_m, best_v});
    return PARENT_OK;
}

// The tuned (method, variant) for this problem, tuning on first use.
int tuned_choice(parent_engine* e, const Traj& tj, BinSet& s, int* method, int* variant) {
    for (auto& t : e->tuned)
        if (t.b == s.bins[3] && t.a == s.bins[4] && t.d == s.bins[5] && t.N == tj.N && t.D == tj.D) {
            *method = t.method;
            *variant = t.variant;
            return PARENT_OK;
        }
    if (int rc = run_autotune(e, tj, s)) return rc;
    const auto& t = e->tuned.back();
    *method = t.method;
    *variant = t.variant;
    return PARENT_OK;
}

// The slab kernel serves every bin count; the others need uint8 indices and
// a whole histogram in LDS (the plain kernel's is the largest).
bool needs_slabs(const parent_engine* e, const BinSet& s) {
    return s.idx2 != 1 || plain_lds_bytes(bmax2d(s)) > (size_t)e->lds_max;
}

// LDS of one slab-kernel workgroup: the whole histogram when it fits, else
// what the device allows beside the kernel's static arrays
size_t slab_lds(const parent_engine* e, const BinSet& s) {
    const size_t b = (size_t)bmax2d(s), room = (size_t)e->lds_max - 256;
    return std::min(b * b * 4, room) / 4 * 4;
}

// The kernel a request runs on: bin counts the LDS kernels cannot hold run
// on the slab kernel whatever was asked; AUTO takes the autotuner's choice,
// MFMA falls back to the atomic kernels outside its bin range, ATOMIC falls
// back to the plain kernel when no tile configuration fits the LDS.
int resolve_method(parent_engine* e, const Traj& tj, BinSet& s, int method, int* used, int* variant) {
    if (method < PARENT_METHOD_PLAIN || method > PARENT_METHOD_SLABS) return fail(e, PARENT_E_INVALID, "unknown method");
    *used = method;
    *variant = 0;
    if (method == PARENT_METHOD_SLABS || needs_slabs(e, s)) {
        *used = PARENT_METHOD_SLABS;
        *variant = e->forced_slab_rows;
        return PARENT_OK;
    }
    if (method == PARENT_METHOD_AUTO) return tuned_choice(e, tj, s, used, variant);
    if (method == PARENT_METHOD_MFMA) {
        if (mfma_fits(e, s)) {
            *variant = e->forced_mfma >= 0 && mfma_variant_fits(e->forced_mfma, bmax2d(s), tj.N) ? e->forced_mfma : 0;
            return PARENT_OK;
        }
        *used = PARENT_METHOD_ATOMIC;
    }
    if (*used == PARENT_METHOD_ATOMIC) {
        *variant = choose_atomic_config(e, s, tj.N);
        if (*variant < 0) {
            *used = PARENT_METHOD_PLAIN;
            *variant = 0;
        }
    }
    return PARENT_OK;
}

// Chooses set s's kernel and uploads its tile list (synchronises; done
// before any capture).
int plan_pairs(parent_engine* e, const Traj& t, BinSet& s, int method) {
    int used = method, variant = 0;
    if (int rc = resolve_method(e, t, s, method, &used, &variant)) return rc;
    if (used != PARENT_METHOD_PLAIN && used != PARENT_METHOD_SLABS)
        if (int rc = ensure_tiles(e, t, s, used, variant)) return rc;
    s.method = used;
    s.variant = variant;
    return PARENT_OK;
}

// All pairs of set s with its planned kernel: kernels only (capturable).
hipError_t enqueue_pairs(parent_engine* e, const Traj& t, BinSet& s) {
    const PairInputs in = with_nlogn(e, pair_inputs(t, s));
    if (s.method == PARENT_METHOD_PLAIN) return launch_pairs_plain(in, nullptr, e->stream);
    if (s.method == PARENT_METHOD_SLABS) return launch_pairs_slabs(in, slab_lds(e, s), s.variant, nullptr, e->stream);
    const TileQueue q{s.tiles.dev.as<int>(), s.tiles.ntiles, 2, nullptr};
    return launch_tiled(e, in, bmax2d(s), s.method, s.variant, q, nullptr, nullptr, nullptr);
}

int run_pairs(parent_engine* e, const Traj& t, BinSet& s, int method) {
    if (!s.prepared) return fail(e, PARENT_E_INVALID, "parent_engine_prepare has not succeeded");
    if (int rc = plan_pairs(e, t, s, method)) return rc;
    HIPCHK(e, hipEventRecord(e->ev[4], e->stream));
    HIPCHK(e, enqueue_pairs(e, t, s));
    HIPCHK(e, hipEventRecord(e->ev[5], e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->stats.ms_pairs = elapsed(e->ev[4], e->ev[5]);
    e->stats.pair_updates = (double)n_pairs(t.nD) * t.N;
    e->stats.method = s.method;
    e->stats.kernel_variant = s.variant;
    s.s2d_valid = true;
    return PARENT_OK;
}

// MIST over set s's S1D/S2D (synchronises).
int run_mist(parent_engine* e, BinSet& s, int32_t* edges, double* mi, int32_t* left, double* totals) {
    if (!s.s2d_valid || s.ent_nD < 1) return fail(e, PARENT_E_INVALID, "no 2D entropies on the device");
    const int nD = s.ent_nD, V = n_dofs(nD);
    Scratch& x = e->scratch;
    // the dense MI matrix: each Prim step then reads one contiguous row
    // (c2: 162 MB, c5: 1.8 GB); without room for it the kernel gathers MI
    // from S1D / S2D
    const size_t dense_bytes = (size_t)V * V * 8;
    bool dense = x.mdense.ensure(dense_bytes) == hipSuccess;
    if (!dense) (void)hipGetLastError();
    HIPCHK(e, x.morder.ensure((size_t)V * 4));
    HIPCHK(e, x.medges.ensure((size_t)V * 16));
    HIPCHK(e, x.mmi.ensure((size_t)V * 8));
    HIPCHK(e, x.mleft.ensure((size_t)V * 4));
    HIPCHK(e, x.mflags.ensure(64));
    MistArgs a{};
    a.nD = nD;
    a.V = V;
    a.s1d = s.s1d.as<double>();
    a.s2d = s.s2d.as<double>();
    a.mi_dense = dense ? x.mdense.as<double>() : nullptr;
    a.order = x.morder.as<int>();
    a.flags = x.mflags.as<int>();
    a.edges = x.medges.as<int>();
    a.mi = x.mmi.as<double>();
    a.left = x.mleft.as<int>();
    if (mist_keys_in_memory(V, dense)) {
        HIPCHK(e, x.mkey.ensure((size_t)V * 8));
        HIPCHK(e, x.mupos.ensure((size_t)V * 4));
        HIPCHK(e, x.mlive.ensure((size_t)V * 4));
        a.key = x.mkey.as<double>();
        a.upos = x.mupos.as<unsigned>();
        a.live = x.mlive.as<int>();
    }
    // the spanning-tree path first where it applies; the Prim kernel then
    // returns at once if that path wrote the result
    const bool tree = dense && e->mist_path_mode == 0 && V <= mist_reg_max_vertices() &&
                      x.mtree.ensure(mist_tree_bytes(V)) == hipSuccess;
    if (!tree) (void)hipGetLastError();
    MistTreeArgs t{};
    if (tree) {
        t = mist_tree_layout(x.mtree.as<char>(), V, nullptr);
        a.skip = t.state + 3;
    }
    int flags = 0, tree_done = 0;
    HIPCHK(e, hipEventRecord(e->ev[6], e->stream));
    if (dense) HIPCHK(e, launch_mi_dense(nD, a.s1d, a.s2d, x.mdense.as<double>(), e->stream));
    if (tree) HIPCHK(e, launch_mist_tree(a, t, e->stream));
    HIPCHK(e, launch_mist(a, e->stream));
    HIPCHK(e, hipEventRecord(e->ev[7], e->stream));
    if (tree) HIPCHK(e, hipMemcpyAsync(&tree_done, t.state + 3, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    std::vector<double> s1(V);
    HIPCHK(e, hipMemcpyAsync(edges, a.edges, (size_t)(V - 1) * 16, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipMemcpyAsync(mi, a.mi, (size_t)(V - 1) * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipMemcpyAsync(left, a.left, (size_t)(V - 1) * 4, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipMemcpyAsync(s1.data(), a.s1d, (size_t)V * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipMemcpyAsync(&flags, x.mflags.p, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->stats.ms_mist = elapsed(e->ev[6], e->ev[7]);
    e->stats.mist_path = tree_done ? 1 : 0;
    if (flags & kFlagMist) return fail(e, PARENT_E_DATA, "MIST: no finite mutual information left to pick");
    // totals in the reference's summation order (get_PAR_MIST.cpp:276-284, 663-741, 754-765)
    double t1[3] = {0, 0, 0};
    for (int t = 0; t < 3; t++)
        for (int i = 0; i < count_of_type(nD, t); i++) t1[t] += s1[first_of_type(nD, t) + i];
    double cat[6] = {0, 0, 0, 0, 0, 0}, total_mi = 0;
    for (int k = 0; k < V - 1; k++) {
        const int ta = edges[4 * k], tb = edges[4 * k + 1];
        const int lo = std::min(ta, tb), hi = std::max(ta, tb);
        const int c = lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
        total_mi += mi[k];
        cat[c] += mi[k];
    }
    totals[PARENT_TOT_S1D_B] = t1[0];
    totals[PARENT_TOT_S1D_A] = t1[1];
    totals[PARENT_TOT_S1D_D] = t1[2];
    for (int c = 0; c < 6; c++) totals[PARENT_TOT_MI_BB + c] = cat[c];
    totals[PARENT_TOT_TOTAL] = t1[0] + t1[1] + t1[2] - total_mi;
    return PARENT_OK;
}

int read_s1d(parent_engine* e, BinSet& s, double* out) {
    if (s.ent_nD < 1) return fail(e, PARENT_E_INVALID, "no entropies computed");
    HIPCHK(e, hipMemcpyAsync(out, s.s1d.p, (size_t)n_dofs(s.ent_nD) * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return PARENT_OK;
}

int read_s2d(parent_engine* e, BinSet& s, double* out) {
    if (!s.s2d_valid) return fail(e, PARENT_E_INVALID, "parent_engine_pairs has not succeeded");
    HIPCHK(e, hipMemcpyAsync(out, s.s2d.p, (size_t)n_pairs(s.ent_nD) * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return PARENT_OK;
}

void drop_graph(parent_engine* e) {
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->graph_exec = nullptr;
    e->graph = nullptr;
    e->graph_key.clear();
}

int create_stream_and_events(parent_engine* e, int device) {
    hipError_t h;
    if ((h = hipSetDevice(device)) != hipSuccess || (h = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess) {
        set_last_error(std::string("HIP init failed: ") + hipGetErrorString(h));
        return PARENT_E_DEVICE;
    }
    for (auto& ev : e->ev) {
        if ((h = hipEventCreate(&ev)) != hipSuccess) {
            set_last_error(std::string("hipEventCreate failed: ") + hipGetErrorString(h));
            return PARENT_E_DEVICE;
        }
    }
    return PARENT_OK;
}

int bfail(parent_batch* b, int code, const std::string& msg) { return fail(b->eng, code, msg); }

// The element-wise largest bin counts of a batch's trajectories: what the
// one kernel of a merged launch must hold (each tile reads its own
// trajectory's counts from its descriptor).
BinSet batch_envelope(const parent_batch* b) {
    BinSet v;
    for (const auto& it : b->items)
        for (int k = 0; k < 6; k++) v.bins[k] = std::max(v.bins[k], it->set.bins[k]);
    v.idx2 = bmax2d(v) > kMaxBins8 ? 2 : 1;
    return v;
}

// A launch's list of (row, column, item) tiles, padded with empty tiles to a
// multiple of 8 and dealt so that block b of a launch runs entry b of the
// list whatever XCD it lands on (the kernels map block b to list slot
// (b % 8) * chunk + b / 8).  The launchers issue a queue in launches of `seg`
// tiles (a multiple of 8), each with its own block-to-slot map, so each
// segment is dealt alone.  (pad_D, pad_item): a trajectory whose r0 = c0 = D
// tile holds no pair.
std::vector<int> deal_queue(std::vector<int> lst, int pad_D, int pad_item, int seg) {
    while ((lst.size() / 3) % 8) {
        lst.push_back(pad_D);
        lst.push_back(pad_D);
        lst.push_back(pad_item);
    }
    const long long T = (long long)(lst.size() / 3);
    std::vector<int> q(lst.size());
    for (long long s0 = 0; s0 < T; s0 += seg) {
        const long long n_seg = std::min<long long>(seg, T - s0), chunk = n_seg / 8;
        for (long long blk = 0; blk < n_seg; blk++) {
            const long long slot = s0 + (blk % 8) * chunk + blk / 8;
            for (int c = 0; c < 3; c++) q[3 * (size_t)slot + c] = lst[3 * ((size_t)s0 + blk) + c];
        }
    }
    return q;
}

int queue_segment(int method, int variant) {
    return method == PARENT_METHOD_MFMA ? mfma_tiles_per_launch() : atomic_tiles_per_launch(variant);
}

std::vector<int> by_frames_desc(const parent_batch* b, std::vector<int> members) {
    std::stable_sort(members.begin(), members.end(),
                     [&](int x, int y) { return b->items[x]->traj.N > b->items[y]->traj.N; });
    return members;
}

// Uploads the PairInputs of every item (each tile's descriptor).
int ensure_descs(parent_engine* e, parent_batch* b) {
    const int n = (int)b->items.size();
    if (b->descs_items == n && b->descs_nlogn == e->scratch.nlogn.p) return PARENT_OK;
    std::vector<PairInputs> d(n);
    for (int k = 0; k < n; k++) d[k] = with_nlogn(e, pair_inputs(b->items[k]->traj, b->items[k]->set));
    HIPCHK(e, b->descs.ensure(d.size() * sizeof(PairInputs)));
    HIPCHK(e, hipMemcpyAsync(b->descs.p, d.data(), d.size() * sizeof(PairInputs), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    b->descs_items = n;
    b->descs_nlogn = e->scratch.nlogn.p;
    return PARENT_OK;
}

// Uploads the merged tile queue of one launch: its items by decreasing frame
// count (cost per tile), each item's tiles in its band order.
int build_queue(parent_engine* e, parent_batch* b, const BinSet& env, parent_batch::Queue& qu) {
    const parent_batch::Plan& pl = qu.plan;
    int tr, tc;
    tile_shape(env, pl.method, pl.variant, &tr, &tc);
    const std::vector<int> order = by_frames_desc(b, pl.members);
    std::vector<int> lst;
    for (int k : order) {
        const Traj& t = b->items[k]->traj;
        std::vector<int> tl = make_tiles(t.D, t.Np, tr, tc);
        for (size_t i = 0; i + 1 < tl.size(); i += 2) {
            lst.push_back(tl[i]);
            lst.push_back(tl[i + 1]);
            lst.push_back(k);
        }
    }
    if (lst.size() / 3 + 7 > (size_t)INT32_MAX) return bfail(b, PARENT_E_UNSUPPORTED, "batch tile queue too long");
    const std::vector<int> q = deal_queue(std::move(lst), b->items[order[0]]->traj.D, order[0],
                                          queue_segment(pl.method, pl.variant));
    HIPCHK(e, qu.tiles.ensure(q.size() * 4));
    HIPCHK(e, hipMemcpyAsync(qu.tiles.p, q.data(), q.size() * 4, hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    qu.ntiles = (int)(q.size() / 3);
    return PARENT_OK;
}

// Times configuration (m, id) on a sample of the merged queue of `members`
// and returns the full queue's estimated time: kBatchRuns contiguous runs
// spread evenly over the queue (contiguous, so that the tiles of a run share
// index rows in L2 as in the full launch), 1/64 of its tiles and at least
// kFinalTilesPerCu per compute unit (the whole queue when it is smaller).
constexpr int kBatchRuns = 64;
int time_queue_sample(parent_engine* e, parent_batch* b, const std::vector<int>& members, int m, int id,
                      double* est_ms) {
    const BinSet env = batch_envelope(b);
    int tr, tc;
    tile_shape(env, m, id, &tr, &tc);
    const std::vector<int> order = by_frames_desc(b, members);
    std::vector<long long> cnt(order.size());
    long long total = 0;
    for (size_t i = 0; i < order.size(); i++) {
        const Traj& t = b->items[order[i]]->traj;
        cnt[i] = count_tiles(t.D, t.Np, tr, tc);
        total += cnt[i];
    }
    if (total < 1) return bfail(b, PARENT_E_INVALID, "batch has no pairs");
    const long long want = std::min(total, std::max((long long)kFinalTilesPerCu * e->cus, total / 64));
    const long long R = std::min<long long>(kBatchRuns, want), L = want / R;
    std::vector<int> lst;
    long long off = 0;
    for (size_t i = 0; i < order.size(); i++) {
        const Traj& t = b->items[order[i]]->traj;
        const long long a = off, z = off + cnt[i];
        std::vector<int> tl;  // the item's tiles, listed when a run reaches it
        for (long long r = 0; r < R; r++) {
            const long long s0 = std::max(0LL, std::min((2 * r + 1) * total / (2 * R) - L / 2, total - L));
            const long long lo = std::max(s0, a), hi = std::min(s0 + L, z);
            if (lo >= hi) continue;
            if (tl.empty()) tl = make_tiles(t.D, t.Np, tr, tc);
            if ((long long)tl.size() != 2 * cnt[i])
                return bfail(b, PARENT_E_INVALID, "tile list and tile count disagree");
            for (long long x = lo; x < hi; x++) {
                lst.push_back(tl[2 * (x - a)]);
                lst.push_back(tl[2 * (x - a) + 1]);
                lst.push_back(order[i]);
            }
        }
        off = z;
    }
    const double sampled = (double)(lst.size() / 3);
    const std::vector<int> q = deal_queue(std::move(lst), b->items[order[0]]->traj.D, order[0], queue_segment(m, id));
    DevBuf dq;
    HIPCHK(e, dq.ensure(q.size() * 4));
    HIPCHK(e, hipMemcpyAsync(dq.p, q.data(), q.size() * 4, hipMemcpyHostToDevice, e->stream));
    if (int rc = ensure_descs(e, b)) return rc;
    const TileQueue tq{dq.as<int>(), (int)(q.size() / 3), 3, b->descs.as<PairInputs>()};
    const Traj& t0 = b->items[order[0]]->traj;
    const PairInputs in0 = with_nlogn(e, pair_inputs(t0, b->items[order[0]]->set));  // unused by the batch kernels
    double ms = 0;
    if (int rc = time_launch(e, [&] { return launch_tiled(e, in0, bmax2d(env), m, id, tq, nullptr, nullptr, nullptr); }, &ms))
        return rc;
    *est_ms = ms * (double)total / sampled;
    return PARENT_OK;
}

// Kernel choice for a set of a batch's items (`members`) that share one
// launch.  One configuration runs every tile of their merged queue, and the
// right one depends on the mix: per pair, a tile costs a fixed part
// (clearing its LDS histograms, the epilogue) and a part per frame.  Each
// candidate is timed on a tuning sample (tune_sample: 8 runs of rows) of the
// members' longest and shortest trajectories; the two timings give the model
// ms = alpha * tiles + beta * updates, which predicts the full queue.  That
// model ranks the candidates only roughly (two configurations 1% apart in
// prediction ran 880 and 1,060 ms on the long class of a 64-trajectory
// batch), so the kBatchFinalists best are then timed on a sample of the
// merged queue itself (time_queue_sample), and the fastest there is kept
// (*pred_ms: its estimate for the full queue).
constexpr int kBatchFinalists = 3;
int batch_autotune(parent_engine* e, parent_batch* b, const std::vector<int>& members, int* method, int* variant,
                   double* pred_ms) {
    int kl = members[0], ks = members[0];
    for (int k : members) {
        const Traj& t = b->items[k]->traj;
        if (t.N > b->items[kl]->traj.N || (t.N == b->items[kl]->traj.N && t.D > b->items[kl]->traj.D)) kl = k;
        if (t.N < b->items[ks]->traj.N || (t.N == b->items[ks]->traj.N && t.D > b->items[ks]->traj.D)) ks = k;
    }
    const BinSet s0 = batch_envelope(b);
    std::vector<std::pair<int, int>> cand;
    for (int id = 0; id < atomic_config_count(); id++)
        if (atomic_tuned(id) && variant_fits(e, s0, id, b->items[kl]->traj.N)) cand.push_back({PARENT_METHOD_ATOMIC, id});
    if (mfma_fits(e, s0)) cand.push_back({PARENT_METHOD_MFMA, 0});
    if (cand.empty()) return bfail(b, PARENT_E_UNSUPPORTED, "no tiled kernel fits these bins");
    double total_upd = 0;
    for (int k : members) total_upd += (double)n_pairs(b->items[k]->traj.nD) * b->items[k]->traj.N;
    e->last_tune.clear();
    DevBuf dt;
    double best = -1;
    for (auto& mv : cand) {
        const int m = mv.first, id = mv.second;
        int tr, tc;
        tile_shape(s0, m, id, &tr, &tc);
        double ms[2] = {0, 0}, tiles[2] = {0, 0}, upd[2] = {0, 0};
        const int probes[2] = {kl, ks};
        const int np = kl == ks ? 1 : 2;
        for (int p = 0; p < np; p++) {
            const Traj& t = b->items[probes[p]]->traj;
            BinSet& s = b->items[probes[p]]->set;
            std::vector<int> w = tune_sample(t.D, tr, tc, 8, (size_t)kProbeTilesPerCu * e->cus);
            if (w.empty()) continue;
            tiles[p] = (double)(w.size() / 2);
            upd[p] = (double)tile_pairs(t.D, w, tr, tc) * t.N;
            HIPCHK(e, dt.ensure(w.size() * sizeof(int)));
            HIPCHK(e, hipMemcpyAsync(dt.p, w.data(), w.size() * sizeof(int), hipMemcpyHostToDevice, e->stream));
            const PairInputs in = with_nlogn(e, pair_inputs(t, s));
            const TileQueue q{dt.as<int>(), (int)(w.size() / 2), 2, nullptr};
            if (int rc = time_launch(e, [&] { return launch_tiled(e, in, bmax2d(s0), m, id, q, nullptr, nullptr, nullptr); }, &ms[p]))
                return rc;
        }
        double alpha = 0, beta = 0;
        const double det = tiles[0] * upd[1] - tiles[1] * upd[0];
        const bool fit = np == 2 && std::fabs(det) > 1e-9 * std::fabs(tiles[0] * upd[1]);
        if (fit) {
            alpha = (ms[0] * upd[1] - ms[1] * upd[0]) / det;
            beta = (tiles[0] * ms[1] - tiles[1] * ms[0]) / det;
        }
        if (!fit || alpha < 0 || beta < 0) {
            // no separable fit: charge both timings to one term
            if (fit && beta < 0 && alpha >= 0) {
                beta = 0;
                alpha = (ms[0] + ms[1]) / (tiles[0] + tiles[1]);
            } else {
                alpha = 0;
                beta = (ms[0] + ms[1]) / (upd[0] + upd[1]);
            }
        }
        double total_tiles = 0;
        for (int k : members)
            total_tiles += (double)count_tiles(b->items[k]->traj.D, b->items[k]->traj.Np, tr, tc);
        const double pred = alpha * total_tiles + beta * total_upd;
        const int block = m == PARENT_METHOD_MFMA ? 256 : atomic_config(id).block;
        e->last_tune.push_back(parent_tune_entry{m, id, tr, tc, block, pred, total_upd, total_upd / (pred * 1e-3)});
        if (best < 0 || pred < best) {
            best = pred;
            *method = m;
            *variant = id;
        }
    }
    // second stage: the kBatchFinalists best predictions timed on a sample
    // of the members' own merged queue; their table rows take that estimate
    std::vector<size_t> rank(e->last_tune.size());
    std::iota(rank.begin(), rank.end(), 0);
    std::stable_sort(rank.begin(), rank.end(), [&](size_t x, size_t y) { return e->last_tune[x].ms < e->last_tune[y].ms; });
    best = -1;
    for (size_t f = 0; f < rank.size() && f < (size_t)kBatchFinalists; f++) {
        parent_tune_entry& te = e->last_tune[rank[f]];
        double est = 0;
        if (int rc = time_queue_sample(e, b, members, te.method, te.variant, &est)) return rc;
        te.ms = est;
        te.updates_per_s = total_upd / (est * 1e-3);
        if (best < 0 || est < best) {
            best = est;
            *method = te.method;
            *variant = te.variant;
        }
    }
    *pred_ms = best;
    return PARENT_OK;
}

// Trajectories with fewer frames than this many per cell of the batch's
// largest 2D histogram form the short class (18,000 frames at 30 bins).
constexpr int kShortFramesPerCell = 20;

// The launches PARENT_METHOD_AUTO runs.  One launch of the whole merged
// queue, or, when the batch holds short trajectories and longer ones, one
// launch per class, each tuned on its own: a short trajectory's tiles spend
// much of their time on the fixed part (clear, epilogue), and their counts
// stay small enough for narrower counters, so the two classes favour
// different configurations.  The split is kept when the two classes'
// estimated times add up to less than the merged queue's (classes 1), always
// (2) or never (0).  e->last_tune keeps the merged queue's table.
int plan_batch(parent_engine* e, parent_batch* b) {
    const int n = (int)b->items.size();
    std::vector<int> all(n);
    std::iota(all.begin(), all.end(), 0);
    parent_batch::Plan merged;
    merged.members = all;
    double p_all = 0;
    if (int rc = batch_autotune(e, b, all, &merged.method, &merged.variant, &p_all)) return rc;
    const std::vector<parent_tune_entry> table = e->last_tune;
    std::vector<parent_batch::Plan> plan{merged};
    const double bm = bmax2d(batch_envelope(b));
    parent_batch::Plan lng, shrt;
    for (int k = 0; k < n; k++)
        (b->items[k]->traj.N < kShortFramesPerCell * bm * bm ? shrt : lng).members.push_back(k);
    if (b->classes != 0 && !lng.members.empty() && !shrt.members.empty()) {
        double p_long = 0, p_short = 0;
        if (int rc = batch_autotune(e, b, lng.members, &lng.method, &lng.variant, &p_long)) return rc;
        if (int rc = batch_autotune(e, b, shrt.members, &shrt.method, &shrt.variant, &p_short)) return rc;
        if (b->classes == 2 || p_long + p_short < p_all) plan = {lng, shrt};
    }
    e->last_tune = table;
    b->tuned_plan = plan;
    b->tuned_items = n;
    return PARENT_OK;
}

}  // namespace

extern "C" {

int parent_engine_create(int device, parent_engine** out) {
    if (!out) return PARENT_E_INVALID;
    *out = nullptr;
    int ndev = 0;
    hipError_t h = hipGetDeviceCount(&ndev);
    if (h != hipSuccess || ndev == 0) {
        set_last_error(std::string("no HIP device: ") + hipGetErrorString(h));
        return PARENT_E_DEVICE;
    }
    if (device < 0 || device >= ndev) {
        set_last_error("device index out of range");
        return PARENT_E_INVALID;
    }
    parent_engine* e = new parent_engine();
    e->device = device;
    e->sets.emplace_back(new BinSet());
    if (int rc = create_stream_and_events(e, device)) {
        parent_engine_destroy(e);
        return rc;
    }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        e->lds_max = lds;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        e->cus = cus;
    // the device address of this GPU's static scratch of
    // parent_engine_mist_noalloc (looked up once per device)
    const hipError_t hs = mist_noalloc_scratch(device, &e->noalloc_scratch);
    if (hs != hipSuccess) {
        set_last_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(hs));
        parent_engine_destroy(e);
        return PARENT_E_DEVICE;
    }
    *out = e;
    return PARENT_OK;
}

void parent_engine_destroy(parent_engine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    e->sets.clear();
    e->traj.values_own.release();
    e->scratch = Scratch();
    for (auto& ev : e->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char* parent_engine_error(const parent_engine* e) { return e ? e->err.c_str() : "null engine"; }

}  // extern "C"

namespace {

// Forget the previous trajectory and every result computed from it.
Traj& reset_trajectory(parent_engine* e, int nD, int N, int dp) {
    drop_graph(e);
    e->traj_serial++;
    Traj& t = e->traj;
    t.nD = nD;
    t.N = N;
    t.dp = dp;
    t.D = n_dofs(nD);
    t.Np = (int)(((long long)N + kFrameAlign - 1) / kFrameAlign * kFrameAlign);
    t.values = nullptr;
    e->sets.resize(1);
    e->active = 0;
    e->sets[0]->prepared = false;
    e->sets[0]->s2d_valid = false;
    e->sets[0]->ent_nD = -1;
    e->stats = parent_stats{};
    return t;
}

// Host side of the streamed .bat load: two pinned record buffers, two events.
// The destructor waits for the stream first, so no DMA still reads a buffer
// when it is freed (also on error paths).
struct BatStream {
    hipStream_t stream;
    void* pinned[2] = {nullptr, nullptr};
    hipEvent_t copied[2] = {nullptr, nullptr};
    explicit BatStream(hipStream_t s) : stream(s) {}
    ~BatStream() {
        (void)hipStreamSynchronize(stream);
        for (int b = 0; b < 2; b++) {
            if (pinned[b]) (void)hipHostFree(pinned[b]);
            if (copied[b]) (void)hipEventDestroy(copied[b]);
        }
    }
};

}  // namespace

extern "C" {

int parent_engine_load_bat(parent_engine* e, parent_bat* bat) {
    if (!e || !bat) return PARENT_E_INVALID;
    int nD = 0, N = 0, dp = 0;
    if (parent_bat_info(bat, &nD, &N, &dp) != PARENT_OK) return PARENT_E_INVALID;
    if (nD < 1) return fail(e, PARENT_E_INVALID, "need n_dihedrals >= 1");
    const long long rec = parent_bat_record_bytes(bat);
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    Traj& t = reset_trajectory(e, nD, N, dp);
    const size_t vsz = dp ? 8 : 4;
    HIPCHK(e, t.values_own.ensure((size_t)N * t.D * vsz));
    // chunks of ~32 MB of records: the file read of chunk k+1 overlaps the
    // DMA and the de-interleave kernel of chunk k
    const int chunk = (int)std::max(1LL, std::min<long long>(N, (32LL << 20) / rec));
    DevBuf stage[2];
    BatStream hs(e->stream);
    for (int b = 0; b < 2; b++) {
        HIPCHK(e, stage[b].ensure((size_t)chunk * rec));
        HIPCHK(e, hipHostMalloc(&hs.pinned[b], (size_t)chunk * rec, hipHostMallocDefault));
        HIPCHK(e, hipEventCreateWithFlags(&hs.copied[b], hipEventDisableTiming));
    }
    char* out = static_cast<char*>(t.values_own.p);
    for (int f0 = 0, k = 0; f0 < N; f0 += chunk, k++) {
        const int b = k & 1, nf = std::min(chunk, N - f0);
        if (k >= 2) HIPCHK(e, hipEventSynchronize(hs.copied[b]));
        if (int rc = parent_bat_read_raw(bat, f0, nf, hs.pinned[b])) return fail(e, rc, parent_last_error());
        HIPCHK(e, hipMemcpyAsync(stage[b].p, hs.pinned[b], (size_t)nf * rec, hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipEventRecord(hs.copied[b], e->stream));
        HIPCHK(e, launch_bat_records(stage[b].p, rec, nf, nD, dp, out + (size_t)f0 * t.D * vsz, e->stream));
    }
    HIPCHK(e, hipStreamSynchronize(e->stream));
    t.values = t.values_own.p;
    e->stats.ms_upload = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PARENT_OK;
}

int parent_engine_set_trajectory(parent_engine* e, int n_dihedrals, int n_frames, int double_prec,
                                 const void* values, int on_device) {
    if (!e || !values) return PARENT_E_INVALID;
    if (n_dihedrals < 1 || n_frames < 1 || (double_prec != 0 && double_prec != 1))
        return fail(e, PARENT_E_INVALID, "need n_dihedrals >= 1, n_frames >= 1, double_prec in {0,1}");
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    Traj& t = reset_trajectory(e, n_dihedrals, n_frames, double_prec);
    const size_t bytes = (size_t)t.N * t.D * (double_prec ? 8 : 4);
    if (on_device) {
        t.values_own.release();
        t.values = values;
    } else {
        HIPCHK(e, t.values_own.ensure(bytes));
        HIPCHK(e, hipEventRecord(e->ev[0], e->stream));
        HIPCHK(e, hipMemcpyAsync(t.values_own.p, values, bytes, hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipEventRecord(e->ev[1], e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
        e->stats.ms_upload = elapsed(e->ev[0], e->ev[1]);
        t.values = t.values_own.p;
    }
    return PARENT_OK;
}

int parent_engine_prepare(parent_engine* e, const int bins[6]) {
    if (!e || !bins) return PARENT_E_INVALID;
    if (!e->traj.values) return fail(e, PARENT_E_INVALID, "no trajectory set");
    if (int rc = check_bins(e, bins)) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    drop_graph(e);  // the captured sweep refers to the sets' buffers
    e->sets.resize(1);
    e->active = 0;
    BinSet& s = *e->sets[0];
    s.ranges_from = nullptr;
    memcpy(s.bins, bins, sizeof(s.bins));
    return run_prepare(e, e->traj, s);
}

int parent_engine_pairs(parent_engine* e, int method) {
    if (!e) return PARENT_E_INVALID;
    HIPCHK(e, hipSetDevice(e->device));
    drop_graph(e);  // may re-plan a set the captured sweep refers to
    return run_pairs(e, e->traj, cur(e), method);
}

int parent_engine_sweep(parent_engine* e, int n_sets, const int* bins, int method) {
    if (!e || n_sets < 1 || !bins) return PARENT_E_INVALID;
    if (!e->traj.values) return fail(e, PARENT_E_INVALID, "no trajectory set");
    for (int k = 0; k < n_sets; k++)
        if (int rc = check_bins(e, bins + 6 * k)) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    const Traj& t = e->traj;
    std::vector<int> key(bins, bins + 6 * n_sets);
    key.push_back(method);
    key.push_back((int)e->traj_serial);
    key.push_back(e->forced_atomic);
    key.push_back(e->forced_mfma);
    if (key != e->graph_key) drop_graph(e);
    while ((int)e->sets.size() < n_sets) e->sets.emplace_back(new BinSet());
    e->sets.resize(n_sets);
    e->active = 0;
    // eager pass: buffers, data checks, kernel choice (autotuning) and tile
    // lists, all of which synchronise and so stay outside the capture
    if (!e->graph_exec) {
        for (int k = 0; k < n_sets; k++) {
            BinSet& s = *e->sets[k];
            s.ranges_from = k ? e->sets[0].get() : nullptr;
            memcpy(s.bins, bins + 6 * k, sizeof(s.bins));
            if (int rc = run_prepare(e, t, s)) return rc;
            if (int rc = plan_pairs(e, t, s, method)) return rc;
        }
        HIPCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
        hipError_t h = hipSuccess;
        for (int k = 0; k < n_sets && h == hipSuccess; k++) {
            h = enqueue_prepare(e, t, *e->sets[k]);
            if (h == hipSuccess) h = enqueue_pairs(e, t, *e->sets[k]);
        }
        hipGraph_t g = nullptr;
        const hipError_t he = hipStreamEndCapture(e->stream, &g);
        if (h != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return hip_fail(e, h, "sweep capture");
        }
        HIPCHK(e, he);
        e->graph = g;
        HIPCHK(e, hipGraphInstantiate(&e->graph_exec, e->graph, nullptr, nullptr, 0));
        e->graph_key = key;
    }
    HIPCHK(e, hipEventRecord(e->ev[4], e->stream));
    HIPCHK(e, hipGraphLaunch(e->graph_exec, e->stream));
    HIPCHK(e, hipEventRecord(e->ev[5], e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    double upd = 0;
    for (int k = 0; k < n_sets; k++) {
        BinSet& s = *e->sets[k];
        if (int rc = check_flags(e, s)) return rc;
        s.prepared = true;
        s.s2d_valid = true;
        s.ent_nD = t.nD;
        upd += (double)n_pairs(t.nD) * t.N;
    }
    e->stats.ms_sweep = elapsed(e->ev[4], e->ev[5]);
    e->stats.sweep_sets = n_sets;
    e->stats.pair_updates = upd;
    return PARENT_OK;
}

int parent_engine_select(parent_engine* e, int set) {
    if (!e) return PARENT_E_INVALID;
    if (set < 0 || set >= (int)e->sets.size()) return fail(e, PARENT_E_INVALID, "bin set index out of range");
    e->active = set;
    e->stats.method = e->sets[set]->method;
    e->stats.kernel_variant = e->sets[set]->variant;
    return PARENT_OK;
}

int parent_engine_read_s1d(parent_engine* e, double* out) {
    if (!e || !out) return PARENT_E_INVALID;
    HIPCHK(e, hipSetDevice(e->device));
    return read_s1d(e, cur(e), out);
}

int parent_engine_read_s2d(parent_engine* e, double* out) {
    if (!e || !out) return PARENT_E_INVALID;
    HIPCHK(e, hipSetDevice(e->device));
    return read_s2d(e, cur(e), out);
}

int parent_engine_read_ranges(parent_engine* e, double* vmin, double* vmax, double* modfit) {
    if (!e) return PARENT_E_INVALID;
    BinSet& s = cur(e);
    if (!s.prepared) return fail(e, PARENT_E_INVALID, "parent_engine_prepare has not succeeded");
    HIPCHK(e, hipSetDevice(e->device));
    const size_t D = e->traj.D;
    const BinSet& r = s.ranges();
    if (vmin) HIPCHK(e, hipMemcpyAsync(vmin, r.vmin.p, D * 8, hipMemcpyDeviceToHost, e->stream));
    if (vmax) HIPCHK(e, hipMemcpyAsync(vmax, r.vmax.p, D * 8, hipMemcpyDeviceToHost, e->stream));
    if (modfit) HIPCHK(e, hipMemcpyAsync(modfit, r.modfit.p, D * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return PARENT_OK;
}

int parent_engine_read_bins(parent_engine* e, uint8_t* out) {
    if (!e || !out) return PARENT_E_INVALID;
    BinSet& s = cur(e);
    if (!s.prepared) return fail(e, PARENT_E_INVALID, "parent_engine_prepare has not succeeded");
    if (s.idx2 != 1) return fail(e, PARENT_E_UNSUPPORTED, "the indices are 16-bit (2D bins > 255): use parent_engine_read_bins16");
    HIPCHK(e, hipSetDevice(e->device));
    const Traj& t = e->traj;
    HIPCHK(e, hipMemcpy2DAsync(out, t.N, s.bins2.p, t.Np, t.N, t.D, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return PARENT_OK;
}

int parent_engine_read_bins16(parent_engine* e, uint16_t* out) {
    if (!e || !out) return PARENT_E_INVALID;
    BinSet& s = cur(e);
    if (!s.prepared) return fail(e, PARENT_E_INVALID, "parent_engine_prepare has not succeeded");
    HIPCHK(e, hipSetDevice(e->device));
    const Traj& t = e->traj;
    if (s.idx2 == 2) {
        HIPCHK(e, hipMemcpy2DAsync(out, (size_t)t.N * 2, s.bins2.p, (size_t)t.Np * 2, (size_t)t.N * 2, t.D,
                                   hipMemcpyDeviceToHost, e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
        return PARENT_OK;
    }
    std::vector<uint8_t> b8((size_t)t.N * t.D);
    HIPCHK(e, hipMemcpy2DAsync(b8.data(), t.N, s.bins2.p, t.Np, t.N, t.D, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    for (size_t k = 0; k < b8.size(); k++) out[k] = b8[k];
    return PARENT_OK;
}

int parent_engine_pair_counts(parent_engine* e, int method, int n_pairs_req, const int32_t* pairs,
                              int32_t* counts) {
    if (!e || n_pairs_req < 0 || (n_pairs_req > 0 && (!pairs || !counts))) return PARENT_E_INVALID;
    BinSet& s = cur(e);
    const Traj& tj = e->traj;
    if (!s.prepared) return fail(e, PARENT_E_INVALID, "parent_engine_prepare has not succeeded");
    if (n_pairs_req == 0) return PARENT_OK;
    HIPCHK(e, hipSetDevice(e->device));
    const Bins3 B = bins2d(s);
    std::vector<long long> off(n_pairs_req);
    long long total = 0;
    for (int k = 0; k < n_pairs_req; k++) {
        const int gi = pairs[2 * k], gj = pairs[2 * k + 1];
        if (gi < 0 || gj >= tj.D || gi >= gj) return fail(e, PARENT_E_INVALID, "pairs must satisfy 0 <= gi < gj < D");
        off[k] = total;
        total += (long long)B.of(dof_type(tj.nD, gi)) * B.of(dof_type(tj.nD, gj));
    }
    DevBuf dpairs, doff, dout;
    HIPCHK(e, dout.ensure(total * 4));
    const PairInputs in = with_nlogn(e, pair_inputs(tj, s));
    int used = method, variant = 0;
    if (int rc = resolve_method(e, tj, s, method, &used, &variant)) return rc;
    if (used == PARENT_METHOD_PLAIN || used == PARENT_METHOD_SLABS) {
        HIPCHK(e, dpairs.ensure((size_t)n_pairs_req * 8));
        HIPCHK(e, doff.ensure((size_t)n_pairs_req * 8));
        HIPCHK(e, hipMemcpyAsync(dpairs.p, pairs, (size_t)n_pairs_req * 8, hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipMemcpyAsync(doff.p, off.data(), (size_t)n_pairs_req * 8, hipMemcpyHostToDevice, e->stream));
        CountsDump dump{n_pairs_req, dpairs.as<int>(), doff.as<long long>(), dout.as<int>()};
        if (used == PARENT_METHOD_SLABS) HIPCHK(e, launch_pairs_slabs(in, slab_lds(e, s), variant, &dump, e->stream));
        else HIPCHK(e, launch_pairs_plain(in, &dump, e->stream));
    } else {
        // the tiles holding the requested pairs, each pair as (slot, offset)
        int tr, tc;
        tile_shape(s, used, variant, &tr, &tc);
        std::vector<std::vector<int>> per_tile;
        std::vector<int> tl;
        std::map<long long, int> key;  // tile key -> index
        for (int k = 0; k < n_pairs_req; k++) {
            const int gi = pairs[2 * k], gj = pairs[2 * k + 1];
            const int r0 = gi / tr * tr, c0 = gj / tc * tc;
            const long long tk = (long long)r0 * tj.D + c0;
            auto it = key.find(tk);
            int idx;
            if (it == key.end()) {
                idx = (int)tl.size() / 2;
                key[tk] = idx;
                tl.push_back(r0);
                tl.push_back(c0);
                per_tile.emplace_back();
            } else {
                idx = it->second;
            }
            per_tile[idx].push_back(k);
        }
        std::vector<int> begin(1, 0), entries;
        for (size_t t = 0; t < per_tile.size(); t++) {
            for (int k : per_tile[t]) {
                const int gi = pairs[2 * k], gj = pairs[2 * k + 1];
                const int r0 = tl[2 * t], c0 = tl[2 * t + 1];
                entries.push_back((gi - r0) * tc + (gj - c0));
                entries.push_back((int)(off[k] >> 31));
                entries.push_back((int)(off[k] & 0x7FFFFFFF));
            }
            begin.push_back((int)entries.size() / 3);
        }
        DevBuf dtiles, dbegin, dent;
        HIPCHK(e, dtiles.ensure(tl.size() * 4));
        HIPCHK(e, dbegin.ensure(begin.size() * 4));
        HIPCHK(e, dent.ensure(entries.size() * 4));
        HIPCHK(e, hipMemcpyAsync(dtiles.p, tl.data(), tl.size() * 4, hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipMemcpyAsync(dbegin.p, begin.data(), begin.size() * 4, hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipMemcpyAsync(dent.p, entries.data(), entries.size() * 4, hipMemcpyHostToDevice, e->stream));
        const TileQueue q{dtiles.as<int>(), (int)(tl.size() / 2), 2, nullptr};
        HIPCHK(e, launch_tiled(e, in, bmax2d(s), used, variant, q, dbegin.as<int>(), dent.as<int>(), dout.as<int>()));
        HIPCHK(e, hipStreamSynchronize(e->stream));
    }
    HIPCHK(e, hipMemcpyAsync(counts, dout.p, total * 4, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return PARENT_OK;
}

int parent_engine_set_entropies(parent_engine* e, int n_dihedrals, const double* s1d, const double* s2d) {
    if (!e || !s1d || !s2d || n_dihedrals < 1) return PARENT_E_INVALID;
    HIPCHK(e, hipSetDevice(e->device));
    drop_graph(e);
    BinSet& s = cur(e);
    const size_t D = n_dofs(n_dihedrals), P = n_pairs(n_dihedrals);
    HIPCHK(e, s.s1d.ensure(D * 8));
    HIPCHK(e, s.s2d.ensure(std::max<size_t>(P, 1) * 8));
    HIPCHK(e, hipMemcpyAsync(s.s1d.p, s1d, D * 8, hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemcpyAsync(s.s2d.p, s2d, P * 8, hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    s.ent_nD = n_dihedrals;
    s.s2d_valid = true;
    s.prepared = false;
    return PARENT_OK;
}

int parent_engine_mist(parent_engine* e, int32_t* edges, double* mi, int32_t* left, double* totals) {
    if (!e || !edges || !mi || !left || !totals) return PARENT_E_INVALID;
    HIPCHK(e, hipSetDevice(e->device));
    return run_mist(e, cur(e), edges, mi, left, totals);
}

int parent_engine_mist_noalloc(parent_engine* e, int n_dihedrals, const double* s1d, const double* s2d,
                               int32_t* edges, double* mi, int32_t* left, double* totals, int32_t* status,
                               void* stream) {
    if (!e || !s1d || !s2d || !edges || !mi || !left || !totals || !status) return PARENT_E_INVALID;
    if (n_dihedrals < 1) return fail(e, PARENT_E_INVALID, "need n_dihedrals >= 1");
    if (n_dofs(n_dihedrals) > kMistNoAllocMaxVertices)
        return fail(e, PARENT_E_UNSUPPORTED, "parent_engine_mist_noalloc: more than 65,536 dofs");
    HIPCHK(e, hipSetDevice(e->device));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : e->stream;
    HIPCHK(e, launch_mist_noalloc(n_dihedrals, s1d, s2d, edges, mi, left, totals, status, e->noalloc_scratch, s));
    return PARENT_OK;
}

int parent_engine_autotune(parent_engine* e, parent_tune_entry* entries, int max_entries, int* n_entries) {
    if (!e) return PARENT_E_INVALID;
    BinSet& s = cur(e);
    const Traj& t = e->traj;
    if (!s.prepared) return fail(e, PARENT_E_INVALID, "parent_engine_prepare has not succeeded");
    HIPCHK(e, hipSetDevice(e->device));
    e->tuned.erase(std::remove_if(e->tuned.begin(), e->tuned.end(), [&](const parent_engine::Tuned& x) {
        return x.b == s.bins[3] && x.a == s.bins[4] && x.d == s.bins[5] && x.N == t.N && x.D == t.D;
    }), e->tuned.end());
    drop_graph(e);
    if (needs_slabs(e, s)) {
        // nothing to choose: only the slab kernel holds these bins
        e->last_tune.clear();
    } else if (int rc = run_autotune(e, t, s)) {
        return rc;
    }
    const int n = (int)e->last_tune.size();
    if (n_entries) *n_entries = n;
    for (int k = 0; k < n && k < max_entries && entries; k++) entries[k] = e->last_tune[k];
    return PARENT_OK;
}

int parent_engine_set_variant(parent_engine* e, int method, int variant) {
    if (!e) return PARENT_E_INVALID;
    if (method == PARENT_METHOD_SLABS) {
        if (variant < -1) return fail(e, PARENT_E_INVALID, "slab rows must be >= 0 (or -1 for the default)");
        e->forced_slab_rows = variant < 0 ? 0 : variant;
        drop_graph(e);
        return PARENT_OK;
    }
    if (method == PARENT_METHOD_MFMA) {
        if (variant < -1 || variant > 1) return fail(e, PARENT_E_INVALID, "MFMA variant must be -1, 0 or 1");
        e->forced_mfma = variant;
        drop_graph(e);
        return PARENT_OK;
    }
    if (method != PARENT_METHOD_ATOMIC)
        return fail(e, PARENT_E_INVALID, "only PARENT_METHOD_ATOMIC, PARENT_METHOD_MFMA and PARENT_METHOD_SLABS have variants");
    if (variant < -1 || variant >= atomic_config_count()) return fail(e, PARENT_E_INVALID, "variant out of range");
    e->forced_atomic = variant;
    return PARENT_OK;
}

int parent_batch_set_classes(parent_batch* b, int mode) {
    if (!b) return PARENT_E_INVALID;
    if (mode < 0 || mode > 2) return bfail(b, PARENT_E_INVALID, "classes mode must be 0, 1 or 2");
    b->classes = mode;
    b->tuned_items = -1;  // plan again on the next PARENT_METHOD_AUTO run
    return PARENT_OK;
}

int parent_batch_set_variant(parent_batch* b, int method, int variant) {
    if (!b) return PARENT_E_INVALID;
    const int rc = parent_engine_set_variant(b->eng, method, variant);
    b->queues.clear();  // a merged queue is dealt for its kernel's launch size
    return rc;
}

int parent_engine_set_mist_path(parent_engine* e, int mode) {
    if (!e) return PARENT_E_INVALID;
    if (mode != 0 && mode != 1) return fail(e, PARENT_E_INVALID, "MIST path mode must be 0 or 1");
    e->mist_path_mode = mode;
    return PARENT_OK;
}

int parent_engine_variant_count(parent_engine* e, int method, int* count) {
    if (!e || !count) return PARENT_E_INVALID;
    *count = method == PARENT_METHOD_ATOMIC ? atomic_config_count()
             : method == PARENT_METHOD_MFMA   ? 2
                                              : (method == PARENT_METHOD_PLAIN || method == PARENT_METHOD_SLABS ? 1 : 0);
    return PARENT_OK;
}

int parent_engine_sync(parent_engine* e) {
    if (!e) return PARENT_E_INVALID;
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return PARENT_OK;
}

int parent_engine_stats(parent_engine* e, parent_stats* out) {
    if (!e || !out) return PARENT_E_INVALID;
    *out = e->stats;
    return PARENT_OK;
}

/* ---------------------------------------------------------------------- */
/* batch of trajectories                                                   */
/* ---------------------------------------------------------------------- */

int parent_batch_create(int device, const int bins[6], parent_batch** out) {
    if (!out || !bins) return PARENT_E_INVALID;
    *out = nullptr;
    for (int k = 0; k < 6; k++)
        if (bins[k] < 1 || bins[k] > kMaxBins) {
            set_last_error(bins[k] < 1 ? "number of bins < 1" : "more than 4096 bins per dof is not supported");
            return bins[k] < 1 ? PARENT_E_INVALID : PARENT_E_UNSUPPORTED;
        }
    parent_engine* e = nullptr;
    if (int rc = parent_engine_create(device, &e)) return rc;
    parent_batch* b = new parent_batch();
    b->eng = e;
    memcpy(b->bins, bins, sizeof(b->bins));
    *out = b;
    return PARENT_OK;
}

void parent_batch_destroy(parent_batch* b) {
    if (!b) return;
    if (b->eng) {
        (void)hipSetDevice(b->eng->device);
        (void)hipStreamSynchronize(b->eng->stream);
    }
    b->items.clear();
    b->queues.clear();
    b->descs.release();
    parent_engine_destroy(b->eng);
    delete b;
}

const char* parent_batch_error(const parent_batch* b) { return b && b->eng ? b->eng->err.c_str() : "null batch"; }

int parent_batch_add(parent_batch* b, int n_dihedrals, int n_frames, int double_prec, const void* values,
                     int on_device, int* index) {
    if (!b) return PARENT_E_INVALID;
    return parent_batch_add_binned(b, b->bins, n_dihedrals, n_frames, double_prec, values, on_device, index);
}

int parent_batch_add_binned(parent_batch* b, const int bins[6], int n_dihedrals, int n_frames, int double_prec,
                            const void* values, int on_device, int* index) {
    if (!b || !values || !bins) return PARENT_E_INVALID;
    parent_engine* e = b->eng;
    if (int rc = check_bins(e, bins)) return rc;
    if (n_dihedrals < 1 || n_frames < 1 || (double_prec != 0 && double_prec != 1))
        return fail(e, PARENT_E_INVALID, "need n_dihedrals >= 1, n_frames >= 1, double_prec in {0,1}");
    HIPCHK(e, hipSetDevice(e->device));
    std::unique_ptr<parent_batch::Item> it(new parent_batch::Item());
    Traj& t = it->traj;
    t.nD = n_dihedrals;
    t.N = n_frames;
    t.dp = double_prec;
    t.D = n_dofs(n_dihedrals);
    t.Np = (int)(((long long)n_frames + kFrameAlign - 1) / kFrameAlign * kFrameAlign);
    if (on_device) {
        t.values = values;
    } else {
        const size_t bytes = (size_t)t.N * t.D * (double_prec ? 8 : 4);
        HIPCHK(e, t.values_own.ensure(bytes));
        HIPCHK(e, hipMemcpyAsync(t.values_own.p, values, bytes, hipMemcpyHostToDevice, e->stream));
        t.values = t.values_own.p;
    }
    memcpy(it->set.bins, bins, sizeof(it->set.bins));
    // run_prepare may reallocate the shared n ln n table (a longer
    // trajectory), which the uploaded queue descriptors point to: drop the
    // queues and descriptors before, so that a rejected trajectory leaves no
    // stale pointer
    b->queues.clear();
    b->descs_items = -1;
    const int rc = run_prepare(e, t, it->set);
    // the index matrix and per-dof constants are all the pairs need: the
    // uploaded values go right away
    t.values_own.release();
    t.values = nullptr;
    if (rc) return rc;
    if (index) *index = (int)b->items.size();
    b->items.push_back(std::move(it));
    return PARENT_OK;
}

int parent_batch_run(parent_batch* b, int method) {
    if (!b) return PARENT_E_INVALID;
    parent_engine* e = b->eng;
    if (b->items.empty()) return bfail(b, PARENT_E_INVALID, "empty batch");
    HIPCHK(e, hipSetDevice(e->device));
    const int n = (int)b->items.size();
    // the item with the most pair-frame updates: the PairInputs argument of
    // the launches (the batch instantiations read each tile's own descriptor)
    int big = 0;
    for (int k = 1; k < n; k++) {
        const Traj &a = b->items[k]->traj, &c = b->items[big]->traj;
        if ((double)n_pairs(a.nD) * a.N > (double)n_pairs(c.nD) * c.N) big = k;
    }
    // the n ln n table must cover the longest trajectory
    int nmax = 0;
    for (auto& it : b->items) nmax = std::max(nmax, it->traj.N);
    HIPCHK(e, e->scratch.nlogn.ensure(((size_t)nmax + 1) * 8));
    HIPCHK(e, launch_nlogn(nmax, e->scratch.nlogn.as<double>(), e->stream));
    BinSet& sb = b->items[big]->set;
    const Traj& tb = b->items[big]->traj;
    const BinSet env = batch_envelope(b);  // the one kernel must hold every trajectory's bins
    if (method < PARENT_METHOD_PLAIN || method > PARENT_METHOD_SLABS) return bfail(b, PARENT_E_INVALID, "unknown method");
    if (method == PARENT_METHOD_SLABS || needs_slabs(e, env)) {
        // bins the LDS kernels cannot hold: the slab kernel, one launch per
        // trajectory (it has no tile queue)
        HIPCHK(e, hipEventRecord(e->ev[4], e->stream));
        double upd = 0;
        for (auto& it : b->items) {
            const PairInputs in = with_nlogn(e, pair_inputs(it->traj, it->set));
            HIPCHK(e, launch_pairs_slabs(in, slab_lds(e, it->set), e->forced_slab_rows, nullptr, e->stream));
            upd += (double)n_pairs(it->traj.nD) * it->traj.N;
        }
        HIPCHK(e, hipEventRecord(e->ev[5], e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
        for (auto& it : b->items) {
            it->set.s2d_valid = true;
            it->set.method = PARENT_METHOD_SLABS;
            it->set.variant = e->forced_slab_rows;
        }
        b->method = PARENT_METHOD_SLABS;
        b->variant = e->forced_slab_rows;
        e->stats.ms_pairs = elapsed(e->ev[4], e->ev[5]);
        e->stats.pair_updates = upd;
        e->stats.method = PARENT_METHOD_SLABS;
        e->stats.kernel_variant = e->forced_slab_rows;
        e->stats.batch_launches = (int)b->items.size();
        return PARENT_OK;
    }
    std::vector<parent_batch::Plan> plan;
    if (method == PARENT_METHOD_AUTO) {
        if (b->tuned_items != n)
            if (int rc = plan_batch(e, b)) return rc;
        plan = b->tuned_plan;
    } else {
        parent_batch::Plan one;
        one.members.resize(n);
        std::iota(one.members.begin(), one.members.end(), 0);
        if (method == PARENT_METHOD_MFMA && mfma_fits(e, env)) {
            one.method = PARENT_METHOD_MFMA;
            one.variant = e->forced_mfma >= 0 && mfma_variant_fits(e->forced_mfma, bmax2d(env), nmax) ? e->forced_mfma : 0;
        } else {
            // ATOMIC, MFMA outside its bin range, and PLAIN (which has no tile
            // queue): the tiled atomic kernel, in a configuration whose counters
            // hold the longest trajectory
            one.method = PARENT_METHOD_ATOMIC;
            one.variant = choose_atomic_config(e, env, nmax);
            if (one.variant < 0) return bfail(b, PARENT_E_UNSUPPORTED, "no tiled kernel fits these bins");
        }
        plan = {one};
    }
    if (int rc = ensure_descs(e, b)) return rc;
    if (b->queues.size() != plan.size()) {
        b->queues.clear();
        for (size_t i = 0; i < plan.size(); i++) b->queues.emplace_back(new parent_batch::Queue());
    }
    for (size_t i = 0; i < plan.size(); i++) {
        parent_batch::Queue& qu = *b->queues[i];
        if (qu.ntiles > 0 && qu.plan.members == plan[i].members && qu.plan.method == plan[i].method &&
            qu.plan.variant == plan[i].variant)
            continue;
        qu.plan = plan[i];
        qu.ntiles = 0;
        if (int rc = build_queue(e, b, env, qu)) return rc;
    }
    const PairInputs in0 = with_nlogn(e, pair_inputs(tb, sb));  // unused by the batch instantiations
    HIPCHK(e, hipEventRecord(e->ev[4], e->stream));
    for (auto& qu : b->queues) {
        const TileQueue q{qu->tiles.as<int>(), qu->ntiles, 3, b->descs.as<PairInputs>()};
        HIPCHK(e, launch_tiled(e, in0, bmax2d(env), qu->plan.method, qu->plan.variant, q, nullptr, nullptr, nullptr));
    }
    HIPCHK(e, hipEventRecord(e->ev[5], e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    // stats name the configuration of the launch with the most updates
    double upd = 0, top = -1;
    for (auto& qu : b->queues) {
        double u = 0;
        for (int k : qu->plan.members) {
            BinSet& s = b->items[k]->set;
            s.s2d_valid = true;
            s.method = qu->plan.method;
            s.variant = qu->plan.variant;
            u += (double)n_pairs(b->items[k]->traj.nD) * b->items[k]->traj.N;
        }
        upd += u;
        if (u > top) {
            top = u;
            b->method = qu->plan.method;
            b->variant = qu->plan.variant;
        }
    }
    e->stats.ms_pairs = elapsed(e->ev[4], e->ev[5]);
    e->stats.pair_updates = upd;
    e->stats.method = b->method;
    e->stats.kernel_variant = b->variant;
    e->stats.batch_launches = (int)b->queues.size();
    return PARENT_OK;
}

int parent_batch_autotune(parent_batch* b, parent_tune_entry* entries, int max_entries, int* n_entries) {
    if (!b) return PARENT_E_INVALID;
    parent_engine* e = b->eng;
    if (b->items.empty()) return bfail(b, PARENT_E_INVALID, "empty batch");
    HIPCHK(e, hipSetDevice(e->device));
    int nmax = 0;
    for (auto& it : b->items) nmax = std::max(nmax, it->traj.N);
    HIPCHK(e, e->scratch.nlogn.ensure(((size_t)nmax + 1) * 8));
    HIPCHK(e, launch_nlogn(nmax, e->scratch.nlogn.as<double>(), e->stream));
    if (int rc = plan_batch(e, b)) return rc;
    const int n = (int)e->last_tune.size();
    if (n_entries) *n_entries = n;
    for (int k = 0; k < n && k < max_entries && entries; k++) entries[k] = e->last_tune[k];
    return PARENT_OK;
}

int parent_batch_read(parent_batch* b, int index, double* s1d, double* s2d) {
    if (!b) return PARENT_E_INVALID;
    parent_engine* e = b->eng;
    if (index < 0 || index >= (int)b->items.size()) return fail(e, PARENT_E_INVALID, "batch index out of range");
    HIPCHK(e, hipSetDevice(e->device));
    BinSet& s = b->items[index]->set;
    if (s1d)
        if (int rc = read_s1d(e, s, s1d)) return rc;
    if (s2d)
        if (int rc = read_s2d(e, s, s2d)) return rc;
    return PARENT_OK;
}

int parent_batch_mist(parent_batch* b, int index, int32_t* edges, double* mi, int32_t* left, double* totals) {
    if (!b || !edges || !mi || !left || !totals) return PARENT_E_INVALID;
    parent_engine* e = b->eng;
    if (index < 0 || index >= (int)b->items.size()) return fail(e, PARENT_E_INVALID, "batch index out of range");
    HIPCHK(e, hipSetDevice(e->device));
    return run_mist(e, b->items[index]->set, edges, mi, left, totals);
}

int parent_batch_stats(parent_batch* b, parent_stats* out) {
    if (!b || !out) return PARENT_E_INVALID;
    *out = b->eng->stats;
    return PARENT_OK;
}

}  // extern "C"
