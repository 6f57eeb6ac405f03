// Engine: device buffers, phase sequencing and the C ABI of parent_hip.h.
//
// One engine = one HIP stream on one GPU.  The trajectory lives in HBM as
// frame-major values; prepare() turns it into dof-major uint8 bin indices
// plus per-dof constants; pairs() runs one of the pair kernels over all
// D(D-1)/2 dof pairs and leaves S2D in .par body order on the device.
#include "../../include/parent_hip.h"
#include "host_common.h"
#include "kernels.h"

#include <algorithm>
#include <map>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using namespace parent;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= n && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (bytes == 0) return hipSuccess;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = bytes;
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
    int tr = -1, tc = -1, D = -1;
    int ntiles = 0;
    DevBuf dev;
};

}  // namespace

struct parent_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    std::string err;
    int lds_max = 65536;

    // trajectory
    int nD = -1, N = 0, dp = 0, D = 0, Np = 0;
    const void* values = nullptr;
    DevBuf values_own;

    // prepare outputs
    int bins[6] = {0, 0, 0, 0, 0, 0};
    bool prepared = false;
    bool s2d_valid = false;
    DevBuf emin, emax, mfhist, flags, vmin, vmax, modfit, bins2, bins1, s1d, c2, jtab, nlogn, s2d;

    // MIST scratch
    DevBuf mkey, mupar, minpos, mstate, medges, mmi, mleft;
    int ent_nD = -1;  // nD of the entropies currently in s1d/s2d

    TileList tiles;
    parent_stats stats{};

    // autotuner: the fastest pair-kernel configuration per problem key
    struct Tuned {
        int b, a, d, N, D;
        int method, variant;
    };
    std::vector<Tuned> tuned;
    std::vector<parent_tune_entry> last_tune;
    int forced_atomic = -1;  // parent_engine_set_variant
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

Bins3 bins2d(const parent_engine* e) { return Bins3{e->bins[3], e->bins[4], e->bins[5]}; }
Bins3 bins1d(const parent_engine* e) { return Bins3{e->bins[0], e->bins[1], e->bins[2]}; }

int bmax2d(const parent_engine* e) { return std::max(e->bins[3], std::max(e->bins[4], e->bins[5])); }

PairInputs pair_inputs(parent_engine* e) {
    PairInputs in{};
    in.bins = e->bins2.as<uint8_t>();
    in.nD = e->nD;
    in.D = e->D;
    in.N = e->N;
    in.Np = e->Np;
    in.B = bins2d(e);
    in.vmin = e->vmin.as<double>();
    in.vmax = e->vmax.as<double>();
    in.jtab = e->jtab.as<double>();
    in.c2 = e->c2.as<double>();
    in.nlogn = e->nlogn.as<double>();
    in.logN = log((double)e->N);
    in.s2d = e->s2d.as<double>();
    return in;
}

// Tiles of a TR x TC configuration over the upper triangle: row blocks at
// multiples of TR, column blocks at multiples of TC, kept when some column of
// the block lies above the block's first row.  Order: bands of ~16 rows, and
// within a band column by column, so the workgroups running at the same time
// (one contiguous run of the list per XCD) read the same column-dof index rows
// from L2 while their band's row-dof rows stay resident there.
void push_band(std::vector<int>& t, int D, int tr, int tc, int b0, int b1) {
    const int cfirst = ((b0 + 1) / tc) * tc;
    for (int c0 = cfirst; c0 < D; c0 += tc)
        for (int r0 = b0; r0 < b1; r0 += tr)
            if (c0 + tc - 1 > r0) {
                t.push_back(r0);
                t.push_back(c0);
            }
}

std::vector<int> make_tiles(int D, int tr, int tc) {
    std::vector<int> t;
    const int band = std::max(tr, 16 / tr * tr);
    for (int b0 = 0; b0 < D - 1; b0 += band) push_band(t, D, tr, tc, b0, std::min(b0 + band, D - 1));
    return t;
}

// Tile shape (row dofs x column dofs) of a tiled method's variant.
void tile_shape(const parent_engine* e, int method, int variant, int* tr, int* tc) {
    if (method == PARENT_METHOD_MFMA) {
        *tr = *tc = mfma_tile_dofs(bmax2d(e));
    } else {
        const AtomicConfig c = atomic_config(variant);
        *tr = c.tr;
        *tc = c.tc;
    }
}

int ensure_tiles(parent_engine* e, int method, int variant) {
    int tr, tc;
    tile_shape(e, method, variant, &tr, &tc);
    if (e->tiles.tr == tr && e->tiles.tc == tc && e->tiles.D == e->D) return PARENT_OK;
    std::vector<int> t = make_tiles(e->D, tr, tc);
    HIPCHK(e, e->tiles.dev.ensure(std::max<size_t>(t.size(), 3) * sizeof(int)));
    if (!t.empty())
        HIPCHK(e, hipMemcpyAsync(e->tiles.dev.p, t.data(), t.size() * sizeof(int), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->tiles.tr = tr;
    e->tiles.tc = tc;
    e->tiles.D = e->D;
    e->tiles.ntiles = (int)(t.size() / 2);
    return PARENT_OK;
}

// Largest tile configuration whose histograms fit the LDS of one workgroup.
bool variant_fits(const parent_engine* e, int id) { return atomic_fits(id, bmax2d(e), (size_t)e->lds_max); }

bool mfma_fits(const parent_engine* e) {
    return mfma_tile_dofs(bmax2d(e)) > 0 && mfma_lds_bytes() <= (size_t)e->lds_max;
}

// one launch of a tiled method's variant over a work list
hipError_t launch_variant(parent_engine* e, const PairInputs& in, int method, int id, const int* work, int nwork,
                          const int* db, const int* de, int* dout) {
    if (method == PARENT_METHOD_MFMA) return launch_pairs_mfma(in, bmax2d(e), work, nwork, db, de, dout, e->stream);
    return launch_pairs_atomic(in, id, work, nwork, db, de, dout, e->stream);
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

int choose_atomic_config(const parent_engine* e) {
    const int bmax = bmax2d(e);
    if (e->forced_atomic >= 0) return variant_fits(e, e->forced_atomic) ? e->forced_atomic : -1;
    // fallback order when not autotuning: the best of the autotuner's tables
    // at 50 bins first, then configurations with larger histogram slots
    static const int order[] = {2, 3, 1, 9, 0, 11, 13, 10, 15, 16, 14, 18};
    for (int id : order)
        if (variant_fits(e, id)) return id;
    (void)bmax;
    return -1;
}

int check_ready(parent_engine* e) {
    if (!e->prepared) return fail(e, PARENT_E_INVALID, "parent_engine_prepare has not succeeded");
    return PARENT_OK;
}

// Rows sampled by the autotuner: 8 runs of 12 rows spread over the dof range.
// Each configuration's rate is its own work list's pair count (tile heights
// that do not divide 12 cover a few more rows) over its time.
std::vector<std::pair<int, int>> tune_rows(int D) {
    std::vector<std::pair<int, int>> rows;
    for (int k = 0; k < 8; k++) {
        int r = (int)((k + 0.5) * (D - 1) / 8.0) / 12 * 12;
        int r1 = std::min(r + 12, D - 1);
        if (r < r1 && (rows.empty() || rows.back().second <= r)) rows.push_back({r, r1});
    }
    return rows;
}

std::vector<int> tune_work(int D, int tr, int tc, const std::vector<std::pair<int, int>>& rows) {
    std::vector<int> t;
    for (auto& rr : rows) push_band(t, D, tr, tc, rr.first, rr.second);
    return t;
}

// Times `launch` twice (first run warms caches) with events on the engine stream.
template <class F>
int time_launch(parent_engine* e, F launch, double* ms) {
    for (int rep = 0; rep < 2; rep++) {
        HIPCHK(e, hipEventRecord(e->ev[4], e->stream));
        HIPCHK(e, launch());
        HIPCHK(e, hipEventRecord(e->ev[5], e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
    }
    *ms = elapsed(e->ev[4], e->ev[5]);
    return PARENT_OK;
}

int run_autotune(parent_engine* e) {
    e->last_tune.clear();
    const int D = e->D;
    const auto rows = tune_rows(D);
    PairInputs in = pair_inputs(e);
    DevBuf dt;
    int best_m = PARENT_METHOD_PLAIN, best_v = 0;
    double best_rate = -1;
    // candidates: every fitting atomic configuration, then the MFMA kernel
    std::vector<std::pair<int, int>> cand;
    for (int id = 0; id < atomic_config_count(); id++)
        if (variant_fits(e, id)) cand.push_back({PARENT_METHOD_ATOMIC, id});
    if (mfma_fits(e)) cand.push_back({PARENT_METHOD_MFMA, 0});
    for (auto& mv : cand) {
        const int m = mv.first, id = mv.second;
        int tr, tc;
        tile_shape(e, m, id, &tr, &tc);
        const int block = m == PARENT_METHOD_MFMA ? 256 : atomic_config(id).block;
        std::vector<int> t = tune_work(D, tr, tc, rows);
        const double updates = (double)tile_pairs(D, t, tr, tc) * e->N;
        HIPCHK(e, dt.ensure(std::max<size_t>(t.size(), 2) * sizeof(int)));
        HIPCHK(e, hipMemcpyAsync(dt.p, t.data(), t.size() * sizeof(int), hipMemcpyHostToDevice, e->stream));
        double ms = 0;
        const int nt = (int)(t.size() / 2);
        if (int rc = time_launch(e, [&] { return launch_variant(e, in, m, id, dt.as<int>(), nt, nullptr, nullptr, nullptr); }, &ms))
            return rc;
        parent_tune_entry te{m, id, tr, tc, block, ms, updates, updates / (ms * 1e-3)};
        e->last_tune.push_back(te);
        if (te.updates_per_s > best_rate) {
            best_rate = te.updates_per_s;
            best_m = m;
            best_v = id;
        }
    }
    dt.release();
    e->tuned.push_back({e->bins[3], e->bins[4], e->bins[5], e->N, D, best_m, best_v});
    return PARENT_OK;
}

// The tuned (method, variant) for the current problem, tuning on first use.
int tuned_choice(parent_engine* e, int* method, int* variant) {
    for (auto& t : e->tuned)
        if (t.b == e->bins[3] && t.a == e->bins[4] && t.d == e->bins[5] && t.N == e->N && t.D == e->D) {
            *method = t.method;
            *variant = t.variant;
            return PARENT_OK;
        }
    if (int rc = run_autotune(e)) return rc;
    const auto& t = e->tuned.back();
    *method = t.method;
    *variant = t.variant;
    return PARENT_OK;
}

// The kernel a request runs on: AUTO takes the autotuner's choice, MFMA falls
// back to the atomic kernels outside its bin range, ATOMIC falls back to the
// plain kernel when no tile configuration fits the LDS.
int resolve_method(parent_engine* e, int method, int* used, int* variant) {
    *used = method;
    *variant = 0;
    if (method == PARENT_METHOD_AUTO) return tuned_choice(e, used, variant);
    if (method == PARENT_METHOD_MFMA) {
        if (mfma_fits(e)) return PARENT_OK;
        *used = PARENT_METHOD_ATOMIC;
    }
    if (*used == PARENT_METHOD_ATOMIC) {
        *variant = choose_atomic_config(e);
        if (*variant < 0) {
            *used = PARENT_METHOD_PLAIN;
            *variant = 0;
        }
    }
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
    if ((h = hipSetDevice(device)) != hipSuccess || (h = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess) {
        set_last_error(std::string("HIP init failed: ") + hipGetErrorString(h));
        delete e;
        return PARENT_E_DEVICE;
    }
    for (auto& ev : e->ev) {
        if ((h = hipEventCreate(&ev)) != hipSuccess) {
            set_last_error(std::string("hipEventCreate failed: ") + hipGetErrorString(h));
            parent_engine_destroy(e);
            return PARENT_E_DEVICE;
        }
    }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        e->lds_max = lds;
    *out = e;
    return PARENT_OK;
}

void parent_engine_destroy(parent_engine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    DevBuf* bufs[] = {&e->values_own, &e->emin, &e->emax, &e->mfhist, &e->flags, &e->vmin, &e->vmax,
                      &e->modfit, &e->bins2, &e->bins1, &e->s1d, &e->c2, &e->jtab, &e->nlogn, &e->s2d,
                      &e->mkey, &e->mupar, &e->minpos, &e->mstate, &e->medges, &e->mmi, &e->mleft,
                      &e->tiles.dev};
    for (DevBuf* b : bufs) b->release();
    for (auto& ev : e->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char* parent_engine_error(const parent_engine* e) { return e ? e->err.c_str() : "null engine"; }

int parent_engine_set_trajectory(parent_engine* e, int n_dihedrals, int n_frames, int double_prec,
                                 const void* values, int on_device) {
    if (!e || !values) return PARENT_E_INVALID;
    if (n_dihedrals < 1 || n_frames < 1 || (double_prec != 0 && double_prec != 1))
        return fail(e, PARENT_E_INVALID, "need n_dihedrals >= 1, n_frames >= 1, double_prec in {0,1}");
    HIPCHK(e, hipSetDevice(e->device));
    e->nD = n_dihedrals;
    e->N = n_frames;
    e->dp = double_prec;
    e->D = n_dofs(n_dihedrals);
    e->Np = (int)(((long long)n_frames + kFrameAlign - 1) / kFrameAlign * kFrameAlign);
    e->prepared = false;
    e->s2d_valid = false;
    e->stats = parent_stats{};
    const size_t bytes = (size_t)e->N * e->D * (double_prec ? 8 : 4);
    if (on_device) {
        e->values = values;
    } else {
        HIPCHK(e, e->values_own.ensure(bytes));
        HIPCHK(e, hipEventRecord(e->ev[0], e->stream));
        HIPCHK(e, hipMemcpyAsync(e->values_own.p, values, bytes, hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipEventRecord(e->ev[1], e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
        e->stats.ms_upload = elapsed(e->ev[0], e->ev[1]);
        e->values = e->values_own.p;
    }
    return PARENT_OK;
}

int parent_engine_prepare(parent_engine* e, const int bins[6]) {
    if (!e || !bins) return PARENT_E_INVALID;
    if (!e->values) return fail(e, PARENT_E_INVALID, "no trajectory set");
    for (int k = 0; k < 6; k++) {
        if (bins[k] < 1) return fail(e, PARENT_E_INVALID, "number of bins < 1");
        if (bins[k] > kMaxBins) return fail(e, PARENT_E_UNSUPPORTED, "more than 255 bins per dof is not supported");
    }
    HIPCHK(e, hipSetDevice(e->device));
    memcpy(e->bins, bins, sizeof(e->bins));
    e->prepared = false;
    e->s2d_valid = false;
    const int D = e->D, nD = e->nD, N = e->N;
    const bool need1 = bins[0] != bins[3] || bins[1] != bins[4] || bins[2] != bins[5];
    HIPCHK(e, e->emin.ensure(D * 8));
    HIPCHK(e, e->emax.ensure(D * 8));
    HIPCHK(e, e->mfhist.ensure((size_t)nD * kModfitBins * 4));
    HIPCHK(e, e->flags.ensure(64));
    HIPCHK(e, e->vmin.ensure(D * 8));
    HIPCHK(e, e->vmax.ensure(D * 8));
    HIPCHK(e, e->modfit.ensure(D * 8));
    HIPCHK(e, e->bins2.ensure((size_t)D * e->Np));
    if (need1) HIPCHK(e, e->bins1.ensure((size_t)D * e->Np));
    HIPCHK(e, e->s1d.ensure(D * 8));
    HIPCHK(e, e->c2.ensure(D * 8));
    HIPCHK(e, e->jtab.ensure((size_t)D * 256 * 8));
    HIPCHK(e, e->nlogn.ensure(((size_t)N + 1) * 8));
    PrepareArgs a{};
    a.values = e->values;
    a.double_prec = e->dp;
    a.nD = nD;
    a.D = D;
    a.N = N;
    a.Np = e->Np;
    a.pi = acos(-1.0);  // PARENT.cpp:112
    a.B1 = bins1d(e);
    a.B2 = bins2d(e);
    a.emin = e->emin.as<unsigned long long>();
    a.emax = e->emax.as<unsigned long long>();
    a.mfhist = e->mfhist.as<unsigned>();
    a.flags = e->flags.as<int>();
    a.vmin = e->vmin.as<double>();
    a.vmax = e->vmax.as<double>();
    a.modfit = e->modfit.as<double>();
    a.bins2 = e->bins2.as<uint8_t>();
    a.bins1 = need1 ? e->bins1.as<uint8_t>() : nullptr;
    a.s1d = e->s1d.as<double>();
    a.c2 = e->c2.as<double>();
    a.jtab = e->jtab.as<double>();
    a.nlogn = e->nlogn.as<double>();
    HIPCHK(e, hipEventRecord(e->ev[2], e->stream));
    HIPCHK(e, launch_prepare(a, e->stream));
    HIPCHK(e, hipEventRecord(e->ev[3], e->stream));
    int flags = 0;
    HIPCHK(e, hipMemcpyAsync(&flags, e->flags.p, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->stats.ms_prepare = elapsed(e->ev[2], e->ev[3]);
    if (flags & kFlagNegBond) return fail(e, PARENT_E_DATA, "ERROR: a bond length is smaller than 0.0");
    if (flags & kFlagRange)
        return fail(e, PARENT_E_RANGE, "ERROR: a value falls outside its histogram range (angle >= pi or dihedral outside [0, 2pi+5e-9))");
    e->prepared = true;
    e->ent_nD = nD;
    return PARENT_OK;
}

int parent_engine_pairs(parent_engine* e, int method) {
    if (!e) return PARENT_E_INVALID;
    if (int rc = check_ready(e)) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    const long long P = n_pairs(e->nD);
    HIPCHK(e, e->s2d.ensure((size_t)std::max<long long>(P, 1) * 8));
    const int bmax = bmax2d(e);
    if ((size_t)bmax * bmax * 4 > (size_t)e->lds_max)
        return fail(e, PARENT_E_UNSUPPORTED, "2D histogram does not fit in LDS");
    PairInputs in = pair_inputs(e);
    int used = method, variant = 0;
    if (int rc = resolve_method(e, method, &used, &variant)) return rc;
    if (used != PARENT_METHOD_PLAIN) {
        if (int rc = ensure_tiles(e, used, variant)) return rc;
    }
    HIPCHK(e, hipEventRecord(e->ev[4], e->stream));
    if (used == PARENT_METHOD_PLAIN) {
        HIPCHK(e, launch_pairs_plain(in, nullptr, e->stream));
    } else if (used == PARENT_METHOD_ATOMIC || used == PARENT_METHOD_MFMA) {
        HIPCHK(e, launch_variant(e, in, used, variant, e->tiles.dev.as<int>(), e->tiles.ntiles, nullptr, nullptr, nullptr));
    } else {
        return fail(e, PARENT_E_INVALID, "unknown method");
    }
    HIPCHK(e, hipEventRecord(e->ev[5], e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->stats.ms_pairs = elapsed(e->ev[4], e->ev[5]);
    e->stats.pair_updates = (double)P * e->N;
    e->stats.method = used;
    e->stats.kernel_variant = variant;
    e->s2d_valid = true;
    return PARENT_OK;
}

int parent_engine_read_s1d(parent_engine* e, double* out) {
    if (!e || !out) return PARENT_E_INVALID;
    if (e->ent_nD < 0) return fail(e, PARENT_E_INVALID, "no entropies computed");
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipMemcpyAsync(out, e->s1d.p, (size_t)n_dofs(e->ent_nD) * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return PARENT_OK;
}

int parent_engine_read_s2d(parent_engine* e, double* out) {
    if (!e || !out) return PARENT_E_INVALID;
    if (!e->s2d_valid) return fail(e, PARENT_E_INVALID, "parent_engine_pairs has not succeeded");
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipMemcpyAsync(out, e->s2d.p, (size_t)n_pairs(e->ent_nD) * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return PARENT_OK;
}

int parent_engine_read_ranges(parent_engine* e, double* vmin, double* vmax, double* modfit) {
    if (!e) return PARENT_E_INVALID;
    if (int rc = check_ready(e)) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    if (vmin) HIPCHK(e, hipMemcpyAsync(vmin, e->vmin.p, e->D * 8, hipMemcpyDeviceToHost, e->stream));
    if (vmax) HIPCHK(e, hipMemcpyAsync(vmax, e->vmax.p, e->D * 8, hipMemcpyDeviceToHost, e->stream));
    if (modfit) HIPCHK(e, hipMemcpyAsync(modfit, e->modfit.p, e->D * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return PARENT_OK;
}

int parent_engine_read_bins(parent_engine* e, uint8_t* out) {
    if (!e || !out) return PARENT_E_INVALID;
    if (int rc = check_ready(e)) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipMemcpy2DAsync(out, e->N, e->bins2.p, e->Np, e->N, e->D, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return PARENT_OK;
}

int parent_engine_pair_counts(parent_engine* e, int method, int n_pairs_req, const int32_t* pairs,
                              int32_t* counts) {
    if (!e || n_pairs_req < 0 || (n_pairs_req > 0 && (!pairs || !counts))) return PARENT_E_INVALID;
    if (int rc = check_ready(e)) return rc;
    if (n_pairs_req == 0) return PARENT_OK;
    HIPCHK(e, hipSetDevice(e->device));
    const Bins3 B = bins2d(e);
    std::vector<long long> off(n_pairs_req);
    long long total = 0;
    for (int k = 0; k < n_pairs_req; k++) {
        const int gi = pairs[2 * k], gj = pairs[2 * k + 1];
        if (gi < 0 || gj >= e->D || gi >= gj) return fail(e, PARENT_E_INVALID, "pairs must satisfy 0 <= gi < gj < D");
        off[k] = total;
        total += (long long)B.of(dof_type(e->nD, gi)) * B.of(dof_type(e->nD, gj));
    }
    DevBuf dpairs, doff, dout;
    HIPCHK(e, dout.ensure(total * 4));
    PairInputs in = pair_inputs(e);
    int used = method, variant = 0;
    if (int rc = resolve_method(e, method, &used, &variant)) return rc;
    int rc = PARENT_OK;
    if (used == PARENT_METHOD_PLAIN) {
        HIPCHK(e, dpairs.ensure((size_t)n_pairs_req * 8));
        HIPCHK(e, doff.ensure((size_t)n_pairs_req * 8));
        HIPCHK(e, hipMemcpyAsync(dpairs.p, pairs, (size_t)n_pairs_req * 8, hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipMemcpyAsync(doff.p, off.data(), (size_t)n_pairs_req * 8, hipMemcpyHostToDevice, e->stream));
        CountsDump dump{n_pairs_req, dpairs.as<int>(), doff.as<long long>(), dout.as<int>()};
        HIPCHK(e, launch_pairs_plain(in, &dump, e->stream));
    } else if (used == PARENT_METHOD_ATOMIC || used == PARENT_METHOD_MFMA) {
        // the tiles holding the requested pairs, each pair as (slot, offset)
        AtomicConfig c{};
        tile_shape(e, used, variant, &c.tr, &c.tc);
        const int stride = 2;
        std::vector<std::vector<int>> per_tile;
        std::vector<int> tl;
        std::map<long long, int> key;  // tile key -> index
        for (int k = 0; k < n_pairs_req; k++) {
            const int gi = pairs[2 * k], gj = pairs[2 * k + 1];
            const int r0 = gi / c.tr * c.tr, c0 = gj / c.tc * c.tc;
            const long long tk = (long long)r0 * e->D + c0;
            auto it = key.find(tk);
            int idx;
            if (it == key.end()) {
                idx = (int)tl.size() / stride;
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
                const int r0 = tl[stride * t], c0 = tl[stride * t + 1];
                entries.push_back((gi - r0) * c.tc + (gj - c0));
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
        HIPCHK(e, launch_variant(e, in, used, variant, dtiles.as<int>(), (int)(tl.size() / stride), dbegin.as<int>(), dent.as<int>(), dout.as<int>()));
        HIPCHK(e, hipStreamSynchronize(e->stream));
        dtiles.release();
        dbegin.release();
        dent.release();
    } else {
        rc = fail(e, PARENT_E_INVALID, "unknown method");
    }
    if (rc == PARENT_OK) {
        HIPCHK(e, hipMemcpyAsync(counts, dout.p, total * 4, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
    }
    dpairs.release();
    doff.release();
    dout.release();
    return rc;
}

int parent_engine_set_entropies(parent_engine* e, int n_dihedrals, const double* s1d, const double* s2d) {
    if (!e || !s1d || !s2d || n_dihedrals < 1) return PARENT_E_INVALID;
    HIPCHK(e, hipSetDevice(e->device));
    const size_t D = n_dofs(n_dihedrals), P = n_pairs(n_dihedrals);
    HIPCHK(e, e->s1d.ensure(D * 8));
    HIPCHK(e, e->s2d.ensure(P * 8));
    HIPCHK(e, hipMemcpyAsync(e->s1d.p, s1d, D * 8, hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemcpyAsync(e->s2d.p, s2d, P * 8, hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->ent_nD = n_dihedrals;
    e->s2d_valid = true;
    e->prepared = false;
    return PARENT_OK;
}

int parent_engine_mist(parent_engine* e, int32_t* edges, double* mi, int32_t* left, double* totals) {
    if (!e || !edges || !mi || !left || !totals) return PARENT_E_INVALID;
    if (!e->s2d_valid) return fail(e, PARENT_E_INVALID, "no 2D entropies on the device");
    HIPCHK(e, hipSetDevice(e->device));
    const int nD = e->ent_nD, V = n_dofs(nD);
    HIPCHK(e, e->mkey.ensure(V * 8));
    HIPCHK(e, e->mupar.ensure(V * 4));
    HIPCHK(e, e->minpos.ensure(V * 4));
    HIPCHK(e, e->mstate.ensure(V * 4));
    HIPCHK(e, e->medges.ensure((size_t)V * 16));
    HIPCHK(e, e->mmi.ensure(V * 8));
    HIPCHK(e, e->mleft.ensure(V * 4));
    HIPCHK(e, e->flags.ensure(64));
    MistArgs a{};
    a.nD = nD;
    a.V = V;
    a.s1d = e->s1d.as<double>();
    a.s2d = e->s2d.as<double>();
    a.key = e->mkey.as<double>();
    a.upar = e->mupar.as<int>();
    a.inpos = e->minpos.as<int>();
    a.state = e->mstate.as<int>();
    a.edges = e->medges.as<int>();
    a.mi = e->mmi.as<double>();
    a.left = e->mleft.as<int>();
    a.flags = e->flags.as<int>();
    int flags = 0;
    HIPCHK(e, hipMemcpyAsync(e->flags.p, &flags, sizeof(int), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipEventRecord(e->ev[6], e->stream));
    HIPCHK(e, launch_mist(a, e->stream));
    HIPCHK(e, hipEventRecord(e->ev[7], e->stream));
    std::vector<double> s1(V);
    HIPCHK(e, hipMemcpyAsync(edges, a.edges, (size_t)(V - 1) * 16, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipMemcpyAsync(mi, a.mi, (size_t)(V - 1) * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipMemcpyAsync(left, a.left, (size_t)(V - 1) * 4, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipMemcpyAsync(s1.data(), a.s1d, (size_t)V * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipMemcpyAsync(&flags, e->flags.p, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->stats.ms_mist = elapsed(e->ev[6], e->ev[7]);
    if (flags & kFlagMist) return fail(e, PARENT_E_DATA, "MIST: no finite mutual information left to pick");
    // totals in the reference's summation order (get_PAR_MIST.cpp:276-284, 663-741, 754-765)
    double t1[3] = {0, 0, 0};
    for (int t = 0; t < 3; t++)
        for (int i = 0; i < count_of_type(nD, t); i++) t1[t] += s1[first_of_type(nD, t) + i];
    double cat[6] = {0, 0, 0, 0, 0, 0}, total_mi = 0;
    for (int s = 0; s < V - 1; s++) {
        const int ta = edges[4 * s], tb = edges[4 * s + 1];
        const int lo = std::min(ta, tb), hi = std::max(ta, tb);
        const int c = lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
        total_mi += mi[s];
        cat[c] += mi[s];
    }
    totals[PARENT_TOT_S1D_B] = t1[0];
    totals[PARENT_TOT_S1D_A] = t1[1];
    totals[PARENT_TOT_S1D_D] = t1[2];
    for (int c = 0; c < 6; c++) totals[PARENT_TOT_MI_BB + c] = cat[c];
    totals[PARENT_TOT_TOTAL] = t1[0] + t1[1] + t1[2] - total_mi;
    return PARENT_OK;
}

int parent_engine_autotune(parent_engine* e, parent_tune_entry* entries, int max_entries, int* n_entries) {
    if (!e) return PARENT_E_INVALID;
    if (int rc = check_ready(e)) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    const long long P = n_pairs(e->nD);
    HIPCHK(e, e->s2d.ensure((size_t)std::max<long long>(P, 1) * 8));
    e->tuned.erase(std::remove_if(e->tuned.begin(), e->tuned.end(), [&](const parent_engine::Tuned& t) {
        return t.b == e->bins[3] && t.a == e->bins[4] && t.d == e->bins[5] && t.N == e->N && t.D == e->D;
    }), e->tuned.end());
    if (int rc = run_autotune(e)) return rc;
    const int n = (int)e->last_tune.size();
    if (n_entries) *n_entries = n;
    for (int k = 0; k < n && k < max_entries && entries; k++) entries[k] = e->last_tune[k];
    return PARENT_OK;
}

int parent_engine_set_variant(parent_engine* e, int method, int variant) {
    if (!e) return PARENT_E_INVALID;
    if (method != PARENT_METHOD_ATOMIC) return fail(e, PARENT_E_INVALID, "only PARENT_METHOD_ATOMIC has variants");
    if (variant < -1 || variant >= atomic_config_count()) return fail(e, PARENT_E_INVALID, "variant out of range");
    e->forced_atomic = variant;
    return PARENT_OK;
}

int parent_engine_variant_count(parent_engine* e, int method, int* count) {
    if (!e || !count) return PARENT_E_INVALID;
    *count = method == PARENT_METHOD_ATOMIC ? atomic_config_count()
                                            : (method == PARENT_METHOD_PLAIN || method == PARENT_METHOD_MFMA ? 1 : 0);
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

}  // extern "C"
