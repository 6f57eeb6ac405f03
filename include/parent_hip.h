/*
 * parent_hip.h -- C ABI of the MI355X (gfx950) all-pairs entropy / mutual
 * information engine.
 *
 * The reference (PARENT, markusfleck/PARENT) has no library interface for this
 * path: PARENT.x and get_PAR_MIST.x are monolithic MPI programs behind a
 * process + file boundary (.bat in, .par / MIST text out).  Each entry point
 * below replaces one block of those programs or of their IO library; the
 * comment above it names the reference code it stands for.
 *
 * Conventions
 *   - every function returns int: 0 = PARENT_OK, nonzero = a PARENT_E_* code;
 *     parent_engine_error() / parent_last_error() give the message.
 *   - caller-owned buffers, plain pointers and sizes; no exceptions cross
 *     the boundary.
 *   - one engine owns one HIP stream and its device allocations on one GPU;
 *     an engine is not re-entrant (one thread at a time).
 *   - global dof id: bonds 0..nB-1, angles nB..nB+nA-1, dihedrals
 *     nB+nA..D-1 with nB = nD+2, nA = nD+1, D = 3nD+3 (PARENT.cpp:298-299).
 *   - "frame-major values": value[f*D + g] for frame f and global dof g.
 *   - bins[6] = {bonds1D, angles1D, dihedrals1D, bonds2D, angles2D,
 *     dihedrals2D}, the six counts of PARENT.x's command line
 *     (PARENT.cpp:215-238) and of the .par header (io_binary.cpp:385-396).
 *   - 2D results ("s2d") are D(D-1)/2 doubles in .par body order: the
 *     bonds-bonds half matrix in reverse order, bonds-angles, bonds-dihedrals,
 *     angles-angles (reverse), angles-dihedrals, dihedrals-dihedrals
 *     (reverse) (io_binary.cpp:606-631, PARENT.cpp:1107-1112).
 */
#ifndef PARENT_HIP_H
#define PARENT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PARENT_OK = 0,
    PARENT_E_INVALID = 1,     /* bad argument or corrupt header (reference: return 1 / MPI_Abort(..,1)) */
    PARENT_E_NOMEM = 2,       /* host or device allocation failed (read_PAR_body returns 2) */
    PARENT_E_IO = 3,          /* file cannot be opened, read or written */
    PARENT_E_DATA = 4,        /* a bond length < 0 (PARENT.cpp:522-525) */
    PARENT_E_RANGE = 5,       /* a value falls outside its histogram (reference: out-of-bounds write) */
    PARENT_E_DEVICE = 6,      /* HIP runtime error */
    PARENT_E_UNSUPPORTED = 7  /* request outside what this library implements (e.g. > 4096 bins) */
};

/* pair-kernel selection for parent_engine_pairs() */
enum {
    PARENT_METHOD_PLAIN = 0,  /* one dof pair per workgroup, LDS atomic histogram, per-cell FP64 epilogue */
    PARENT_METHOD_ATOMIC = 1, /* tiled: many pairs per workgroup, LDS atomic histograms, fused n*ln(n) epilogue */
    PARENT_METHOD_MFMA = 2,   /* one-hot int8 x int8 -> int32 MFMA contraction, fused epilogue */
    PARENT_METHOD_AUTO = 3,   /* autotuned choice per bin count */
    PARENT_METHOD_SLABS = 4   /* one dof pair per workgroup, the histogram in row slabs that fit LDS:
                                 every bin count up to 4096; the kernel any request runs on when the
                                 2D histogram does not fit LDS or a 2D bin count exceeds 255 */
};

/* MIST totals, in the order get_PAR_MIST.x prints them (get_PAR_MIST.cpp:754-765) */
enum {
    PARENT_TOT_S1D_B = 0, PARENT_TOT_S1D_A, PARENT_TOT_S1D_D,
    PARENT_TOT_MI_BB, PARENT_TOT_MI_BA, PARENT_TOT_MI_BD,
    PARENT_TOT_MI_AA, PARENT_TOT_MI_AD, PARENT_TOT_MI_DD,
    PARENT_TOT_TOTAL, PARENT_TOT_COUNT
};

/* ------------------------------------------------------------------ */
/* .bat / .par files (host)                                            */
/* ------------------------------------------------------------------ */

typedef struct parent_bat parent_bat;

/* Opens a .bat and reads its header.  Replaces read_BAT_header
 * (io_binary.cpp:107-200), with the same validation: version >= 0,
 * double_prec in {0,1}, nDihedrals >= 0, numFrames >= 1. */
int parent_bat_open(const char* path, parent_bat** out);

/* Header fields of an opened .bat. */
int parent_bat_info(const parent_bat* bat, int* n_dihedrals, int* n_frames, int* double_prec);

/* Reads all frames into frame-major values (float if double_prec == 0,
 * double otherwise), D values per frame in global dof order.  Replaces the
 * read_BAT_frame loop and the per-frame scatter (io_binary.cpp:281-368,
 * PARENT.cpp:443-472).  Fails with PARENT_E_IO on a truncated file. */
int parent_bat_read_frames(parent_bat* bat, void* values);

/* Bytes of one frame record: 11 floats (time, xtcPrec, box), then 6
 * external coordinates and 3 nD + 3 values in the file's precision
 * (io_binary.cpp:281-368). */
int64_t parent_bat_record_bytes(const parent_bat* bat);

/* Copies frame records [first_frame, first_frame + n_frames) verbatim into
 * `records` (n_frames * parent_bat_record_bytes bytes; pinned memory lets the
 * caller DMA it straight to the device).  The streaming half of
 * read_BAT_frame (io_binary.cpp:281-368): the values are not reordered. */
int parent_bat_read_raw(parent_bat* bat, int first_frame, int n_frames, void* records);

void parent_bat_close(parent_bat* bat);

/* Writes a .par v4: header copied from the .bat's (write_PAR_header,
 * io_binary.cpp:371-442) then the body (write_PAR_body, io_binary.cpp:606-631):
 * s1d[D] then s2d[D(D-1)/2]. */
int parent_par_write(const char* path, const parent_bat* header, const int bins[6],
                     const double* s1d, const double* s2d);

/* Reads a .par: header fields and body arrays.  Replaces read_PAR_header +
 * read_PAR_body (io_binary.cpp:446-673), with the same validation.  s1d/s2d
 * may be NULL to query sizes; otherwise they hold D and D(D-1)/2 doubles. */
int parent_par_read(const char* path, int* n_dihedrals, int* n_frames, int bins[6],
                    double* s1d, double* s2d);

/* Writes the MIE listing of a .par body, byte-identical to the stdout of
 * the reference's get_PAR_MIE.x (get_PAR_MIE.cpp:54-199): the 1D entropies,
 * one "<t1>-<t2> i  j   S2D   MI" line per pair, then the 16 totals, all
 * at precision 12.  Replaces that program's listing and summing loops;
 * MI follows get_mutual (util.cpp:47-81).
 *   path      output file, "-" for stdout, or NULL to compute totals only
 *   s1d, s2d  the .par body (parent_par_read layout)
 *   nthreads  formatting threads (<= 0: OpenMP default)
 *   totals    NULL or 16 doubles, in the order of the printed TOTAL lines
 *             (each a sequential FP64 sum in listing order) */
int parent_mie_write(const char* path, int n_dihedrals, const double* s1d, const double* s2d,
                     int nthreads, double* totals);

/* Builds a .bat from a GROMACS topology and a compressed (.xtc) trajectory:
 * the .xtc -> .bat direction of the reference's BAT_builder.x (bat.cpp:82-176;
 * topology.cpp, BAT_topology.cpp, BAT_trajectory.cpp:77-299, io_binary.cpp:41-278),
 * with the same dihedral tree, phase angles and bytes.
 *   backbone_names  blank-separated atom names for the phase-angle rule (-bb)
 *   double_prec     1: FP64 values (the reference's default), 0: FP32
 *   nthreads        threads that decode and convert frames (<= 0: OpenMP default)
 *   n_frames        NULL or the number of frames written */
int parent_bat_build(const char* top_path, const char* xtc_path, const char* bat_path,
                     const char* backbone_names, int double_prec, int nthreads, int* n_frames);

/* Message of the last failed parent_bat_* / parent_par_* / parent_synth_* /
 * parent_mie_* call on this thread. */
const char* parent_last_error(void);

/* ------------------------------------------------------------------ */
/* engine (device)                                                     */
/* ------------------------------------------------------------------ */

typedef struct parent_engine parent_engine;

typedef struct {
    double ms_upload;     /* H2D of the trajectory (0 if it was already on the device) */
    double ms_prepare;    /* ranges + modfit + pre-binning + 1D entropies */
    double ms_pairs;      /* all 2D histograms + entropy epilogue */
    double ms_mist;       /* device Prim */
    double pair_updates;  /* sum over pairs of frames */
    int method;           /* PARENT_METHOD_* actually used by the last parent_engine_pairs */
    int kernel_variant;   /* tile configuration id chosen for that method */
    double ms_sweep;      /* last parent_engine_sweep: the graph launch (prepare + pairs of every set) */
    int sweep_sets;       /* number of bin sets in that sweep */
    int mist_path;        /* last parent_engine_mist: 1 the spanning-tree path, 0 the O(V^2) Prim */
    int batch_launches;   /* last parent_batch_run: pair kernel launches (tile queues; one per
                             trajectory on the slab kernel) */
} parent_stats;

/* One row of the autotuner's table. */
typedef struct {
    int method;            /* PARENT_METHOD_* */
    int variant;           /* configuration id within the method */
    int tile_rows, tile_cols, block;
    double ms;             /* kernel time on the tuning sample */
    double updates;        /* pair-frame updates in the sample */
    double updates_per_s;
} parent_tune_entry;

/* Creates an engine on HIP device `device`.  No counterpart in the
 * reference (MPI_Init + buffers, PARENT.cpp:131-141, 300-440). */
int parent_engine_create(int device, parent_engine** out);
void parent_engine_destroy(parent_engine* eng);
const char* parent_engine_error(const parent_engine* eng);

/* Hands the engine a trajectory.  values: frame-major, float when
 * double_prec == 0 else double.  on_device == 0: host pointer, copied to the
 * device now; on_device == 1: device pointer the engine borrows until the
 * next set_trajectory/destroy.  Replaces the read-in + scatter of
 * PARENT.cpp:443-493 (the data then lives in HBM, not per-rank RAM). */
/* Streams the frames of an opened .bat into the engine: the file is read in
 * ~32 MB chunks of raw frame records into pinned buffers, each chunk is
 * DMA'd to the device while the next one is read, and a kernel reorders the
 * records into frame-major values in HBM.  Same result as
 * parent_bat_read_frames + parent_engine_set_trajectory(on_device = 0),
 * without the host copy of the trajectory.  Replaces the read-in loop of
 * PARENT.cpp:443-493 (read_BAT_frame, io_binary.cpp:281-368).
 * parent_stats.ms_upload reports the whole load (file read included). */
int parent_engine_load_bat(parent_engine* eng, parent_bat* bat);

int parent_engine_set_trajectory(parent_engine* eng, int n_dihedrals, int n_frames,
                                 int double_prec, const void* values, int on_device);

/* Extrema, dihedral modfit, bin indices and 1D entropies: PARENT.cpp:503-772
 * (rules at 511-637, 1D at 644-772).  Bins must be in [1, 4096] (the
 * reference takes any positive count, PARENT.cpp:211-238); indices are
 * stored as uint8 while a matrix's counts are <= 255, else as uint16. */
int parent_engine_prepare(parent_engine* eng, const int bins[6]);

/* All 2D entropies S_ij for the bins of the last prepare:
 * PARENT.cpp:774-1479 (histogram fill at 802-804 and its 14 twins, epilogues
 * at 805-822, 838-855, 870-885, 1158-1175, 1190-1205, 1400-1413). */
int parent_engine_pairs(parent_engine* eng, int method);

/* Bin-count sweep over the resident trajectory: n_sets bin sets (bins: n_sets
 * x 6 ints, each as for parent_engine_prepare), each pre-binned and paired
 * with `method` (PARENT_METHOD_AUTO: the autotuner's choice per set, so small
 * bin counts run on the MFMA kernel and large ones on the atomic kernels).
 * The first call per (bin sets, method, trajectory) prepares every set,
 * tunes, uploads the tile lists and captures all prepare and pair kernels of
 * all sets into one hipGraph; every call launches that graph once
 * (stats.ms_sweep).  Afterwards parent_engine_select(eng, k) makes set k the
 * one that the read, pair_counts and mist calls use.  The reference runs
 * PARENT.x once per bin count (run.sh:40); no counterpart of the batching. */
int parent_engine_sweep(parent_engine* eng, int n_sets, const int* bins, int method);

/* Selects the bin set (0 after prepare; 0..n_sets-1 after a sweep) that the
 * read_*, pair_counts, mist and autotune calls act on; stats.method and
 * stats.kernel_variant then name the pair kernel that set ran on. */
int parent_engine_select(parent_engine* eng, int set);

/* Copies results to the host (synchronises the engine's stream). */
int parent_engine_read_s1d(parent_engine* eng, double* s1d /* D */);
int parent_engine_read_s2d(parent_engine* eng, double* s2d /* D(D-1)/2 */);

/* Per-dof ranges after prepare: min, max (FP64, PARENT.cpp:528-636) and the
 * dihedral shift modFit (PARENT.cpp:609; 0 for bonds/angles). */
int parent_engine_read_ranges(parent_engine* eng, double* vmin, double* vmax, double* modfit);

/* The 2D bin indices, dof-major: out[g*n_frames + f] (PARENT.cpp:803 index
 * expressions, here computed once per dof and frame).  PARENT_E_UNSUPPORTED
 * when a 2D bin count exceeds 255: then use parent_engine_read_bins16. */
int parent_engine_read_bins(parent_engine* eng, uint8_t* out);

/* The same for any bin count, as uint16. */
int parent_engine_read_bins16(parent_engine* eng, uint16_t* out);

/* Histogram counts of listed pairs through `method`'s fill path, without the
 * epilogue: pairs[2k], pairs[2k+1] = (gi < gj); counts of pair k are
 * bins(gi) x bins(gj) int32, row = gi, concatenated.  For parity checks of
 * the long long histo[][] of PARENT.cpp:790-804. */
int parent_engine_pair_counts(parent_engine* eng, int method, int n_pairs, const int32_t* pairs,
                              int32_t* counts);

/* MIST on the engine's current S1D/S2D (get_PAR_MIST.cpp:396-766), as an
 * O(V^2) device Prim that reproduces the reference's single-thread scan order
 * and tie rule.  edges: (D-1) x 4 int32 {type_in, type_out, index_in,
 * index_out} (0-based within type), mi: D-1 doubles, left: D-1 int32 (the
 * "left to process" count printed per line), totals: PARENT_TOT_COUNT
 * doubles. */
int parent_engine_mist(parent_engine* eng, int32_t* edges, double* mi, int32_t* left,
                       double* totals);

/* MIST of entropies already in device memory, allocating nothing: the same
 * Prim as parent_engine_mist (edges, MI bits, left counts and totals are
 * identical), with its scratch in a static device array of 1.25 MiB per GPU
 * instead of engine buffers.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).
 *   s1d [D], s2d [D(D-1)/2] (.par body order): inputs in device memory
 *   edges [(D-1) x 4], mi [D-1], left [D-1], totals [PARENT_TOT_COUNT],
 *   status [1]: outputs in device memory; status = 0, or PARENT_E_DATA when
 *   no finite MI was left to pick (the edges are then incomplete)
 *   stream: a hipStream_t of the engine's device, or NULL for the engine's
 * The kernels are enqueued on the stream and the call returns without
 * waiting for them (nothing synchronises).  D may be at most 65,536
 * (PARENT_E_UNSUPPORTED above).  Every call on one GPU shares that GPU's
 * scratch array: calls that overlap on one GPU, from any engine, race. */
int parent_engine_mist_noalloc(parent_engine* eng, int n_dihedrals, const double* s1d, const double* s2d,
                               int32_t* edges, double* mi, int32_t* left, double* totals, int32_t* status,
                               void* stream);

/* Loads S1D/S2D from host arrays (e.g. a .par read with parent_par_read) so
 * parent_engine_mist can run without a trajectory (the get_PAR_MIST.x path,
 * get_PAR_MIST.cpp:211-327). */
int parent_engine_set_entropies(parent_engine* eng, int n_dihedrals, const double* s1d,
                                const double* s2d);

/* Times every pair-kernel configuration that fits the device on one run of
 * row dofs against all their partners (12 rows, more when the configuration's
 * tiles are large, so that the sample holds 16 tiles per compute unit),
 * re-times the 3 fastest on 32 such runs spread over the dofs (at least 32
 * tiles per compute unit) and keeps the fastest of those for
 * PARENT_METHOD_AUTO on this (bins, frames, dofs).  Rows of the table carry
 * their own sample's updates; the 3 finalists' rows are the 32-run ones.
 * PARENT_METHOD_AUTO tunes on first use by itself; this call re-tunes and
 * returns the table (up to max_entries rows; *n_entries = rows measured).
 * No counterpart in the reference (its single loop nest has no variants). */
int parent_engine_autotune(parent_engine* eng, parent_tune_entry* entries, int max_entries,
                           int* n_entries);

/* Pins the configuration `method` uses when called explicitly (not through
 * PARENT_METHOD_AUTO): variant in [0, count) or -1 for the default.  A
 * variant whose histograms do not fit the bins makes the method fall back to
 * PARENT_METHOD_PLAIN.  For PARENT_METHOD_SLABS the variant caps the
 * histogram rows per slab (0 or -1: as many as the LDS holds), which
 * exercises several slabs at small bin counts.  For PARENT_METHOD_MFMA,
 * variant 0 multiplies int8 one-hot operands and variant 1 FP4 ones (FP32
 * sums; 8 to 12 bins and fewer than 2^24 frames, else variant 0 runs).  For
 * testing every configuration the autotuner may pick; no counterpart in the
 * reference. */
int parent_engine_set_variant(parent_engine* eng, int method, int variant);
/* How parent_engine_mist finds the tree: 0 (default) the spanning-tree path
 * where it applies (dense MI matrix fits, V <= 16,384, no equal candidates
 * in any Borůvka choice, every MI finite and > -1e200), else the O(V^2)
 * Prim; 1 always the O(V^2) Prim.  Both give the same result.  For testing;
 * no counterpart in the reference. */
int parent_engine_set_mist_path(parent_engine* eng, int mode);
int parent_engine_variant_count(parent_engine* eng, int method, int* count);

/* Waits for the engine's stream. */
int parent_engine_sync(parent_engine* eng);

/* Phase timings of the last calls (HIP events on the engine's stream);
 * replaces the gettimeofday stopwatches of PARENT.cpp:95-97,498,1489,1501-1505. */
int parent_engine_stats(parent_engine* eng, parent_stats* out);

/* ------------------------------------------------------------------ */
/* batch of independent trajectories (device)                          */
/* ------------------------------------------------------------------ */

/* Many trajectories of uneven dof and frame counts with one set of bins.
 * The reference runs PARENT.x once per trajectory (run.sh:40); here each
 * trajectory is pre-binned as it is added (PARENT.cpp:503-772, its values
 * are then dropped from the device) and parent_batch_run computes every pair
 * of every trajectory (PARENT.cpp:774-1479) in one launch over a merged tile
 * queue ordered by cost.  No counterpart of the batching in the reference. */
typedef struct parent_batch parent_batch;

int parent_batch_create(int device, const int bins[6], parent_batch** out);
void parent_batch_destroy(parent_batch* b);
const char* parent_batch_error(const parent_batch* b);

/* Adds one trajectory (frame-major values as for
 * parent_engine_set_trajectory; a device pointer is only read during this
 * call) and returns its index.  Fails like parent_engine_prepare on bad data. */
int parent_batch_add(parent_batch* b, int n_dihedrals, int n_frames, int double_prec,
                     const void* values, int on_device, int* index);

/* The same with this trajectory's own bins (as for parent_engine_prepare):
 * trajectories of different bin counts share the one launch of
 * parent_batch_run, whose kernel is chosen to hold the largest counts of the
 * batch (each tile reads its trajectory's counts).  The reference runs
 * PARENT.x once per trajectory and bin setting (run.sh:40). */
int parent_batch_add_binned(parent_batch* b, const int bins[6], int n_dihedrals, int n_frames, int double_prec,
                            const void* values, int on_device, int* index);

/* All pairs of all added trajectories over merged tile queues (method as
 * for parent_engine_pairs; the tile shape is chosen for the batch's largest
 * bin counts; PARENT_METHOD_PLAIN runs the tiled atomic kernel, as the plain
 * kernel has no tile queue).  One launch, except that PARENT_METHOD_AUTO may
 * split the batch into two frame-count classes (see
 * parent_batch_set_classes), each with its own configuration and launch.
 * stats.ms_pairs covers the launches, stats.batch_launches counts them and
 * stats.method / kernel_variant name the configuration of the one with the
 * most updates. */
int parent_batch_run(parent_batch* b, int method);
/* Frame-count classes of PARENT_METHOD_AUTO: trajectories of fewer than 20
 * frames per cell of the batch's largest 2D histogram (18,000 at 30 bins)
 * against the rest, each class tuned apart and run as its own launch.
 * mode 1 (default): split when the two classes' estimated times add up to
 * less than one merged launch's; 2: whenever both classes are non-empty;
 * 0: never.  The next PARENT_METHOD_AUTO run tunes again.  No
 * counterpart in the reference (it runs one trajectory per process). */
int parent_batch_set_classes(parent_batch* b, int mode);
/* parent_engine_set_variant for the batch's kernels (method
 * PARENT_METHOD_ATOMIC, PARENT_METHOD_MFMA or PARENT_METHOD_SLABS): pins the
 * configuration that parent_batch_run(b, method) uses.  For testing; no
 * counterpart in the reference. */
int parent_batch_set_variant(parent_batch* b, int method, int variant);

/* The batch autotuner (what PARENT_METHOD_AUTO runs on first use): every
 * configuration that fits is timed on the tuning rows of the longest and the
 * shortest trajectory, the two timings fit ms = a * tiles + b * updates for
 * the whole merged queue, the 3 lowest predictions are timed again on a
 * sample of that queue (64 contiguous runs, 1/64 of its tiles), and the
 * fastest there is kept.  The same is done for each frame-count class
 * (parent_batch_set_classes) to decide on a split.  Returns the merged
 * queue's table (entries[k].ms = predicted ms; for the 3 finalists, the
 * estimate from the queue sample). */
int parent_batch_autotune(parent_batch* b, parent_tune_entry* entries, int max_entries, int* n_entries);

/* S1D / S2D of trajectory `index` (either pointer may be NULL). */
int parent_batch_read(parent_batch* b, int index, double* s1d, double* s2d);

/* MIST of trajectory `index`, outputs as parent_engine_mist. */
int parent_batch_mist(parent_batch* b, int index, int32_t* edges, double* mi, int32_t* left,
                      double* totals);

int parent_batch_stats(parent_batch* b, parent_stats* out);

/* ------------------------------------------------------------------ */
/* synthetic trajectories (no counterpart in the reference: test and   */
/* benchmark input generator)                                          */
/* ------------------------------------------------------------------ */

/* Fills frame-major float values [n_frames][3*atoms-6] of a linear chain of
 * `atoms` atoms: Gaussian bonds, wrapped-normal angles, von Mises mixture
 * dihedrals in [0, 2*pi), 20% of dofs sharing one of 10 latent factors.
 * Deterministic in (atoms, n_frames, seed) for any thread count. */
int parent_synth_values(int atoms, int n_frames, uint64_t seed, float* values, int nthreads);

/* Writes the same trajectory as a single-precision .bat v3 file. */
int parent_synth_write_bat(const char* path, int atoms, int n_frames, uint64_t seed, int nthreads);

#ifdef __cplusplus
}
#endif

#endif /* PARENT_HIP_H */
