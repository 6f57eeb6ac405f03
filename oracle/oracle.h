/*
 * CPU oracle for the all-pairs entropy / mutual-information path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++ restatement of the
 * reference's algorithm (PARENT.cpp single-rank path, io_binary.cpp formats,
 * util.cpp get_mutual, get_PAR_MIST.cpp Prim) used to check the HIP engine.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load it.  The product library and executables never link it.
 *
 * Every function keeps the reference's loop order and FP64 expression order
 * (built with -ffp-contract=off), so that oracle_parent.x writes a .par that is
 * byte-identical to the reference PARENT.x on the same input.
 *
 * Dof numbering ("global dof id") used throughout: bonds 0..nB-1, then angles
 * nB..nB+nA-1, then dihedrals nB+nA..D-1, with nB = nD+2, nA = nD+1,
 * D = 3*nD+3 (PARENT.cpp:298-299).
 */
#ifndef PARENT_ORACLE_H
#define PARENT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* dof types (util.h:33-35) */
#define ORACLE_B 0
#define ORACLE_A 1
#define ORACLE_D 2

/* Reads a .bat (io_binary.cpp:107-200 header, 281-368 frames).  On success
 * *values is a malloc'd dof-major FP64 array [D][N] (the reference's
 * xxxChunk[dof][frame] layout, PARENT.cpp:346-351, after the FP32->FP64
 * widening of io_binary.cpp:322-366).  Returns 0, or nonzero on failure. */
int oracle_bat_read(const char* path, int* n_dihedrals, int* n_frames,
                    int* double_prec, double** values);
void oracle_free(void* p);

/* Range rule of one dof (PARENT.cpp:511-637).  For dihedrals the values are
 * shifted in place by the modfit rule and *modfit receives the shift; for
 * bonds/angles *modfit = 0.  Returns 0, or 1 for a negative bond
 * (PARENT.cpp:522-525), or 2 if a dihedral falls outside the modfit
 * histogram (index outside [0,100), undefined behaviour in the reference). */
int oracle_dof_range(int type, int n_frames, double* x, double pi,
                     double* vmin, double* vmax, double* modfit);

/* Bin indices int((x-min)/((max-min)/bins)) (PARENT.cpp:654-656 and the 2D
 * fills).  Returns the number of indices outside [0,bins). */
int oracle_bin_index(int n_frames, const double* x, double vmin, double vmax,
                     int bins, int32_t* out);

/* 1D entropy of one dof (PARENT.cpp:650-671 bonds, 694-715 angles,
 * 738-757 dihedrals). */
double oracle_entropy1d(int type, int n_frames, const double* x, double vmin,
                        double vmax, int bins);

/* 2D entropy of one pair; t1 <= t2 in B<A<D order, and for equal types dof 1
 * has the lower index (PARENT.cpp:805-822 BB, 838-855 BA, 870-885 BD,
 * 1158-1175 AA, 1190-1205 AD, 1400-1413 DD).  If counts != NULL the bins1 x
 * bins2 histogram is stored there row-major (dof 1 = row). */
double oracle_entropy2d(int t1, const double* x1, double min1, double max1, int bins1,
                        int t2, const double* x2, double min2, double max2, int bins2,
                        int n_frames, int64_t* counts);

/* Position of pair (gi < gj, global dof ids) in the .par body's 2D section,
 * i.e. offset into the concatenation bb|ba|bd|aa|ad|dd
 * (PARENT.cpp:1106-1115, util.cpp:52-78, io_binary.cpp:606-631). */
int64_t oracle_par_pos(int n_dihedrals, int gi, int gj);

/* Full single-rank computation: ranges, 1D and all 2D entropies.
 * values: dof-major [D][N] FP64, modified in place (dihedral shift).
 * bins: {b1D, a1D, d1D, b2D, a2D, d2D}.  s1d: [D]; s2d: [D(D-1)/2] in .par
 * body order.  vmin/vmax/modfit: [D] (may be NULL).  nthreads <= 0: OpenMP
 * default.  Returns 0 or the first nonzero oracle_dof_range code. */
int oracle_compute(int n_dihedrals, int n_frames, double* values, const int bins[6],
                   double* s1d, double* s2d, double* vmin, double* vmax,
                   double* modfit, int nthreads);

/* The same computation with the reference's OpenMP schedule (rows one after
 * another, each row's partners split statically among the threads, one
 * histogram buffer per thread and row; PARENT.cpp:774-1479 on one rank):
 * the form the CPU baseline times.  Same outputs as oracle_compute; returns
 * 2 when a value falls outside its histogram. */
int oracle_compute_rowwise(int n_dihedrals, int n_frames, double* values, const int bins[6],
                           double* s1d, double* s2d, int nthreads);

/* The same computation restricted to a selection of dofs and a list of
 * pairs among them, for problems too large to run every pair on the host.
 * cols: dof-major [n_sel][N] FP64 values of the selected global dofs
 * dofs[0..n_sel) (shifted in place for dihedrals); pairs: [n_pairs][2]
 * indices into the selection, dofs[a] < dofs[b].  s1d_sel: [n_sel];
 * s2d_pairs: [n_pairs].  Returns 0 or the first nonzero oracle_dof_range
 * code. */
int oracle_compute_subset(int n_dihedrals, int n_frames, int n_sel, const int32_t* dofs, double* cols,
                          const int bins[6], int n_pairs, const int32_t* pairs, double* s1d_sel,
                          double* s2d_pairs, int nthreads);

/* MI of a pair from .par arrays (util.cpp:47-81): (S1D[lo] + S1D[hi]) - S2D. */
double oracle_mutual(int n_dihedrals, const double* s1d, const double* s2d, int gi, int gj);

/* MIST (get_PAR_MIST.cpp:400-766), the reference's O(V^3) scan run on one
 * thread.  edges: (V-1) x 4 ints {type_in, type_out, idx_in, idx_out}
 * (0-based within type), mi: V-1 doubles, left: V-1 ints ("left to process").
 * totals: {S1D_B, S1D_A, S1D_D, MI_BB, MI_BA, MI_BD, MI_AA, MI_AD, MI_DD,
 * total}. */
int oracle_mist(int n_dihedrals, const double* s1d, const double* s2d,
                int32_t* edges, double* mi, int32_t* left, double* totals);

#ifdef __cplusplus
}
#endif

#endif
