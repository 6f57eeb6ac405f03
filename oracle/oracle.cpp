// CPU oracle: restatement of the reference's entropy path.  TEST INFRASTRUCTURE
// ONLY (see oracle.h).  Build with -ffp-contract=off: every FP64 expression
// below is written in the reference's evaluation order so results are
// bit-identical to the reference binaries built without -march.
#include "oracle.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <array>
#include <vector>
#include <omp.h>

namespace {

const int kModfitBins = 100;  // PARENT.cpp MODFITNBINS

int n_of_type(int nD, int t) { return t == ORACLE_B ? nD + 2 : (t == ORACLE_A ? nD + 1 : nD); }
int first_of_type(int nD, int t) { return t == ORACLE_B ? 0 : (t == ORACLE_A ? nD + 2 : 2 * nD + 3); }
int type_of(int nD, int g) { return g < nD + 2 ? ORACLE_B : (g < 2 * nD + 3 ? ORACLE_A : ORACLE_D); }

template <class T>
bool rd(std::ifstream& f, T* v) {
    f.read(reinterpret_cast<char*>(v), sizeof(T));
    return !f.fail();
}

}  // namespace

extern "C" {

void oracle_free(void* p) { free(p); }

// io_binary.cpp:107-200 (header) and 281-368 (frames); PARENT.cpp:443-472
// stores each frame's values dof-major.
int oracle_bat_read(const char* path, int* n_dihedrals, int* n_frames, int* double_prec,
                    double** values) {
    std::ifstream f(path, std::ios::binary | std::ios::in);
    if (!f.is_open()) return 1;
    int version, dp, nD, N;
    if (!rd(f, &version) || !rd(f, &dp) || !rd(f, &nD) || !rd(f, &N)) return 1;
    if (version < 0 || (dp != 0 && dp != 1) || nD < 0 || N < 1) return 1;
    if (version >= 3) f.seekg((std::streamoff)(nD + 3) * (8 + 4 + 8 + 31), std::ios::cur);
    f.seekg((std::streamoff)nD * 6 * 4 + (std::streamoff)(nD + 3) * 4, std::ios::cur);
    if (f.fail()) return 1;
    const int nB = nD + 2, nA = nD + 1;
    const size_t D = (size_t)3 * nD + 3;
    double* v = (double*)malloc(D * (size_t)N * sizeof(double));
    if (!v) return 2;
    double* b = v;
    double* a = v + (size_t)nB * N;
    double* d = v + (size_t)(nB + nA) * N;
    for (int fr = 0; fr < N; fr++) {
        float hdr[11];
        f.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
        if (dp == 1) {
            double ext[6], x;
            f.read(reinterpret_cast<char*>(ext), sizeof(ext));
            rd(f, &x); b[(size_t)0 * N + fr] = x;
            rd(f, &x); b[(size_t)1 * N + fr] = x;
            rd(f, &x); a[(size_t)0 * N + fr] = x;
            for (int i = 0; i < nD; i++) {
                rd(f, &x); b[(size_t)(i + 2) * N + fr] = x;
                rd(f, &x); a[(size_t)(i + 1) * N + fr] = x;
                rd(f, &x); d[(size_t)i * N + fr] = x;
            }
        } else {
            float ext[6], x;
            f.read(reinterpret_cast<char*>(ext), sizeof(ext));
            rd(f, &x); b[(size_t)0 * N + fr] = x;
            rd(f, &x); b[(size_t)1 * N + fr] = x;
            rd(f, &x); a[(size_t)0 * N + fr] = x;
            for (int i = 0; i < nD; i++) {
                rd(f, &x); b[(size_t)(i + 2) * N + fr] = x;
                rd(f, &x); a[(size_t)(i + 1) * N + fr] = x;
                rd(f, &x); d[(size_t)i * N + fr] = x;
            }
        }
        if (f.fail()) {
            free(v);
            return 1;
        }
    }
    *n_dihedrals = nD;
    *n_frames = N;
    *double_prec = dp;
    *values = v;
    return 0;
}

// PARENT.cpp:511-533 (bonds), 536-562 (angles), 564-637 (dihedrals).
int oracle_dof_range(int type, int N, double* x, double pi, double* vmin, double* vmax,
                     double* modfit) {
    double tmpMin, tmpMax;
    *modfit = 0.0;
    if (type == ORACLE_D) {
        long long histo[kModfitBins];
        for (int k = 0; k < kModfitBins; k++) histo[k] = 0;
        double binsize = (2 * pi + 0.000000005) / kModfitBins;
        for (int i = 0; i < N; i++) {
            int k = int((x[i]) / binsize);
            if (k < 0 || k >= kModfitBins) return 2;  // reference: out-of-bounds write
            histo[k] += 1;
        }
        bool zeroExists = false;
        for (int k = 0; k < kModfitBins; k++) zeroExists = zeroExists || (histo[k] == 0);
        long long longest;
        int longestPos, current, currentPos = 0;
        if (zeroExists) {
            longest = 0;
            current = 0;
            longestPos = -1;
            for (int k = 0; k < 2 * kModfitBins; k++) {
                int l = k % kModfitBins;
                if ((current == 0) && (histo[l] == 0)) {
                    current = 1;
                    currentPos = k;
                }
                if ((current > 0) && (histo[l] == 0)) current += 1;
                if ((current > 0) && (histo[l] != 0)) {
                    if (current > longest) {
                        longest = current;
                        longestPos = currentPos;
                    }
                    current = 0;
                }
            }
        } else {
            longestPos = 0;
            longest = histo[0];
            for (int k = 0; k < kModfitBins; k++) {
                if (histo[k] < longest) {
                    longest = histo[k];
                    longestPos = k;
                }
            }
        }
        double modFit = 2 * pi - (longestPos + 0.5) * binsize;
        for (int k = 0; k < N; k++) x[k] = x[k] + modFit - 2 * pi * int((x[k] + modFit) / (2 * pi));
        *modfit = modFit;
    }
    tmpMax = -1.0e100;
    tmpMin = 1.0e100;
    for (int i = 0; i < N; i++) {
        if (x[i] > tmpMax) tmpMax = x[i];
        if (x[i] < tmpMin) tmpMin = x[i];
    }
    if (type == ORACLE_B) {
        if ((tmpMin < 0.0) || (tmpMax < 0.0)) return 1;
        tmpMin -= 0.000000005;
        tmpMax += 0.000000005;
        if (tmpMin < 0.0) tmpMin = 0.0;
        if ((tmpMax - tmpMin) < 1.0e-4) tmpMax += 0.05;
    } else if (type == ORACLE_A) {
        tmpMin -= 0.000000005;
        tmpMax += 0.000000005;
        if (tmpMin < 0) tmpMin = 0;
        if (tmpMax > pi) tmpMax = pi;
        if ((tmpMax - tmpMin) < 1.0e-4) tmpMax += 0.05;
    } else {
        tmpMin -= 0.000000005;
        tmpMax += 0.000000005;
        if (tmpMin < 0.0) tmpMin = 0.0;
        if (tmpMax > 2 * pi) tmpMax = 2 * pi;
        if ((tmpMax - tmpMin) < 1.0e-4) tmpMax += 0.05;
    }
    *vmin = tmpMin;
    *vmax = tmpMax;
    return 0;
}

int oracle_bin_index(int N, const double* x, double vmin, double vmax, int bins, int32_t* out) {
    double binsize = (vmax - vmin) / bins;
    int bad = 0;
    for (int i = 0; i < N; i++) {
        int k = int((x[i] - vmin) / binsize);
        if (k < 0 || k >= bins) bad++;
        out[i] = k;
    }
    return bad;
}

// PARENT.cpp:650-671, 694-715, 738-757.
double oracle_entropy1d(int type, int numFrames, const double* x, double vmin, double vmax,
                        int bins) {
    std::vector<long long> hv(bins, 0);
    long long* histo = hv.data();  // plain pointer: unoptimised builds index it like the reference's array
    double binsize = (vmax - vmin) / (bins);
    for (int i = 0; i < numFrames; i++) {
        int k = int((x[i] - vmin) / binsize);
        if (k < 0 || k >= bins) return NAN;  // reference: out-of-bounds write
        histo[k] += 1;
    }
    int occupbins = 0;
    double plnpsum = 0, probDens;
    if (type == ORACLE_B) {
        double blen = vmin + (binsize / 2.0);
        for (int k = 0; k < bins; k++) {
            probDens = histo[k] / (numFrames * binsize * blen * blen);
            if (probDens > 0) {
                plnpsum = plnpsum + blen * blen * probDens * log(probDens);
                occupbins = occupbins + 1;
            }
            blen += binsize;
        }
    } else if (type == ORACLE_A) {
        double theta = vmin + (binsize / 2.0);
        for (int k = 0; k < bins; k++) {
            probDens = histo[k] / (numFrames * binsize * sin(theta));
            if (probDens > 0) {
                plnpsum = plnpsum + sin(theta) * probDens * log(probDens);
                occupbins = occupbins + 1;
            }
            theta += binsize;
        }
    } else {
        for (int k = 0; k < bins; k++) {
            probDens = histo[k] / (numFrames * binsize);
            if (probDens > 0) {
                plnpsum = plnpsum + probDens * log(probDens);
                occupbins = occupbins + 1;
            }
        }
    }
    plnpsum = -plnpsum * binsize;
    return plnpsum + (occupbins - 1.0) / (2.0 * numFrames);
}

// The six 2D epilogues of PARENT.cpp (rank==i variants, the ones a single
// rank executes): BB 805-822, BA 838-855, BD 870-885, AA 1158-1175,
// AD 1190-1205, DD 1400-1413.
}  // extern "C"

namespace {

// The 2D entropy of the pair (row dof a, column dof b) of the dof-major
// value rows `rows` with ranges mins/maxs, into a caller's histogram buffer
// of bins1 rows of `stride` >= bins2 long longs, zeroed here over
// [bins1][bins2] as the reference zeroes its per-thread long long
// histo[B][biggest] before each pair (PARENT.cpp:790-797).  The fill indexes
// rows[a][i] and mins[a] like the reference's chunk[j][iOmp] and min[j]
// (PARENT.cpp:803), so that an unoptimised build (the reference's shipped
// flags) runs the same memory accesses as the reference's.  CHECK: a value
// outside its histogram gives NaN (the reference writes out of bounds);
// without it the caller has checked every index beforehand.
template <bool CHECK>
double entropy2d_core(int t1, const double* const* rows, const double* mins, const double* maxs, int a,
                      int bins1, int t2, int b, int bins2, int numFrames, long long* h, int stride,
                      int64_t* counts) {
    for (int k = 0; k < bins1; k++)
        for (int l = 0; l < bins2; l++) h[(size_t)k * stride + l] = 0;
    const double min1 = mins[a], min2 = mins[b];
    double binsize1 = (maxs[a] - mins[a]) / bins1;
    double binsize2 = (maxs[b] - mins[b]) / bins2;
    for (int i = 0; i < numFrames; i++) {
        if (CHECK) {
            int k = int((rows[a][i] - mins[a]) / binsize1), l = int((rows[b][i] - mins[b]) / binsize2);
            if (k < 0 || k >= bins1 || l < 0 || l >= bins2) return NAN;  // reference: out-of-bounds write
            h[(size_t)k * stride + l] += 1;
        } else {
            h[(size_t)int((rows[a][i] - mins[a]) / binsize1) * stride + int((rows[b][i] - mins[b]) / binsize2)] += 1;
        }
    }
    if (counts)
        for (int k = 0; k < bins1; k++)
            for (int l = 0; l < bins2; l++) counts[(size_t)k * bins2 + l] = h[(size_t)k * stride + l];
    int occupbins = 0;
    double plnpsum = 0, sum1, probDens;
#define H(k, l) ((double)h[(size_t)(k) * stride + (l)])
    if (t1 == ORACLE_B && t2 == ORACLE_B) {
        double blen1 = min1 + binsize1 / 2.0, blen2;
        for (int k = 0; k < bins1; k++) {
            sum1 = 0.0;
            blen2 = min2 + binsize2 / 2.0;
            for (int l = 0; l < bins2; l++) {
                probDens = H(k, l) / (numFrames * binsize1 * binsize2 * blen1 * blen1 * blen2 * blen2);
                if (probDens > 0) {
                    sum1 = sum1 + blen1 * blen1 * blen2 * blen2 * probDens * log(probDens) * binsize1 * binsize2;
                    occupbins = occupbins + 1;
                }
                blen2 = blen2 + binsize2;
            }
            plnpsum = plnpsum + sum1;
            blen1 = blen1 + binsize1;
        }
    } else if (t1 == ORACLE_B && t2 == ORACLE_A) {
        double blen1 = min1 + binsize1 / 2.0, theta;
        for (int k = 0; k < bins1; k++) {
            sum1 = 0.0;
            theta = min2 + binsize2 / 2.0;
            for (int l = 0; l < bins2; l++) {
                probDens = H(k, l) / (numFrames * binsize1 * binsize2 * blen1 * blen1 * sin(theta));
                if (probDens > 0) {
                    sum1 = sum1 + blen1 * blen1 * sin(theta) * probDens * log(probDens) * binsize1 * binsize2;
                    occupbins = occupbins + 1;
                }
                theta = theta + binsize2;
            }
            plnpsum = plnpsum + sum1;
            blen1 = blen1 + binsize1;
        }
    } else if (t1 == ORACLE_B && t2 == ORACLE_D) {
        double blen1 = min1 + binsize1 / 2.0;
        for (int k = 0; k < bins1; k++) {
            sum1 = 0.0;
            for (int l = 0; l < bins2; l++) {
                probDens = H(k, l) / (numFrames * binsize1 * binsize2 * blen1 * blen1);
                if (probDens > 0) {
                    sum1 = sum1 + blen1 * blen1 * probDens * log(probDens) * binsize1 * binsize2;
                    occupbins = occupbins + 1;
                }
            }
            plnpsum = plnpsum + sum1;
            blen1 = blen1 + binsize1;
        }
    } else if (t1 == ORACLE_A && t2 == ORACLE_A) {
        double theta1 = min1 + binsize1 / 2.0, theta2;
        for (int k = 0; k < bins1; k++) {
            sum1 = 0.0;
            theta2 = min2 + binsize2 / 2.0;
            for (int l = 0; l < bins2; l++) {
                probDens = H(k, l) / (numFrames * binsize1 * binsize2 * sin(theta1) * sin(theta2));
                if (probDens > 0) {
                    sum1 = sum1 + sin(theta1) * sin(theta2) * probDens * log(probDens) * binsize1 * binsize2;
                    occupbins = occupbins + 1;
                }
                theta2 = theta2 + binsize2;
            }
            plnpsum = plnpsum + sum1;
            theta1 = theta1 + binsize1;
        }
    } else if (t1 == ORACLE_A && t2 == ORACLE_D) {
        double theta1 = min1 + binsize1 / 2.0;
        for (int k = 0; k < bins1; k++) {
            sum1 = 0.0;
            for (int l = 0; l < bins2; l++) {
                probDens = H(k, l) / (numFrames * binsize1 * binsize2 * sin(theta1));
                if (probDens > 0) {
                    sum1 = sum1 + sin(theta1) * probDens * log(probDens) * binsize1 * binsize2;
                    occupbins = occupbins + 1;
                }
            }
            plnpsum = plnpsum + sum1;
            theta1 = theta1 + binsize1;
        }
    } else {  // DD
        for (int k = 0; k < bins1; k++) {
            sum1 = 0.0;
            for (int l = 0; l < bins2; l++) {
                probDens = H(k, l) / (numFrames * binsize1 * binsize2);
                if (probDens > 0) {
                    sum1 = sum1 + probDens * log(probDens) * binsize1 * binsize2;
                    occupbins = occupbins + 1;
                }
            }
            plnpsum = plnpsum + sum1;
        }
    }
#undef H
    return -plnpsum + (occupbins - 1.0) / (2.0 * numFrames);
}

}  // namespace

extern "C" {

double oracle_entropy2d(int t1, const double* x1, double min1, double max1, int bins1, int t2,
                        const double* x2, double min2, double max2, int bins2, int numFrames,
                        int64_t* counts) {
    std::vector<long long> histo((size_t)bins1 * bins2);
    const double* rows[2] = {x1, x2};
    const double mins[2] = {min1, min2}, maxs[2] = {max1, max2};
    return entropy2d_core<true>(t1, rows, mins, maxs, 0, bins1, t2, 1, bins2, numFrames, histo.data(), bins2,
                                counts);
}

// Offsets of the six 2D blocks in the .par body (io_binary.cpp:606-631) and
// the reverse-order half-matrix index (util.cpp:52-57).
int64_t oracle_par_pos(int nD, int gi, int gj) {
    const int64_t nB = nD + 2, nA = nD + 1, nDD = nD;
    const int64_t oBB = 0, oBA = nB * (nB - 1) / 2, oBD = oBA + nB * nA, oAA = oBD + nB * nDD,
                  oAD = oAA + nA * (nA - 1) / 2, oDD = oAD + nA * nDD;
    int t1 = type_of(nD, gi), t2 = type_of(nD, gj);
    int64_t i = gi - first_of_type(nD, t1), j = gj - first_of_type(nD, t2);
    auto rev = [](int64_t n, int64_t s, int64_t g) { return (n - s) * (n - s - 1) / 2 + s - g; };
    if (t1 == ORACLE_B && t2 == ORACLE_B) return oBB + rev(nB, i, j);
    if (t1 == ORACLE_B && t2 == ORACLE_A) return oBA + i * nA + j;
    if (t1 == ORACLE_B && t2 == ORACLE_D) return oBD + i * nDD + j;
    if (t1 == ORACLE_A && t2 == ORACLE_A) return oAA + rev(nA, i, j);
    if (t1 == ORACLE_A && t2 == ORACLE_D) return oAD + i * nDD + j;
    return oDD + rev(nDD, i, j);
}

int oracle_compute(int nD, int N, double* values, const int bins[6], double* s1d, double* s2d,
                   double* vmin_out, double* vmax_out, double* modfit_out, int nthreads) {
    const int D = 3 * nD + 3;
    const double pi = acos(-1);  // PARENT.cpp:112
    std::vector<double> vmin(D), vmax(D), mf(D);
    int err = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(max : err)
    for (int g = 0; g < D; g++) {
        int r = oracle_dof_range(type_of(nD, g), N, values + (size_t)g * N, pi, &vmin[g], &vmax[g], &mf[g]);
        if (r > err) err = r;
    }
    if (err) return err;
    if (vmin_out) memcpy(vmin_out, vmin.data(), D * sizeof(double));
    if (vmax_out) memcpy(vmax_out, vmax.data(), D * sizeof(double));
    if (modfit_out) memcpy(modfit_out, mf.data(), D * sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
    for (int g = 0; g < D; g++) {
        int t = type_of(nD, g);
        s1d[g] = oracle_entropy1d(t, N, values + (size_t)g * N, vmin[g], vmax[g], bins[t]);
    }
    // all pairs gi < gj; each pair is independent, so the thread split does
    // not change any result bit.
#pragma omp parallel for schedule(dynamic, 1)
    for (int gi = 0; gi < D - 1; gi++) {
        int t1 = type_of(nD, gi);
        for (int gj = gi + 1; gj < D; gj++) {
            int t2 = type_of(nD, gj);
            s2d[oracle_par_pos(nD, gi, gj)] =
                oracle_entropy2d(t1, values + (size_t)gi * N, vmin[gi], vmax[gi], bins[3 + t1], t2,
                                 values + (size_t)gj * N, vmin[gj], vmax[gj], bins[3 + t2], N, nullptr);
        }
    }
    return 0;
}

// The reference's schedule on one rank (PARENT.cpp:774-1479 with one MPI
// process): the row dofs one after another in global order, and for each row
// one OpenMP parallel region whose threads split that row's partners of each
// type with the default (static) schedule and share nothing but the inputs;
// every thread fills one histogram buffer of B_row x biggest long longs per
// row (PARENT.cpp:786-792).  Same results as oracle_compute (pairs are
// independent); this is the form the CPU baseline times.
int oracle_compute_rowwise(int nD, int N, double* values, const int bins[6], double* s1d, double* s2d,
                           int nthreads) {
    const int D = 3 * nD + 3;
    const double pi = acos(-1);  // PARENT.cpp:112
    std::vector<double> vmin(D), vmax(D), mf(D);
    int err = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for reduction(max : err)
    for (int g = 0; g < D; g++) {
        int r = oracle_dof_range(type_of(nD, g), N, values + (size_t)g * N, pi, &vmin[g], &vmax[g], &mf[g]);
        if (r > err) err = r;
    }
    if (err) return err;
#pragma omp parallel for
    for (int g = 0; g < D; g++) {
        int t = type_of(nD, g);
        s1d[g] = oracle_entropy1d(t, N, values + (size_t)g * N, vmin[g], vmax[g], bins[t]);
    }
    // every value's index at its type's 2D bin count, checked once here so
    // that the pair loop needs no test (the reference has none)
    int bad = 0;
#pragma omp parallel for reduction(max : bad)
    for (int g = 0; g < D; g++) {
        const int t = type_of(nD, g), nb = bins[3 + t];
        const double* x = values + (size_t)g * N;
        const double binsize = (vmax[g] - vmin[g]) / nb;
        for (int i = 0; i < N; i++) {
            const int k = int((x[i] - vmin[g]) / binsize);
            if (k < 0 || k >= nb) bad = 1;
        }
    }
    if (bad) return 2;
    std::vector<const double*> rows(D);
    for (int g = 0; g < D; g++) rows[g] = values + (size_t)g * N;
    int biggest = 1;
    for (int t = 0; t < 3; t++) biggest = bins[3 + t] > biggest ? bins[3 + t] : biggest;
    for (int gi = 0; gi < D - 1; gi++) {
        const int t1 = type_of(nD, gi);
#pragma omp parallel
        {
            std::vector<long long> histo((size_t)bins[3 + t1] * biggest);
            long long* h = histo.data();
            for (int t2 = t1; t2 < 3; t2++) {
                const int lo = t2 == t1 ? gi + 1 : first_of_type(nD, t2);
                const int hi = first_of_type(nD, t2) + n_of_type(nD, t2);
#pragma omp for
                for (int gj = lo; gj < hi; gj++)
                    s2d[oracle_par_pos(nD, gi, gj)] =
                        entropy2d_core<false>(t1, rows.data(), vmin.data(), vmax.data(), gi, bins[3 + t1], t2, gj,
                                              bins[3 + t2], N, h, biggest, nullptr);
            }
        }
    }
    return 0;
}

int oracle_compute_subset(int nD, int N, int n_sel, const int32_t* dofs, double* cols, const int bins[6],
                          int n_pairs, const int32_t* pairs, double* s1d_sel, double* s2d_pairs, int nthreads) {
    const double pi = acos(-1);  // PARENT.cpp:112
    std::vector<double> vmin(n_sel), vmax(n_sel), mf(n_sel);
    int err = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(max : err)
    for (int k = 0; k < n_sel; k++) {
        int r = oracle_dof_range(type_of(nD, dofs[k]), N, cols + (size_t)k * N, pi, &vmin[k], &vmax[k], &mf[k]);
        if (r > err) err = r;
    }
    if (err) return err;
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < n_sel; k++) {
        const int t = type_of(nD, dofs[k]);
        s1d_sel[k] = oracle_entropy1d(t, N, cols + (size_t)k * N, vmin[k], vmax[k], bins[t]);
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (int p = 0; p < n_pairs; p++) {
        const int a = pairs[2 * p], b = pairs[2 * p + 1];
        const int t1 = type_of(nD, dofs[a]), t2 = type_of(nD, dofs[b]);
        s2d_pairs[p] = oracle_entropy2d(t1, cols + (size_t)a * N, vmin[a], vmax[a], bins[3 + t1], t2,
                                        cols + (size_t)b * N, vmin[b], vmax[b], bins[3 + t2], N, nullptr);
    }
    return 0;
}

double oracle_mutual(int nD, const double* s1d, const double* s2d, int gi, int gj) {
    int lo = gi < gj ? gi : gj, hi = gi < gj ? gj : gi;
    return s1d[lo] + s1d[hi] - s2d[oracle_par_pos(nD, lo, hi)];
}

// get_PAR_MIST.cpp:396-766 on one rank and one thread: Out holds each type's
// remaining dofs in increasing order, In the processed ones in insertion
// order, the root is bond 0, and every step scans the nine (In type, Out type)
// categories in the order BB BA BD AB AA AD DB DA DD with a strict '>'.
int oracle_mist(int nD, const double* s1d, const double* s2d, int32_t* edges, double* mi,
                int32_t* left, double* totals) {
    std::vector<int> In[3], Out[3];
    for (int t = 0; t < 3; t++)
        for (int i = 0; i < n_of_type(nD, t); i++) Out[t].push_back(i);
    In[ORACLE_B].push_back(Out[ORACLE_B][0]);
    Out[ORACLE_B].erase(Out[ORACLE_B].begin());
    double tot1[3] = {0, 0, 0};
    for (int t = 0; t < 3; t++)
        for (int i = 0; i < n_of_type(nD, t); i++) tot1[t] += s1d[first_of_type(nD, t) + i];
    double totMI[6] = {0, 0, 0, 0, 0, 0}, totalMutual = 0;
    auto cat = [](int ta, int tb) {
        int lo = ta < tb ? ta : tb, hi = ta < tb ? tb : ta;
        return lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);  // BB BA BD AA AD DD
    };
    int step = 0;
    // The argmax over In x Out scans (t1, t2, i, j) lexicographically and keeps
    // the first maximum (strict >).  The scan is split into contiguous runs of
    // (t1, t2, i) rows, one per thread; each run keeps its own first maximum
    // and the runs are merged in order with strict >, which selects the same
    // element as the sequential scan.
    struct Best {
        double v = -1e200;
        int t1 = -1, t2 = -1, i = -1, j = -1;
    };
    std::vector<std::array<int, 3>> rows;
    while (Out[0].size() + Out[1].size() + Out[2].size() > 0) {
        rows.clear();
        for (int t1 = 0; t1 < 3; t1++)
            for (int t2 = 0; t2 < 3; t2++)
                for (size_t i = 0; i < In[t1].size(); i++) rows.push_back({t1, t2, (int)i});
        std::vector<Best> part(omp_get_max_threads());
#pragma omp parallel num_threads((int)part.size())
        {
            const int me = omp_get_thread_num(), nth = omp_get_num_threads();
            const size_t r0 = rows.size() * me / nth, r1 = rows.size() * (me + 1) / nth;
            Best b;
            for (size_t r = r0; r < r1; r++) {
                const int t1 = rows[r][0], t2 = rows[r][1], i = rows[r][2];
                for (size_t j = 0; j < Out[t2].size(); j++) {
                    const double v = oracle_mutual(nD, s1d, s2d, first_of_type(nD, t1) + In[t1][i],
                                                   first_of_type(nD, t2) + Out[t2][j]);
                    if (v > b.v) b = {v, t1, t2, i, (int)j};
                }
            }
            part[me] = b;
        }
        Best w;
        for (const Best& b : part)
            if (b.v > w.v) w = b;
        const double best = w.v;
        const int bt1 = w.t1, bt2 = w.t2, bi = w.i, bj = w.j;
        if (bt1 < 0) return 1;
        totalMutual += best;
        edges[4 * step + 0] = bt1;
        edges[4 * step + 1] = bt2;
        edges[4 * step + 2] = In[bt1][bi];
        edges[4 * step + 3] = Out[bt2][bj];
        mi[step] = best;
        left[step] = (int)(Out[0].size() + Out[1].size() + Out[2].size());
        In[bt2].push_back(Out[bt2][bj]);
        Out[bt2].erase(Out[bt2].begin() + bj);
        totMI[cat(bt1, bt2)] += best;
        step++;
    }
    totals[0] = tot1[0];
    totals[1] = tot1[1];
    totals[2] = tot1[2];
    for (int c = 0; c < 6; c++) totals[3 + c] = totMI[c];
    totals[9] = tot1[0] + tot1[1] + tot1[2] - totalMutual;
    return 0;
}

}  // extern "C"
