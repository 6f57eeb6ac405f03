// MIE text listing of a .par: the output of the reference's get_PAR_MIE.x,
// byte for byte (get_PAR_MIE.cpp:54-199).
//
// Layout of the listing (every number at precision 12, i.e. printf "%.12g"):
//   "bond k   S"  / "angle k   S" / "dihedral k   S"          one per dof
//   "<t1>-<t2> i  j   S2D   MI"                              one per pair
//      blocks bond-bond, bond-angle, bond-dihedral, angle-angle,
//      angle-dihedral, dihedral-dihedral; i < j inside a type, 1-based.
//      The half-matrix blocks are stored in reverse in the .par, so the
//      k-th listed pair of such a block reads entry count-1-k.
//   16 "TOTAL ..." lines.
// Each total is a sequential FP64 sum in listing order, as the reference
// accumulates it, and MI = (S1D_i + S1D_j) - S2D_ij (util.cpp:47-81).
//
// The reference flushes every line (std::endl).  Here the listing is cut into
// tasks of whole rows, formatted in parallel into per-task buffers, and
// written in order with large writes.
#include "../../include/parent_hip.h"
#include "host_common.h"

#include <charconv>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <omp.h>

namespace {

using parent::set_last_error;

// printf("%.12g") -- std::to_chars with an explicit precision formats as
// printf does in the "C" locale, which is what the reference's ostream
// (precision 12, default float field) produces.
inline char* put_double(char* p, double v) { return std::to_chars(p, p + 40, v, std::chars_format::general, 12).ptr; }
inline char* put_int(char* p, long long v) { return std::to_chars(p, p + 24, v).ptr; }
inline char* put_str(char* p, const char* s, size_t n) {
    memcpy(p, s, n);
    return p + n;
}

struct Block {
    const char* label;  // "bond-bond" ...
    int t1, t2;         // dof types
    bool half;          // i < j half matrix stored in reverse
    int64_t n1, n2;     // rows, columns (n2 == n1 for half blocks)
    int64_t off;        // offset of the block in s2d
};

// One task: rows [r0, r1) of a block.
struct Task {
    int block;
    int64_t r0, r1;
};

}  // namespace

extern "C" int parent_mie_write(const char* path, int nD, const double* s1d, const double* s2d, int nthreads,
                                double* totals) {
    if (nD < 0 || !s1d || !s2d) {
        set_last_error("parent_mie_write: invalid argument");
        return PARENT_E_INVALID;
    }
    const int64_t nB = nD + 2, nA = nD + 1, nd = nD;
    const double* S1[3] = {s1d, s1d + nB, s1d + nB + nA};
    const int64_t cnt[3] = {nB, nA, nd};
    static const char* kName[3] = {"bond", "angle", "dihedral"};
    // io_binary.cpp:606-631 block order
    Block blocks[6];
    {
        int64_t off = 0;
        const int bt[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
        static const char* kLabel[6] = {"bond-bond", "bond-angle", "bond-dihedral",
                                        "angle-angle", "angle-dihedral", "dihedral-dihedral"};
        for (int b = 0; b < 6; b++) {
            const int t1 = bt[b][0], t2 = bt[b][1];
            blocks[b] = {kLabel[b], t1, t2, t1 == t2, cnt[t1], cnt[t2], off};
            off += t1 == t2 ? cnt[t1] * (cnt[t1] - 1) / 2 : cnt[t1] * cnt[t2];
        }
    }
    auto entries_of_row = [&](const Block& b, int64_t i) { return b.half ? b.n1 - 1 - i : b.n2; };
    // first listing position of row i inside a half block
    auto half_row_start = [](int64_t n, int64_t i) { return i * (2 * n - i - 1) / 2; };

    // ---- totals: sequential sums in listing order -------------------------
    double tot[16];
    for (int t = 0; t < 3; t++) {
        double s = 0.0;
        for (int64_t k = 0; k < cnt[t]; k++) s += S1[t][k];
        tot[t] = s;
    }
    for (int bi = 0; bi < 6; bi++) {
        const Block& b = blocks[bi];
        const double* A = S1[b.t1];
        const double* Bv = S1[b.t2];
        const double* E = s2d + b.off;
        double se = 0.0, sm = 0.0;
        if (b.half) {
            int64_t counter = b.n1 * (b.n1 - 1) / 2 - 1;
            for (int64_t i = 0; i + 1 < b.n1; i++)
                for (int64_t j = i + 1; j < b.n1; j++, counter--) {
                    const double mi = A[i] + A[j] - E[counter];
                    se += E[counter];
                    sm += mi;
                }
        } else {
            for (int64_t i = 0; i < b.n1; i++)
                for (int64_t j = 0; j < b.n2; j++) {
                    const double e = E[i * b.n2 + j];
                    const double mi = A[i] + Bv[j] - e;
                    se += e;
                    sm += mi;
                }
        }
        tot[3 + 2 * bi] = se;
        tot[4 + 2 * bi] = sm;
    }
    tot[15] = tot[0] + tot[1] + tot[2] - tot[4] - tot[6] - tot[8] - tot[10] - tot[12] - tot[14];
    if (totals) memcpy(totals, tot, sizeof(tot));
    if (!path) return PARENT_OK;  // totals only

    FILE* f = strcmp(path, "-") == 0 ? stdout : fopen(path, "wb");
    if (!f) {
        set_last_error(std::string("ERROR: COULD NOT OPEN FILE ") + path + " !");
        return PARENT_E_IO;
    }
    bool ok = true;
    std::string buf;
    buf.resize(64 * (size_t)(nB + nA + nd) + 64);
    {
        char* p = buf.data();
        for (int t = 0; t < 3; t++)
            for (int64_t k = 0; k < cnt[t]; k++) {
                p = put_str(p, kName[t], strlen(kName[t]));
                *p++ = ' ';
                p = put_int(p, k + 1);
                p = put_str(p, "   ", 3);
                p = put_double(p, S1[t][k]);
                *p++ = '\n';
            }
        ok = ok && fwrite(buf.data(), 1, p - buf.data(), f) == (size_t)(p - buf.data());
    }

    // ---- pair lines --------------------------------------------------------
    constexpr int64_t kLinesPerTask = 16384;
    std::vector<Task> tasks;
    for (int bi = 0; bi < 6; bi++) {
        const Block& b = blocks[bi];
        int64_t r0 = 0, lines = 0;
        for (int64_t i = 0; i < b.n1; i++) {
            lines += entries_of_row(b, i);
            if (lines >= kLinesPerTask || i + 1 == b.n1) {
                if (lines > 0) tasks.push_back({bi, r0, i + 1});
                r0 = i + 1;
                lines = 0;
            }
        }
    }
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    const size_t group = (size_t)nt * 4;
    std::vector<std::string> out(group);
    // longest line: label + 2 ints + 2 doubles + separators < 128 bytes
    for (size_t g0 = 0; g0 < tasks.size() && ok; g0 += group) {
        const size_t g1 = std::min(tasks.size(), g0 + group);
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
        for (long long ti = (long long)g0; ti < (long long)g1; ti++) {
            const Task& tk = tasks[ti];
            const Block& b = blocks[tk.block];
            const double* A = S1[b.t1];
            const double* Bv = S1[b.t2];
            const double* E = s2d + b.off;
            int64_t lines = 0;
            for (int64_t i = tk.r0; i < tk.r1; i++) lines += entries_of_row(b, i);
            std::string& s = out[ti - g0];
            s.resize((size_t)lines * 128);
            char* p = s.data();
            const size_t ll = strlen(b.label);
            const int64_t P = b.half ? b.n1 * (b.n1 - 1) / 2 : 0;
            for (int64_t i = tk.r0; i < tk.r1; i++) {
                const int64_t j0 = b.half ? i + 1 : 0;
                int64_t k = b.half ? half_row_start(b.n1, i) : 0;  // listing position in the block
                for (int64_t j = j0; j < b.n2; j++, k++) {
                    const double e = b.half ? E[P - 1 - k] : E[i * b.n2 + j];
                    const double mi = A[i] + Bv[j] - e;
                    p = put_str(p, b.label, ll);
                    *p++ = ' ';
                    p = put_int(p, i + 1);
                    p = put_str(p, "  ", 2);
                    p = put_int(p, j + 1);
                    p = put_str(p, "   ", 3);
                    p = put_double(p, e);
                    p = put_str(p, "   ", 3);
                    p = put_double(p, mi);
                    *p++ = '\n';
                }
            }
            s.resize(p - s.data());
        }
        for (size_t ti = g0; ti < g1 && ok; ti++) {
            const std::string& s = out[ti - g0];
            ok = fwrite(s.data(), 1, s.size(), f) == s.size();
        }
    }

    // ---- totals ------------------------------------------------------------
    static const char* kTot[16] = {
        "TOTAL 1D BONDS ENTROPY = ", "TOTAL 1D ANGLES ENTROPY = ", "TOTAL 1D DIHEDRALS ENTROPY = ",
        "TOTAL 2D BONDS-BONDS ENTROPY = ", "TOTAL 2D BONDS-BONDS MUTUAL INFORMATION = ",
        "TOTAL 2D BONDS-ANGLES ENTROPY = ", "TOTAL 2D BONDS-ANGLES MUTUAL INFORMATION = ",
        "TOTAL 2D BONDS-DIHEDRALS ENTROPY = ", "TOTAL 2D BONDS-DIHEDRALS MUTUAL INFORMATION = ",
        "TOTAL 2D ANGLES-ANGLES ENTROPY = ", "TOTAL 2D ANGLES-ANGLES MUTUAL INFORMATION = ",
        "TOTAL 2D ANGLES-DIHEDRALS ENTROPY = ", "TOTAL 2D ANGLES-DIHEDRALS MUTUAL INFORMATION = ",
        "TOTAL 2D DIHEDRALS-DIHEDRALS ENTROPY = ", "TOTAL 2D DIHEDRALS-DIHEDRALS MUTUAL INFORMATION = ",
        "TOTAL CONFIGURATIONAL ENTROPY = "};
    if (ok) {
        char line[160];
        for (int t = 0; t < 16 && ok; t++) {
            char* p = put_str(line, kTot[t], strlen(kTot[t]));
            p = put_double(p, tot[t]);
            *p++ = '\n';
            ok = fwrite(line, 1, p - line, f) == (size_t)(p - line);
        }
    }
    if (fflush(f) != 0) ok = false;
    if (f != stdout && fclose(f) != 0) ok = false;
    if (!ok) {
        set_last_error(std::string("ERROR WRITING ") + path);
        return PARENT_E_IO;
    }
    return PARENT_OK;
}
