// .bat reader and .par reader/writer for the engine's C ABI.
//
// Formats follow the reference IO library byte for byte (io_binary.cpp):
//   .bat v3 header 107-200, frames 281-368; .par v4 header 371-442 /
//   446-604, body 606-673.
// The reference reads one value per istream::read call; here frames are read
// in large blocks and de-interleaved in memory, which does not change any byte.
#include "../../include/parent_hip.h"
#include "host_common.h"

#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace parent {

thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

}  // namespace parent

using parent::set_last_error;

struct parent_bat {
    FILE* f = nullptr;
    int version = 0, double_prec = 0, nD = 0, N = 0;
    std::vector<std::string> residues, atoms, molecules;
    std::vector<int32_t> resnums;
    std::vector<int32_t> dihedrals;  // nD x 6
    std::vector<float> masses;
    long frames_offset = 0;
};

namespace {

bool rd(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

// Strings are taken up to the first NUL of a persistent 31-byte buffer, as the
// reference's push_back(dummystring) does (io_binary.cpp:147-160).
int read_atoms(FILE* f, int natoms, parent_bat* b) {
    char buf[32];
    memset(buf, 0, sizeof(buf));
    for (int i = 0; i < natoms; i++) {
        int32_t rn;
        if (!rd(f, buf, 8)) return 1;
        b->residues.emplace_back(buf);
        if (!rd(f, &rn, 4)) return 1;
        b->resnums.push_back(rn);
        if (!rd(f, buf, 8)) return 1;
        b->atoms.emplace_back(buf);
        if (!rd(f, buf, 31)) return 1;
        b->molecules.emplace_back(buf);
    }
    return 0;
}

void put_str(std::vector<char>& out, const std::string& s, size_t width) {
    // sprintf("%.7s") / sprintf("%.30s") into a zeroed buffer (io_binary.cpp:400-417)
    size_t n = s.size() < width - 1 ? s.size() : width - 1;
    size_t at = out.size();
    out.resize(at + width, 0);
    memcpy(out.data() + at, s.data(), n);
}

template <class T>
void put(std::vector<char>& out, const T& v) {
    size_t at = out.size();
    out.resize(at + sizeof(T));
    memcpy(out.data() + at, &v, sizeof(T));
}

}  // namespace

namespace {

// Frame record (io_binary.cpp:281-368): time, xtcPrec, box[9] as float; then
// 6 external coordinates, bond0, bond1, angle0 and nD triples
// (bond[i+2], angle[i+1], dihedral[i]) in the trajectory's precision.
template <class T>
void deinterleave(const char* rec, int nD, T* row) {
    // the record's values start 11 floats + 6 values in and need not be
    // aligned for T, so they are copied out bytewise
    const char* v = rec + 11 * sizeof(float) + 6 * sizeof(T);
    auto at = [&](size_t k) {
        T x;
        memcpy(&x, v + k * sizeof(T), sizeof(T));
        return x;
    };
    const int nB = nD + 2, nA = nD + 1;
    T* b = row;
    T* a = row + nB;
    T* d = row + nB + nA;
    b[0] = at(0);
    b[1] = at(1);
    a[0] = at(2);
    for (int i = 0; i < nD; i++) {
        b[i + 2] = at(3 + 3 * (size_t)i + 0);
        a[i + 1] = at(3 + 3 * (size_t)i + 1);
        d[i] = at(3 + 3 * (size_t)i + 2);
    }
}

}  // namespace

extern "C" {

const char* parent_last_error(void) { return parent::g_last_error.c_str(); }

int parent_bat_open(const char* path, parent_bat** out) {
    if (!path || !out) {
        set_last_error("parent_bat_open: null argument");
        return PARENT_E_INVALID;
    }
    *out = nullptr;
    FILE* f = fopen(path, "rb");
    if (!f) {
        set_last_error(std::string("ERROR: Could not open file ") + path + " !");
        return PARENT_E_IO;
    }
    parent_bat* b = new parent_bat();
    b->f = f;
    int32_t hdr[4];
    if (!rd(f, hdr, sizeof(hdr))) {
        set_last_error("ERROR: FILE HEADER TRUNCATED");
        parent_bat_close(b);
        return PARENT_E_IO;
    }
    b->version = hdr[0];
    b->double_prec = hdr[1];
    b->nD = hdr[2];
    b->N = hdr[3];
    // io_binary.cpp:127-145
    char msg[160];
    if (b->version < 0) {
        snprintf(msg, sizeof msg, "ERROR: FILE HEADER CORRUPTED. VERSION NUMBER (%d) < 0!", b->version);
    } else if (b->double_prec != 0 && b->double_prec != 1) {
        snprintf(msg, sizeof msg, "ERROR: FILE HEADER CORRUPTED. DOUBLE PRECISION VALUE (%d) NEITHER 0 NOR 1!",
                 b->double_prec);
    } else if (b->nD < 0) {
        snprintf(msg, sizeof msg, "ERROR: FILE HEADER CORRUPTED. NUMBER OF DIHEDRALS (%d) < 0!", b->nD);
    } else if (b->N < 1) {
        snprintf(msg, sizeof msg, "ERROR: FILE HEADER CORRUPTED. NUMBER OF FRAMES (%d) < 1!", b->N);
    } else {
        msg[0] = 0;
    }
    if (msg[0]) {
        set_last_error(msg);
        parent_bat_close(b);
        return PARENT_E_INVALID;
    }
    int natoms = b->nD + 3;
    int fail = 0;
    if (b->version >= 3) fail |= read_atoms(f, natoms, b);
    b->dihedrals.resize((size_t)6 * b->nD);
    b->masses.resize(natoms);
    if (!fail && b->nD > 0) fail |= !rd(f, b->dihedrals.data(), b->dihedrals.size() * 4);
    if (!fail) fail |= !rd(f, b->masses.data(), b->masses.size() * 4);
    if (fail) {
        set_last_error("ERROR: FILE HEADER TRUNCATED");
        parent_bat_close(b);
        return PARENT_E_IO;
    }
    b->frames_offset = ftell(f);
    *out = b;
    return PARENT_OK;
}

int parent_bat_info(const parent_bat* b, int* n_dihedrals, int* n_frames, int* double_prec) {
    if (!b) return PARENT_E_INVALID;
    if (n_dihedrals) *n_dihedrals = b->nD;
    if (n_frames) *n_frames = b->N;
    if (double_prec) *double_prec = b->double_prec;
    return PARENT_OK;
}

int parent_bat_read_frames(parent_bat* b, void* values) {
    if (!b || !values) {
        set_last_error("parent_bat_read_frames: null argument");
        return PARENT_E_INVALID;
    }
    const int nD = b->nD;
    const size_t D = (size_t)parent::n_dofs(nD);
    const size_t vsz = b->double_prec ? 8 : 4;
    const size_t rec = 11 * 4 + (6 + 3 + 3 * (size_t)nD) * vsz;
    if (fseeko(b->f, (off_t)b->frames_offset, SEEK_SET) != 0) {
        set_last_error("seek failed");
        return PARENT_E_IO;
    }
    const size_t chunk = rec > (64u << 20) ? 1 : (64u << 20) / rec;
    std::vector<char> buf(chunk * rec);
    for (size_t f0 = 0; f0 < (size_t)b->N; f0 += chunk) {
        size_t nf = std::min(chunk, (size_t)b->N - f0);
        if (fread(buf.data(), rec, nf, b->f) != nf) {
            set_last_error("An ERROR has occurred while reading the frames (file truncated).");
            return PARENT_E_IO;
        }
        for (size_t k = 0; k < nf; k++) {
            if (b->double_prec)
                deinterleave<double>(buf.data() + k * rec, nD, (double*)values + (f0 + k) * D);
            else
                deinterleave<float>(buf.data() + k * rec, nD, (float*)values + (f0 + k) * D);
        }
    }
    return PARENT_OK;
}

int64_t parent_bat_record_bytes(const parent_bat* b) {
    if (!b) return -1;
    const int64_t vsz = b->double_prec ? 8 : 4;
    return 11 * 4 + (6 + 3 + 3 * (int64_t)b->nD) * vsz;
}

int parent_bat_read_raw(parent_bat* b, int first_frame, int n_frames, void* records) {
    if (!b || !records || first_frame < 0 || n_frames < 0 || (int64_t)first_frame + n_frames > b->N) {
        set_last_error("parent_bat_read_raw: invalid argument");
        return PARENT_E_INVALID;
    }
    const int64_t rec = parent_bat_record_bytes(b);
    const int64_t at = b->frames_offset + (int64_t)first_frame * rec;
    const int64_t bytes = (int64_t)n_frames * rec;
    // positioned reads, split over up to 8 threads: a single reader is bound
    // by one core's copy out of the page cache
    const int fd = fileno(b->f);
    const int64_t kPart = 4 << 20;
    const int parts = (int)std::min<int64_t>(8, std::max<int64_t>(1, bytes / kPart));
    int bad = 0;
#pragma omp parallel for num_threads(parts) schedule(static, 1) reduction(| : bad)
    for (int k = 0; k < parts; k++) {
        const int64_t lo = bytes * k / parts, hi = bytes * (k + 1) / parts;
        char* dst = static_cast<char*>(records);
        for (int64_t off = lo; off < hi;) {
            const ssize_t got = pread(fd, dst + off, (size_t)(hi - off), (off_t)(at + off));
            if (got <= 0) {
                bad = 1;
                break;
            }
            off += got;
        }
    }
    if (bad) {
        set_last_error("An ERROR has occurred while reading the frames (file truncated).");
        return PARENT_E_IO;
    }
    return PARENT_OK;
}

void parent_bat_close(parent_bat* b) {
    if (!b) return;
    if (b->f) fclose(b->f);
    delete b;
}

int parent_par_write(const char* path, const parent_bat* h, const int bins[6], const double* s1d,
                     const double* s2d) {
    if (!path || !h || !bins || !s1d || !s2d) {
        set_last_error("parent_par_write: null argument");
        return PARENT_E_INVALID;
    }
    std::vector<char> hdr;
    put<int32_t>(hdr, 4);
    put<int32_t>(hdr, h->double_prec);
    put<int32_t>(hdr, h->nD);
    put<int32_t>(hdr, h->N);
    for (int k = 0; k < 6; k++) put<int32_t>(hdr, bins[k]);
    const int natoms = h->nD + 3;
    for (int i = 0; i < natoms; i++) {
        put_str(hdr, i < (int)h->residues.size() ? h->residues[i] : std::string(), 8);
        put<int32_t>(hdr, i < (int)h->resnums.size() ? h->resnums[i] : 0);
        put_str(hdr, i < (int)h->atoms.size() ? h->atoms[i] : std::string(), 8);
        put_str(hdr, i < (int)h->molecules.size() ? h->molecules[i] : std::string(), 31);
    }
    for (int32_t v : h->dihedrals) put<int32_t>(hdr, v);
    for (float m : h->masses) put<float>(hdr, m);
    FILE* f = fopen(path, "wb");
    if (!f) {
        set_last_error(std::string("ERROR: Could not open file ") + path + " !");
        return PARENT_E_IO;
    }
    const size_t D = (size_t)parent::n_dofs(h->nD);
    const size_t P = (size_t)parent::n_pairs(h->nD);
    bool ok = fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
    ok = ok && fwrite(s1d, sizeof(double), D, f) == D;
    ok = ok && fwrite(s2d, sizeof(double), P, f) == P;
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        set_last_error(std::string("AN ERROR HAS OCCURED WHILE WRITING THE FILE ") + path + " .");
        return PARENT_E_IO;
    }
    return PARENT_OK;
}

int parent_par_read(const char* path, int* n_dihedrals, int* n_frames, int bins[6], double* s1d,
                    double* s2d) {
    FILE* f = fopen(path, "rb");
    if (!f) {
        set_last_error(std::string("ERROR: COULD NOT OPEN FILE ") + (path ? path : "") + " !");
        return PARENT_E_IO;
    }
    auto fail = [&](int code, const std::string& m) {
        fclose(f);
        set_last_error(m);
        return code;
    };
    int32_t h[10];
    if (!rd(f, h, sizeof(h))) return fail(PARENT_E_IO, "ERROR: FILE HEADER TRUNCATED");
    const int version = h[0], dp = h[1], nD = h[2], N = h[3];
    // io_binary.cpp:474-519
    if (version < 0) return fail(PARENT_E_INVALID, "ERROR: FILE HEADER CORRUPTED. VERSION NUMBER < 0!");
    if (dp != 0 && dp != 1) return fail(PARENT_E_INVALID, "ERROR: FILE HEADER CORRUPTED. DOUBLE PRECISION VALUE NEITHER 0 NOR 1!");
    if (nD < 0) return fail(PARENT_E_INVALID, "ERROR: FILE HEADER CORRUPTED. NUMBER OF DIHEDRALS < 0!");
    if (N < 1) return fail(PARENT_E_INVALID, "ERROR: FILE HEADER CORRUPTED. NUMBER OF FRAMES < 1!");
    for (int k = 0; k < 6; k++)
        if (h[4 + k] < 1) return fail(PARENT_E_INVALID, "ERROR: FILE HEADER CORRUPTED. NUMBER BINS < 1!");
    const long natoms = nD + 3;
    if (version >= 3 && fseek(f, natoms * (8 + 4 + 8 + 31), SEEK_CUR) != 0)
        return fail(PARENT_E_IO, "ERROR: FILE HEADER TRUNCATED");
    // io_binary.cpp:554-591: dihedral atom ids >= 1, type in {-1,0,1}, parent >= -1
    std::vector<int32_t> dih((size_t)6 * nD);
    if (nD > 0 && !rd(f, dih.data(), dih.size() * 4)) return fail(PARENT_E_IO, "ERROR: FILE HEADER TRUNCATED");
    for (int i = 0; i < nD; i++) {
        const int32_t* r = &dih[(size_t)6 * i];
        if (r[0] < 1 || r[1] < 1 || r[2] < 1 || r[3] < 1)
            return fail(PARENT_E_INVALID, "ERROR: FILE HEADER CORRUPTED. DIHEDRAL ID < 1 !");
        if (!(r[4] == 1 || r[4] == 0 || r[4] == -1))
            return fail(PARENT_E_INVALID, "ERROR: FILE HEADER CORRUPTED. DIHEDRAL TYPE NEITHER 0, 1 NOR -1 !");
        if (r[5] < -1)
            return fail(PARENT_E_INVALID, "ERROR: FILE HEADER CORRUPTED. DIHEDRAL IS PHASEANGLE OF A DIHEDRAL WITH NEGATIVE ID !");
    }
    if (fseek(f, natoms * 4, SEEK_CUR) != 0) return fail(PARENT_E_IO, "ERROR: FILE HEADER TRUNCATED");
    if (n_dihedrals) *n_dihedrals = nD;
    if (n_frames) *n_frames = N;
    if (bins)
        for (int k = 0; k < 6; k++) bins[k] = h[4 + k];
    const size_t D = (size_t)parent::n_dofs(nD);
    const size_t P = (size_t)parent::n_pairs(nD);
    if (s1d && fread(s1d, sizeof(double), D, f) != D) return fail(PARENT_E_IO, "ERROR READING FILE BODY");
    if (!s1d && fseek(f, (long)(D * sizeof(double)), SEEK_CUR) != 0) return fail(PARENT_E_IO, "ERROR READING FILE BODY");
    if (s2d && fread(s2d, sizeof(double), P, f) != P) return fail(PARENT_E_IO, "ERROR READING FILE BODY");
    fclose(f);
    return PARENT_OK;
}

}  // extern "C"
