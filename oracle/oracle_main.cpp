// oracle_parent.x: single-rank restatement of the reference PARENT.x
// (PARENT.cpp:93-1512) on top of the oracle functions.  TEST INFRASTRUCTURE
// ONLY: used to check that the oracle writes a .par byte-identical to the
// reference binary, and as the "port" CPU baseline.
//
//   oracle_parent.x in.bat out.par b1D a1D d1D b2D a2D d2D
#include "oracle.h"

#include <sys/time.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {

struct BatHeader {
    int version = 0, double_prec = 0, nD = 0, N = 0;
    std::vector<std::string> residues, atoms, molecules;
    std::vector<int> resnums;
    std::vector<int> dihedrals;  // nD x 6
    std::vector<float> masses;
};

// io_binary.cpp:107-200.  Strings are read into one persistent 31-byte buffer
// and taken up to the first NUL, as the reference does.
int read_header(std::ifstream& f, BatHeader& h) {
    auto rd = [&](void* p, size_t n) {
        f.read(reinterpret_cast<char*>(p), n);
        return f.fail() ? 1 : 0;
    };
    int fail = 0;
    fail |= rd(&h.version, 4);
    fail |= rd(&h.double_prec, 4);
    fail |= rd(&h.nD, 4);
    fail |= rd(&h.N, 4);
    if (fail) return fail;
    if (h.version < 0) {
        std::cerr << "ERROR: FILE HEADER CORRUPTED. VERSION NUMBER (" << h.version << ") < 0!" << std::endl;
        return 1;
    }
    if (h.double_prec != 0 && h.double_prec != 1) {
        std::cerr << "ERROR: FILE HEADER CORRUPTED. DOUBLE PRECISION VALUE (" << h.double_prec << ") NEITHER 0 NOR 1!" << std::endl;
        return 1;
    }
    if (h.nD < 0) {
        std::cerr << "ERROR: FILE HEADER CORRUPTED. NUMBER OF DIHEDRALS (" << h.nD << ") < 0!" << std::endl;
        return 1;
    }
    if (h.N < 1) {
        std::cerr << "ERROR: FILE HEADER CORRUPTED. NUMBER OF FRAMES (" << h.N << ") < 1!" << std::endl;
        return 1;
    }
    char buf[32] = {0};
    if (h.version >= 3) {
        for (int i = 0; i < h.nD + 3; i++) {
            fail |= rd(buf, 8);
            h.residues.push_back(buf);
            int rn = 0;
            fail |= rd(&rn, 4);
            h.resnums.push_back(rn);
            fail |= rd(buf, 8);
            h.atoms.push_back(buf);
            fail |= rd(buf, 31);
            h.molecules.push_back(buf);
        }
    }
    if (fail) return fail;
    h.dihedrals.resize((size_t)6 * h.nD);
    for (int i = 0; i < 6 * h.nD; i++) fail |= rd(&h.dihedrals[i], 4);
    h.masses.resize(h.nD + 3);
    for (int i = 0; i < h.nD + 3; i++) fail |= rd(&h.masses[i], 4);
    return fail;
}

// io_binary.cpp:371-442.
int write_par_header(std::ofstream& o, const BatHeader& h, const int bins[6]) {
    int version = 4;
    o.write((const char*)&version, 4);
    o.write((const char*)&h.double_prec, 4);
    o.write((const char*)&h.nD, 4);
    o.write((const char*)&h.N, 4);
    for (int k = 0; k < 6; k++) o.write((const char*)&bins[k], 4);
    char s[32];
    for (int i = 0; i < h.nD + 3; i++) {
        memset(s, 0, sizeof(s));
        snprintf(s, 8, "%.7s", i < (int)h.residues.size() ? h.residues[i].c_str() : "");
        o.write(s, 8);
        int rn = i < (int)h.resnums.size() ? h.resnums[i] : 0;
        o.write((const char*)&rn, 4);
        memset(s, 0, sizeof(s));
        snprintf(s, 8, "%.7s", i < (int)h.atoms.size() ? h.atoms[i].c_str() : "");
        o.write(s, 8);
        memset(s, 0, sizeof(s));
        snprintf(s, 31, "%.30s", i < (int)h.molecules.size() ? h.molecules[i].c_str() : "");
        o.write(s, 31);
    }
    o.write((const char*)h.dihedrals.data(), h.dihedrals.size() * 4);
    o.write((const char*)h.masses.data(), h.masses.size() * 4);
    return o.fail() ? 1 : 0;
}

double since(const timeval& a, const timeval& b) {
    return b.tv_sec + 0.000001 * b.tv_usec - a.tv_sec - 0.000001 * a.tv_usec;
}

const char* last_token(const std::string& s) {
    size_t p = s.rfind('.');
    return p == std::string::npos ? s.c_str() : s.c_str() + p + 1;
}

}  // namespace

int main(int argc, char* argv[]) {
    timeval tv1, tv2, tv3, tv4;
    gettimeofday(&tv1, NULL);
    int bins[6];
    // PARENT.cpp:190-244
    if (argc != 9) {
        std::cerr << "USAGE: " << argv[0] << " input.bat entropy.par #bondsbins1D #anglesbins1D #dihedralsbins1D #bondsbins2D #anglesbins2D #dihedralsbins2D\n";
        return 1;
    }
    if (strcmp(last_token(argv[1]), "bat") || strcmp(last_token(argv[2]), "par")) {
        std::cerr << "USAGE: " << argv[0] << " input.bat entropy.par #bondsbins1D #anglesbins1D #dihedralsbins1D #bondsbins2D #anglesbins2D #dihedralsbins2D\n";
        return 1;
    }
    for (int k = 0; k < 6; k++) {
        if (sscanf(argv[3 + k], "%d", &bins[k]) != 1) {
            std::cerr << "ERROR: Could not read number of bins from command line! Aborting" << std::endl;
            return 1;
        }
    }
    std::ifstream in(argv[1], std::ios::binary | std::ios::in);
    std::ofstream par(argv[2], std::ios::binary | std::ios::out);
    if (!in.is_open()) {
        std::cerr << "ERROR: Could not open file " << argv[1] << " ! Aborting." << std::endl;
        return 1;
    }
    if (!par.is_open()) {
        std::cerr << "ERROR: Could not open file " << argv[2] << " ! Aborting." << std::endl;
        return 1;
    }
    BatHeader h;
    if (read_header(in, h) != 0) {
        std::cerr << "AN ERROR HAS OCCURED WHILE READING THE HEADER OF THE FILE " << argv[1] << " . QUITTING PROGRAM.\n";
        return 1;
    }
    in.close();
    if (write_par_header(par, h, bins) != 0) {
        std::cerr << "AN ERROR HAS OCCURED WHILE WRITING THE HEADER OF THE FILE " << argv[2] << " . QUITTING PROGRAM.\n";
        return 1;
    }
    int nD, N, dp;
    double* v = nullptr;
    if (oracle_bat_read(argv[1], &nD, &N, &dp, &v) != 0) {
        std::cerr << "An ERROR has occurred while reading the file " << argv[1] << " . Quitting Program.\n";
        return 1;
    }
    gettimeofday(&tv2, NULL);
    const int D = 3 * nD + 3;
    const long long P = (long long)D * (D - 1) / 2;
    std::vector<double> s1d(D), s2d(P);
    int rc = oracle_compute_rowwise(nD, N, v, bins, s1d.data(), s2d.data(), 0);
    oracle_free(v);
    if (rc == 1) {
        std::cerr << "ERROR: Bond is smaller than 0.0" << std::endl;
        return 1;
    }
    if (rc != 0) {
        std::cerr << "ERROR: value outside the histogram range" << std::endl;
        return 1;
    }
    gettimeofday(&tv3, NULL);
    // io_binary.cpp:606-631: the 1D blocks then the 2D blocks, which
    // oracle_compute already placed in body order.
    par.write((const char*)s1d.data(), D * sizeof(double));
    par.write((const char*)s2d.data(), P * sizeof(double));
    if (par.fail()) std::cerr << "AN ERROR HAS OCCURED WHILE WRITING THE FILE " << argv[2] << " .\n";
    par.close();
    gettimeofday(&tv4, NULL);
    std::cout << "Initialization + readin time: " << since(tv1, tv2) << std::endl;
    std::cout << "Calculation time: " << since(tv2, tv3) << std::endl;
    std::cout << "Output time: " << since(tv3, tv4) << std::endl;
    std::cout << "Total execution time: " << since(tv1, tv4) << std::endl;
    std::cout << "PROGRAM FINISHED SUCCESSFULLY." << std::endl;
    return 0;
}
