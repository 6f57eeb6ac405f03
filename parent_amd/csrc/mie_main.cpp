// get_PAR_MIE_hip.x -- drop-in replacement for the reference's get_PAR_MIE.x.
//
//   get_PAR_MIE_hip.x input.par > output_MIE.txt
//
// Prints the 1D entropies, every pair's 2D entropy and mutual information,
// and the 16 totals to stdout, byte-identical to the reference
// (get_PAR_MIE.cpp:54-199), with its error messages and exit codes
// (get_PAR_MIE.cpp:93-103, 205-209).  OMP_NUM_THREADS threads format the
// listing.
#include "../../include/parent_hip.h"

#include <cstdio>
#include <iostream>
#include <vector>

int main(int argc, char* argv[]) {
    FILE* probe = argc > 1 ? fopen(argv[1], "rb") : nullptr;
    if (!probe) {
        std::cerr << "ERROR: COULD NOT OPEN FILE !\n\nUSAGE:\n" << argv[0] << " input.par" << std::endl;
        return 1;
    }
    fclose(probe);
    int nD = 0, N = 0, bins[6];
    if (parent_par_read(argv[1], &nD, &N, bins, nullptr, nullptr) != PARENT_OK) {
        std::cerr << "ERROR READING HEADER OF FILE " << argv[1] << " !" << std::endl;
        return 1;
    }
    const size_t D = 3 * (size_t)nD + 3, P = D * (D - 1) / 2;
    std::vector<double> s1d(D), s2d(P);
    if (parent_par_read(argv[1], &nD, &N, bins, s1d.data(), s2d.data()) != PARENT_OK) {
        std::cerr << "ERROR READING FILE " << argv[1] << " !" << std::endl;
        return 1;
    }
    if (parent_mie_write("-", nD, s1d.data(), s2d.data(), 0, nullptr) != PARENT_OK) {
        std::cerr << parent_last_error() << std::endl;
        return 1;
    }
    return 0;
}
