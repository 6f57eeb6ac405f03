// get_PAR_MIST_hip.x -- drop-in replacement for the reference's get_PAR_MIST.x.
//
//   get_PAR_MIST_hip.x input.par output.txt
//
// Reads the .par, runs the MIST Prim on the GPU and writes the same text as
// the reference: one line per tree edge
//   "<t_in> <t_out> <in+1> <out+1>   MI: <mi>   left to process: <n>"
// then the ten totals, all at precision(12) (get_PAR_MIST.cpp:161-162,
// 663-741, 754-765).
#include "../../include/parent_hip.h"

#include <sys/time.h>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

int main(int argc, char* argv[]) {
    timeval tv1, tv2, tv3;
    gettimeofday(&tv1, nullptr);
    if (argc != 3) {
        std::cerr << "USAGE:\n" << argv[0] << " input.par output.txt" << std::endl;
        return 1;
    }
    int nD = 0, N = 0, bins[6];
    if (parent_par_read(argv[1], &nD, &N, bins, nullptr, nullptr) != PARENT_OK) {
        std::cerr << parent_last_error() << std::endl;
        std::cerr << "An ERROR has occurred while reading the file " << argv[1] << " . Quitting Program.\n";
        return 1;
    }
    if (nD < 1) {
        std::cerr << "ERROR: at least one dihedral is needed." << std::endl;
        return 1;
    }
    const size_t D = 3 * (size_t)nD + 3, P = D * (D - 1) / 2;
    std::vector<double> s1d(D), s2d(P);
    if (parent_par_read(argv[1], &nD, &N, bins, s1d.data(), s2d.data()) != PARENT_OK) {
        std::cerr << parent_last_error() << std::endl;
        return 1;
    }
    std::ofstream out(argv[2], std::ios::out);
    if (!out.is_open()) {
        std::cerr << "ERROR: COULD NOT OPEN FILE " << argv[2] << " !" << std::endl;
        return 1;
    }
    out.precision(12);
    const char* dev_env = getenv("PARENT_HIP_DEVICE");
    parent_engine* eng = nullptr;
    if (parent_engine_create(dev_env ? atoi(dev_env) : 0, &eng) != PARENT_OK) {
        std::cerr << "ERROR: " << parent_last_error() << std::endl;
        return 1;
    }
    gettimeofday(&tv2, nullptr);
    std::vector<int32_t> edges(4 * (D - 1)), left(D - 1);
    std::vector<double> mi(D - 1), tot(PARENT_TOT_COUNT);
    if (parent_engine_set_entropies(eng, nD, s1d.data(), s2d.data()) != PARENT_OK ||
        parent_engine_mist(eng, edges.data(), mi.data(), left.data(), tot.data()) != PARENT_OK) {
        std::cerr << "ERROR: " << parent_engine_error(eng) << std::endl;
        parent_engine_destroy(eng);
        return 1;
    }
    parent_engine_destroy(eng);
    static const char kType[3] = {'b', 'a', 'd'};
    for (size_t s = 0; s + 1 < D; s++) {
        out << kType[edges[4 * s]] << " " << kType[edges[4 * s + 1]] << " " << edges[4 * s + 2] + 1 << " "
            << edges[4 * s + 3] + 1 << "   MI: " << mi[s] << "   left to process: " << left[s] << std::endl;
    }
    out << "TOTAL 1D BONDS ENTROPY = " << tot[PARENT_TOT_S1D_B] << std::endl;
    out << "TOTAL 1D ANGLES ENTROPY = " << tot[PARENT_TOT_S1D_A] << std::endl;
    out << "TOTAL 1D DIHEDRALS ENTROPY = " << tot[PARENT_TOT_S1D_D] << std::endl;
    out << "TOTAL 2D BONDS-BONDS MUTUAL INFORMATION = " << tot[PARENT_TOT_MI_BB] << std::endl;
    out << "TOTAL 2D BONDS-ANGLES MUTUAL INFORMATION = " << tot[PARENT_TOT_MI_BA] << std::endl;
    out << "TOTAL 2D BONDS-DIHEDRALS MUTUAL INFORMATION = " << tot[PARENT_TOT_MI_BD] << std::endl;
    out << "TOTAL 2D ANGLES-ANGLES MUTUAL INFORMATION = " << tot[PARENT_TOT_MI_AA] << std::endl;
    out << "TOTAL 2D ANGLES-DIHEDRALS MUTUAL INFORMATION = " << tot[PARENT_TOT_MI_AD] << std::endl;
    out << "TOTAL 2D DIHEDRALS-DIHEDRALS MUTUAL INFORMATION = " << tot[PARENT_TOT_MI_DD] << std::endl;
    out << "TOTAL CONFIGURATIONAL ENTROPY = " << tot[PARENT_TOT_TOTAL] << std::endl;
    out.close();
    gettimeofday(&tv3, nullptr);
    std::cout << "Initialization + readin time: " << tv2.tv_sec + 0.000001 * tv2.tv_usec - tv1.tv_sec - 0.000001 * tv1.tv_usec << std::endl;
    std::cout << "Calculation time: " << tv3.tv_sec + 0.000001 * tv3.tv_usec - tv2.tv_sec - 0.000001 * tv2.tv_usec << std::endl;
    std::cout << "Total execution time: " << tv3.tv_sec + 0.000001 * tv3.tv_usec - tv1.tv_sec - 0.000001 * tv1.tv_usec << std::endl;
    std::cout << "PROGRAM FINISHED SUCCESSFULLY." << std::endl;
    return 0;
}
