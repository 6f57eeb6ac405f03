// PARENT_hip.x -- drop-in replacement for the reference's PARENT.x on one GPU.
//
//   PARENT_hip.x input.bat entropy.par #b1D #a1D #d1D #b2D #a2D #d2D
//
// Same command line, input/output formats, stdout tail (the four timing lines
// and "PROGRAM FINISHED SUCCESSFULLY.", PARENT.cpp:1501-1509) and exit codes
// (1 on any error, PARENT.cpp:55 RC_ABORT) as the reference, so run.sh:40 and
// the .par consumers work unchanged.  Optional environment knob:
//   PARENT_HIP_METHOD = auto | plain | atomic | mfma | slabs   (pair kernel, default auto)
//   PARENT_HIP_DEVICE = <n>                            (HIP device, default 0)
#include "../../include/parent_hip.h"

#include <sys/time.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

namespace {

// Last '.'-separated token, skipping empty tokens like strtok (PARENT.cpp:196-210).
std::string extension(const char* path) {
    std::string s(path), last;
    size_t i = 0;
    while (i <= s.size()) {
        size_t j = s.find('.', i);
        if (j == std::string::npos) j = s.size();
        if (j > i) last = s.substr(i, j - i);
        i = j + 1;
    }
    return last;
}

double since(const timeval& a, const timeval& b) {
    return b.tv_sec + 0.000001 * b.tv_usec - a.tv_sec - 0.000001 * a.tv_usec;
}

void usage(const char* prog) {
    std::cerr << "USAGE: " << prog
              << " input.bat entropy.par #bondsbins1D #anglesbins1D #dihedralsbins1D #bondsbins2D #anglesbins2D #dihedralsbins2D\n";
}

int method_from_env() {
    const char* m = getenv("PARENT_HIP_METHOD");
    if (!m || !strcmp(m, "auto")) return PARENT_METHOD_AUTO;
    if (!strcmp(m, "plain")) return PARENT_METHOD_PLAIN;
    if (!strcmp(m, "atomic")) return PARENT_METHOD_ATOMIC;
    if (!strcmp(m, "mfma")) return PARENT_METHOD_MFMA;
    if (!strcmp(m, "slabs")) return PARENT_METHOD_SLABS;
    return PARENT_METHOD_AUTO;
}

}  // namespace

int main(int argc, char* argv[]) {
    timeval tv1, tv2, tv3, tv4;
    gettimeofday(&tv1, nullptr);
    std::cout << "PARENT_hip: configurational entropy (MIE) of a BAT trajectory on one MI355X GPU" << std::endl;
    if (argc != 9) {
        usage(argv[0]);
        return 1;
    }
    if (extension(argv[1]) != "bat" || extension(argv[2]) != "par") {
        usage(argv[0]);
        return 1;
    }
    static const char* what[6] = {"1D bonds", "1D angles", "1D dihedrals", "2D bonds", "2D angles", "2D dihedrals"};
    int bins[6];
    for (int k = 0; k < 6; k++) {
        if (sscanf(argv[3 + k], "%d", &bins[k]) != 1) {
            std::cerr << "ERROR: Could not read number of bins for " << what[k] << " from command line! Aborting" << std::endl;
            return 1;
        }
    }
    std::cout << "Reading file " << argv[1] << " ." << std::endl;
    parent_bat* bat = nullptr;
    int rc = parent_bat_open(argv[1], &bat);
    if (rc != PARENT_OK) {
        std::cerr << parent_last_error() << std::endl;
        std::cerr << "AN ERROR HAS OCCURED WHILE READING THE HEADER OF THE FILE " << argv[1] << " . QUITTING PROGRAM.\n";
        return 1;
    }
    int nD, N, dp;
    parent_bat_info(bat, &nD, &N, &dp);
    std::cout << argv[1] << " specs:" << std::endl;
    std::cout << "Precision: " << (dp == 1 ? "double" : "single") << " #Atoms: " << nD + 3 << " #Frames: " << N << std::endl;
    if (nD < 1) {
        std::cerr << "ERROR: at least 4 atoms (one dihedral) are needed." << std::endl;
        parent_bat_close(bat);
        return 1;
    }
    const size_t D = 3 * (size_t)nD + 3;
    const char* dev_env = getenv("PARENT_HIP_DEVICE");
    parent_engine* eng = nullptr;
    if ((rc = parent_engine_create(dev_env ? atoi(dev_env) : 0, &eng)) != PARENT_OK) {
        std::cerr << "ERROR: " << parent_last_error() << std::endl;
        parent_bat_close(bat);
        return 1;
    }
    auto die = [&](const char* where) {
        std::cerr << "ERROR (" << where << "): " << parent_engine_error(eng) << std::endl;
        parent_engine_destroy(eng);
        parent_bat_close(bat);
        return 1;
    };
    // frames stream from the file through pinned buffers straight into HBM
    if ((rc = parent_engine_load_bat(eng, bat)) != PARENT_OK) {
        if (rc == PARENT_E_IO) {
            std::cerr << parent_engine_error(eng) << std::endl;
            std::cerr << "An ERROR has occurred while reading the file " << argv[1] << " . Quitting Program.\n";
            parent_engine_destroy(eng);
            parent_bat_close(bat);
            return 1;
        }
        return die("read-in");
    }
    std::cout << "READIN AND DISTRIBUTION COMPLETED." << std::endl;
    gettimeofday(&tv2, nullptr);
    if (parent_engine_prepare(eng, bins) != PARENT_OK) return die("extrema / pre-binning / 1D");
    std::cout << "CALCULATION OF EXTREMA AND 1D ENTROPIES COMPLETED." << std::endl;
    if (parent_engine_pairs(eng, method_from_env()) != PARENT_OK) return die("2D entropies");
    std::cout << "2D CALCULATION COMPLETED." << std::endl;
    const long long P = (long long)D * (D - 1) / 2;
    std::vector<double> s1d(D), s2d(P);
    if (parent_engine_read_s1d(eng, s1d.data()) != PARENT_OK || parent_engine_read_s2d(eng, s2d.data()) != PARENT_OK)
        return die("copy-back");
    gettimeofday(&tv3, nullptr);
    if (parent_par_write(argv[2], bat, bins, s1d.data(), s2d.data()) != PARENT_OK) {
        std::cerr << parent_last_error() << std::endl;
        std::cerr << "AN ERROR HAS OCCURED WHILE WRITING THE FILE " << argv[2] << " .\n";
    }
    gettimeofday(&tv4, nullptr);
    parent_stats st{};
    parent_engine_stats(eng, &st);
    std::cout << "GPU phases [ms]: upload " << st.ms_upload << ", prepare " << st.ms_prepare << ", pairs "
              << st.ms_pairs << " (method " << st.method << ", variant " << st.kernel_variant << ")" << std::endl;
    parent_engine_destroy(eng);
    parent_bat_close(bat);
    std::cout << "Initialization + readin time: " << since(tv1, tv2) << std::endl;
    std::cout << "Calculation time: " << since(tv2, tv3) << std::endl;
    std::cout << "Output time: " << since(tv3, tv4) << std::endl;
    std::cout << "Total execution time: " << since(tv1, tv4) << std::endl;
    std::cout << "PROGRAM FINISHED SUCCESSFULLY." << std::endl;
    return 0;
}
