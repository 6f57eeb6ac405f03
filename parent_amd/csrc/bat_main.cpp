// BAT_builder_hip.x -- the .xtc -> .bat direction of the reference's
// BAT_builder.x, same command line (bat.cpp:82-176):
//
//   BAT_builder_hip.x -t input.top -x input.xtc -o output.bat -bb "CA C N ..." [--single_precision]
//
// Writes the same .bat bytes (double precision unless --single_precision).
// The reverse direction (-b input.bat -o output.xtc) is not provided.
// OMP_NUM_THREADS threads decode and convert frames.
#include "../../include/parent_hip.h"

#include <cstring>
#include <iostream>
#include <string>

namespace {

const char* option(int argc, char** argv, const char* name) {
    for (int i = 1; i + 1 < argc; i++)
        if (strcmp(argv[i], name) == 0) return argv[i + 1];
    return nullptr;
}
bool has(int argc, char** argv, const char* name) {
    for (int i = 1; i < argc; i++)
        if (strcmp(argv[i], name) == 0) return true;
    return false;
}
std::string extension(const std::string& f) {
    const size_t p = f.rfind('.');
    return p == std::string::npos ? f : f.substr(p + 1);
}
int usage(const char* prog) {
    std::cerr << "USAGE: " << prog
              << " -t input.top -x input.xtc -o output.bat -bb \"BackboneAtomName1 BackboneAtomName2 "
                 "BackboneAtomName3 ...\" [--single_precision]\n";
    return 1;
}

}  // namespace

int main(int argc, char* argv[]) {
    if (argc == 5 && has(argc, argv, "-b")) {
        std::cerr << "ERROR: the .bat -> .xtc direction is not provided by this program.\n";
        return 1;
    }
    if (argc != 9 && argc != 10) return usage(argv[0]);
    const char *top = option(argc, argv, "-t"), *xtc = option(argc, argv, "-x"), *bat = option(argc, argv, "-o"),
               *bb = option(argc, argv, "-bb");
    if (!top || !xtc || !bat || !bb || (argc == 10 && !has(argc, argv, "--single_precision"))) return usage(argv[0]);
    if (extension(top) != "top" || extension(xtc) != "xtc" || extension(bat) != "bat") return usage(argv[0]);
    std::cout << "Reading topology, building the BAT topology and writing the .bat trajectory." << std::endl;
    int frames = 0;
    if (parent_bat_build(top, xtc, bat, bb, argc == 10 ? 0 : 1, 0, &frames) != PARENT_OK) {
        std::cerr << parent_last_error() << std::endl;
        return 1;
    }
    std::cout << "Processed " << frames << " frames." << std::endl;
    return 0;
}
