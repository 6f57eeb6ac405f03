// Internal host-side helpers shared by the IO, synthetic-data and engine
// translation units.  Not part of the C ABI.
#pragma once

#include <cstdint>
#include <string>

namespace parent {

// global dof layout: bonds | angles | dihedrals (PARENT.cpp:298-299)
inline int n_bonds(int nD) { return nD + 2; }
inline int n_angles(int nD) { return nD + 1; }
inline int n_dofs(int nD) { return 3 * nD + 3; }
inline int first_of_type(int nD, int t) { return t == 0 ? 0 : (t == 1 ? nD + 2 : 2 * nD + 3); }
inline int count_of_type(int nD, int t) { return t == 0 ? nD + 2 : (t == 1 ? nD + 1 : nD); }
inline int type_of(int nD, int g) { return g < nD + 2 ? 0 : (g < 2 * nD + 3 ? 1 : 2); }
inline int64_t n_pairs(int nD) {
    int64_t D = n_dofs(nD);
    return D * (D - 1) / 2;
}

// Offset of pair (gi < gj) inside the .par body's 2D section
// (io_binary.cpp:606-631 block order; reverse half-matrix index of
// PARENT.cpp:1107-1112 / util.cpp:52-57).
inline int64_t par_pos(int nD, int gi, int gj) {
    const int64_t nB = nD + 2, nA = nD + 1, nd = nD;
    const int t1 = type_of(nD, gi), t2 = type_of(nD, gj);
    const int64_t i = gi - first_of_type(nD, t1), j = gj - first_of_type(nD, t2);
    const int64_t oBA = nB * (nB - 1) / 2, oBD = oBA + nB * nA, oAA = oBD + nB * nd,
                  oAD = oAA + nA * (nA - 1) / 2, oDD = oAD + nA * nd;
    if (t1 == 0 && t2 == 0) return (nB - i) * (nB - i - 1) / 2 + i - j;
    if (t1 == 0 && t2 == 1) return oBA + i * nA + j;
    if (t1 == 0 && t2 == 2) return oBD + i * nd + j;
    if (t1 == 1 && t2 == 1) return oAA + (nA - i) * (nA - i - 1) / 2 + i - j;
    if (t1 == 1 && t2 == 2) return oAD + i * nd + j;
    return oDD + (nd - i) * (nd - i - 1) / 2 + i - j;
}

void set_last_error(const std::string& msg);

}  // namespace parent
