// BAT builder: GROMACS topology (.top) + compressed trajectory (.xtc) ->
// bond-angle-torsion trajectory (.bat), the input of the entropy engine.
// Restates the reference's BAT_builder.x (.xtc -> .bat direction):
//   topology      topology.cpp:41-714    (sections, molecules, bonds table, root atoms)
//   BAT tree      BAT_topology.cpp:45-384 (breadth-first numbering, dihedrals, phase angles)
//   conversion    BAT_trajectory.cpp:77-299 (Cartesian -> BAT per frame, phase angles, [0, 2pi))
//   .bat writer   io_binary.cpp:41-102, 205-278
// The same rules give the same dihedral list, and the per-frame arithmetic is
// evaluated in the reference's precision and order (coordinates in float,
// geometry in double, libm acos/atan2/sqrt, no contraction), so the output is
// byte-identical.  Unlike the reference, frames are decoded and converted in
// parallel (OpenMP) from an in-memory copy of the .xtc.
//
// The .xtc decoder follows the published XTC format (xdrfile's
// xdr3dfcoord): big-endian XDR header fields, then per frame a bit stream of
// integer coordinates -- full-size triples, or runs of small deltas whose
// width follows the "magic integer" table, the first two atoms of a run
// swapped -- scaled by 1 / precision in single precision.
#include "../../include/parent_hip.h"
#include "host_common.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>
#include <omp.h>

using parent::set_last_error;

namespace {

// ---------------------------------------------------------------------------
// topology (topology.cpp)
// ---------------------------------------------------------------------------

// blanks only (not tabs) are stripped, as strip_line / strip_blanks do
// (io_text.cpp:37-92)
std::string strip_line(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && s[a] == ' ') a++;
    while (b > a && s[b - 1] == ' ') b--;
    return s.substr(a, b - a);
}
std::string strip_blanks(const std::string& s) {
    std::string o;
    for (char c : s)
        if (c != ' ') o += c;
    return o;
}
// 1: comment (first non-blank is ';'), 0: content, 2: blank (topology.cpp:283-296)
int comment_kind(const std::string& s) {
    for (char c : s)
        if (c != ' ') return c == ';' ? 1 : 0;
    return 2;
}
std::string dir_of(const std::string& f) {
    const size_t p = f.rfind('/');
    return p == std::string::npos ? std::string("./") : f.substr(0, p + 1);
}
bool file_exists(const std::string& f) {
    std::ifstream t(f.c_str());
    return t.is_open();
}

struct Topology {
    // per moleculetype
    std::vector<std::string> mt_names;
    std::vector<std::vector<int>> mt_atoms;               // atoms in order of first appearance in [ bonds ]
    std::vector<std::vector<std::array<int, 2>>> mt_bonds;
    std::vector<std::vector<float>> mt_masses;
    std::vector<std::vector<std::string>> mt_res, mt_atomnames;
    std::vector<std::vector<int>> mt_resnum, mt_backbone;
    bool new_bonds = false, new_masses = false, new_atoms = false, new_backbone = false;
    // system
    std::vector<std::string> mol_names;
    std::vector<std::vector<int>> molecules;
    std::vector<std::vector<std::array<int, 2>>> bonds;
    std::vector<float> masses;
    std::vector<std::string> residues, atom_names, belongs;
    std::vector<int> resnums, backbone;
    int total_atoms = 0;
    std::vector<std::vector<int>> table;  // bonds table: neighbours of atom k (1-based) at [k-1]
    std::vector<std::array<int, 4>> roots;
    std::string err;
};

// #include expansion into a list of content lines (topology.cpp:298-375)
bool preprocess(const std::string& file, std::vector<std::string>* out, std::string* err) {
    std::ifstream in(file.c_str());
    if (!in.is_open()) {
        *err = "ERROR: COULD NOT OPEN FILE \"" + file + "\"";
        return false;
    }
    const std::string dir = dir_of(file);
    std::string raw;
    while (std::getline(in, raw)) {
        const std::string line = strip_line(raw);
        char c1[1000] = {0}, c2[1000] = {0};
        sscanf(line.c_str(), "%999s %999s", c1, c2);
        if (std::string(c1) == "#include") {
            std::string name(c2);
            if (name.size() >= 2) name = name.substr(1, name.size() - 2);  // drop the quotes
            bool found = false;
            for (const std::string& cand : {dir + name, std::string("/usr/local/gromacs/share/gromacs/top/") + name}) {
                if (!line.empty() && file_exists(cand)) {
                    if (!preprocess(cand, out, err)) return false;
                    found = true;
                    break;
                }
            }
            if (!found && comment_kind(line) == 0) {
                out->push_back(line);
                fprintf(stderr, "WARNING FROM PREPROCESSOR: file %s NOT FOUND. If the file does not contain any "
                                "bond-structure information of your molecule(s), this should be fine.\n",
                        name.c_str());
            }
        } else if (comment_kind(line) == 0 && c1[0] != '\0') {
            out->push_back(line);
        }
    }
    return true;
}

void new_moleculetype(Topology& t, const std::string& name) {
    t.mt_names.push_back(name);
    t.new_bonds = t.new_masses = t.new_atoms = t.new_backbone = true;
    // placeholders overwritten by the first bond / backbone atom (topology.cpp:152-174)
    t.mt_backbone.push_back({-10});
    t.mt_atoms.push_back({-10, -10});
    t.mt_bonds.push_back({{-10, -10}});
}

void add_bond(Topology& t, int a, int b) {
    const int cur = (int)t.mt_names.size() - 1;
    if (t.new_bonds) {
        t.mt_atoms[cur][0] = a;
        t.mt_atoms[cur][1] = b;
        t.mt_bonds[cur][0] = {a, b};
        t.new_bonds = false;
        return;
    }
    t.mt_bonds[cur].push_back({a, b});
    for (int x : {a, b})
        if (std::find(t.mt_atoms[cur].begin(), t.mt_atoms[cur].end(), x) == t.mt_atoms[cur].end())
            t.mt_atoms[cur].push_back(x);
}

bool read_topology(const std::string& top, const std::vector<std::string>& bb_names, Topology& t) {
    std::vector<std::string> lines;
    if (!preprocess(top, &lines, &t.err)) return false;
    bool atomsf = false, bondsf = false, mtf = false, molsf = false;
    int has_atoms = 1, has_bonds = 1;
    // fields keep their previous values when a line has fewer of them, as the
    // reference's scanf targets (declared once, topology.cpp:385-390) do
    int n1 = 0, n2 = 0, n3 = 0;
    float charge = 0.f, mass = 0.f;
    char c1[1000] = {0}, c2[1000] = {0}, c3[1000] = {0};
    auto check_prev = [&](int idx) {
        if (has_atoms == 0) {
            t.err = "ERROR: NO \"[ atoms ]\" SECTION FOUND FOR MOLECULETYPE " + t.mt_names[idx];
            return false;
        }
        if (has_bonds == 0 && t.mt_masses.size() > (size_t)idx && t.mt_masses[idx].size() > 1)
            fprintf(stderr, "WARNING: NO \"[ bonds ]\" SECTION FOUND FOR MOLECULETYPE %s\n", t.mt_names[idx].c_str());
        if (has_atoms > 1 || has_bonds > 1) {
            t.err = "ERROR: SEVERAL \"[ atoms ]\" OR \"[ bonds ]\" SECTIONS FOUND FOR MOLECULETYPE " + t.mt_names[idx];
            return false;
        }
        return true;
    };
    for (const std::string& raw : lines) {
        const std::string line = strip_line(raw);
        const char c0 = line.empty() ? '\0' : line[0];
        if (mtf && c0 != '[' && c0 != '#' && comment_kind(line) == 0) {
            sscanf(line.c_str(), "%999s %d", c1, &n2);
            new_moleculetype(t, c1);
            if (t.mt_names.size() > 1 && !check_prev((int)t.mt_names.size() - 2)) return false;
            has_atoms = has_bonds = 0;
        }
        if (bondsf && c0 != '[' && c0 != '#' && comment_kind(line) == 0) {
            if (t.mt_names.empty()) {
                t.err = "ERROR: ENCOUNTERED [ bonds ] SECTION BEFORE [ moleculetype ] SECTION!";
                return false;
            }
            sscanf(line.c_str(), "%d %d", &n1, &n2);
            add_bond(t, n1, n2);
        }
        if (atomsf && c0 != '[' && c0 != '#' && comment_kind(line) == 0) {
            if (t.mt_names.empty()) {
                t.err = "ERROR: ENCOUNTERED [ atoms ] SECTION BEFORE [ moleculetype ] SECTION!";
                return false;
            }
            sscanf(line.c_str(), "%d %999s %d %999s %999s %d %f %f", &n1, c1, &n2, c2, c3, &n3, &charge, &mass);
            const int cur = (int)t.mt_names.size() - 1;
            if (t.new_masses) {
                t.mt_masses.push_back({mass});
                t.new_masses = false;
            } else {
                t.mt_masses[cur].push_back(mass);
            }
            if (t.new_atoms) {
                t.mt_res.push_back({c2});
                t.mt_resnum.push_back({n2});
                t.mt_atomnames.push_back({c3});
                t.new_atoms = false;
            } else {
                t.mt_res[cur].push_back(c2);
                t.mt_resnum[cur].push_back(n2);
                t.mt_atomnames[cur].push_back(c3);
            }
            bool is_bb = false;
            for (const std::string& b : bb_names) is_bb = is_bb || b == c3;
            if (is_bb) {
                if (t.new_backbone) {
                    t.mt_backbone[cur][0] = n1;
                    t.new_backbone = false;
                } else {
                    t.mt_backbone[cur].push_back(n1);
                }
            }
        }
        if (molsf && c0 != '[' && c0 != '#' && comment_kind(line) == 0) {
            sscanf(line.c_str(), "%999s %d", c1, &n2);
            for (int i = 0; i < n2; i++) t.mol_names.push_back(c1);
            if (!t.mt_names.empty() && !check_prev((int)t.mt_names.size() - 1)) return false;
        }
        if (c0 == '[') atomsf = bondsf = mtf = molsf = false;
        if (c0 == '[' && line.back() == ']') {
            const std::string h = strip_blanks(line);
            if (h == "[atoms]") {
                atomsf = true;
                has_atoms++;
            } else if (h == "[bonds]") {
                bondsf = true;
                has_bonds++;
            } else if (h == "[moleculetype]") {
                mtf = true;
            } else if (h == "[molecules]") {
                molsf = true;
            }
        }
    }
    if (t.mt_names.empty()) {
        t.err = "ERROR: NO MOLECULETYPES FOUND IN (PREPROCESSED) TOPOLOGY FILE \"" + top + "\"!";
        return false;
    }
    if (t.mol_names.empty()) {
        t.err = "ERROR: NO MOLECULES FOUND IN THE \"[ molecules ]\" SECTION OF THE (PREPROCESSED) TOPOLOGY FILE \"" +
                top + "\"!";
        return false;
    }
    if (t.mt_masses.size() < t.mt_names.size() || t.mt_res.size() < t.mt_names.size()) {
        t.err = "ERROR: A MOLECULETYPE HAS NO [ atoms ] ENTRIES";
        return false;
    }
    // assemble the system in [ molecules ] order (topology.cpp:178-266)
    std::vector<int> copies(t.mt_names.size(), 0), seen(t.mt_names.size(), 1);
    for (size_t i = 0; i < t.mt_names.size(); i++)
        for (const std::string& m : t.mol_names) copies[i] += m == t.mt_names[i];
    int offset = 0;
    for (const std::string& m : t.mol_names) {
        size_t j = 0;
        while (j < t.mt_names.size() && t.mt_names[j] != m) j++;
        if (j == t.mt_names.size()) {
            t.err = "ERROR: MOLECULETYPE \"" + m + "\" DOES NOT EXIST!";
            return false;
        }
        std::vector<int> mol = t.mt_atoms[j];
        for (int& a : mol) a += offset;
        t.molecules.push_back(mol);
        for (float mm : t.mt_masses[j]) t.masses.push_back(mm);
        for (int b : t.mt_backbone[j]) t.backbone.push_back(b + offset);
        std::vector<std::array<int, 2>> bl = t.mt_bonds[j];
        for (auto& b : bl) {
            b[0] += offset;
            b[1] += offset;
        }
        t.bonds.push_back(bl);
        for (size_t k = 0; k < t.mt_res[j].size(); k++) {
            t.residues.push_back(t.mt_res[j][k]);
            t.resnums.push_back(t.mt_resnum[j][k]);
            t.atom_names.push_back(t.mt_atomnames[j][k]);
            t.belongs.push_back(copies[j] <= 1 ? m : m + "_" + std::to_string(seen[j]));
        }
        seen[j]++;
        offset += (int)t.mt_atoms[j].size();
    }
    t.total_atoms = offset;
    // bonds table: for atom k the partners of every bond containing it, in
    // molecule and bond order (topology.cpp:566-590)
    t.table.assign(t.total_atoms, {});
    for (const auto& bl : t.bonds)
        for (const auto& b : bl) {
            if (b[0] >= 1 && b[0] <= t.total_atoms) t.table[b[0] - 1].push_back(b[1]);
            if (b[1] >= 1 && b[1] <= t.total_atoms) t.table[b[1] - 1].push_back(b[0]);
        }
    // root atoms (topology.cpp:594-646): a terminal atom r0 whose partner r1
    // has exactly one non-terminal partner r2 (two as a fallback), the one
    // with the fewest other terminal partners
    auto deg = [&](int a) { return (a >= 1 && a <= t.total_atoms) ? (int)t.table[a - 1].size() : 0; };
    for (size_t i = 0; i < t.molecules.size(); i++) {
        const std::vector<int>& mol = t.molecules[i];
        int lowest = (int)mol.size() + 1;
        std::array<int, 4> r = {0, 0, 0, (int)mol.size()};
        bool done = false;
        for (int minc = 1; minc <= 2 && !done; minc++) {
            for (int a : mol) {
                if (deg(a) != 1) continue;
                const int b = t.table[a - 1][0];
                if (b < 1 || b > t.total_atoms) continue;
                int counter = 0, impropers = -1, sav = -1;
                for (int c : t.table[b - 1]) {
                    if (deg(c) > 1) {
                        counter++;
                        sav = c;
                    }
                    if (deg(c) == 1) impropers++;
                }
                if (counter == minc && impropers < lowest) {
                    lowest = impropers;
                    r[0] = a;
                    r[1] = b;
                    r[2] = sav;
                }
            }
            if (r[0] == 0 || r[1] == 0 || r[2] == 0) {
                if (minc < 2) continue;
                t.err = "ERROR: COULD NOT FIND ROOT ATOMS FOR MOLECULE " + std::to_string(i + 1) + ".";
                return false;
            }
            if (minc == 2)
                fprintf(stderr, "WARNING: COULD NOT FIND A CANONICAL ROOT ATOM SET FOR MOLECULE %zu. USING A DIHEDRAL "
                                "ANGLE WITHOUT A PHYSICAL BONDS BASIS.\n", i);
            t.roots.push_back(r);
            done = true;
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
// BAT tree (BAT_topology.cpp)
// ---------------------------------------------------------------------------

struct BatTree {
    std::vector<std::array<int, 6>> dih;  // atoms 1..4, type (0 physical, 1 pseudo, -1 improper), parent (-1 or 1-based)
    std::string err;
};

bool build_tree(const Topology& t, BatTree& out) {
    const int N = (int)t.table.size();
    std::vector<int> to_real(N, -1), to_bat(N, -1);
    int cur = 1;  // next BAT number
    auto take = [&](int real, std::vector<int>* list) {
        // remove `real` from the list, scanning forward as the reference does
        // (an erased element's successor is not re-examined), then number its
        // unnumbered partners (BAT_topology.cpp:144-163)
        for (size_t i = 0; i < list->size(); i++)
            if ((*list)[i] == real) list->erase(list->begin() + i);
        for (int c : t.table[real - 1])
            if (to_bat[c - 1] == -1) {
                to_bat[c - 1] = cur;
                to_real[cur - 1] = c;
                cur++;
                list->push_back(c);
            }
    };
    for (size_t j = 0; j < t.roots.size(); j++) {
        const auto& r = t.roots[j];
        const int first = cur;
        for (int k = 0; k < 3; k++) {
            to_real[cur - 1 + k] = r[k];
            to_bat[r[k] - 1] = cur + k;
        }
        cur += 3;
        std::vector<int> todo = {r[1], r[2]};
        take(r[1], &todo);
        take(r[2], &todo);
        while (!todo.empty()) {
            std::vector<int> next = todo;
            for (int x : todo) take(x, &next);
            todo = next;
        }
        if (cur - first != r[3]) {
            out.err = "ERROR BUILDING BAT TOPOLOGY. THE TOPOLOGY (MOLECULETYPE) OF MOLECULE NUMBER " + std::to_string(j) +
                      " IS NOT FULLY CONNECTED BY COVALENT BONDS (2 OR MORE MOLECULES?).";
            return false;
        }
    }
    // partner of `real` with the smallest BAT number, outside `excl` (BAT_topology.cpp:166-190)
    auto closest = [&](int real, const std::vector<int>& excl, bool* ok) {
        int best = N + 1;
        for (int c : t.table[real - 1])
            if (to_bat[c - 1] < best && std::find(excl.begin(), excl.end(), c) == excl.end()) best = to_bat[c - 1];
        if (best == N + 1 || best < 1) {
            *ok = false;
            return 0;
        }
        return to_real[best - 1];
    };
    int offset = 0;
    for (size_t j = 0; j < t.roots.size(); j++) {
        const auto& r = t.roots[j];
        if (j > 0) {
            const auto& p = t.roots[j - 1];
            offset += p[3];
            // three pseudo dihedrals chain the molecule's roots to the previous ones
            out.dih.push_back({p[0], p[1], p[2], r[0], 1, -1});
            out.dih.push_back({p[1], p[2], r[0], r[1], 1, -1});
            out.dih.push_back({p[2], r[0], r[1], r[2], 1, -1});
        }
        for (int c : t.table[r[1] - 1])
            if (c != r[0] && c != r[2]) out.dih.push_back({r[2], r[0], r[1], c, -1, -1});  // impropers
        const int start = offset + 3 + (int)t.table[r[1] - 1].size() - 2;
        for (int i = start; i < r[3] + offset; i++) {
            bool ok = true;
            const int d3 = to_real[i];
            const int d2 = closest(d3, {}, &ok);
            const int d1 = ok ? closest(d2, {d3}, &ok) : 0;
            const int d0 = ok ? closest(d1, {d3, d2}, &ok) : 0;
            if (!ok) {
                out.err = "ERROR BUILDING BAT TOPOLOGY.";
                return false;
            }
            out.dih.push_back({d0, d1, d2, d3, 0, -1});
        }
    }
    offset += t.roots.back()[3];
    if ((int)out.dih.size() != offset - 3) {
        out.err = "ERROR: " + std::to_string(offset - 3) + " DIHEDRALS/IMPROPERS SHOULD HAVE BEEN FOUND, BUT ONLY " +
                  std::to_string(out.dih.size()) + " WERE ACTUALLY FOUND.";
        return false;
    }
    // phase angles (BAT_topology.cpp:297-380): dihedrals sharing their first
    // three atoms form a group; its first all-backbone member (else its first
    // member) is the parent of the others
    const int n = (int)out.dih.size();
    std::vector<char> used(n, 0);
    std::vector<std::vector<int>> groups;
    for (int i = 0; i < n; i++) {
        if (used[i]) continue;
        bool first = true;
        for (int j = i + 1; j < n; j++) {
            if (used[j]) continue;
            if (out.dih[i][0] == out.dih[j][0] && out.dih[i][1] == out.dih[j][1] && out.dih[i][2] == out.dih[j][2]) {
                if (first) {
                    groups.push_back({i, j});
                    used[i] = used[j] = 1;
                    first = false;
                } else {
                    groups.back().push_back(j);
                    used[j] = 1;
                }
            }
        }
    }
    std::vector<char> bb(N, 0);
    for (int b : t.backbone)
        if (b >= 1 && b <= N) bb[b - 1] = 1;
    for (const auto& g : groups) {
        int full = -1;
        for (size_t k = 0; k < g.size(); k++) {
            const auto& d = out.dih[g[k]];
            if (bb[d[0] - 1] && bb[d[1] - 1] && bb[d[2] - 1] && bb[d[3] - 1]) {
                full = (int)k;
                break;
            }
        }
        const int parent = g[full < 0 ? 0 : full];
        for (int m : g) out.dih[m][5] = m == parent ? -1 : parent + 1;
    }
    return true;
}

// ---------------------------------------------------------------------------
// .xtc decoding
// ---------------------------------------------------------------------------

// widths of the small-delta runs (the format's "magic integers", about
// 2^(i/3)); index 9 is the smallest usable one
const int kMagic[] = {0,       0,       0,       0,       0,       0,       0,       0,       0,       8,
                      10,      12,      16,      20,      25,      32,      40,      50,      64,      80,
                      101,     128,     161,     203,     256,     322,     406,     512,     645,     812,
                      1024,    1290,    1625,    2048,    2580,    3250,    4096,    5060,    6501,    8192,
                      10321,   13003,   16384,   20642,   26007,   32768,   41285,   52015,   65536,   82570,
                      104031,  131072,  165140,  208063,  262144,  330280,  416127,  524287,  660561,  832255,
                      1048576, 1321122, 1664510, 2097152, 2642245, 3329021, 4194304, 5284491, 6658042, 8388607,
                      10568983, 13316085, 16777216};
constexpr int kMagicFirst = 9;
constexpr int kMagicLast = (int)(sizeof(kMagic) / sizeof(kMagic[0]));

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }
float bef(const uint8_t* p) {
    const uint32_t u = be32(p);
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// bits of `v` (smallest n with v < 2^n)
int bits_of(unsigned v) {
    int n = 0;
    unsigned long long num = 1;
    while (v >= num && n < 32) {
        n++;
        num <<= 1;
    }
    return n;
}
// bits of the mixed-radix product sizes[0] * sizes[1] * sizes[2]
int bits_of_product(const unsigned sizes[3]) {
    unsigned bytes[32] = {1};
    int nbytes = 1;
    for (int i = 0; i < 3; i++) {
        unsigned carry = 0;
        int k = 0;
        for (; k < nbytes; k++) {
            carry = bytes[k] * sizes[i] + carry;
            bytes[k] = carry & 0xFF;
            carry >>= 8;
        }
        while (carry) {
            bytes[k++] = carry & 0xFF;
            carry >>= 8;
        }
        nbytes = k;
    }
    int bits = 0;
    unsigned num = 1;
    while (bytes[nbytes - 1] >= num) {
        bits++;
        num *= 2;
    }
    return bits + (nbytes - 1) * 8;
}

// MSB-first bit reader over the frame's compressed bytes
struct Bits {
    const uint8_t* p;
    size_t n, at = 0;
    unsigned have = 0, acc = 0;  // bits pending in acc
    bool bad = false;
    uint8_t byte() {
        if (at >= n) {
            bad = true;
            return 0;
        }
        return p[at++];
    }
    int get(int nbits) {
        const int want = nbits;
        int v = 0;
        while (nbits >= 8) {
            acc = (acc << 8) | byte();
            v |= (int)((acc >> have) << (nbits - 8));
            nbits -= 8;
        }
        if (nbits > 0) {
            if ((int)have < nbits) {
                have += 8;
                acc = (acc << 8) | byte();
            }
            have -= nbits;
            v |= (int)((acc >> have) & ((1u << nbits) - 1));
        }
        return want >= 32 ? v : v & (int)((1u << want) - 1);
    }
    // three integers packed in mixed radix sizes[0..2] over `nbits` bits
    void get3(int nbits, const unsigned sizes[3], int out[3]) {
        int bytes[32] = {0};
        int nb = 0;
        while (nbits > 8) {
            bytes[nb++] = get(8);
            nbits -= 8;
        }
        if (nbits > 0) bytes[nb++] = get(nbits);
        for (int i = 2; i > 0; i--) {
            unsigned rem = 0;
            for (int j = nb - 1; j >= 0; j--) {
                rem = (rem << 8) | (unsigned)bytes[j];
                const unsigned q = rem / sizes[i];
                bytes[j] = (int)q;
                rem -= q * sizes[i];
            }
            out[i] = (int)rem;
        }
        out[0] = bytes[0] | (bytes[1] << 8) | (bytes[2] << 16) | (bytes[3] << 24);
    }
};

struct XtcFrame {
    size_t off;       // file offset of the frame
    int natoms, step;
    float time, box[9], prec;
};

// frame boundaries of an in-memory .xtc (a cheap sequential pass over the headers)
bool scan_xtc(const std::vector<uint8_t>& f, std::vector<XtcFrame>* frames, std::string* err) {
    size_t at = 0;
    while (at + 4 <= f.size()) {
        XtcFrame fr{};
        fr.off = at;
        if (at + 56 > f.size()) break;  // truncated header: the reference stops here too
        if (be32(&f[at]) != 1995) {
            *err = "ERROR: BAD MAGIC NUMBER IN .xtc FRAME " + std::to_string(frames->size() + 1);
            return false;
        }
        fr.natoms = (int)be32(&f[at + 4]);
        fr.step = (int)be32(&f[at + 8]);
        fr.time = bef(&f[at + 12]);
        for (int k = 0; k < 9; k++) fr.box[k] = bef(&f[at + 16 + 4 * k]);
        const int lsize = (int)be32(&f[at + 52]);
        size_t end;
        if (lsize <= 9) {
            end = at + 56 + 12 * (size_t)lsize;
        } else {
            if (at + 56 + 36 > f.size()) break;
            fr.prec = bef(&f[at + 56]);
            const size_t nbytes = be32(&f[at + 88]);
            end = at + 92 + ((nbytes + 3) & ~(size_t)3);
        }
        if (end > f.size()) break;  // a partial last frame is not read
        frames->push_back(fr);
        at = end;
    }
    return true;
}

// coordinates of one frame (nm, float) into x[3 * natoms]
bool decode_xtc_frame(const std::vector<uint8_t>& f, const XtcFrame& fr, float* x) {
    const uint8_t* p = &f[fr.off + 52];
    const int lsize = (int)be32(p);
    if (lsize != fr.natoms) return false;
    p += 4;
    if (lsize <= 9) {
        for (int k = 0; k < 3 * lsize; k++) x[k] = bef(p + 4 * k);
        return true;
    }
    const float prec = bef(p);
    int minint[3], maxint[3];
    for (int k = 0; k < 3; k++) {
        minint[k] = (int)be32(p + 4 + 4 * k);
        maxint[k] = (int)be32(p + 16 + 4 * k);
    }
    int smallidx = (int)be32(p + 28);
    const size_t nbytes = be32(p + 32);
    unsigned sizeint[3], bitsizeint[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++) sizeint[k] = (unsigned)(maxint[k] - minint[k] + 1);
    int bitsize;
    if ((sizeint[0] | sizeint[1] | sizeint[2]) > 0xFFFFFF) {
        for (int k = 0; k < 3; k++) bitsizeint[k] = bits_of(sizeint[k]);
        bitsize = 0;  // each coordinate on its own bit width
    } else {
        bitsize = bits_of_product(sizeint);
    }
    if (smallidx < kMagicFirst || smallidx >= kMagicLast) return false;
    int smaller = kMagic[std::max(kMagicFirst, smallidx - 1)] / 2;
    int smallnum = kMagic[smallidx] / 2;
    unsigned sizesmall[3] = {(unsigned)kMagic[smallidx], (unsigned)kMagic[smallidx], (unsigned)kMagic[smallidx]};
    Bits in{p + 36, nbytes};
    const float inv = (float)(1.0 / (double)prec);  // inv_precision is a float
    std::vector<int> ic(3 * (size_t)lsize + 3);
    float* out = x;
    int run = 0, i = 0;
    while (i < lsize) {
        int* cur = &ic[3 * (size_t)i];
        if (bitsize == 0) {
            for (int k = 0; k < 3; k++) cur[k] = in.get((int)bitsizeint[k]);
        } else {
            in.get3(bitsize, sizeint, cur);
        }
        i++;
        int prev[3];
        for (int k = 0; k < 3; k++) {
            cur[k] += minint[k];
            prev[k] = cur[k];
        }
        int is_smaller = 0;
        if (in.get(1) == 1) {
            run = in.get(5);
            is_smaller = run % 3;
            run -= is_smaller;
            is_smaller--;
        }
        if (run > 0) {
            if (i + run / 3 > lsize) return false;
            cur += 3;
            for (int k = 0; k < run; k += 3) {
                in.get3(smallidx, sizesmall, cur);
                i++;
                for (int d = 0; d < 3; d++) cur[d] += prev[d] - smallnum;
                if (k == 0) {
                    // the first two atoms of a run are stored swapped
                    for (int d = 0; d < 3; d++) std::swap(cur[d], prev[d]);
                    for (int d = 0; d < 3; d++) *out++ = (float)prev[d] * inv;
                } else {
                    for (int d = 0; d < 3; d++) prev[d] = cur[d];
                }
                for (int d = 0; d < 3; d++) *out++ = (float)cur[d] * inv;
            }
        } else {
            for (int d = 0; d < 3; d++) *out++ = (float)cur[d] * inv;
        }
        smallidx += is_smaller;
        if (smallidx < kMagicFirst || smallidx >= kMagicLast) return false;
        if (is_smaller < 0) {
            smallnum = smaller;
            smaller = smallidx > kMagicFirst ? kMagic[smallidx - 1] / 2 : 0;
        } else if (is_smaller > 0) {
            smaller = smallnum;
            smallnum = kMagic[smallidx] / 2;
        }
        sizesmall[0] = sizesmall[1] = sizesmall[2] = (unsigned)kMagic[smallidx];
    }
    return !in.bad && out == x + 3 * (size_t)lsize;
}

// ---------------------------------------------------------------------------
// Cartesian -> BAT (BAT_trajectory.cpp:208-299, 303-309, 196-201)
// ---------------------------------------------------------------------------

// x: float coordinates in Angstrom.  Differences of coordinates are taken
// in float (rvec arithmetic), the geometry in double.
struct FrameOut {
    double ext[6];  // root atom x, y, z; theta; phi; external dihedral
    std::vector<double> vals;  // bond0, bond1, angle0, then (bond, angle, dihedral) per dihedral
};

void to_bat(const std::vector<std::array<int, 6>>& top, const float* x, double pi, double* ext, double* vals) {
    const int nd = (int)top.size();
    auto X = [&](int atom, int k) { return x[3 * (size_t)(atom - 1) + k]; };
    double b1[3], b2[3], b3[3], A[3], C[3];
    const auto& t0 = top[0];
    for (int k = 0; k < 3; k++) {
        ext[k] = X(t0[0], k);
        b1[k] = ext[k];
        b2[k] = (double)(X(t0[1], k) - X(t0[0], k));
        b3[k] = (double)(X(t0[2], k) - X(t0[1], k));
    }
    const double phi = atan2(b2[1], b2[0]);
    const double rb0 = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
    const double rb1 = sqrt(b3[0] * b3[0] + b3[1] * b3[1] + b3[2] * b3[2]);
    const double theta = acos(b2[2] / rb0);
    auto cross = [](const double* u, const double* v, double* w) {
        w[0] = u[1] * v[2] - u[2] * v[1];
        w[1] = u[2] * v[0] - u[0] * v[2];
        w[2] = u[0] * v[1] - u[1] * v[0];
    };
    auto torsion = [](const double* A, const double* C, const double* b2, double len) {
        return atan2((A[2] * C[0] - A[0] * C[2]) * b2[1] / len - (A[2] * C[1] - A[1] * C[2]) * b2[0] / len -
                         (A[1] * C[0] - A[0] * C[1]) * b2[2] / len,
                     A[0] * C[0] + A[1] * C[1] + A[2] * C[2]);
    };
    cross(b1, b2, A);
    cross(b2, b3, C);
    ext[3] = theta;
    ext[4] = phi;
    ext[5] = torsion(A, C, b2, rb0);
    const double rangle = acos((b2[0] * b3[0] + b2[1] * b3[1] + b2[2] * b3[2]) / rb0 / rb1);
    vals[0] = rb0;
    vals[1] = rb1;
    vals[2] = rangle;
    double* v = vals + 3;
    for (int i = 0; i < nd; i++) {
        const auto& d = top[i];
        for (int k = 0; k < 3; k++) {
            b1[k] = (double)(X(d[1], k) - X(d[0], k));
            b2[k] = (double)(X(d[2], k) - X(d[1], k));
            b3[k] = (double)(X(d[3], k) - X(d[2], k));
        }
        const double bond = sqrt(b3[0] * b3[0] + b3[1] * b3[1] + b3[2] * b3[2]);
        const double prev = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
        v[3 * i] = bond;
        v[3 * i + 1] = acos((b2[0] * b3[0] + b2[1] * b3[1] + b2[2] * b3[2]) / bond / prev);
        cross(b1, b2, A);
        cross(b2, b3, C);
        v[3 * i + 2] = torsion(A, C, b2, prev);
    }
    // phase angles: difference to the parent dihedral (parents are never children)
    for (int i = 0; i < nd; i++)
        if (top[i][5] != -1) v[3 * i + 2] -= v[3 * (top[i][5] - 1) + 2];
    // [-pi, pi] -> [0, 2 pi)
    for (int i = 0; i < nd; i++) {
        const double d = v[3 * i + 2];
        v[3 * i + 2] = d + 2 * pi - int((d + 2 * pi) / (2 * pi)) * 2 * pi;
    }
}

// .bat header (io_binary.cpp:41-102)
void put_bytes(std::vector<char>& o, const void* p, size_t n) {
    o.insert(o.end(), static_cast<const char*>(p), static_cast<const char*>(p) + n);
}
void put_name(std::vector<char>& o, const std::string& s, size_t width) {
    std::vector<char> b(width, 0);
    memcpy(b.data(), s.data(), std::min(s.size(), width - 1));
    put_bytes(o, b.data(), width);
}

}  // namespace

extern "C" int parent_bat_build(const char* top_path, const char* xtc_path, const char* bat_path,
                                const char* backbone_names, int double_prec, int nthreads, int* n_frames_out) {
    if (!top_path || !xtc_path || !bat_path || (double_prec != 0 && double_prec != 1)) {
        set_last_error("parent_bat_build: invalid argument");
        return PARENT_E_INVALID;
    }
    std::vector<std::string> bb;
    {
        std::istringstream ss(backbone_names ? backbone_names : "");
        std::string w;
        while (ss >> w) bb.push_back(w);
    }
    Topology topo;
    if (!read_topology(top_path, bb, topo)) {
        set_last_error(topo.err);
        return PARENT_E_INVALID;
    }
    BatTree tree;
    if (!build_tree(topo, tree)) {
        set_last_error(tree.err);
        return PARENT_E_INVALID;
    }
    const int nd = (int)tree.dih.size(), natoms = nd + 3;
    if ((int)topo.masses.size() < natoms || (int)topo.residues.size() < natoms) {
        set_last_error("ERROR: THE TOPOLOGY'S [ atoms ] DO NOT COVER ITS BONDED ATOMS");
        return PARENT_E_INVALID;
    }
    std::vector<uint8_t> xtc;
    {
        FILE* f = fopen(xtc_path, "rb");
        if (!f) {
            set_last_error(std::string("ERROR: UNABLE TO OPEN FILE ") + xtc_path + " !");
            return PARENT_E_IO;
        }
        fseeko(f, 0, SEEK_END);
        const off_t sz = ftello(f);
        fseeko(f, 0, SEEK_SET);
        xtc.resize((size_t)sz);
        const bool ok = sz == 0 || fread(xtc.data(), 1, (size_t)sz, f) == (size_t)sz;
        fclose(f);
        if (!ok) {
            set_last_error(std::string("ERROR: UNABLE TO READ FILE ") + xtc_path + " !");
            return PARENT_E_IO;
        }
    }
    std::vector<XtcFrame> frames;
    std::string err;
    if (!scan_xtc(xtc, &frames, &err)) {
        set_last_error(err);
        return PARENT_E_INVALID;
    }
    if (!frames.empty() && frames[0].natoms != natoms) {
        set_last_error("ERROR: NUMBER OF ATOMS IN THE TRAJECTORY(" + std::to_string(frames[0].natoms) +
                       ") DOES NOT FIT THE TOPOLOGY(" + std::to_string(natoms) + ")!");
        return PARENT_E_INVALID;
    }
    FILE* out = fopen(bat_path, "wb");
    if (!out) {
        set_last_error(std::string("ERROR: UNABLE TO OPEN FILE ") + bat_path + " !");
        return PARENT_E_IO;
    }
    std::vector<char> hdr;
    const int version = 3, zero = 0;
    put_bytes(hdr, &version, 4);
    put_bytes(hdr, &double_prec, 4);
    put_bytes(hdr, &nd, 4);
    put_bytes(hdr, &zero, 4);  // frame count, patched at the end
    for (int i = 0; i < natoms; i++) {
        put_name(hdr, topo.residues[i], 8);
        put_bytes(hdr, &topo.resnums[i], 4);
        put_name(hdr, topo.atom_names[i], 8);
        put_name(hdr, topo.belongs[i], 31);
    }
    for (const auto& d : tree.dih) put_bytes(hdr, d.data(), 6 * sizeof(int));
    put_bytes(hdr, topo.masses.data(), (size_t)natoms * sizeof(float));
    bool ok = fwrite(hdr.data(), 1, hdr.size(), out) == hdr.size();

    const double pi = acos(-1.0);
    const size_t vsz = double_prec ? 8 : 4;
    const size_t rec = 11 * 4 + (6 + 3 + 3 * (size_t)nd) * vsz;
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    const int chunk = 256;  // frames decoded and converted per parallel batch
    std::vector<char> buf((size_t)chunk * rec);
    int written = 0;
    bool bad_frame = false;
    for (size_t f0 = 0; f0 < frames.size() && ok && !bad_frame; f0 += chunk) {
        const int n = (int)std::min<size_t>(chunk, frames.size() - f0);
        std::vector<char> good(n, 1);
#pragma omp parallel num_threads(nt)
        {
            std::vector<float> x(3 * (size_t)natoms);
            std::vector<double> vals(3 + 3 * (size_t)nd);
            double ext[6];
#pragma omp for schedule(dynamic, 4)
            for (int k = 0; k < n; k++) {
                const XtcFrame& fr = frames[f0 + k];
                if (fr.natoms != natoms || !decode_xtc_frame(xtc, fr, x.data())) {
                    good[k] = 0;
                    continue;
                }
                for (float& c : x) c *= 10.0;  // nm -> Angstrom, kept in float
                to_bat(tree.dih, x.data(), pi, ext, vals.data());
                char* r = &buf[(size_t)k * rec];
                memcpy(r, &fr.time, 4);
                memcpy(r + 4, &fr.prec, 4);
                memcpy(r + 8, fr.box, 36);
                char* v = r + 44;
                if (double_prec) {
                    memcpy(v, ext, 48);
                    memcpy(v + 48, vals.data(), vals.size() * 8);
                } else {
                    for (int q = 0; q < 6; q++) {
                        const float fv = (float)ext[q];
                        memcpy(v + 4 * q, &fv, 4);
                    }
                    for (size_t q = 0; q < vals.size(); q++) {
                        const float fv = (float)vals[q];
                        memcpy(v + 24 + 4 * q, &fv, 4);
                    }
                }
            }
        }
        // frames are written in order up to the first one that does not decode
        int m = 0;
        while (m < n && good[m]) m++;
        if (m < n) bad_frame = true;
        ok = ok && fwrite(buf.data(), 1, (size_t)m * rec, out) == (size_t)m * rec;
        written += m;
    }
    if (ok) {
        ok = fseeko(out, 12, SEEK_SET) == 0 && fwrite(&written, 4, 1, out) == 1;
    }
    if (fclose(out) != 0) ok = false;
    if (!ok) {
        set_last_error(std::string("ERROR WRITING ") + bat_path);
        return PARENT_E_IO;
    }
    if (n_frames_out) *n_frames_out = written;
    return PARENT_OK;
}
