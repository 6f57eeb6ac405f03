// Synthetic BAT trajectories of a linear chain (test and benchmark input; no
// counterpart in the reference, whose inputs come from BAT_builder.x).
//
// Distributions per dof (drawn once from the seed):
//   bonds      N(mu, 0.003) nm,        mu ~ U(0.10, 0.15)
//   angles     wrapped normal, sd 0.05, mu ~ U(1.8, 2.1), folded into (0, pi)
//   dihedrals  3-component von Mises mixture, kappa ~ U(2, 20),
//              weights ~ Dirichlet(1,1,1), means ~ U(-pi, pi]
// 20% of the dofs share one of 10 latent N(0,1) factors per frame: their
// standardised deviate is 0.5*z + sqrt(0.75)*eps.  Correlated dihedrals use
// the wrapped-normal approximation N(0, 1/kappa) of their von Mises
// component so the latent factor can drive them; the others sample von Mises
// exactly (Best & Fisher 1979 rejection).
//
// Value conventions of the reference's .bat (BAT_trajectory.cpp:209-214):
// dihedrals in [0, 2*pi).  Two FP32 limits of PARENT.cpp are respected:
// dihedrals stay <= 6.283185005f (float32(2*pi) would index past the modfit
// histogram, PARENT.cpp:569-572) and angles stay <= 3.1415925f
// (float32(pi) > pi would index past the angle histogram, PARENT.cpp:554-556).
//
// Randomness is counter based (splitmix64 of (seed, stream, dof, frame,
// draw)), so every value depends only on its coordinates, not on the thread
// layout.
#include "../../include/parent_hip.h"
#include "host_common.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <omp.h>

namespace {

const double kPi = 3.14159265358979323846;

inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t key(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b, uint64_t c) {
    uint64_t h = mix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
    h = mix64(h ^ a);
    h = mix64(h ^ (b * 0x9E3779B97F4A7C15ull));
    return mix64(h ^ (c * 0xC2B2AE3D27D4EB4Full));
}

// uniform in (0, 1)
inline double uni(uint64_t k) { return ((k >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

inline double normal(uint64_t k) {
    double u1 = uni(mix64(k ^ 0x1234567ull)), u2 = uni(mix64(k ^ 0x89ABCDEFull));
    return sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
}

enum { S_PARAM = 1, S_EPS, S_LATENT, S_COMP, S_VM };

struct DofParam {
    int type;      // 0 bond, 1 angle, 2 dihedral
    int factor;    // latent factor id, -1 if independent
    double mu, sd;
    double w[3], mean[3], kappa[3];
    double vm_r[3];  // Best-Fisher r per component
};

DofParam make_param(uint64_t seed, int nD, int g) {
    DofParam p{};
    p.type = parent::type_of(nD, g);
    auto u = [&](int k) { return uni(key(seed, S_PARAM, (uint64_t)g, (uint64_t)k, 0)); };
    p.factor = u(0) < 0.2 ? (int)(u(1) * 10.0) : -1;
    if (p.factor > 9) p.factor = 9;
    if (p.type == 0) {
        p.mu = 0.10 + 0.05 * u(2);
        p.sd = 0.003;
    } else if (p.type == 1) {
        p.mu = 1.8 + 0.3 * u(2);
        p.sd = 0.05;
    } else {
        double s = 0;
        for (int c = 0; c < 3; c++) {
            p.w[c] = -log(u(10 + c));  // Dirichlet(1,1,1) via normalised Exp(1)
            s += p.w[c];
            p.mean[c] = -kPi + 2.0 * kPi * u(20 + c);
            p.kappa[c] = 2.0 + 18.0 * u(30 + c);
            double tau = 1.0 + sqrt(1.0 + 4.0 * p.kappa[c] * p.kappa[c]);
            double rho = (tau - sqrt(2.0 * tau)) / (2.0 * p.kappa[c]);
            p.vm_r[c] = (1.0 + rho * rho) / (2.0 * rho);
        }
        for (int c = 0; c < 3; c++) p.w[c] /= s;
    }
    return p;
}

// von Mises(0, kappa) by Best & Fisher's wrapped-Cauchy rejection.
double von_mises(uint64_t seed, int g, int f, double kappa, double r) {
    for (uint64_t attempt = 0; attempt < 1000; attempt++) {
        double u1 = uni(key(seed, S_VM, (uint64_t)g, (uint64_t)f, 3 * attempt + 0));
        double u2 = uni(key(seed, S_VM, (uint64_t)g, (uint64_t)f, 3 * attempt + 1));
        double u3 = uni(key(seed, S_VM, (uint64_t)g, (uint64_t)f, 3 * attempt + 2));
        double z = cos(kPi * u1);
        double fz = (1.0 + r * z) / (r + z);
        double c = kappa * (r - fz);
        if (c * (2.0 - c) - u2 > 0.0 || log(c / u2) + 1.0 - c >= 0.0) {
            double t = acos(fmin(1.0, fmax(-1.0, fz)));
            return u3 < 0.5 ? -t : t;
        }
    }
    return 0.0;
}

float value_of(uint64_t seed, const DofParam& p, int g, int f, const double* latent) {
    double eps = normal(key(seed, S_EPS, (uint64_t)g, (uint64_t)f, 0));
    double dev = p.factor >= 0 ? 0.5 * latent[p.factor] + sqrt(0.75) * eps : eps;
    if (p.type == 0) {
        double v = p.mu + p.sd * dev;
        return (float)(v < 0.0 ? 0.0 : v);
    }
    if (p.type == 1) {
        double t = p.mu + p.sd * dev;
        for (int k = 0; k < 8 && (t < 0.0 || t > kPi); k++) t = t > kPi ? 2.0 * kPi - t : -t;
        float tf = (float)t;
        const float kMaxAngle = 3.1415925f;  // largest float below pi
        if (tf > kMaxAngle) tf = kMaxAngle;
        if (tf < 1e-6f) tf = 1e-6f;
        return tf;
    }
    double uc = uni(key(seed, S_COMP, (uint64_t)g, (uint64_t)f, 0));
    int c = uc < p.w[0] ? 0 : (uc < p.w[0] + p.w[1] ? 1 : 2);
    double d = p.factor >= 0 ? dev / sqrt(p.kappa[c]) : von_mises(seed, g, f, p.kappa[c], p.vm_r[c]);
    double phi = fmod(p.mean[c] + d, 2.0 * kPi);
    if (phi < 0.0) phi += 2.0 * kPi;
    float pf = (float)phi;
    const float kMaxDihedral = 6.283185005187988f;  // largest float below 2*pi
    if (pf > kMaxDihedral) pf = kMaxDihedral;
    if (pf < 0.0f) pf = 0.0f;
    return pf;
}

int check_args(int atoms, int n_frames) {
    if (atoms < 4 || n_frames < 1) {
        parent::set_last_error("synthetic chain needs atoms >= 4 and n_frames >= 1");
        return PARENT_E_INVALID;
    }
    return PARENT_OK;
}

// Generates frames [f0, f0+nf) into rows of D floats.
void gen_frames(uint64_t seed, int nD, const std::vector<DofParam>& prm, int64_t f0, int64_t nf,
                float* out, int nthreads) {
    const int D = parent::n_dofs(nD);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t k = 0; k < nf; k++) {
        const int f = (int)(f0 + k);
        double latent[10];
        for (int q = 0; q < 10; q++) latent[q] = normal(key(seed, S_LATENT, (uint64_t)q, (uint64_t)f, 0));
        float* row = out + (size_t)k * D;
        for (int g = 0; g < D; g++) row[g] = value_of(seed, prm[g], g, f, latent);
    }
}

}  // namespace

extern "C" {

int parent_synth_values(int atoms, int n_frames, uint64_t seed, float* values, int nthreads) {
    if (int rc = check_args(atoms, n_frames)) return rc;
    if (!values) return PARENT_E_INVALID;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    const int nD = atoms - 3, D = parent::n_dofs(nD);
    std::vector<DofParam> prm(D);
    for (int g = 0; g < D; g++) prm[g] = make_param(seed, nD, g);
    gen_frames(seed, nD, prm, 0, n_frames, values, nthreads);
    return PARENT_OK;
}

// Writes a single-precision .bat v3 (io_binary.cpp:43-102 header, 203-278
// frames) for the chain atoms 1..a: dihedral i spans atoms (i+1..i+4),
// physical (type 0), no parent.
int parent_synth_write_bat(const char* path, int atoms, int n_frames, uint64_t seed, int nthreads) {
    if (int rc = check_args(atoms, n_frames)) return rc;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    FILE* f = fopen(path, "wb");
    if (!f) {
        parent::set_last_error(std::string("cannot open ") + path);
        return PARENT_E_IO;
    }
    const int nD = atoms - 3, D = parent::n_dofs(nD), natoms = atoms;
    const int nB = nD + 2, nA = nD + 1;
    std::vector<char> h;
    auto put32 = [&](int32_t v) { h.insert(h.end(), (char*)&v, (char*)&v + 4); };
    auto putf = [&](float v) { h.insert(h.end(), (char*)&v, (char*)&v + 4); };
    auto putstr = [&](const char* s, size_t w) {
        std::vector<char> b(w, 0);
        strncpy(b.data(), s, w - 1);
        h.insert(h.end(), b.begin(), b.end());
    };
    put32(3);
    put32(0);
    put32(nD);
    put32(n_frames);
    for (int i = 0; i < natoms; i++) {
        char name[16];
        snprintf(name, sizeof name, "C%d", i + 1);
        putstr("SYN", 8);
        put32(i / 10 + 1);
        putstr(name, 8);
        putstr("synthetic_chain", 31);
    }
    for (int i = 0; i < nD; i++) {
        put32(i + 1);
        put32(i + 2);
        put32(i + 3);
        put32(i + 4);
        put32(0);
        put32(-1);
    }
    for (int i = 0; i < natoms; i++) putf(12.011f);
    bool ok = fwrite(h.data(), 1, h.size(), f) == h.size();

    std::vector<DofParam> prm(D);
    for (int g = 0; g < D; g++) prm[g] = make_param(seed, nD, g);
    const int64_t rec_floats = 11 + 6 + 3 + 3 * (int64_t)nD;
    const int64_t chunk = 4096;
    std::vector<float> vals((size_t)chunk * D), recs((size_t)chunk * rec_floats);
    for (int64_t f0 = 0; ok && f0 < n_frames; f0 += chunk) {
        int64_t nf = std::min<int64_t>(chunk, n_frames - f0);
        gen_frames(seed, nD, prm, f0, nf, vals.data(), nthreads);
        for (int64_t k = 0; k < nf; k++) {
            const float* v = &vals[(size_t)k * D];
            float* r = &recs[(size_t)k * rec_floats];
            memset(r, 0, sizeof(float) * 17);
            r[0] = (float)(f0 + k);  // time
            r[1] = 1000.0f;          // xtc precision
            const float* b = v;
            const float* a = v + nB;
            const float* d = v + nB + nA;
            r[17] = b[0];
            r[18] = b[1];
            r[19] = a[0];
            for (int i = 0; i < nD; i++) {
                r[20 + 3 * i + 0] = b[i + 2];
                r[20 + 3 * i + 1] = a[i + 1];
                r[20 + 3 * i + 2] = d[i];
            }
        }
        ok = fwrite(recs.data(), sizeof(float) * rec_floats, nf, f) == (size_t)nf;
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        parent::set_last_error(std::string("write failed: ") + path);
        return PARENT_E_IO;
    }
    return PARENT_OK;
}

}  // extern "C"
