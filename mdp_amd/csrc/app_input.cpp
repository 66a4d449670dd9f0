// The binary inputs of the upstream C++ application, as the Julia preprocessing writes them (mdp(app):
// src/Julia/Preprocessing/Preprocessing.jl:60-321 with writeapp.jl, writeconfig.jl, writelattice.jl,
// writeregion.jl, writedomain.jl) and the upstream reader reads them (src/C++/Configuration/readdata.cpp:
// 11-290), and the system the application builds from them (readdata.cpp:490-700): atoms from config.bin,
// or lattice points inside the region (the LAMMPS-style lattice of src/C++/Core/cpuLattice.cpp, the box of
// cpuDomain.cpp, cpuAtomLattice/cpuAtomAdd of cpuAtom.cpp), initial velocities, and the MD engine set up
// with the application's SNAP potential, reference pair potential, ensemble, time step and neighbour skin.
//
// Every file is a stream of float64: the number of size entries, the size entries, then each field's
// values in a fixed order (integers stored as doubles).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "mdpsnap.h"
#include "runtime.hpp"

using namespace mdp;

// lattice points must be the upstream application's bit for bit: no fused multiply-adds
#pragma clang fp contract(off)

namespace {

// app.bin fields in file order, by their index in the size table (readdata.cpp:24-86; index 49 unused)
const char* const kAppFields[60] = {
    "ndims", "flag", "bcs", "pbc", "boxoffset", "atomnumbers", "atommasses", "atomcharges", "simparam", "solparam",
    "eta", "kappa", "muml", "mu1a", "mu1b", "mu2a", "mu2b", "mu2c", "mu3a", "mu3b",
    "mu3c", "mu4a", "mu4b", "pot1a", "pot1b", "pot2a", "pot2b", "pot2c", "pot3a", "pot3b",
    "pot3c", "pot4a", "pot4b", "rcutsqml", "rcutsq2a", "rcutsq2b", "rcutsq2c", "rcutsq3a", "rcutsq3b", "rcutsq3c",
    "rcutsq4a", "rcutsq4b", "rcutsqmax", "atom1b", "atom2b", "atom2c", "atom3b", "atom3c", "atom4b", "",
    "traininglist", "validatelist", "nveparam", "nvtparam", "snaparam", "snapelemradius", "snapelemweight",
    "snapcoeff", "createvelocity", "setatomtypefraction"};
// config.bin fields (readdata.cpp:150-167)
const char* const kConfigFields[18] = {"cfg_ndims", "natoms", "a", "b", "c", "e", "t", "x", "q", "v",
                                       "f", "we", "wf", "ws", "stress", "cfg_pbc", "move", "group"};

struct Reader {
    std::ifstream in;
    std::string name;
    explicit Reader(const std::string& path) : in(path.c_str(), std::ios::in | std::ios::binary), name(path) {}
    bool ok() const { return static_cast<bool>(in); }
    std::vector<double> read(long n) {
        if (n < 0 || n > (1L << 34)) throw Error(name + ": corrupt size table");
        std::vector<double> v(static_cast<size_t>(n));
        if (n) in.read(reinterpret_cast<char*>(v.data()), sizeof(double) * v.size());
        if (n && !in) throw Error(name + ": file ends before its size table says");
        return v;
    }
    std::vector<long> sizes() {
        const std::vector<double> l = read(1);
        const std::vector<double> s = read(static_cast<long>(l[0]));
        std::vector<long> out(s.size());
        for (size_t k = 0; k < s.size(); ++k) out[k] = static_cast<long>(s[k]);
        return out;
    }
};

constexpr double kBig = 1.0e30;

}  // namespace

struct mdp_app {
    std::vector<std::vector<double>> app;    // by size-table index
    std::vector<std::vector<double>> cfg;    // config.bin fields
    bool has_config = false, has_lattice = false, has_region = false, has_domain = false;
    // lattice.bin (readdata.cpp:180-228)
    int lat_style = 0, lat_nbasis = 0, lat_spaceflag = 0;
    double lat_scale = 1.0, lat_origin[3] = {0, 0, 0}, lat_spacing[3] = {1, 1, 1};
    double lat_a[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    int lat_orient[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    std::vector<double> lat_basis;
    std::vector<int> lat_type;
    // region.bin / domain.bin (readdata.cpp:230-290)
    int reg_triclinic = 0, dom_triclinic = 0;
    double reg_lo[3] = {0, 0, 0}, reg_hi[3] = {0, 0, 0}, reg_tilt[3] = {0, 0, 0};
    double dom_lo[3] = {0, 0, 0}, dom_hi[3] = {0, 0, 0}, dom_tilt[3] = {0, 0, 0};

    const std::vector<double>& field(const std::string& name) const {
        for (int k = 0; k < 60; ++k)
            if (name == kAppFields[k] && kAppFields[k][0]) return app[k];
        for (int k = 0; k < 18; ++k)
            if (name == kConfigFields[k]) return cfg[k];
        throw Error("mdp_app_field: unknown field " + name);
    }
    double get(const std::string& name, size_t k, double dflt) const {
        const std::vector<double>& v = field(name);
        return k < v.size() ? v[k] : dflt;
    }

    void read(const std::string& prefix) {
        {
            Reader r(prefix + "app.bin");
            if (!r.ok()) throw Error("mdp_app_read: cannot open " + prefix + "app.bin");
            const std::vector<long> ns = r.sizes();
            if (ns.size() < 60) throw Error("mdp_app_read: app.bin size table has fewer than 60 entries");
            app.assign(60, {});
            for (int k = 0; k < 60; ++k) {
                if (k == 49) continue;   // never written
                app[k] = r.read(ns[k]);
            }
        }
        cfg.assign(18, {});
        {
            Reader r(prefix + "config.bin");
            if (r.ok()) {
                const std::vector<long> ns = r.sizes();
                for (int k = 0; k < 18 && k < static_cast<int>(ns.size()); ++k) cfg[k] = r.read(ns[k]);
                has_config = true;
            }
        }
        {
            Reader r(prefix + "lattice.bin");
            if (r.ok()) {
                const std::vector<long> ns = r.sizes();
                if (ns.size() < 11) throw Error("mdp_app_read: lattice.bin size table has fewer than 11 entries");
                const std::vector<double> tmp = r.read(ns[0]);
                if (tmp.size() < 4) throw Error("mdp_app_read: lattice.bin header too short");
                auto three = [&](long n, double* dst) {
                    const std::vector<double> v = r.read(n);
                    for (size_t d = 0; d < 3 && d < v.size(); ++d) dst[d] = v[d];
                };
                auto three_i = [&](long n, int* dst) {
                    const std::vector<double> v = r.read(n);
                    for (size_t d = 0; d < 3 && d < v.size(); ++d) dst[d] = static_cast<int>(v[d]);
                };
                three(ns[1], lat_origin);
                three_i(ns[2], lat_orient[0]);
                three_i(ns[3], lat_orient[1]);
                three_i(ns[4], lat_orient[2]);
                three(ns[5], lat_spacing);
                three(ns[6], lat_a[0]);
                three(ns[7], lat_a[1]);
                three(ns[8], lat_a[2]);
                lat_basis = r.read(ns[9]);
                const std::vector<double> t = r.read(ns[10]);
                lat_type.assign(t.begin(), t.end());
                lat_style = static_cast<int>(tmp[0]);
                lat_nbasis = static_cast<int>(tmp[1]);
                lat_spaceflag = static_cast<int>(tmp[2]);
                lat_scale = tmp[3];
                has_lattice = true;
            }
        }
        auto box3 = [&](const char* file, int& tri, double* lo, double* hi, double* tilt, bool& has) {
            Reader r(prefix + file);
            if (!r.ok()) return;
            const std::vector<long> ns = r.sizes();
            if (ns.size() < 4) throw Error(std::string("mdp_app_read: ") + file + " size table too short");
            const std::vector<double> tmp = r.read(ns[0]);
            const std::vector<double> a = r.read(ns[1]), b = r.read(ns[2]), c = r.read(ns[3]);
            for (int d = 0; d < 3; ++d) {
                lo[d] = d < static_cast<int>(a.size()) ? a[d] : 0.0;
                hi[d] = d < static_cast<int>(b.size()) ? b[d] : 0.0;
                tilt[d] = d < static_cast<int>(c.size()) ? c[d] : 0.0;
            }
            tri = tmp.empty() ? 0 : static_cast<int>(tmp[0]);
            has = true;
        };
        box3("region.bin", reg_triclinic, reg_lo, reg_hi, reg_tilt, has_region);
        box3("domain.bin", dom_triclinic, dom_lo, dom_hi, dom_tilt, has_domain);
    }

    // ---- the lattice of the application (cpuLattice.cpp): primitive cell, rotation, spacing --------
    struct Lattice {
        double basis[8 * 3] = {0};
        int nbasis = 0;
        double P[9] = {0}, Pinv[9] = {0}, Rrow[9] = {0}, Rcol[9] = {0}, spacing[3] = {0, 0, 0}, scale = 1.0;
    };

    static void lattice_to_box(double* x, const Lattice& L, const double* origin) {
        const double x1 = (L.P[0] * x[0] + L.P[1] * x[1] + L.P[2] * x[2]) * L.scale;
        const double y1 = (L.P[3] * x[0] + L.P[4] * x[1] + L.P[5] * x[2]) * L.scale;
        const double z1 = (L.P[6] * x[0] + L.P[7] * x[1] + L.P[8] * x[2]) * L.scale;
        const double xn = L.Rrow[0] * x1 + L.Rrow[1] * y1 + L.Rrow[2] * z1;
        const double yn = L.Rrow[3] * x1 + L.Rrow[4] * y1 + L.Rrow[5] * z1;
        const double zn = L.Rrow[6] * x1 + L.Rrow[7] * y1 + L.Rrow[8] * z1;
        x[0] = xn + L.spacing[0] * origin[0];
        x[1] = yn + L.spacing[1] * origin[1];
        x[2] = zn + L.spacing[2] * origin[2];
    }
    static void box_to_lattice(double* x, const Lattice& L, const double* origin) {
        x[0] -= L.spacing[0] * origin[0];
        x[1] -= L.spacing[1] * origin[1];
        x[2] -= L.spacing[2] * origin[2];
        const double x1 = (L.Rcol[0] * x[0] + L.Rcol[1] * x[1] + L.Rcol[2] * x[2]) / L.scale;
        const double y1 = (L.Rcol[3] * x[0] + L.Rcol[4] * x[1] + L.Rcol[5] * x[2]) / L.scale;
        const double z1 = (L.Rcol[6] * x[0] + L.Rcol[7] * x[1] + L.Rcol[8] * x[2]) / L.scale;
        x[0] = L.Pinv[0] * x1 + L.Pinv[1] * y1 + L.Pinv[2] * z1;
        x[1] = L.Pinv[3] * x1 + L.Pinv[4] * y1 + L.Pinv[5] * z1;
        x[2] = L.Pinv[6] * x1 + L.Pinv[7] * y1 + L.Pinv[8] * z1;
    }
    static void grow(double* lo, double* hi, const double* x) {
        for (int d = 0; d < 3; ++d) {
            lo[d] = x[d] < lo[d] ? x[d] : lo[d];
            hi[d] = x[d] > hi[d] ? x[d] : hi[d];
        }
    }

    Lattice lattice() const {
        enum { NONE = 0, SC, BCC, FCC, HCP, DIAMOND, SQ, SQ2, HEX, CUSTOM };
        Lattice L;
        L.scale = lat_scale;
        const int style = lat_style;
        if (style == NONE) throw Error("lattice style none: no lattice atoms to create");
        double a1[3], a2[3], a3[3];
        for (int d = 0; d < 3; ++d) {
            a1[d] = lat_a[0][d];
            a2[d] = lat_a[1][d];
            a3[d] = lat_a[2][d];
        }
        auto add = [&](double x, double y, double z) {
            L.basis[3 * L.nbasis] = x;
            L.basis[3 * L.nbasis + 1] = y;
            L.basis[3 * L.nbasis + 2] = z;
            ++L.nbasis;
        };
        if (style == CUSTOM) {
            const int nb = static_cast<int>(lat_basis.size() / 3);
            if (nb < 1 || nb > 8) throw Error("custom lattice: 1 to 8 basis atoms supported");
            for (int m = 0; m < nb; ++m) add(lat_basis[3 * m], lat_basis[3 * m + 1], lat_basis[3 * m + 2]);
        } else {
            const double I[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
            for (int d = 0; d < 3; ++d) {
                a1[d] = I[0][d];
                a2[d] = I[1][d];
                a3[d] = I[2][d];
            }
            if (style == HEX) a2[1] = std::sqrt(3.0);
            if (style == HCP) {
                a2[1] = std::sqrt(3.0);
                a3[2] = std::sqrt(8.0 / 3.0);
            }
            switch (style) {
                case SC: case SQ: add(0, 0, 0); break;
                case BCC: add(0, 0, 0); add(0.5, 0.5, 0.5); break;
                case FCC: add(0, 0, 0); add(0.5, 0.5, 0); add(0.5, 0, 0.5); add(0, 0.5, 0.5); break;
                case HCP: add(0, 0, 0); add(0.5, 0.5, 0); add(0.5, 5.0 / 6.0, 0.5); add(0, 1.0 / 3.0, 0.5); break;
                case SQ2: case HEX: add(0, 0, 0); add(0.5, 0.5, 0); break;
                case DIAMOND:
                    add(0, 0, 0); add(0.5, 0.5, 0); add(0.5, 0, 0.5); add(0, 0.5, 0.5);
                    add(0.25, 0.25, 0.25); add(0.25, 0.75, 0.75); add(0.75, 0.25, 0.75); add(0.75, 0.75, 0.25);
                    break;
                default: throw Error("lattice: unknown style " + std::to_string(style));
            }
        }
        // P: primitive vectors as columns; Pinv by cofactors over the determinant
        const double P[9] = {a1[0], a2[0], a3[0], a1[1], a2[1], a3[1], a1[2], a2[2], a3[2]};
        std::memcpy(L.P, P, sizeof(P));
        const double det = P[0] * P[4] * P[8] + P[1] * P[5] * P[6] + P[2] * P[3] * P[7] - P[0] * P[5] * P[7] -
                           P[1] * P[3] * P[8] - P[2] * P[4] * P[6];
        L.Pinv[0] = (P[4] * P[8] - P[5] * P[7]) / det;
        L.Pinv[3] = (P[5] * P[6] - P[3] * P[8]) / det;
        L.Pinv[6] = (P[3] * P[7] - P[4] * P[6]) / det;
        L.Pinv[1] = (P[2] * P[7] - P[1] * P[8]) / det;
        L.Pinv[4] = (P[0] * P[8] - P[2] * P[6]) / det;
        L.Pinv[7] = (P[1] * P[6] - P[0] * P[7]) / det;
        L.Pinv[2] = (P[1] * P[5] - P[2] * P[4]) / det;
        L.Pinv[5] = (P[2] * P[3] - P[0] * P[5]) / det;
        L.Pinv[8] = (P[0] * P[4] - P[1] * P[3]) / det;
        // Rrow: normalised orientation vectors as rows; Rcol its transpose
        for (int r = 0; r < 3; ++r) {
            const int* o = lat_orient[r];
            const double len = std::sqrt(static_cast<double>(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]));
            for (int c = 0; c < 3; ++c) L.Rrow[3 * r + c] = o[c] / len;
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) L.Rcol[3 * r + c] = L.Rrow[3 * c + r];
        if (lat_spaceflag == 0) {
            // extent of the 8 corners of the unit cell in box space (spacing zero while they are mapped)
            double lo[3] = {kBig, kBig, kBig}, hi[3] = {-kBig, -kBig, -kBig};
            const double corners[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                          {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
            for (const auto& cn : corners) {
                double x[3] = {cn[0], cn[1], cn[2]};
                lattice_to_box(x, L, lat_origin);
                grow(lo, hi, x);
            }
            for (int d = 0; d < 3; ++d) L.spacing[d] = hi[d] - lo[d];
        } else {
            for (int d = 0; d < 3; ++d) L.spacing[d] = lat_spacing[d] * L.scale;
        }
        return L;
    }

    // atoms of the application: x[3 n], 1-based types, the orthogonal box lengths
    void atoms(std::vector<double>& x, std::vector<int>& t, double box[3]) const {
        x.clear();
        t.clear();
        if (get("flag", 13, 2) != 2) throw Error("mdp_app: only metal units (unitstyle 2) are supported");
        if (!field("setatomtypefraction").empty()) throw Error("mdp_app: setatomtypefraction is not supported");
        if (has_config) {
            const std::vector<double>& na = field("natoms");
            if (na.empty()) throw Error("mdp_app: config.bin holds no configuration");
            const long n = static_cast<long>(na[0]);
            const std::vector<double>&a = field("a"), &b = field("b"), &c = field("c");
            if (a.size() < 3 || b.size() < 3 || c.size() < 3) throw Error("mdp_app: config.bin without cell vectors");
            if (std::fabs(b[0]) > 1e-10 || std::fabs(c[0]) > 1e-10 || std::fabs(c[1]) > 1e-10)
                throw Error("mdp_app: the MD engine needs an orthogonal box (config.bin cell is tilted)");
            box[0] = a[0];
            box[1] = b[1];
            box[2] = c[2];
            const std::vector<double>&xs = field("x"), &ts = field("t");
            if (static_cast<long>(xs.size()) < 3 * n || static_cast<long>(ts.size()) < n)
                throw Error("mdp_app: config.bin holds fewer positions or types than natoms");
            x.assign(xs.begin(), xs.begin() + 3 * n);
            for (long i = 0; i < n; ++i) t.push_back(static_cast<int>(ts[i]));
            return;
        }
        if (!has_lattice || !has_region) throw Error("mdp_app: neither config.bin nor lattice.bin + region.bin");
        if (reg_triclinic) throw Error("mdp_app: a tilted region is not supported (the MD engine box is orthogonal)");
        const Lattice L = lattice();
        // the box (readdata.cpp:523-551): region in lattice spacings, sub-box = box on one process
        double lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            hi[d] = L.spacing[d] * reg_hi[d];
            lo[d] = L.spacing[d] * reg_lo[d];
        }
        double hinv[3];
        for (int d = 0; d < 3; ++d) hinv[d] = 1.0 / (hi[d] - lo[d]);
        double sublo[3], subhi[3];
        for (int d = 0; d < 3; ++d) {
            sublo[d] = lo[d] + (hi[d] - lo[d]) * 0.0;
            subhi[d] = lo[d] + (hi[d] - lo[d]) * 1.0;
        }
        // lattice-space bounding box of the 8 sub-box corners, then the index range (latticebounds)
        double lmin[3] = {kBig, kBig, kBig}, lmax[3] = {-kBig, -kBig, -kBig};
        for (int k = 0; k < 8; ++k) {
            double p[3] = {(k & 1) ? subhi[0] : sublo[0], (k & 2) ? subhi[1] : sublo[1], (k & 4) ? subhi[2] : sublo[2]};
            box_to_lattice(p, L, lat_origin);
            grow(lmin, lmax, p);
        }
        int ilo = static_cast<int>(lmin[0]) - 1, jlo = static_cast<int>(lmin[1]) - 1, klo = static_cast<int>(lmin[2]) - 1;
        const int ihi = static_cast<int>(lmax[0]) + 1, jhi = static_cast<int>(lmax[1]) + 1, khi = static_cast<int>(lmax[2]) + 1;
        if (lmin[0] < 0.0) --ilo;
        if (lmin[1] < 0.0) --jlo;
        if (lmin[2] < 0.0) --klo;
        // every lattice point of the range, kept when its box fraction lies in [0, 1) (cpuAtomLattice, cpuAtomAdd)
        for (int k = klo; k <= khi; ++k)
            for (int j = jlo; j <= jhi; ++j)
                for (int i = ilo; i <= ihi; ++i)
                    for (int m = 0; m < L.nbasis; ++m) {
                        double p[3] = {i + L.basis[3 * m], j + L.basis[3 * m + 1], k + L.basis[3 * m + 2]};
                        lattice_to_box(p, L, lat_origin);
                        const double l0 = hinv[0] * (p[0] - lo[0]), l1 = hinv[1] * (p[1] - lo[1]),
                                     l2 = hinv[2] * (p[2] - lo[2]);
                        if (l0 >= 0.0 && l0 < 1.0 && l1 >= 0.0 && l1 < 1.0 && l2 >= 0.0 && l2 < 1.0) {
                            x.push_back(p[0]);
                            x.push_back(p[1]);
                            x.push_back(p[2]);
                            t.push_back(m < static_cast<int>(lat_type.size()) ? lat_type[m] : 1);
                        }
                    }
        for (int d = 0; d < 3; ++d) box[d] = hi[d] - lo[d];
        // positions relative to the box origin (the engine's box starts at 0)
        for (size_t q = 0; q < x.size(); ++q) x[q] -= lo[q % 3];
    }
};

namespace {
template <class F>
int api(F&& f) {
    clear_error();
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    }
}
}  // namespace

extern "C" {

int mdp_app_read(mdp_app** h, const char* prefix) {
    return api([&] {
        if (!h || !prefix) throw Error("mdp_app_read: null argument");
        *h = nullptr;
        auto a = std::make_unique<mdp_app>();
        a->read(prefix);
        *h = a.release();
    });
}

int mdp_app_destroy(mdp_app* h) {
    return api([&] { delete h; });
}

long mdp_app_field(const mdp_app* h, const char* name, double* out, long cap) {
    long n = -1;
    const int rc = api([&] {
        if (!h || !name) throw Error("mdp_app_field: null argument");
        const std::vector<double>& v = h->field(name);
        n = static_cast<long>(v.size());
        if (out)
            for (long k = 0; k < n && k < cap; ++k) out[k] = v[k];
    });
    return rc ? -1 : n;
}

int mdp_app_has(const mdp_app* h, int* flags4) {
    return api([&] {
        if (!h || !flags4) throw Error("mdp_app_has: null argument");
        flags4[0] = h->has_config;
        flags4[1] = h->has_lattice;
        flags4[2] = h->has_region;
        flags4[3] = h->has_domain;
    });
}

long mdp_app_atoms(const mdp_app* h, double* x, int* type, double* box3, long cap) {
    long n = -1;
    const int rc = api([&] {
        if (!h) throw Error("mdp_app_atoms: null handle");
        std::vector<double> xs;
        std::vector<int> ts;
        double box[3];
        h->atoms(xs, ts, box);
        n = static_cast<long>(ts.size());
        if (box3)
            for (int d = 0; d < 3; ++d) box3[d] = box[d];
        if (x && type) {
            if (cap < n) throw Error("mdp_app_atoms: capacity below the atom count");
            std::memcpy(x, xs.data(), sizeof(double) * xs.size());
            std::memcpy(type, ts.data(), sizeof(int) * ts.size());
        }
    });
    return rc ? -1 : n;
}

int mdp_md_create_app(mdp_md** md, const mdp_app* h) {
    return api([&] {
        if (!md || !h) throw Error("mdp_md_create_app: null argument");
        *md = nullptr;
        if (h->get("flag", 0, -1) != 1) throw Error("mdp_md_create_app: the application's descriptor is not SNAP");
        const std::vector<double>& sp = h->field("snaparam");
        if (sp.size() < 12) throw Error("mdp_md_create_app: snaparam needs 12 entries");
        const int nelem = static_cast<int>(sp[0]), ncoeffall = static_cast<int>(sp[1]), J = static_cast<int>(sp[2]);
        if (sp[8] != 0 || sp[9] != 0) throw Error("mdp_md_create_app: quadratic and chemflag SNAP are not supported");
        for (const char* other : {"pot1a", "pot1b", "pot2b", "pot2c", "pot3a", "pot3b", "pot3c", "pot4a", "pot4b"})
            if (!h->field(other).empty()) throw Error(std::string("mdp_md_create_app: potential ") + other + " not supported");
        const std::vector<double>& coeff = h->field("snapcoeff");
        if (static_cast<long>(coeff.size()) < static_cast<long>(nelem) * ncoeffall)
            throw Error("mdp_md_create_app: snapcoeff holds fewer than nelem * ncoeff values");
        std::vector<double> x;
        std::vector<int> t;
        double box[3];
        h->atoms(x, t, box);
        const int n = static_cast<int>(t.size());
        const int ntypes = std::max(1, static_cast<int>(h->get("ndims", 8, 1)));
        for (int ti : t)
            if (ti < 1 || ti > ntypes) throw Error("mdp_md_create_app: atom type outside 1..natomtype");
        const std::vector<double>&rad = h->field("snapelemradius"), &wj = h->field("snapelemweight");
        const std::vector<double>& am = h->field("atommasses");
        std::vector<double> r(ntypes), w(ntypes), m(ntypes);
        for (int k = 0; k < ntypes; ++k) {
            r[k] = k < static_cast<int>(rad.size()) ? rad[k] : rad.back();
            w[k] = k < static_cast<int>(wj.size()) ? wj[k] : wj.back();
            m[k] = k + 1 < static_cast<int>(am.size()) ? am[k + 1] : 0.0;   // atommasses[0] is a placeholder
            if (!(m[k] > 0.0)) throw Error("mdp_md_create_app: every atom type needs a positive mass");
        }
        // initial velocities: config.bin's, then createvelocity (readdata.cpp:609-640)
        std::vector<double> v(3 * static_cast<size_t>(n), 0.0);
        if (h->has_config && h->field("v").size() >= v.size())
            std::copy(h->field("v").begin(), h->field("v").begin() + v.size(), v.begin());
        const std::vector<double>& cv = h->field("createvelocity");
        if (!cv.empty()) {
            if (cv.size() < 7) throw Error("mdp_md_create_app: createvelocity needs 7 entries");
            if (cv[3] != 0) throw Error("mdp_md_create_app: createvelocity with sum_flag 1 is not supported");
            std::vector<double> mass1(ntypes + 1, 0.0);
            for (int k = 0; k < ntypes; ++k) mass1[k + 1] = m[k];
            if (mdp_velocity_create(x.data(), t.data(), mass1.data(), n, cv[0], static_cast<int>(cv[1]),
                                    static_cast<int>(cv[2]), static_cast<int>(cv[4]), 1.0364269e-4, 8.617343e-5,
                                    v.data()))
                throw Error(std::string("mdp_md_create_app: ") + mdp_snap_last_error());
        }
        mdp_md_params p{};
        p.twojmax = J;
        p.ntypes = ntypes;
        p.rcutfac = sp[3];
        p.rfac0 = sp[4];
        p.rmin0 = sp[5];
        p.bzeroflag = static_cast<int>(sp[6]);
        p.switchflag = static_cast<int>(sp[7]);
        p.bnormflag = static_cast<int>(sp[10]);
        p.wselfallflag = static_cast<int>(sp[11]);
        p.elemradius = r.data();
        p.elemweight = w.data();
        p.coeff = coeff.data();
        p.mass = m.data();
        p.dt = h->get("simparam", 1, 0.0);
        p.skin = h->get("solparam", 1, 0.0);
        p.rebuild_every = std::max(1, static_cast<int>(h->get("ndims", 11, 1)));
        mdp_md* e = nullptr;
        if (mdp_md_create(&e, &p, n, box, x.data(), v.data(), t.data()))
            throw Error(std::string("mdp_md_create_app: ") + mdp_snap_last_error());
        auto fail = [&](const char* what) {
            const std::string msg = std::string("mdp_md_create_app: ") + what + ": " + mdp_snap_last_error();
            mdp_md_destroy(e);
            throw Error(msg);
        };
        const std::vector<double>& pot2a = h->field("pot2a");
        if (!pot2a.empty()) {
            const std::vector<double>&mu = h->field("mu2a"), &rc2 = h->field("rcutsq2a");
            if (pot2a.size() != 1 || pot2a[0] != 1 || mu.size() < 3) {
                mdp_md_destroy(e);
                throw Error("mdp_md_create_app: the pair potential must be the ZBL potential (pot2a = [1], mu2a = "
                            "[cut_inner, cut_outer, z])");
            }
            if (!rc2.empty() && std::fabs(std::sqrt(rc2[0]) - mu[1]) > 1e-12 * mu[1]) {
                mdp_md_destroy(e);
                throw Error("mdp_md_create_app: rcut2a differs from the ZBL cut_outer mu2a[2]");
            }
            if (mdp_md_set_zbl(e, mu[0], mu[1], mu[2])) fail("ZBL");
        }
        const int ens = static_cast<int>(h->get("flag", 14, 0));
        if (ens == 2) {
            const std::vector<double>& nv = h->field("nvtparam");
            if (nv.size() < 3) fail("nvtparam needs [t_start, t_stop, t_damp]");
            const long steps = static_cast<long>(h->get("ndims", 3, 0));
            const double time0 = h->get("simparam", 0, 0.0);
            const long cur = p.dt > 0 ? static_cast<long>(time0 / p.dt) : 0;
            if (mdp_md_set_nvt(e, nv[0], nv[1], nv[2], nv.size() > 3 ? nv[3] : 0.0, 0.0,
                               nv.size() > 4 ? static_cast<int>(nv[4]) : 3, nv.size() > 5 ? static_cast<int>(nv[5]) : 1,
                               0, std::max(1L, steps), cur))
                fail("NVT");
        } else if (ens != 0) {
            mdp_md_destroy(e);
            throw Error("mdp_md_create_app: only the nve and nvt ensembles are supported");
        }
        *md = e;
    });
}

}  // extern "C"
