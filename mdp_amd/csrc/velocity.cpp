// Initial velocities exactly as the upstream app creates them (host code).
//
// readdata.cpp:609-640: cpuVelocityCreate (cpuVelocityIntegration.cpp:54-135) in its "geom" loop mode,
// the mode of the upstream examples (snapTA06A.jl:77), then zero total momentum and rescale to the
// target temperature with tdof = 3 (N - 1).  In that mode every atom's random stream is re-seeded from
// the seed and the bytes of its coordinates, so the velocities depend on nothing but the inputs.
// The generator: Park & Miller's minimal standard multiplicative congruential generator (a = 16807,
// m = 2^31 - 1) in Schrage's overflow-free form; normal deviates by Marsaglia's polar method with the
// second deviate of each pair kept for the next call; re-seeding by Jenkins' one-at-a-time hash over the
// bytes (as signed chars) of the seed and the three coordinates, masked to 27 bits, then five draws.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/mdpsnap.h"
#include "runtime.hpp"

#pragma clang fp contract(off)   // the same roundings as the upstream loops, operation by operation

namespace {

struct Stream {
    int s = 1;
    bool have = false;   // a second normal deviate is waiting
    double second = 0.0;

    double uniform() {
        constexpr int a = 16807, m = 2147483647, q = 127773, r = 2836;   // q = m / a, r = m % a
        const int k = s / q;
        s = a * (s - k * q) - r * k;
        if (s < 0) s += m;
        return (1.0 / m) * s;
    }
    double normal() {
        if (have) {
            have = false;
            return second;
        }
        double v1, v2, rsq;
        do {
            v1 = 2.0 * uniform() - 1.0;
            v2 = 2.0 * uniform() - 1.0;
            rsq = v1 * v1 + v2 * v2;
        } while (rsq >= 1.0 || rsq == 0.0);
        const double fac = std::sqrt(-2.0 * std::log(rsq) / rsq);
        second = v1 * fac;
        have = true;
        return v2 * fac;
    }
    void reseed(int base, const double* coord) {
        std::uint32_t h = 0;
        auto mix = [&](const void* p, size_t n) {
            const signed char* c = static_cast<const signed char*>(p);
            for (size_t i = 0; i < n; ++i) {
                h += static_cast<std::uint32_t>(static_cast<int>(c[i]));
                h += h << 10;
                h ^= h >> 6;
            }
        };
        mix(&base, sizeof(int));
        mix(coord, 3 * sizeof(double));
        h += h << 3;
        h ^= h >> 11;
        h += h << 15;
        s = static_cast<int>(h & 0x7ffffffu);
        if (!s) s = 1;
        for (int i = 0; i < 5; ++i) uniform();
        have = false;
    }
};

// cpuVelocityCreate's draws (cpuVelocityIntegration.cpp:54-135) on a caller-held generator state:
// g.s = seed[0], the waiting normal deviate = (save[0], second[0]); written back by the caller
void draw(Stream& g, int dist_flag, double r[3]) {
    for (int d = 0; d < 3; ++d) r[d] = dist_flag == 0 ? g.uniform() - 0.5 : g.normal();
}

}  // namespace

extern "C" int mdp_velocity_create_loop(const double* x, double* v, const double* mass, double* second, int* seed,
                                        int* save, const int* map, const int* type, int seed0, int sum_flag,
                                        int dist_flag, int loop_flag, int dim, int mpi_rank, int nlocal, int natoms) {
    mdp::clear_error();
    try {
        if (!v || !mass || !second || !seed || !save || !type) throw mdp::Error("mdp_velocity_create_loop: null argument");
        if (dim != 2 && dim != 3) throw mdp::Error("mdp_velocity_create_loop: dim must be 2 or 3");
        if (nlocal < 0 || natoms < 0) throw mdp::Error("mdp_velocity_create_loop: negative atom count");
        if (loop_flag < 0 || loop_flag > 2) throw mdp::Error("mdp_velocity_create_loop: loop_flag must be 0, 1 or 2");
        if (loop_flag == 0 && !map) throw mdp::Error("mdp_velocity_create_loop: loop mode 0 needs map");
        if (loop_flag == 2 && (!x || dim != 3))
            throw mdp::Error("mdp_velocity_create_loop: loop mode 2 needs x and dim = 3 (each atom's stream is seeded "
                             "from three coordinates)");
        if (sum_flag == 0)
            for (long i = 0; i < long(dim) * nlocal; ++i) v[i] = 0.0;
        Stream g;
        g.s = seed[0];
        g.have = false;            // every loop mode starts with save = 0
        g.second = second[0];
        auto add = [&](int i, const double r[3]) {
            const double factor = 1.0 / std::sqrt(mass[type[i]]);
            v[i * dim + 0] += r[0] * factor;
            v[i * dim + 1] += r[1] * factor;
            if (dim == 3) v[i * dim + 2] += r[2] * factor;
        };
        double r[3];
        if (loop_flag == 0) {
            // every atom of the system draws in global-ID order (IDs 1..natoms); map[id] is its local
            // index, and only the atoms held here (0 <= map[id] < nlocal) take their draw
            for (int id = 1; id <= natoms; ++id) {
                draw(g, dist_flag, r);
                const int m = map[id];
                if (m >= 0 && m < nlocal) add(m, r);
            }
        } else if (loop_flag == 1) {
            // each rank's own stream: seed + rank, 100 discarded draws, then its atoms in local order
            g.s = g.s + mpi_rank;
            for (int i = 0; i < 100; ++i) g.uniform();
            for (int i = 0; i < nlocal; ++i) {
                draw(g, dist_flag, r);
                add(i, r);
            }
        } else {
            g.s = 1;
            for (int i = 0; i < nlocal; ++i) {
                g.reseed(seed0, x + 3 * i);
                draw(g, dist_flag, r);
                add(i, r);
            }
        }
        seed[0] = g.s;
        save[0] = g.have ? 1 : 0;
        second[0] = g.second;
        return 0;
    } catch (const std::exception& e) {
        mdp::set_error(e.what());
        return 1;
    }
}

extern "C" int mdp_velocity_create(const double* x, const int* type, const double* mass, int natoms, double t_desired,
                                   int seed0, int dist_flag, int loop_flag, double mvv2e, double boltz, double* v) {
    mdp::clear_error();
    try {
        if (!x || !type || !mass || !v) throw mdp::Error("mdp_velocity_create: null argument");
        if (natoms < 2) throw mdp::Error("mdp_velocity_create: needs at least two atoms");
        if (loop_flag != 2)
            throw mdp::Error("mdp_velocity_create: only loop mode 2 (geom) is supported here: loop modes 0 (all) and "
                             "1 (local) draw from the application's generator state, which the upstream app never "
                             "sets; mdp_velocity_create_loop takes that state explicitly");
        if (!(t_desired > 0.0)) throw mdp::Error("mdp_velocity_create: the temperature must be positive");
        Stream g;
        for (int i = 0; i < natoms; ++i) {
            g.reseed(seed0, x + 3 * i);
            double r[3];
            for (int d = 0; d < 3; ++d) r[d] = dist_flag == 0 ? g.uniform() - 0.5 : g.normal();
            const double factor = 1.0 / std::sqrt(mass[type[i]]);
            for (int d = 0; d < 3; ++d) v[3 * i + d] = 0.0 + r[d] * factor;
        }
        // zero total momentum (cpuComputeMass, cpuComputeVCM, cpuVelocityZeroMomentum), atom order
        double mtot = mass[type[0]];
        for (int i = 1; i < natoms; ++i) mtot = mtot + mass[type[i]];
        double vcm[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < natoms; ++i)
            for (int d = 0; d < 3; ++d) vcm[d] += v[3 * i + d] * mass[type[i]];
        for (int d = 0; d < 3; ++d) vcm[d] /= mtot;
        for (int i = 0; i < natoms; ++i)
            for (int d = 0; d < 3; ++d) v[3 * i + d] -= vcm[d];
        // rescale to t_desired (cpuComputeTempScalar, cpuArrayMultiplyScalar)
        const double tfactor = mvv2e / ((natoms - 1) * 3.0 * boltz);
        double t = 0.0;
        for (int i = 0; i < natoms; ++i)
            t += (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]) * mass[type[i]];
        t = tfactor * t;
        const double scale = std::sqrt(t_desired / t);
        for (int i = 0; i < 3 * natoms; ++i) v[i] = scale * v[i];
        return 0;
    } catch (const std::exception& e) {
        mdp::set_error(e.what());
        return 1;
    }
}
