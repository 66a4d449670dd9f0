// Reference pair potential of the SNAP application (ZBL) as backend-3 entry points of the upstream
// empirical-potential route: the generated pair gradient (src/C++/Potentials/cpuPairaGradient.cpp,
// reached through gpuPairaGradient, src/C++/Common/mdpgpucore.h:373), the full-list energy/force
// decomposition (src/C++/Core/cpuEnergyForceTally.cpp:216-227) and the pair virial tally
// (src/C++/Core/cpuVirialTally.cpp:34-57), as composed by cpuFullNeighPairEnergyForceVirial
// (src/C++/Potentials/cpuPotentials.cpp:1587-1649).  Device arrays in and out.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "mdpsnap.h"
#include "runtime.hpp"
#include "zbl.hpp"

using namespace mdp;

namespace {

// one thread per pair: u and du/dxij of the pair vector xij = x_j - x_i
__global__ void k_paira_zbl(double* __restrict__ u, double* __restrict__ u_xij, const double* __restrict__ xij,
                            ZblParams p, int ng) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ng) return;
    const double x = xij[3 * static_cast<size_t>(q)], y = xij[3 * static_cast<size_t>(q) + 1],
                 z = xij[3 * static_cast<size_t>(q) + 2];
    const double r = sqrt(x * x + y * y + z * z);
    double uu, dudr;
    zbl_eval(p, r, uu, dudr);
    const double g = dudr / r;
    u[q] = uu;
    u_xij[3 * static_cast<size_t>(q)] = g * x;
    u_xij[3 * static_cast<size_t>(q) + 1] = g * y;
    u_xij[3 * static_cast<size_t>(q) + 2] = g * z;
}

// e[ai] += eij / 2, f[ai] -= fij (a full list holds each pair from both atoms)
__global__ void k_full_pair_decomp(double* __restrict__ e, double* __restrict__ f, const double* __restrict__ eij,
                                   const double* __restrict__ fij, const int* __restrict__ ai, int ijnum, int dim) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ijnum) return;
    const int i = ai[q];
    atomicAdd(&e[i], 0.5 * eij[q]);
    for (int d = 0; d < dim; ++d) atomicAdd(&f[static_cast<size_t>(dim) * i + d], -fij[static_cast<size_t>(dim) * q + d]);
}

// v[k*inum + ai] += factor * (rij (x) fij)_k, k = xx, yy, zz, xy, xz, yz
__global__ void k_virial_pair_tally(double* __restrict__ v, const double* __restrict__ fij,
                                    const double* __restrict__ rij, double factor, const int* __restrict__ ai, int dim,
                                    int inum, int ijnum) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ijnum) return;
    const int i = ai[q];
    const size_t b = static_cast<size_t>(dim) * q;
    const double dx = rij[b], dy = rij[b + 1], dz = dim == 3 ? rij[b + 2] : 0.0;
    const double fx = fij[b], fy = fij[b + 1], fz = dim == 3 ? fij[b + 2] : 0.0;
    atomicAdd(&v[0 * static_cast<size_t>(inum) + i], factor * dx * fx);
    atomicAdd(&v[1 * static_cast<size_t>(inum) + i], factor * dy * fy);
    atomicAdd(&v[2 * static_cast<size_t>(inum) + i], factor * dz * fz);
    atomicAdd(&v[3 * static_cast<size_t>(inum) + i], factor * dx * fy);
    atomicAdd(&v[4 * static_cast<size_t>(inum) + i], factor * dx * fz);
    atomicAdd(&v[5 * static_cast<size_t>(inum) + i], factor * dy * fz);
}

template <class F>
void guarded_call(F&& f) {
    std::lock_guard<std::mutex> lk(library_mutex());
    clear_error();
    try {
        f();
    } catch (const std::exception& e) {
        set_error(e.what());
    }
}

void sync(hipStream_t st) {
    MDP_HIP(hipGetLastError());
    MDP_HIP(hipStreamSynchronize(st));
}

inline int blocks(long n, int b) { return static_cast<int>((n + b - 1) / b); }

}  // namespace

extern "C" {

void mdp_hipPairaGradient(double* u, double* u_xij, double* xij, double* qi, double* qj, int* ti, int* tj, int* ai,
                          int* aj, double* mu, double* eta, int* kappa, int dim, int ncq, int nmu, int neta,
                          int nkappa, int ng, int potnum) {
    (void)qi, (void)qj, (void)ti, (void)tj, (void)ai, (void)aj, (void)eta, (void)kappa, (void)ncq, (void)neta,
        (void)nkappa;
    guarded_call([&] {
        if (potnum != 1) throw Error("mdp_hipPairaGradient: potnum must be 1 (the ZBL pair potential)");
        if (dim != 3) throw Error("mdp_hipPairaGradient: dim must be 3");
        if (nmu < 3 || !mu) throw Error("mdp_hipPairaGradient: mu must hold [cut_inner, cut_outer, z]");
        if (ng < 0) throw Error("mdp_hipPairaGradient: negative pair count");
        if (ng == 0) return;
        hipStream_t st = library_stream();
        double m[3];
        MDP_HIP(hipMemcpyAsync(m, mu, sizeof(m), hipMemcpyDeviceToHost, st));
        MDP_HIP(hipStreamSynchronize(st));
        if (!(m[1] > m[0]) || !(m[2] > 0.0)) throw Error("mdp_hipPairaGradient: need cut_outer > cut_inner and z > 0");
        hipLaunchKernelGGL(k_paira_zbl, dim3(blocks(ng, 256)), dim3(256), 0, st, u, u_xij, xij,
                           zbl_params(m[0], m[1], m[2]), ng);
        sync(st);
    });
}

void mdp_hipFullNeighPairDecomposition(double* e, double* f, double* eij, double* fij, int* ai, int ijnum, int dim) {
    guarded_call([&] {
        if (dim != 2 && dim != 3) throw Error("mdp_hipFullNeighPairDecomposition: dim must be 2 or 3");
        if (ijnum <= 0) return;
        hipStream_t st = library_stream();
        hipLaunchKernelGGL(k_full_pair_decomp, dim3(blocks(ijnum, 256)), dim3(256), 0, st, e, f, eij, fij, ai, ijnum,
                           dim);
        sync(st);
    });
}

void mdp_hipVirialPairTally(double* v, double* fij, double* rij, double factor, int* ai, int dim, int inum,
                            int ijnum) {
    guarded_call([&] {
        if (dim != 2 && dim != 3) throw Error("mdp_hipVirialPairTally: dim must be 2 or 3");
        if (ijnum <= 0) return;
        hipStream_t st = library_stream();
        hipLaunchKernelGGL(k_virial_pair_tally, dim3(blocks(ijnum, 256)), dim3(256), 0, st, v, fij, rij, factor, ai,
                           dim, inum, ijnum);
        sync(st);
    });
}

}  // extern "C"
