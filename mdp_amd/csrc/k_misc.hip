// Device kernels of the descriptor path (cpuComputeSij/Uij, Zero/AddUarraytot, Bi2, Dbidrj), the
// tallies, the pair-list builders, AddWself2Ui and a device-wide exclusive scan.  These follow the
// upstream functions one output element (or one pair) per thread with runtime loops; the
// performance-critical force path lives in k_ui/k_yi/k_dedr and the MD engine.
#include <hip/hip_runtime.h>

#include "snap_device.hpp"
#include "snap_launch.hpp"

namespace mdp {

static inline int nblk(long n, int b) { return static_cast<int>((n + b - 1) / b); }

__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }

// ================================================================================================
// exclusive scan (three phases)
// ================================================================================================

constexpr int SCAN_BLOCK = 1024;

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_block(const int* __restrict__ in, int* __restrict__ out,
                                                           int* __restrict__ bsum, int n) {
    __shared__ int sh[SCAN_BLOCK];
    const int i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
    const int v = (i < n) ? in[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < SCAN_BLOCK; d <<= 1) {
        const int t = (threadIdx.x >= static_cast<unsigned>(d)) ? sh[threadIdx.x - d] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = sh[threadIdx.x] - v;
    if (threadIdx.x == SCAN_BLOCK - 1) bsum[blockIdx.x] = sh[threadIdx.x];
}

// exclusive scan of the block sums by one workgroup (chunks of 1024), total to *total
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(int* __restrict__ bsum, int nb, int* __restrict__ total) {
    __shared__ int sh[SCAN_BLOCK];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += SCAN_BLOCK) {
        const int i = base + threadIdx.x;
        const int v = (i < nb) ? bsum[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < SCAN_BLOCK; d <<= 1) {
            const int t = (threadIdx.x >= static_cast<unsigned>(d)) ? sh[threadIdx.x - d] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + sh[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == SCAN_BLOCK - 1) carry += sh[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_add(int* __restrict__ out, const int* __restrict__ bsum, int n) {
    const int i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
    if (i < n) out[i] += bsum[blockIdx.x];
}

size_t scan_tmp_ints(int n) { return static_cast<size_t>((n + SCAN_BLOCK - 1) / SCAN_BLOCK) + 1; }

void exclusive_scan(const int* in, int* out, int n, int* tmp, hipStream_t st) {
    const int nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (nb > 0) hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(SCAN_BLOCK), 0, st, in, out, tmp, n);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, st, tmp, nb, out + n);
    if (nb > 0) hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(SCAN_BLOCK), 0, st, out, tmp, n);
}

// ================================================================================================
// per-pair U and dU (layer-by-layer recursion, cpuSnap.cpp:1412-1833)
// ================================================================================================

__global__ __launch_bounds__(128) void k_uij(SnapDev s, double* __restrict__ Sr, double* __restrict__ Si,
                                             double* dr0, double* dr1, double* dr2, double* di0, double* di1,
                                             double* di2, const double* __restrict__ rij, const int* __restrict__ ti,
                                             const int* __restrict__ tj, int ijnum) {
    const int ij = blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= ijnum) return;
    const int J = s.J, jd = J + 1;
    const double rc = (s.radelem[ti[ij]] + s.radelem[tj[ij]]) * s.rcutfac;
    PairGeom g;
    pair_geometry<true>(rij[3 * ij], rij[3 * ij + 1], rij[3 * ij + 2], rc, s.rmin0, s.rfac0, s.switchflag,
                        s.wjelem[tj[ij]], g);
    double* dR[3] = {dr0, dr1, dr2};
    double* dI[3] = {di0, di1, di2};
    const size_t P = static_cast<size_t>(ijnum);
#define AT(arr, k) arr[ij + P * static_cast<size_t>(k)]
    AT(Sr, 0) = 1.0;
    AT(Si, 0) = 0.0;
    for (int k = 0; k < 3; ++k) { AT(dR[k], 0) = 0.0; AT(dI[k], 0) = 0.0; }
    for (int j = 1; j <= J; ++j) {
        int jju = u_block(j), jjup = u_block(j - 1);
        for (int mb = 0; 2 * mb <= j; ++mb) {
            AT(Sr, jju) = 0.0;
            AT(Si, jju) = 0.0;
            for (int k = 0; k < 3; ++k) { AT(dR[k], jju) = 0.0; AT(dI[k], jju) = 0.0; }
            for (int ma = 0; ma < j; ++ma) {
                double rp = s.rootpq[(j - ma) * jd + (j - mb)];
                const double ur = AT(Sr, jjup), ui = AT(Si, jjup);
                double dur[3], dui[3];
                for (int k = 0; k < 3; ++k) { dur[k] = AT(dR[k], jjup); dui[k] = AT(dI[k], jjup); }
                AT(Sr, jju) += rp * (g.ar * ur + g.ai * ui);
                AT(Si, jju) += rp * (g.ar * ui - g.ai * ur);
                for (int k = 0; k < 3; ++k) {
                    AT(dR[k], jju) += rp * (g.dar[k] * ur + g.dai[k] * ui + g.ar * dur[k] + g.ai * dui[k]);
                    AT(dI[k], jju) += rp * (g.dar[k] * ui - g.dai[k] * ur + g.ar * dui[k] - g.ai * dur[k]);
                }
                rp = s.rootpq[(ma + 1) * jd + (j - mb)];
                AT(Sr, jju + 1) = -rp * (g.br * ur + g.bi * ui);
                AT(Si, jju + 1) = -rp * (g.br * ui - g.bi * ur);
                for (int k = 0; k < 3; ++k) {
                    AT(dR[k], jju + 1) = -rp * (g.dbr[k] * ur + g.dbi[k] * ui + g.br * dur[k] + g.bi * dui[k]);
                    AT(dI[k], jju + 1) = -rp * (g.dbr[k] * ui - g.dbi[k] * ur + g.br * dui[k] - g.bi * dur[k]);
                }
                ++jju;
                ++jjup;
            }
            ++jju;
        }
        int lo = u_block(j), hi = lo + (j + 1) * (j + 1) - 1;
        for (int mb = 0; 2 * mb <= j; ++mb) {
            const bool with_d = (j % 2 == 1 && mb == j / 2);
            for (int ma = 0; ma <= j; ++ma, ++lo, --hi) {
                const bool even = ((ma + mb) & 1) == 0;
                const double vr = AT(Sr, lo), vi = AT(Si, lo);
                AT(Sr, hi) = even ? vr : -vr;
                AT(Si, hi) = even ? -vi : vi;
                if (with_d)
                    for (int k = 0; k < 3; ++k) {
                        const double wr = AT(dR[k], lo), wi = AT(dI[k], lo);
                        AT(dR[k], hi) = even ? wr : -wr;
                        AT(dI[k], hi) = even ? -wi : wi;
                    }
            }
        }
    }
    for (int j = 0; j <= J; ++j) {
        int jju = u_block(j);
        for (int mb = 0; 2 * mb <= j; ++mb)
            for (int ma = 0; ma <= j; ++ma, ++jju) {
                const double vr = AT(Sr, jju), vi = AT(Si, jju);
                for (int k = 0; k < 3; ++k) {
                    AT(dR[k], jju) = g.dsfac * vr * g.u[k] + g.sfac * AT(dR[k], jju);
                    AT(dI[k], jju) = g.dsfac * vi * g.u[k] + g.sfac * AT(dI[k], jju);
                }
            }
    }
    for (int k = 0; k < s.idxu_max; ++k) {
        AT(Sr, k) = g.sfac * AT(Sr, k);
        AT(Si, k) = g.sfac * AT(Si, k);
    }
#undef AT
}

void launch_uij(const SnapDev& s, double* Sr, double* Si, double* const dSr[3], double* const dSi[3],
                const double* rij, const int* ti, const int* tj, int ijnum, hipStream_t st) {
    if (ijnum <= 0) return;
    hipLaunchKernelGGL(k_uij, dim3(nblk(ijnum, 128)), dim3(128), 0, st, s, Sr, Si, dSr[0], dSr[1], dSr[2], dSi[0],
                       dSi[1], dSi[2], rij, ti, tj, ijnum);
}

// ================================================================================================
// Utot assembly
// ================================================================================================

__global__ void k_zero_utot(SnapDev s, double* __restrict__ Sr, double* __restrict__ Si, double wself,
                            const int* __restrict__ type, const int* __restrict__ ai, int wselfall, int chemflag,
                            int nelem, int N) {
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= N) return;
    const int ielem = chemflag ? s.map[type[ai[ii]]] : 0;
    const size_t n = static_cast<size_t>(N);
    for (int je = 0; je < nelem; ++je) {
        const size_t base = ii + n * s.idxu_max * je;
        int jju = 0;
        for (int j = 0; j <= s.J; ++j)
            for (int mb = 0; mb <= j; ++mb)
                for (int ma = 0; ma <= j; ++ma, ++jju) {
                    Sr[base + n * jju] = (ma == mb && (je == ielem || wselfall)) ? wself : 0.0;
                    Si[base + n * jju] = 0.0;
                }
    }
}

void launch_zero_utot(const SnapDev& s, double* Sr, double* Si, double wself, const int* type, const int* ai,
                      int wselfall, int chemflag, int nelem, int N, hipStream_t st) {
    if (N <= 0) return;
    hipLaunchKernelGGL(k_zero_utot, dim3(nblk(N, 128)), dim3(128), 0, st, s, Sr, Si, wself, type, ai, wselfall,
                       chemflag, nelem, N);
}

__global__ void k_add_utot(SnapDev s, double* __restrict__ Tr, double* __restrict__ Ti, const double* __restrict__ Sr,
                           const double* __restrict__ Si, const int* __restrict__ ai, const int* __restrict__ tj,
                           int chemflag, int N, int ijnum) {
    const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<long>(ijnum) * s.idxu_max) return;
    const int ij = static_cast<int>(idx % ijnum);
    const long jju = idx / ijnum;
    const size_t n = static_cast<size_t>(N);
    const size_t d = ai[ij] + n * jju + (chemflag ? n * s.idxu_max * s.map[tj[ij]] : 0);
    atomic_add(&Tr[d], Sr[idx]);
    atomic_add(&Ti[d], Si[idx]);
}

void launch_add_utot(const SnapDev& s, double* Tr, double* Ti, const double* Sr, const double* Si, const int* ai,
                     const int* tj, int chemflag, int N, int ijnum, hipStream_t st) {
    const long tot = static_cast<long>(ijnum) * s.idxu_max;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_add_utot, dim3(nblk(tot, 256)), dim3(256), 0, st, s, Tr, Ti, Sr, Si, ai, tj, chemflag, N, ijnum);
}

// AddWself2Ui (cpuSnap.cpp:492-542): one thread per (atom, element block)
__global__ void k_add_wself(SnapDev s, double* __restrict__ Ur, double* __restrict__ Ui, double wself,
                            const int* __restrict__ type, int wselfall, int chemflag, int nelem, int N) {
    const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<long>(N) * nelem) return;
    const int ii = static_cast<int>(idx % N), je = static_cast<int>(idx / N);
    const int ielem = chemflag ? s.map[type[ii]] : 0;
    const size_t n = static_cast<size_t>(N);
    const size_t base = ii + n * s.idxu_max * je;
    for (int j = 0; j <= s.J; ++j) {
        const int b = u_block(j);
        if (je == ielem || wselfall)
            for (int m = 0; m <= j; ++m) Ur[base + n * (b + (j + 1) * m + m)] += wself;
        int lo = b, hi = b + (j + 1) * (j + 1) - 1;
        for (int mb = 0; 2 * mb < j; ++mb)
            for (int ma = 0; ma <= j; ++ma, ++lo, --hi) {
                const double vr = Ur[base + n * lo], vi = Ui[base + n * lo];
                const bool even = ((ma + mb) & 1) == 0;
                Ur[base + n * hi] = even ? vr : -vr;
                Ui[base + n * hi] = even ? -vi : vi;
            }
    }
}

void launch_add_wself(const SnapDev& s, double* Ur, double* Ui, double wself, const int* type, int wselfall,
                      int chemflag, int nelem, int N, hipStream_t st) {
    const long tot = static_cast<long>(N) * nelem;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_add_wself, dim3(nblk(tot, 128)), dim3(128), 0, st, s, Ur, Ui, wself, type, wselfall, chemflag,
                       nelem, N);
}

// ================================================================================================
// B from a precomputed Z list (cpuComputeBi2, cpuSnap.cpp:1971-2072) and the bzero shift
// ================================================================================================

__global__ void k_bi2(SnapDev s, double* __restrict__ blist, const double* __restrict__ zr,
                      const double* __restrict__ zi, const double* __restrict__ Sr, const double* __restrict__ Si,
                      int nelem, int N) {
    const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long nb = static_cast<long>(N) * s.idxb_max;
    if (idx >= nb * nelem * nelem * nelem) return;
    const int ii = static_cast<int>(idx % N);
    const int jjb = static_cast<int>((idx / N) % s.idxb_max);
    const int jelem = static_cast<int>(idx / nb);
    const int e3 = jelem % nelem, e2 = (jelem / nelem) % nelem, e1 = jelem / (nelem * nelem);
    const int jd = s.J + 1;
    const int j1 = s.idxb[3 * jjb], j2 = s.idxb[3 * jjb + 1], j = s.idxb[3 * jjb + 2];
    const size_t n = static_cast<size_t>(N);
    const size_t zoff = ii + n * s.idxz_max * (e2 + nelem * e1), uoff = ii + n * s.idxu_max * e3;
    int jjz = s.idxz_block[j + j2 * jd + j1 * jd * jd], jju = u_block(j);
    double sum = 0.0;
    for (int mb = 0; 2 * mb < j; ++mb)
        for (int ma = 0; ma <= j; ++ma, ++jjz, ++jju)
            sum += Sr[uoff + n * jju] * zr[zoff + n * jjz] + Si[uoff + n * jju] * zi[zoff + n * jjz];
    if (j % 2 == 0) {
        const int mb = j / 2;
        for (int ma = 0; ma < mb; ++ma, ++jjz, ++jju)
            sum += Sr[uoff + n * jju] * zr[zoff + n * jjz] + Si[uoff + n * jju] * zi[zoff + n * jjz];
        sum += 0.5 * (Sr[uoff + n * jju] * zr[zoff + n * jjz] + Si[uoff + n * jju] * zi[zoff + n * jjz]);
    }
    blist[idx] = 2.0 * sum;
}

// own == 1: subtract from the atom's own-element triple only (cpuSnap.cpp:2022-2035), else from every
// triple with elem1 == elem2 == elem3 (cpuSnap.cpp:699-708); wselfall: from every triple.
__global__ void k_bzero(SnapDev s, double* __restrict__ blist, const double* __restrict__ bzero,
                        const int* __restrict__ ilist, const int* __restrict__ type, int wselfall, int chemflag,
                        int nelem, int N, int own) {
    const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long nb = static_cast<long>(N) * s.idxb_max;
    if (idx >= nb * nelem * nelem * nelem) return;
    const int ii = static_cast<int>(idx % N);
    const int jjb = static_cast<int>((idx / N) % s.idxb_max);
    const int jelem = static_cast<int>(idx / nb);
    const int e3 = jelem % nelem, e2 = (jelem / nelem) % nelem, e1 = jelem / (nelem * nelem);
    bool sub;
    if (wselfall) {
        sub = true;
    } else if (own) {
        const int ie = chemflag ? s.map[type[ilist[ii]]] : 0;
        sub = (e1 == ie && e2 == ie && e3 == ie);
    } else {
        sub = (e1 == e2 && e1 == e3);
    }
    if (sub) blist[idx] -= bzero[s.idxb[3 * jjb + 2]];
}

void launch_bi2(const SnapDev& s, double* blist, const double* zr, const double* zi, const double* Sr, const double* Si,
                const double* bzero, const int* ilist, const int* type, int bzero_flag, int wselfall, int chemflag,
                int nelem, int N, hipStream_t st) {
    const long tot = static_cast<long>(N) * s.idxb_max * nelem * nelem * nelem;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_bi2, dim3(nblk(tot, 128)), dim3(128), 0, st, s, blist, zr, zi, Sr, Si, nelem, N);
    if (bzero_flag)
        hipLaunchKernelGGL(k_bzero, dim3(nblk(tot, 256)), dim3(256), 0, st, s, blist, bzero, ilist, type, wselfall,
                           chemflag, nelem, N, 1);
}

void launch_bzero(const SnapDev& s, double* blist, const double* bzero, const int* ilist, const int* type, int wselfall,
                  int chemflag, int nelem, int N, int own, hipStream_t st) {
    const long tot = static_cast<long>(N) * s.idxb_max * nelem * nelem * nelem;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_bzero, dim3(nblk(tot, 256)), dim3(256), 0, st, s, blist, bzero, ilist, type, wselfall, chemflag,
                       nelem, N, own);
}

// eatom[ilist[ii]] += beta0 + sum_k beta_k B_k, in upstream order (element triple, then jjb)
__global__ void k_ei_from_b(SnapDev s, double* __restrict__ eatom, const double* __restrict__ blist,
                            const double* __restrict__ coeff, const int* __restrict__ ilist,
                            const int* __restrict__ type, int chemflag, int nelem, int ncoeffall, int N) {
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= N) return;
    const int ie = chemflag ? s.map[type[ii]] : 0;
    const size_t n = static_cast<size_t>(N);
    double e = eatom[ilist[ii]];
    const int nt = nelem * nelem * nelem;
    for (int je = 0; je < nt; ++je)
        for (int jjb = 0; jjb < s.idxb_max; ++jjb) {
            const int ic = jjb + s.idxb_max * je;
            if (ic == 0) e += coeff[ie * ncoeffall];
            e += coeff[1 + ic + ie * ncoeffall] * blist[ii + n * jjb + n * s.idxb_max * je];
        }
    eatom[ilist[ii]] = e;
}

void launch_ei_from_b(const SnapDev& s, double* eatom, const double* blist, const double* coeff, const int* ilist,
                      const int* type, int chemflag, int nelem, int ncoeffall, int N, hipStream_t st) {
    if (N <= 0) return;
    hipLaunchKernelGGL(k_ei_from_b, dim3(nblk(N, 128)), dim3(128), 0, st, s, eatom, blist, coeff, ilist, type, chemflag,
                       nelem, ncoeffall, N);
}

// ================================================================================================
// dB/dr (cpuComputeDbidrj, cpuSnap.cpp:2074-2263): one thread per (pair, jjb)
// ================================================================================================

__device__ __forceinline__ void dz_half(double out[3], const double* __restrict__ zr, const double* __restrict__ zi,
                                        const double* __restrict__ dur, const double* __restrict__ dui, size_t z0,
                                        size_t zs, int jjz, int jju, int jl, size_t ij, size_t ps, size_t nu) {
    out[0] = out[1] = out[2] = 0.0;
    for (int mb = 0; 2 * mb < jl; ++mb)
        for (int ma = 0; ma <= jl; ++ma, ++jjz, ++jju) {
            const double a = zr[z0 + zs * jjz], b = zi[z0 + zs * jjz];
            for (int k = 0; k < 3; ++k) out[k] += dur[ij + ps * jju + nu * k] * a + dui[ij + ps * jju + nu * k] * b;
        }
    if (jl % 2 == 0) {
        const int mb = jl / 2;
        for (int ma = 0; ma < mb; ++ma, ++jjz, ++jju) {
            const double a = zr[z0 + zs * jjz], b = zi[z0 + zs * jjz];
            for (int k = 0; k < 3; ++k) out[k] += dur[ij + ps * jju + nu * k] * a + dui[ij + ps * jju + nu * k] * b;
        }
        const double a = zr[z0 + zs * jjz], b = zi[z0 + zs * jjz];
        for (int k = 0; k < 3; ++k) out[k] += (dur[ij + ps * jju + nu * k] * a + dui[ij + ps * jju + nu * k] * b) * 0.5;
    }
}

__global__ void k_dbidrj(SnapDev s, double* __restrict__ dblist, const double* __restrict__ zr,
                         const double* __restrict__ zi, const double* __restrict__ dur, const double* __restrict__ dui,
                         const int* __restrict__ ai, const int* __restrict__ tj, int bnorm, int chemflag, int nelem,
                         int N, int ijnum) {
    const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<long>(ijnum) * s.idxb_max) return;
    const int ij = static_cast<int>(idx % ijnum), jjb = static_cast<int>(idx / ijnum);
    const int jd = s.J + 1;
    const int e3 = chemflag ? s.map[tj[ij]] : 0;
    const int i = ai[ij];
    const int j1 = s.idxb[3 * jjb], j2 = s.idxb[3 * jjb + 1], j = s.idxb[3 * jjb + 2];
    const size_t n = static_cast<size_t>(N), P = static_cast<size_t>(ijnum);
    const size_t nz = n * s.idxz_max, nbm = P * s.idxb_max, nu = P * s.idxu_max;
    const double j1fac = (j + 1) / (j1 + 1.0), j2fac = (j + 1) / (j2 + 1.0);
    for (int e1 = 0; e1 < nelem; ++e1)
        for (int e2 = 0; e2 < nelem; ++e2) {
            const size_t z0 = i + nz * (e1 * nelem + e2);
            double* dbdr = dblist + nbm * 3 * ((e1 * nelem + e2) * nelem + e3);
            const size_t o = ij + P * 3 * jjb;
            double sv[3];
            dz_half(sv, zr, zi, dur, dui, z0, n, s.idxz_block[j + j2 * jd + j1 * jd * jd], u_block(j), j, ij, P, nu);
            for (int k = 0; k < 3; ++k) dbdr[o + P * k] += 2.0 * sv[k];
            dz_half(sv, zr, zi, dur, dui, z0, n, s.idxz_block[j1 + j2 * jd + j * jd * jd], u_block(j1), j1, ij, P, nu);
            for (int k = 0; k < 3; ++k) dbdr[o + P * k] += bnorm ? 2.0 * sv[k] : 2.0 * sv[k] * j1fac;
            dz_half(sv, zr, zi, dur, dui, z0, n, s.idxz_block[j2 + j1 * jd + j * jd * jd], u_block(j2), j2, ij, P, nu);
            for (int k = 0; k < 3; ++k) dbdr[o + P * k] += bnorm ? 2.0 * sv[k] : 2.0 * sv[k] * j2fac;
        }
}

void launch_dbidrj(const SnapDev& s, double* dblist, const double* zr, const double* zi, const double* dur,
                   const double* dui, const int* ai, const int* tj, int bnorm, int chemflag, int nelem, int N, int ijnum,
                   hipStream_t st) {
    const long tot = static_cast<long>(ijnum) * s.idxb_max;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_dbidrj, dim3(nblk(tot, 128)), dim3(128), 0, st, s, dblist, zr, zi, dur, dui, ai, tj, bnorm,
                       chemflag, nelem, N, ijnum);
}

// ================================================================================================
// tallies (cpuSnap.cpp:2265-2470); packed quadratic rows start after the ncoeff linear terms
// ================================================================================================

__device__ __forceinline__ int quad_row(int nc, int ic) { return nc + nc * (nc + 1) / 2 - (nc - ic) * (nc - ic + 1) / 2; }

__global__ void k_tally_bi(double* __restrict__ bi, const double* __restrict__ b, const int* __restrict__ ilist,
                           const int* __restrict__ type, int N, int ncoeff, int nperdim, int quadratic) {
    const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<long>(N) * ncoeff) return;
    const int ii = static_cast<int>(idx % N), ic = static_cast<int>(idx / N);
    const int n = nperdim * (type[ilist[ii]] - 1);
    const double v = b[idx];
    atomic_add(&bi[ic + n], v);
    if (quadratic == 1) {
        int k = n + quad_row(ncoeff, ic);
        atomic_add(&bi[k], 0.5 * v * v);
        // upstream placement: the first cross term shares the slot of the square term
        for (int jc = ic + 1; jc < ncoeff; ++jc) atomic_add(&bi[k++], v * b[ii + static_cast<long>(N) * jc]);
    }
}

void launch_tally_bi(double* bi, const double* b, const int* ilist, const int* type, int N, int ncoeff, int nperdim,
                     int quadratic, hipStream_t st) {
    const long tot = static_cast<long>(N) * ncoeff;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_tally_bi, dim3(nblk(tot, 256)), dim3(256), 0, st, bi, b, ilist, type, N, ncoeff, nperdim,
                       quadratic);
}

__global__ void k_tally_bd(double* __restrict__ db, const double* __restrict__ b, const double* __restrict__ dbdr,
                           const int* __restrict__ aii, const int* __restrict__ ai, const int* __restrict__ aj,
                           const int* __restrict__ ti, int N, int ijnum, int ncoeff, int nperdim, int quadratic) {
    const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<long>(ijnum) * ncoeff) return;
    const int ij = static_cast<int>(idx % ijnum), ic = static_cast<int>(idx / ijnum);
    const int ii = aii[ij], i = ai[ij], j = aj[ij];
    const int n = nperdim * (ti[ij] - 1);
    const size_t P = static_cast<size_t>(ijnum);
    const size_t nij = P * 3 * ic;
    const double bx = dbdr[ij + nij], by = dbdr[ij + P + nij], bz = dbdr[ij + 2 * P + nij];
    size_t nii = static_cast<size_t>(N) * 3 * (ic + n);
    atomic_add(&db[3 * i + nii], bx);
    atomic_add(&db[1 + 3 * i + nii], by);
    atomic_add(&db[2 + 3 * i + nii], bz);
    atomic_add(&db[3 * j + nii], -bx);
    atomic_add(&db[1 + 3 * j + nii], -by);
    atomic_add(&db[2 + 3 * j + nii], -bz);
    if (quadratic) {
        const double bv = b[ii + static_cast<size_t>(N) * ic];
        int k = quad_row(ncoeff, ic);
        double tx = bv * bx, ty = bv * by, tz = bv * bz;
        nii = static_cast<size_t>(N) * 3 * (k + n);
        atomic_add(&db[3 * i + nii], tx);
        atomic_add(&db[1 + 3 * i + nii], ty);
        atomic_add(&db[2 + 3 * i + nii], tz);
        atomic_add(&db[3 * j + nii], -tx);
        atomic_add(&db[1 + 3 * j + nii], -ty);
        atomic_add(&db[2 + 3 * j + nii], -tz);
        for (int jc = ic + 1; jc < ncoeff; ++jc) {
            const size_t nj = P * 3 * jc;
            const double cx = dbdr[ij + nj], cy = dbdr[ij + P + nj], cz = dbdr[ij + 2 * P + nj];
            const double c = b[ii + static_cast<size_t>(N) * jc];
            tx = bv * cx + bx * c;
            ty = bv * cy + by * c;
            tz = bv * cz + bz * c;
            ++k;
            nii = static_cast<size_t>(N) * 3 * (k + n);
            atomic_add(&db[3 * i + nii], tx);
            atomic_add(&db[1 + 3 * i + nii], ty);
            atomic_add(&db[2 + 3 * i + nii], tz);
            atomic_add(&db[3 * j + nii], -tx);
            atomic_add(&db[1 + 3 * j + nii], -ty);
            atomic_add(&db[2 + 3 * j + nii], -tz);
        }
    }
}

void launch_tally_bd(double* db, const double* b, const double* dbdr, const int* aii, const int* ai, const int* aj,
                     const int* ti, int N, int ijnum, int ncoeff, int nperdim, int quadratic, hipStream_t st) {
    const long tot = static_cast<long>(ijnum) * ncoeff;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_tally_bd, dim3(nblk(tot, 256)), dim3(256), 0, st, db, b, dbdr, aii, ai, aj, ti, N, ijnum,
                       ncoeff, nperdim, quadratic);
}

__device__ __forceinline__ void add_vir(double* v, double dx, double dy, double dz, double bx, double by, double bz) {
    atomic_add(&v[0], dx * bx);
    atomic_add(&v[1], dy * by);
    atomic_add(&v[2], dz * bz);
    atomic_add(&v[3], dx * by);
    atomic_add(&v[4], dx * bz);
    atomic_add(&v[5], dy * bz);
}

__global__ void k_tally_bv(double* __restrict__ bv, const double* __restrict__ b, const double* __restrict__ dbdr,
                           const double* __restrict__ rij, const int* __restrict__ aii, const int* __restrict__ ti, int N,
                           int ijnum, int ncoeff, int nperdim, int quadratic) {
    const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<long>(ijnum) * ncoeff) return;
    const int ij = static_cast<int>(idx % ijnum), ic = static_cast<int>(idx / ijnum);
    const int ii = aii[ij];
    const int n = nperdim * (ti[ij] - 1);
    const size_t P = static_cast<size_t>(ijnum);
    const size_t nij = P * 3 * ic;
    const double dx = -rij[3 * ij], dy = -rij[3 * ij + 1], dz = -rij[3 * ij + 2];
    const double bx = dbdr[ij + nij], by = dbdr[ij + P + nij], bz = dbdr[ij + 2 * P + nij];
    add_vir(bv + 6 * (ic + n), dx, dy, dz, bx, by, bz);
    if (quadratic) {
        const double bval = b[ii + static_cast<size_t>(N) * ic];
        int k = quad_row(ncoeff, ic);
        add_vir(bv + 6 * (k + n), dx, dy, dz, bval * bx, bval * by, bval * bz);
        for (int jc = ic + 1; jc < ncoeff; ++jc) {
            const size_t nj = P * 3 * jc;
            const double c = b[ii + static_cast<size_t>(N) * jc];
            ++k;
            add_vir(bv + 6 * (k + n), dx, dy, dz, bval * dbdr[ij + nj] + bx * c, bval * dbdr[ij + P + nj] + by * c,
                    bval * dbdr[ij + 2 * P + nj] + bz * c);
        }
    }
}

void launch_tally_bv(double* bv, const double* b, const double* dbdr, const double* rij, const int* aii, const int* ti,
                     int N, int ijnum, int ncoeff, int nperdim, int quadratic, hipStream_t st) {
    const long tot = static_cast<long>(ijnum) * ncoeff;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_tally_bv, dim3(nblk(tot, 256)), dim3(256), 0, st, bv, b, dbdr, rij, aii, ti, N, ijnum, ncoeff,
                       nperdim, quadratic);
}

__global__ void k_tally_force_atomic(double* __restrict__ f, double* __restrict__ v, const double* __restrict__ dedr,
                                     const double* __restrict__ rij, const int* __restrict__ ai,
                                     const int* __restrict__ aj, int ijnum) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ijnum) return;
    const double dx = dedr[3 * p], dy = dedr[3 * p + 1], dz = dedr[3 * p + 2];
    const double x = rij[3 * p], y = rij[3 * p + 1], z = rij[3 * p + 2];
    const int a = ai[p], b = aj[p];
    atomic_add(&f[3 * a], dx);
    atomic_add(&f[3 * a + 1], dy);
    atomic_add(&f[3 * a + 2], dz);
    atomic_add(&f[3 * b], -dx);
    atomic_add(&f[3 * b + 1], -dy);
    atomic_add(&f[3 * b + 2], -dz);
    atomic_add(&v[6 * a + 0], -x * dx);
    atomic_add(&v[6 * a + 1], -y * dy);
    atomic_add(&v[6 * a + 2], -z * dz);
    atomic_add(&v[6 * a + 3], -x * dy);
    atomic_add(&v[6 * a + 4], -x * dz);
    atomic_add(&v[6 * a + 5], -y * dz);
}

void launch_tally_force_atomic(double* f, double* v, const double* dedr, const double* rij, const int* ai, const int* aj,
                               int ijnum, hipStream_t st) {
    if (ijnum <= 0) return;
    hipLaunchKernelGGL(k_tally_force_atomic, dim3(nblk(ijnum, 256)), dim3(256), 0, st, f, v, dedr, rij, ai, aj, ijnum);
}

// The backend-3 SnapComputeFi tally (src/C++/Core/cpuSnap4.cpp:1176-1209): forces as above, virial in
// the SoA layout v[k * anum + atom] with -r(x)dE/2 added to both atoms of the pair, components
// (xx, yy, zz, xy, xz, yz).
__global__ void k_tally_force_atomic_soa(double* __restrict__ f, double* __restrict__ v,
                                         const double* __restrict__ dedr, const double* __restrict__ rij,
                                         const int* __restrict__ ai, const int* __restrict__ aj, int ijnum, int anum) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ijnum) return;
    const double dx = dedr[3 * p], dy = dedr[3 * p + 1], dz = dedr[3 * p + 2];
    const double x = rij[3 * p], y = rij[3 * p + 1], z = rij[3 * p + 2];
    const int a = ai[p], b = aj[p];
    atomic_add(&f[3 * a], dx);
    atomic_add(&f[3 * a + 1], dy);
    atomic_add(&f[3 * a + 2], dz);
    atomic_add(&f[3 * b], -dx);
    atomic_add(&f[3 * b + 1], -dy);
    atomic_add(&f[3 * b + 2], -dz);
    const double w[6] = {-0.5 * x * dx, -0.5 * y * dy, -0.5 * z * dz, -0.5 * x * dy, -0.5 * x * dz, -0.5 * y * dz};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        atomic_add(&v[static_cast<size_t>(k) * anum + a], w[k]);
        atomic_add(&v[static_cast<size_t>(k) * anum + b], w[k]);
    }
}

void launch_tally_force_atomic_soa(double* f, double* v, const double* dedr, const double* rij, const int* ai,
                                   const int* aj, int ijnum, int anum, hipStream_t st) {
    if (ijnum <= 0) return;
    hipLaunchKernelGGL(k_tally_force_atomic_soa, dim3(nblk(ijnum, 256)), dim3(256), 0, st, f, v, dedr, rij, ai, aj,
                       ijnum, anum);
}

// ================================================================================================
// pair lists from a CSR neighbour list (cpuSnap.cpp:258-310)
// ================================================================================================

__device__ __forceinline__ bool in_pair(const double* __restrict__ x, int i, int j, int dim, double rcsq) {
#pragma clang fp contract(off)
    const double dx = x[j * dim] - x[i * dim];
    const double dy = x[j * dim + 1] - x[i * dim + 1];
    const double dz = x[j * dim + 2] - x[i * dim + 2];
    const double d2 = dx * dx + dy * dy + dz * dz;
    return d2 < rcsq && d2 > 1e-20;
}

__global__ void k_pair_count(int* __restrict__ pairnum, const double* __restrict__ x, const double* __restrict__ rcutsq,
                             const int* __restrict__ ilist, const int* __restrict__ neighlist,
                             const int* __restrict__ neighnum, const int* __restrict__ atomtype,
                             const int* __restrict__ alist, int inum, int dim, int ntypes) {
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= inum) return;
    const int i = ilist[ii], it = atomtype[i];
    int c = 0;
    for (int l = neighnum[i]; l < neighnum[i + 1]; ++l) {
        const int j = neighlist[l];
        if (in_pair(x, i, j, dim, rcutsq[atomtype[alist[j]] + it * (ntypes + 1)])) ++c;
    }
    pairnum[ii] = c;
}

__global__ void k_pair_fill(int* __restrict__ pairlist, const int* __restrict__ off, const double* __restrict__ x,
                            const double* __restrict__ rcutsq, const int* __restrict__ ilist,
                            const int* __restrict__ neighlist, const int* __restrict__ neighnum,
                            const int* __restrict__ atomtype, const int* __restrict__ alist, int inum, int dim,
                            int ntypes) {
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= inum) return;
    const int i = ilist[ii], it = atomtype[i];
    int c = off[ii];
    for (int l = neighnum[i]; l < neighnum[i + 1]; ++l) {
        const int j = neighlist[l];
        if (in_pair(x, i, j, dim, rcutsq[atomtype[alist[j]] + it * (ntypes + 1)])) pairlist[c++] = j;
    }
}

void launch_pair_count(int* pairnum, const double* x, const double* rcutsq, const int* ilist, const int* neighlist,
                       const int* neighnum, const int* atomtype, const int* alist, int inum, int dim, int ntypes,
                       hipStream_t st) {
    if (inum <= 0) return;
    hipLaunchKernelGGL(k_pair_count, dim3(nblk(inum, 128)), dim3(128), 0, st, pairnum, x, rcutsq, ilist, neighlist,
                       neighnum, atomtype, alist, inum, dim, ntypes);
}

void launch_pair_fill(int* pairlist, const int* off, const double* x, const double* rcutsq, const int* ilist,
                      const int* neighlist, const int* neighnum, const int* atomtype, const int* alist, int inum, int dim,
                      int ntypes, hipStream_t st) {
    if (inum <= 0) return;
    hipLaunchKernelGGL(k_pair_fill, dim3(nblk(inum, 128)), dim3(128), 0, st, pairlist, off, x, rcutsq, ilist, neighlist,
                       neighnum, atomtype, alist, inum, dim, ntypes);
}

__global__ void k_pair_arrays(double* __restrict__ xij, const double* __restrict__ x, int* __restrict__ aii,
                              int* __restrict__ ai, int* __restrict__ aj, int* __restrict__ ti, int* __restrict__ tj,
                              const int* __restrict__ pairlist, const int* __restrict__ pairnumsum,
                              const int* __restrict__ ilist, const int* __restrict__ atomtype,
                              const int* __restrict__ alist, int inum, int dim) {
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= inum) return;
    const int i = ilist[ii];
    for (int k = pairnumsum[ii]; k < pairnumsum[ii + 1]; ++k) {
        const int j = pairlist[k];
        aii[k] = ii;
        ai[k] = i;
        aj[k] = alist[j];
        ti[k] = atomtype[i];
        tj[k] = atomtype[alist[j]];
        for (int d = 0; d < dim; ++d) xij[static_cast<size_t>(k) * dim + d] = x[j * dim + d] - x[i * dim + d];
    }
}

void launch_pair_arrays(double* xij, const double* x, int* aii, int* ai, int* aj, int* ti, int* tj, const int* pairlist,
                        const int* pairnumsum, const int* ilist, const int* atomtype, const int* alist, int inum, int dim,
                        hipStream_t st) {
    if (inum <= 0) return;
    hipLaunchKernelGGL(k_pair_arrays, dim3(nblk(inum, 128)), dim3(128), 0, st, xij, x, aii, ai, aj, ti, tj, pairlist,
                       pairnumsum, ilist, atomtype, alist, inum, dim);
}

// ================================================================================================
// fixed-stride pair lists of the upstream C++ core (src/C++/Core/cpuSnap.cpp:1883-1935)
// ================================================================================================

// neighbours of atom i at neighlist[l + jnum*i] (l < neighnum[i]); kept ones to pairlist[c + jnum*ii].
// Typed cutoffs (atomtype != nullptr): 1e-20 < d^2 < rcutsq[tj + ti*(ntypes+1)] (cpuSnap.cpp:1883-1905);
// one cutoff (atomtype == nullptr): d^2 <= cut2 (cpuSnap.cpp:1833-1853).
__global__ void k_pair_list_strided(int* __restrict__ pairnum, int* __restrict__ pairlist, const double* __restrict__ x,
                                    const double* __restrict__ rcutsq, const int* __restrict__ ilist,
                                    const int* __restrict__ neighlist, const int* __restrict__ neighnum,
                                    const int* __restrict__ atomtype, const int* __restrict__ alist, int inum, int jnum,
                                    int dim, int ntypes, double cut2) {
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= inum) return;
    const int i = ilist[ii];
    const int* nl = neighlist + static_cast<size_t>(jnum) * i;
    int* pl = pairlist + static_cast<size_t>(jnum) * ii;
    const int m = neighnum[i];
    int c = 0;
    if (atomtype) {
        const int it = atomtype[i];
        for (int l = 0; l < m; ++l) {
            const int j = nl[l];
            if (in_pair(x, i, j, dim, rcutsq[atomtype[alist[j]] + it * (ntypes + 1)])) pl[c++] = j;
        }
    } else {
        for (int l = 0; l < m; ++l) {
#pragma clang fp contract(off)
            const int j = nl[l];
            const double dx = x[j * dim] - x[i * dim], dy = x[j * dim + 1] - x[i * dim + 1],
                         dz = x[j * dim + 2] - x[i * dim + 2];
            if (dx * dx + dy * dy + dz * dz <= cut2) pl[c++] = j;
        }
    }
    pairnum[ii] = c;
}

void launch_pair_list_strided(int* pairnum, int* pairlist, const double* x, const double* rcutsq, const int* ilist,
                              const int* neighlist, const int* neighnum, const int* atomtype, const int* alist,
                              int inum, int jnum, int dim, int ntypes, hipStream_t st) {
    if (inum <= 0) return;
    hipLaunchKernelGGL(k_pair_list_strided, dim3(nblk(inum, 128)), dim3(128), 0, st, pairnum, pairlist, x, rcutsq,
                       ilist, neighlist, neighnum, atomtype, alist, inum, jnum, dim, ntypes, 0.0);
}

void launch_pair_list_strided_cut(int* pairnum, int* pairlist, const double* x, double cut2, const int* ilist,
                                  const int* neighlist, const int* neighnum, int inum, int jnum, int dim,
                                  hipStream_t st) {
    if (inum <= 0) return;
    hipLaunchKernelGGL(k_pair_list_strided, dim3(nblk(inum, 128)), dim3(128), 0, st, pairnum, pairlist, x, nullptr,
                       ilist, neighlist, neighnum, nullptr, nullptr, inum, jnum, dim, 0, cut2);
}

// pair k = pairnumsum[ii] + l of the ii-th atom, neighbour pairlist[l + jnum*ii] (l < pairnum[ii])
__global__ void k_pair_arrays_strided(double* __restrict__ xij, const double* __restrict__ x, int* __restrict__ aii,
                                      int* __restrict__ ai, int* __restrict__ aj, int* __restrict__ ti,
                                      int* __restrict__ tj, const int* __restrict__ pairnum,
                                      const int* __restrict__ pairlist, const int* __restrict__ pairnumsum,
                                      const int* __restrict__ ilist, const int* __restrict__ atomtype,
                                      const int* __restrict__ alist, int inum, int jnum, int dim) {
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= inum) return;
    const int i = ilist[ii], it = atomtype[i];
    const int m = pairnum[ii], start = pairnumsum[ii];
    const int* pl = pairlist + static_cast<size_t>(jnum) * ii;
    for (int l = 0; l < m; ++l) {
        const int j = pl[l], k = start + l, o = alist[j];
        aii[k] = ii;
        ai[k] = i;
        aj[k] = o;
        ti[k] = it;
        tj[k] = atomtype[o];
        for (int d = 0; d < dim; ++d) xij[static_cast<size_t>(k) * dim + d] = x[j * dim + d] - x[i * dim + d];
    }
}

void launch_pair_arrays_strided(double* xij, const double* x, int* aii, int* ai, int* aj, int* ti, int* tj,
                                const int* pairnum, const int* pairlist, const int* pairnumsum, const int* ilist,
                                const int* atomtype, const int* alist, int inum, int jnum, int dim, hipStream_t st) {
    if (inum <= 0) return;
    hipLaunchKernelGGL(k_pair_arrays_strided, dim3(nblk(inum, 128)), dim3(128), 0, st, xij, x, aii, ai, aj, ti, tj,
                       pairnum, pairlist, pairnumsum, ilist, atomtype, alist, inum, jnum, dim);
}

// Stot[ii + inum*jju (+ inum*idxu_max*elem)] += S of the atom's pairs (src/C++/Core/cpuSnap2.cpp:269-318).
// Thread = (ii, jju) with ii fastest.  Without chemflag the pair sum is formed first and added once,
// with chemflag each pair is added to its element's block in pair order, as upstream.
__global__ void k_add_utot_pairs(double* __restrict__ Stotr, double* __restrict__ Stoti, const double* __restrict__ Sr,
                                 const double* __restrict__ Si, const int* __restrict__ pairnum,
                                 const int* __restrict__ pairnumsum, const int* __restrict__ map,
                                 const int* __restrict__ tj, int idxu_max, int inum, int ijnum, int chemflag) {
    const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long n = static_cast<long>(inum) * idxu_max;
    if (idx >= n) return;
    const int ii = static_cast<int>(idx % inum);
    const long jju = idx / inum;
    const int m = pairnum[ii], start = pairnumsum[ii];
    const double* sr = Sr + static_cast<long>(ijnum) * jju + start;
    const double* si = Si + static_cast<long>(ijnum) * jju + start;
    if (!chemflag) {
        double qr = 0.0, qi = 0.0;
        for (int l = 0; l < m; ++l) {
            qr += sr[l];
            qi += si[l];
        }
        Stotr[idx] += qr;
        Stoti[idx] += qi;
    } else {
        for (int l = 0; l < m; ++l) {
            const long kl = idx + n * map[tj[start + l]];
            Stotr[kl] += sr[l];
            Stoti[kl] += si[l];
        }
    }
}

void launch_add_utot_pairs(double* Stotr, double* Stoti, const double* Sr, const double* Si, const int* pairnum,
                           const int* pairnumsum, const int* map, const int* tj, int idxu_max, int inum, int ijnum,
                           int chemflag, hipStream_t st) {
    const long n = static_cast<long>(inum) * idxu_max;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_add_utot_pairs, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, Stotr, Stoti,
                       Sr, Si, pairnum, pairnumsum, map, tj, idxu_max, inum, ijnum, chemflag);
}

// ================================================================================================
// energy / force / virial tallies of the upstream C++ core (src/C++/Core/cpuSnap.cpp:1746-1831,
// cpuSnap2.cpp:1265-1292)
// ================================================================================================

// eatom[ii] += beta0 + sum_k beta_k B_k (+ the quadratic terms), element map[type[ilist[ii]]];
// B(ii, k) at b[ii*ncoeff + k] (soa = 0) or b[ii + inum*k] (soa = 1); terms in the upstream order
__global__ void k_tally_energy_full(double* __restrict__ eatom, const double* __restrict__ b,
                                    const double* __restrict__ coeff, const int* __restrict__ ilist,
                                    const int* __restrict__ map, const int* __restrict__ type, int inum, int ncoeff,
                                    int ncoeffall, int quadratic, int soa) {
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= inum) return;
    const double* __restrict__ c = coeff + static_cast<size_t>(map[type[ilist[ii]]]) * ncoeffall;
    auto B = [&](int k) { return soa ? b[ii + static_cast<size_t>(inum) * k] : b[static_cast<size_t>(ii) * ncoeff + k]; };
    double e = c[0];
    for (int k = 0; k < ncoeff; ++k) e += c[k + 1] * B(k);
    if (quadratic) {
        int q = ncoeff + 1;
        for (int k = 0; k < ncoeff; ++k) {
            const double bk = B(k);
            e += 0.5 * c[q++] * bk * bk;
            for (int l = k + 1; l < ncoeff; ++l) e += c[q++] * bk * B(l);
        }
    }
    eatom[ii] += e;
}

// fatom[ai] += f_ij, fatom[aj] -= f_ij
__global__ void k_tally_force_full(double* __restrict__ fatom, const double* __restrict__ fij,
                                   const int* __restrict__ ai, const int* __restrict__ aj, int ijnum) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ijnum) return;
    const int i = ai[k], j = aj[k];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        atomicAdd(fatom + 3 * i + d, fij[3 * k + d]);
        atomicAdd(fatom + 3 * j + d, -fij[3 * k + d]);
    }
}

// vatom[c*inum + atom] += 1/2 (-r (x) f) for both atoms of the pair, c = xx, yy, zz, xy, xz, yz
__global__ void k_tally_virial_full(double* __restrict__ vatom, const double* __restrict__ fij,
                                    const double* __restrict__ rij, const int* __restrict__ ai,
                                    const int* __restrict__ aj, int inum, int ijnum) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ijnum) return;
    const double dx = -rij[3 * k], dy = -rij[3 * k + 1], dz = -rij[3 * k + 2];
    const double fx = fij[3 * k], fy = fij[3 * k + 1], fz = fij[3 * k + 2];
    const double w[6] = {0.5 * dx * fx, 0.5 * dy * fy, 0.5 * dz * fz, 0.5 * dx * fy, 0.5 * dx * fz, 0.5 * dy * fz};
    const int i = ai[k], j = aj[k];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        atomicAdd(vatom + static_cast<size_t>(c) * inum + i, w[c]);
        atomicAdd(vatom + static_cast<size_t>(c) * inum + j, w[c]);
    }
}

void launch_tally_energy_full(double* eatom, const double* b, const double* coeff, const int* ilist, const int* map,
                              const int* type, int inum, int ncoeff, int ncoeffall, int quadratic, int soa,
                              hipStream_t st) {
    if (inum <= 0) return;
    hipLaunchKernelGGL(k_tally_energy_full, dim3(nblk(inum, 128)), dim3(128), 0, st, eatom, b, coeff, ilist, map,
                       type, inum, ncoeff, ncoeffall, quadratic, soa);
}

void launch_tally_force_full(double* fatom, const double* fij, const int* ai, const int* aj, int ijnum,
                             hipStream_t st) {
    if (ijnum <= 0) return;
    hipLaunchKernelGGL(k_tally_force_full, dim3(nblk(ijnum, 256)), dim3(256), 0, st, fatom, fij, ai, aj, ijnum);
}

void launch_tally_virial_full(double* vatom, const double* fij, const double* rij, const int* ai, const int* aj,
                              int inum, int ijnum, hipStream_t st) {
    if (ijnum <= 0) return;
    hipLaunchKernelGGL(k_tally_virial_full, dim3(nblk(ijnum, 256)), dim3(256), 0, st, vatom, fij, rij, ai, aj, inum,
                       ijnum);
}

// dB/dr of the per-pair route (cpuSnap.cpp:1050-1266): Z atom-major zr[(iT + ne^2 i) idxz_max + jjz], dU and
// dB pair-major du[(ij idxu_max + jju) 3 + k], db[((ij ne^3 + itriple) idxb_max + jjb) 3 + k].  One thread per
// (pair, bispectrum triple) owns every element triple of its dB entry, so the three symmetric terms add up
// without atomics.
__device__ __forceinline__ void dbidrj_term(int jj, int jjz, int jju, const double* __restrict__ zr,
                                            const double* __restrict__ zi, const double* __restrict__ dur,
                                            const double* __restrict__ dui, double s[3]) {
    s[0] = s[1] = s[2] = 0.0;
    for (int mb = 0; 2 * mb < jj; ++mb)
        for (int ma = 0; ma <= jj; ++ma, ++jjz, ++jju)
            for (int k = 0; k < 3; ++k) s[k] += dur[3 * jju + k] * zr[jjz] + dui[3 * jju + k] * zi[jjz];
    if (jj % 2 == 0) {
        const int mb = jj / 2;
        for (int ma = 0; ma < mb; ++ma, ++jjz, ++jju)
            for (int k = 0; k < 3; ++k) s[k] += dur[3 * jju + k] * zr[jjz] + dui[3 * jju + k] * zi[jjz];
        for (int k = 0; k < 3; ++k) s[k] += (dur[3 * jju + k] * zr[jjz] + dui[3 * jju + k] * zi[jjz]) * 0.5;
    }
}

__global__ __launch_bounds__(256) void k_dbidrj_atom_major(SnapDev s, double* __restrict__ db,
                                                           const double* __restrict__ zr,
                                                           const double* __restrict__ zi,
                                                           const double* __restrict__ dur,
                                                           const double* __restrict__ dui, const int* __restrict__ ai,
                                                           const int* __restrict__ tj, int ne, int bnorm, int chem,
                                                           int ijnum) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int nb = s.idxb_max, jd = s.J + 1, ne2 = ne * ne, ne3 = ne2 * ne;
    if (t >= static_cast<long>(ijnum) * nb) return;
    const int ij = static_cast<int>(t / nb), jjb = static_cast<int>(t % nb);
    double* __restrict__ d = db + static_cast<size_t>(ij) * ne3 * nb * 3;
    for (int it = 0; it < ne3; ++it)
        for (int k = 0; k < 3; ++k) d[(it * nb + jjb) * 3 + k] = 0.0;
    const int e3 = chem ? s.map[tj[ij]] : 0, i = ai[ij];
    const int j1 = s.idxb[3 * jjb], j2 = s.idxb[3 * jjb + 1], j = s.idxb[3 * jjb + 2];
    const double* __restrict__ pr = dur + static_cast<size_t>(ij) * s.idxu_max * 3;
    const double* __restrict__ pi = dui + static_cast<size_t>(ij) * s.idxu_max * 3;
    const double j1fac = (j + 1) / (j1 + 1.0), j2fac = (j + 1) / (j2 + 1.0);
    double sm[3];
    for (int e1 = 0; e1 < ne; ++e1)
        for (int e2 = 0; e2 < ne; ++e2) {
            const size_t z12 = (static_cast<size_t>(i) * ne2 + e1 * ne + e2) * s.idxz_max;
            const size_t z21 = (static_cast<size_t>(i) * ne2 + e2 * ne + e1) * s.idxz_max;
            double* o = d + (((e1 * ne + e2) * ne + e3) * nb + jjb) * 3;
            dbidrj_term(j, s.idxz_block[j + j2 * jd + j1 * jd * jd], s.idxu_block[j], zr + z12, zi + z12, pr, pi, sm);
            for (int k = 0; k < 3; ++k) o[k] += 2.0 * sm[k];
            o = d + (((e3 * ne + e2) * ne + e1) * nb + jjb) * 3;
            dbidrj_term(j1, s.idxz_block[j1 + j2 * jd + j * jd * jd], s.idxu_block[j1], zr + z12, zi + z12, pr, pi,
                        sm);
            for (int k = 0; k < 3; ++k) o[k] += bnorm ? 2.0 * sm[k] : 2.0 * sm[k] * j1fac;
            o = d + (((e1 * ne + e3) * ne + e2) * nb + jjb) * 3;
            dbidrj_term(j2, s.idxz_block[j2 + j1 * jd + j * jd * jd], s.idxu_block[j2], zr + z21, zi + z21, pr, pi,
                        sm);
            for (int k = 0; k < 3; ++k) o[k] += bnorm ? 2.0 * sm[k] : 2.0 * sm[k] * j2fac;
        }
}

void launch_dbidrj_atom_major(const SnapDev& s, double* db, const double* zr, const double* zi, const double* dur,
                              const double* dui, const int* ai, const int* tj, int ne, int bnorm, int chem, int ijnum,
                              hipStream_t st) {
    const long n = static_cast<long>(ijnum) * s.idxb_max;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_dbidrj_atom_major, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, s, db, zr,
                       zi, dur, dui, ai, tj, ne, bnorm, chem, ijnum);
}

// sfw[k] = sfac(r_k) w_j of every pair k of each listed atom, one thread per atom (cpuSnap.cpp:358-395)
__global__ __launch_bounds__(256) void k_pair_sfw(double* __restrict__ sfw, const double* __restrict__ rij,
                                                  const double* __restrict__ wj, const double* __restrict__ rad,
                                                  double rmin0, double rcutfac, const int* __restrict__ ilist,
                                                  const int* __restrict__ type, const int* __restrict__ pairnum,
                                                  const int* __restrict__ pairnumsum, const int* __restrict__ tj,
                                                  int inum, int sw) {
#pragma clang fp contract(off)
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= inum) return;
    const int itype = type[ilist[ii]], m = pairnum[ii], start = pairnumsum[ii];
    for (int l = 0; l < m; ++l) {
        const int k = start + l, jt = tj[k];
        const double x = rij[3 * k], y = rij[3 * k + 1], z = rij[3 * k + 2];
        const double r = sqrt(x * x + y * y + z * z);
        const double rcut = (rad[itype] + rad[jt]) * rcutfac;
        double sf = 0.0;
        if (sw == 0) sf = 1.0;
        else if (sw == 1) sf = r <= rmin0 ? 1.0 : (r > rcut ? 0.0 : 0.5 * (cos((r - rmin0) * (M_PI / (rcut - rmin0))) + 1.0));
        sfw[k] = sf * wj[jt];
    }
}

// Utot += sfw_k U_k over the pairs of each listed atom (the per-pair route, cpuSnap.cpp:358-395): U pair-major
// ulist[k idxu_max + u], Utot atom-major ulisttot[(ii ne + jelem) idxu_max + u].  One thread per (atom, entry)
// adds the atom's pairs in list order; the weights come from k_pair_sfw.
__global__ __launch_bounds__(256) void k_add_utot_rij(double* __restrict__ utr, double* __restrict__ uti,
                                                      const double* __restrict__ ur, const double* __restrict__ ui,
                                                      const double* __restrict__ sfw, const int* __restrict__ pairnum,
                                                      const int* __restrict__ pairnumsum, const int* __restrict__ map,
                                                      const int* __restrict__ tj, int idxu_max, int ne, int inum,
                                                      int chem) {
#pragma clang fp contract(off)
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long>(inum) * idxu_max) return;
    const int ii = static_cast<int>(t / idxu_max), u = static_cast<int>(t % idxu_max);
    const int m = pairnum[ii], start = pairnumsum[ii];
    for (int l = 0; l < m; ++l) {
        const int k = start + l, je = chem ? map[tj[k]] : 0;
        const double sf = sfw[k];
        const size_t kl = (static_cast<size_t>(ii) * ne + je) * idxu_max + u, kr = static_cast<size_t>(k) * idxu_max + u;
        utr[kl] += sf * ur[kr];
        uti[kl] += sf * ui[kr];
    }
}

void launch_add_utot_rij(double* utr, double* uti, const double* ur, const double* ui, const double* rij,
                         const double* wj, const double* rad, double rmin0, double rcutfac, const int* ilist,
                         const int* type, const int* pairnum, const int* pairnumsum, const int* map, const int* tj,
                         int idxu_max, int ne, int inum, int sw, int chem, double* sfw, hipStream_t st) {
    const long n = static_cast<long>(inum) * idxu_max;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pair_sfw, dim3((inum + 255) / 256), dim3(256), 0, st, sfw, rij, wj, rad, rmin0, rcutfac, ilist,
                       type, pairnum, pairnumsum, tj, inum, sw);
    hipLaunchKernelGGL(k_add_utot_rij, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, utr, uti, ur, ui,
                       sfw, pairnum, pairnumsum, map, tj, idxu_max, ne, inum, chem);
}

// ---- cell-list neighbour search of the upstream core (cpuNeighborList3D.cpp:334-443) -------------------

// Cell of every atom: its position mapped into the unit cube by s2rmap (column-major 3x3) and located on the
// cell edges eta1..3; atoms i < inum search only the inner cells [1, nc - 2], ghosts every cell.  Products
// and sums are rounded one by one, as upstream, so that the cell of an atom on an edge is the same.
__global__ __launch_bounds__(256) void k_cell_list3d(int* __restrict__ clist, const double* __restrict__ x,
                                                     const double* __restrict__ eta1, const double* __restrict__ eta2,
                                                     const double* __restrict__ eta3, const double* __restrict__ m9,
                                                     int nc0, int nc1, int nc2, int inum, int natom) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natom) return;
    const int m = i < inum ? 1 : 0;
    const double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
    const double t0 = m9[0] * x0 + m9[3] * x1 + m9[6] * x2;
    const double t1 = m9[1] * x0 + m9[4] * x1 + m9[7] * x2;
    const double t2 = m9[2] * x0 + m9[5] * x1 + m9[8] * x2;
    int j1 = m, j2 = m, j3 = m;
    for (; j1 < nc0 - m; ++j1)
        if (eta1[j1] <= t0 && t0 <= eta1[j1 + 1]) break;
    for (; j2 < nc1 - m; ++j2)
        if (eta2[j2] <= t1 && t1 <= eta2[j2 + 1]) break;
    for (; j3 < nc2 - m; ++j3)
        if (eta3[j3] <= t2 && t2 <= eta3[j3 + 1]) break;
    clist[i] = j1 + nc0 * j2 + nc0 * nc1 * j3;
}

__global__ __launch_bounds__(256) void k_fill_int(int* __restrict__ a, int v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Neighbours of atom i = alist[ii] among the 27 cells around its cell, in upstream's cell and list order,
// at most jnum of them; cells outside the grid are skipped.
__global__ __launch_bounds__(256) void k_full_neigh_cell(int* __restrict__ nl, int* __restrict__ nn,
                                                         const double* __restrict__ x, double rc2,
                                                         const int* __restrict__ alist, const int* __restrict__ clist,
                                                         const int* __restrict__ c2alist,
                                                         const int* __restrict__ c2anumsum, int nc0, int nc1, int nc2,
                                                         int inum, int jnum) {
#pragma clang fp contract(off)
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= inum) return;
    const int i = alist[ii];
    const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
    const int ca = clist[i], cb = ca % (nc0 * nc1);
    const int c3 = (ca - cb) / (nc0 * nc1), c1 = cb % nc0, c2 = (cb - c1) / nc0;
    int inc = 0;
    for (int d1 = -1; d1 <= 1 && inc < jnum; ++d1)
        for (int d2 = -1; d2 <= 1 && inc < jnum; ++d2)
            for (int d3 = -1; d3 <= 1 && inc < jnum; ++d3) {
                const int k1 = c1 + d1, k2 = c2 + d2, k3 = c3 + d3;
                if (k1 < 0 || k1 >= nc0 || k2 < 0 || k2 >= nc1 || k3 < 0 || k3 >= nc2) continue;
                const int k = k1 + nc0 * k2 + nc0 * nc1 * k3;
                for (int l = c2anumsum[k]; l < c2anumsum[k + 1] && inc < jnum; ++l) {
                    const int j = c2alist[l];
                    if (j == i) continue;
                    const double dx = x[3 * j] - xi0, dy = x[3 * j + 1] - xi1, dz = x[3 * j + 2] - xi2;
                    if (dx * dx + dy * dy + dz * dz <= rc2) nl[inc++ + static_cast<size_t>(jnum) * i] = j;
                }
            }
    nn[ii] = inc;
}

// Neighbours of atom i < inum among all anum atoms, in atom order, at most jnum of them.
__global__ __launch_bounds__(256) void k_full_neigh_all(int* __restrict__ nl, int* __restrict__ nn,
                                                        const double* __restrict__ x, double rc2, int anum, int inum,
                                                        int jnum) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= inum) return;
    const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
    int inc = 0;
    for (int j = 0; j < anum && inc < jnum; ++j) {
        if (j == i) continue;
        const double dx = x[3 * j] - xi0, dy = x[3 * j + 1] - xi1, dz = x[3 * j + 2] - xi2;
        if (dx * dx + dy * dy + dz * dz <= rc2) nl[inc++ + static_cast<size_t>(jnum) * i] = j;
    }
    nn[i] = inc;
}

void launch_cell_list3d(int* clist, const double* x, const double* eta1, const double* eta2, const double* eta3,
                        const double* s2rmap, const int nc[3], int inum, int natom, hipStream_t st) {
    if (natom <= 0) return;
    hipLaunchKernelGGL(k_cell_list3d, dim3((natom + 255) / 256), dim3(256), 0, st, clist, x, eta1, eta2, eta3, s2rmap,
                       nc[0], nc[1], nc[2], inum, natom);
}

void launch_full_neigh_cell(int* nl, int* nn, const double* x, double rc2, const int* alist, const int* clist,
                            const int* c2alist, const int* c2anumsum, const int nc[3], int inum, int jnum,
                            hipStream_t st) {
    if (inum <= 0) return;
    hipLaunchKernelGGL(k_full_neigh_cell, dim3((inum + 255) / 256), dim3(256), 0, st, nl, nn, x, rc2, alist, clist,
                       c2alist, c2anumsum, nc[0], nc[1], nc[2], inum, jnum);
}

void launch_full_neigh_all(int* nl, int* nn, const double* x, double rc2, int anum, int inum, int jnum,
                           hipStream_t st) {
    if (inum <= 0) return;
    hipLaunchKernelGGL(k_full_neigh_all, dim3((inum + 255) / 256), dim3(256), 0, st, nl, nn, x, rc2, anum, inum, jnum);
}

}  // namespace mdp
