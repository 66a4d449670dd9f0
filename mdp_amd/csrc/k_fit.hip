// Batched fitting descriptors for gfx950: per-atom B and per-pair dB/dr of many independent periodic
// configurations in one pass.  Upstream runs, per configuration, cpuComputeUij -> cpuZeroUarraytot2 ->
// cpuAddUarraytot -> cpuComputeZi2 -> cpuComputeBi2 -> cpuComputeDbidrj (cpuSnap.cpp:1620-2263, called
// from snapcpp.jl:539-665 and snappotential.cpp:762-910 inside the per-configuration loop of
// Regression.cpp:68-95).  Here all configurations share one atom index space and one pair list, so
// each stage is a single launch over the whole batch; the per-pair derivative lists dU/dr are never
// stored (they are recomputed in registers inside the dB/dr kernel and contracted on the fly).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "fit_launch.hpp"
#include "fit_index.hpp"
#include "snap_device.hpp"
#include "zrow.hpp"

namespace mdp {

// ---- batched pair build ------------------------------------------------------------------------

__device__ __forceinline__ double image_delta(double xj, double s, double L, double xi) {
#pragma clang fp contract(off)
    const double g = xj + s * L;
    return g - xi;
}
// rint(t) for |t| < 2^51 (far beyond the 2^31 image indices the enumerating paths can hold): adding and
// subtracting 1.5 * 2^52 leaves t rounded to the nearest integer, ties to even, exactly as rint does, in
// two full-rate adds instead of the slower v_rndne_f64
__device__ __forceinline__ double rint_add(double t) {
#pragma clang fp contract(off)
    const double m = 6755399441055744.0;
    return (t + m) - m;
}
__device__ __forceinline__ double dist2(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    return dx * dx + dy * dy + dz * dz;
}

// Pairs of atom g: every periodic image x_j + s*L of every atom j of its configuration with
// |x_j + s*L - x_i|^2 <= rc(ti,tj)^2, the atom itself at s = 0 excluded (neighpairlist.jl:52 with the
// ghost construction of fullneighborlist.jl).  Order: j ascending, then (sz, sy, sx) ascending.
// FILL = false counts, FILL = true writes the CSR slot poff[g].. when the total fits cap; otherwise
// it writes nothing, sets bit 0 of flags[0] and stores the needed total in flags[1].
// FIT_PAIR_SPLIT 64-lane workgroups per chunk of 64 consecutive atoms, the chunks taken in the order
// chunk_order gives (most candidate tests first); workgroup part p of a chunk tests, for each of its atoms,
// the p-th of FIT_PAIR_SPLIT consecutive slices of the configuration's atoms, so that an atom's pairs are
// the concatenation of its slices' pairs in slice order.  The count pass writes the slice counts to
// pcnt[N + p*N + g]; the fill pass starts slice p of atom g at poff[g] plus the counts of slices < p.
template <bool FILL>
__global__ __launch_bounds__(64) void k_fit_pairs(const double* __restrict__ x, const int* __restrict__ type,
                                                  const int* __restrict__ cfg_of, const int* __restrict__ cfg_a0,
                                                  const double* __restrict__ cfg_L, const double* __restrict__ cfg_T,
                                                  const double* __restrict__ rc2, int ntypes, int N,
                                                  const int* __restrict__ chunk_order, int* __restrict__ pcnt,
                                                   const int* __restrict__ poff, long cap, double* __restrict__ rij,
                                                   int* __restrict__ pi, int* __restrict__ pj,
                                                   int* __restrict__ pti, int* __restrict__ ptj,
                                                   int* __restrict__ flags) {
    const int part = static_cast<int>(blockIdx.x % FIT_PAIR_SPLIT);
    const int g = chunk_order[blockIdx.x / FIT_PAIR_SPLIT] * 64 + static_cast<int>(threadIdx.x);
    if (g >= N) return;
    const int c = cfg_of[g];
    const int ja = cfg_a0[c], nc = cfg_a0[c + 1] - ja;
    const int j0 = ja + static_cast<int>(static_cast<long>(nc) * part / FIT_PAIR_SPLIT);
    const int j1 = ja + static_cast<int>(static_cast<long>(nc) * (part + 1) / FIT_PAIR_SPLIT);
    const double Lx = cfg_L[3 * c], Ly = cfg_L[3 * c + 1], Lz = cfg_L[3 * c + 2];
    // tilt of a triclinic cell a = (Lx, 0, 0), b = (bx, Ly, 0), c = (cx, cy, Lz); all zero when orthogonal
    const double bx = cfg_T ? cfg_T[3 * c] : 0.0, cx = cfg_T ? cfg_T[3 * c + 1] : 0.0,
                 cy = cfg_T ? cfg_T[3 * c + 2] : 0.0;
    const bool tri = bx != 0.0 || cx != 0.0 || cy != 0.0;
    const double xi = x[3 * g], yi = x[3 * g + 1], zi = x[3 * g + 2];
    const int ti = type[g];
    long q = 0;
    if constexpr (FILL) {
        q = poff[g];
        for (int p = 0; p < part; ++p) q += pcnt[N + p * N + g];
        if (static_cast<long>(poff[N]) > cap) {   // buffers too small: write nothing, report the need
            if (g == 0 && part == 0) {
                atomicOr(flags, 1);
                flags[1] = poff[N];
            }
            return;
        }
    }
    const double iLx = 1.0 / Lx, iLy = 1.0 / Ly, iLz = 1.0 / Lz;
    double rcmax2 = 0.0, rcmin2 = 1e300;
    for (int t = 1; t <= ntypes; ++t) {
        rcmax2 = fmax(rcmax2, rc2[ti * (ntypes + 1) + t]);
        rcmin2 = fmin(rcmin2, rc2[ti * (ntypes + 1) + t]);
    }
    const bool rconst = rcmin2 == rcmax2;   // one cutoff for every partner type (always with one type)
    // With every box length above twice the largest cutoff (with a relative margin far above rounding)
    // two images of j are more than 2 rc apart, so at most one lies within the cutoff, and in every
    // dimension it is the nearest image s = rint((x_i - x_j) / L) (|s L + x_j - x_i| <= rc < L/2).  The
    // single-image test below evaluates exactly the r_ij and |r_ij|^2 the full enumeration would for
    // that image: same pairs, same order, same bits.
    const double wide2 = 4.0 * rcmax2 * (1.0 + 1e-9);
    const bool single = !tri && wide2 < Lx * Lx && wide2 < Ly * Ly && wide2 < Lz * Lz;
    int cnt = 0;
    auto emit = [&](int j, int tj, double dx, double dy, double dz) __attribute__((always_inline)) {
        if constexpr (FILL) {
            rij[3 * q] = dx;
            rij[3 * q + 1] = dy;
            rij[3 * q + 2] = dz;
            pi[q] = g;
            pj[q] = j;
            pti[q] = ti;
            ptj[q] = tj;
            ++q;
        }
        ++cnt;
    };
    // RC: the cutoff is rcmax2 for every partner (a compile-time choice, so no per-candidate table read)
    auto nearest = [&](int j, double xj, double yj, double zj, int tj, auto rc) __attribute__((always_inline)) {
        const double dx = image_delta(xj, rint_add((xi - xj) * iLx), Lx, xi);
        const double dy = image_delta(yj, rint_add((yi - yj) * iLy), Ly, yi);
        const double dz = image_delta(zj, rint_add((zi - zj) * iLz), Lz, zi);
        double r2;
        if constexpr (decltype(rc)::value) r2 = rcmax2;
        else r2 = rc2[ti * (ntypes + 1) + tj];
        // the test is evaluated by every lane without a branch (a branch on j != g would let the compiler
        // sink each candidate's load into it and wait for it there); only the rare hit branches
        const bool hit = (j != g) & (dist2(dx, dy, dz) <= r2);
        if constexpr (FILL) {
            if (hit) emit(j, tj, dx, dy, dz);
        } else {
            cnt += hit;
        }
    };
    // candidates in batches of four, software-pipelined: the next batch's loads are issued before the
    // current batch is tested, and the scheduling barriers keep the compiler from moving them back down
    // to their first use.  UNI: j is wave-uniform, so the candidates sit in scalar registers.
    auto scan = [&](int ja, int jb, auto rc, auto uni) __attribute__((always_inline)) {
        constexpr int B = 4;
        constexpr bool U = decltype(uni)::value;
        int j = ja;
        if (j + B <= jb) {
            double px[B], py[B], pz[B];
            int pt[B];
            auto load = [&](int jj, double* ax, double* ay, double* az, int* at) __attribute__((always_inline)) {
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    ax[u] = x[3 * (jj + u)];
                    ay[u] = x[3 * (jj + u) + 1];
                    az[u] = x[3 * (jj + u) + 2];
                    at[u] = type[jj + u];
                }
            };
            auto pin = [&](double* ax, double* ay, double* az, int* at) __attribute__((always_inline)) {
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    if constexpr (U) {
                        asm volatile("" : "+s"(ax[u]), "+s"(ay[u]), "+s"(az[u]), "+s"(at[u]));
                    } else {
                        asm volatile("" : "+v"(ax[u]), "+v"(ay[u]), "+v"(az[u]), "+v"(at[u]));
                    }
                }
            };
            load(j, px, py, pz, pt);
            pin(px, py, pz, pt);
            for (; j + 2 * B <= jb; j += B) {
                double nx[B], ny[B], nz[B];
                int nt[B];
                load(j + B, nx, ny, nz, nt);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < B; ++u) nearest(j + u, px[u], py[u], pz[u], pt[u], rc);
                __builtin_amdgcn_sched_barrier(0);
                pin(nx, ny, nz, nt);
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    px[u] = nx[u];
                    py[u] = ny[u];
                    pz[u] = nz[u];
                    pt[u] = nt[u];
                }
            }
#pragma unroll
            for (int u = 0; u < B; ++u) nearest(j + u, px[u], py[u], pz[u], pt[u], rc);
            j += B;
        }
        for (; j < jb; ++j) nearest(j, x[3 * j], x[3 * j + 1], x[3 * j + 2], type[j], rc);
    };
    auto scan_rc = [&](int ja, int jb, auto uni) __attribute__((always_inline)) {
        if (rconst) scan(ja, jb, std::true_type{}, uni);
        else scan(ja, jb, std::false_type{}, uni);
    };
    if (tri) {
        // every image x_j + s1 a + s2 b + s3 c within the cutoff, ordered by j then (s3, s2, s1): the z
        // component fixes the s3 range, then y (given s3) the s2 range, then x (given s2, s3) the s1
        // range, each from the real-number bound rounded outwards (the exact test decides)
        for (int j = j0; j < j1; ++j) {
            const double xj = x[3 * j], yj = x[3 * j + 1], zj = x[3 * j + 2];
            const int tj = type[j];
            const double r2 = rc2[ti * (ntypes + 1) + tj];
            const double rc = sqrt(r2);
            const int s30 = static_cast<int>(floor((zi - zj - rc) * iLz)), s31 = static_cast<int>(ceil((zi - zj + rc) * iLz));
            for (int s3 = s30; s3 <= s31; ++s3) {
                const double dz = image_delta(zj, s3, Lz, zi);
                const double yo = yj + s3 * cy, xo3 = s3 * cx;
                const int s20 = static_cast<int>(floor((yi - yo - rc) * iLy)), s21 = static_cast<int>(ceil((yi - yo + rc) * iLy));
                for (int s2 = s20; s2 <= s21; ++s2) {
                    const double dy = image_delta(yo, s2, Ly, yi);
                    const double xo = xj + (s2 * bx + xo3);
                    const int s10 = static_cast<int>(floor((xi - xo - rc) * iLx)), s11 = static_cast<int>(ceil((xi - xo + rc) * iLx));
                    for (int s1 = s10; s1 <= s11; ++s1) {
                        if (j == g && s1 == 0 && s2 == 0 && s3 == 0) continue;
                        const double dx = image_delta(xo, s1, Lx, xi);
                        if (dist2(dx, dy, dz) <= r2) emit(j, tj, dx, dy, dz);
                    }
                }
            }
        }
    } else if (single) {
        const int cu = __builtin_amdgcn_readfirstlane(c);
        if (__all(c == cu))   // the wave's atoms share one configuration: j, x_j and t_j are wave-uniform
            scan_rc(__builtin_amdgcn_readfirstlane(j0), __builtin_amdgcn_readfirstlane(j1), std::true_type{});
        else
            scan_rc(j0, j1, std::false_type{});
    } else {
        // every image x_j + s L within the cutoff, s from the real-number bound rounded outwards (the
        // exact test decides, so the reciprocal, an ulp off at worst, only ever adds a candidate)
        for (int j = j0; j < j1; ++j) {
            const double xj = x[3 * j], yj = x[3 * j + 1], zj = x[3 * j + 2];
            const int tj = type[j];
            const double r2 = rc2[ti * (ntypes + 1) + tj];
            const double rc = sqrt(r2);
            const int sx0 = static_cast<int>(floor((xi - xj - rc) * iLx)), sx1 = static_cast<int>(ceil((xi - xj + rc) * iLx));
            const int sy0 = static_cast<int>(floor((yi - yj - rc) * iLy)), sy1 = static_cast<int>(ceil((yi - yj + rc) * iLy));
            const int sz0 = static_cast<int>(floor((zi - zj - rc) * iLz)), sz1 = static_cast<int>(ceil((zi - zj + rc) * iLz));
            for (int sz = sz0; sz <= sz1; ++sz) {
                const double dz = image_delta(zj, sz, Lz, zi);
                for (int sy = sy0; sy <= sy1; ++sy) {
                    const double dy = image_delta(yj, sy, Ly, yi);
                    for (int sx = sx0; sx <= sx1; ++sx) {
                        if (j == g && sx == 0 && sy == 0 && sz == 0) continue;
                        const double dx = image_delta(xj, sx, Lx, xi);
                        if (dist2(dx, dy, dz) <= r2) emit(j, tj, dx, dy, dz);
                    }
                }
            }
        }
    }
    if constexpr (!FILL) pcnt[N + part * N + g] = cnt;
}

// pcnt[g] = the sum of atom g's slice counts
__global__ void k_fit_pair_total(int* __restrict__ pcnt, int N) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    int t = 0;
    for (int p = 0; p < FIT_PAIR_SPLIT; ++p) t += pcnt[N + p * N + g];
    pcnt[g] = t;
}

// first pair of every configuration: cfg_p0[c] = poff[cfg_a0[c]]
__global__ void k_fit_cfg_p0(const int* __restrict__ poff, const int* __restrict__ cfg_a0, int* __restrict__ cfg_p0,
                             int C) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c <= C) cfg_p0[c] = poff[cfg_a0[c]];
}

// after a capacity overflow of the fill pass, empty every atom's pair range so that the remaining
// stages of the batch touch no pair (the host then resizes and runs the batch again)
__global__ void k_fit_guard(const int* __restrict__ flags, int* __restrict__ poff, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= N && (flags[0] & 1)) poff[i] = 0;
}

__global__ void k_fill_zero_int(int* __restrict__ p, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}

// ---- Z (atom-major) and B ----------------------------------------------------------------------

// Z of one coupling triple (blockIdx.y) for 64 atoms, all rows mb <= j/2 and columns ma, stored
// atom-major and complex-interleaved: zt[(atom*idxz_max + jjz)*2 + {0,1}], so that the dB/dr
// kernel streams one atom's Z contiguously.  For bispectrum triples (j >= j1) also
// B = 2 sum' U*(j).Z(j1,j2,j) - bzero[j] (cpuSnap.cpp:1971-2072), written in the per-configuration
// layout blist[c][ii + N_c*jjb].  U is the full-layout Utot (element jju of atom g at jju*N + g).
template <int J>
__global__ __launch_bounds__(64) void k_fit_zb(const double* __restrict__ cg, const double* __restrict__ Ur,
                                               const double* __restrict__ Ui, double* __restrict__ zt,
                                               double* __restrict__ blist, const double* __restrict__ bzero,
                                               const int* __restrict__ cfg_of, const int* __restrict__ cfg_a0,
                                               int N, int bnorm, int bzeroflag) {
    const int atom = blockIdx.x * 64 + threadIdx.x;
    if (atom >= N) return;
    constexpr int IZ = n_z(J), IB = n_b(J);
    const int t = blockIdx.y;
    const size_t n = static_cast<size_t>(N);
    static_for<0, J + 1>([&](auto J1c) {
        constexpr int J1 = decltype(J1c)::value;
        static_for<0, J1 + 1>([&](auto J2c) {
            constexpr int J2 = decltype(J2c)::value;
            static_for<0, J + 1>([&](auto JOc) {
                constexpr int JO = decltype(JOc)::value;
                if constexpr (coupled(J1, J2, JO)) {
                    if (t == triple_id(J, J1, J2, JO)) {
                        constexpr int ZB = z_block(J, J1, J2, JO);
                        const double scale = bnorm ? 1.0 / (JO + 1) : 1.0;
                        double bsum = 0.0;
                        double* __restrict__ zo = zt + (static_cast<size_t>(atom) * IZ + ZB) * 2;
                        for (int mb = 0; 2 * mb <= JO; ++mb) {
                            double zr[JO + 1], zi[JO + 1];
#pragma unroll
                            for (int m = 0; m <= JO; ++m) { zr[m] = 0.0; zi[m] = 0.0; }
                            z_row<J, J1, J2, JO>(mb, JO, Ur, Ui, Ur, Ui, n, atom, cg, scale, zr, zi);
#pragma unroll
                            for (int ma = 0; ma <= JO; ++ma) {
                                zo[2 * ((JO + 1) * mb + ma)] = zr[ma];
                                zo[2 * ((JO + 1) * mb + ma) + 1] = zi[ma];
                            }
                            if constexpr (JO >= J1) {
                                const size_t o = static_cast<size_t>(u_block(JO) + (JO + 1) * mb) * n + atom;
                                const int malim = (2 * mb == JO) ? mb : JO;
#pragma unroll
                                for (int ma = 0; ma <= JO; ++ma) {
                                    if (ma <= malim) {
                                        const double w = (2 * mb < JO || ma < mb) ? 1.0 : 0.5;
                                        bsum += w * (Ur[o + ma * n] * zr[ma] + Ui[o + ma * n] * zi[ma]);
                                    }
                                }
                            }
                        }
                        if constexpr (JO >= J1) {
                            constexpr int jjb = b_index(J, J1, J2, JO);
                            static_assert(jjb >= 0 && jjb < IB, "bispectrum triple");
                            const int c = cfg_of[atom];
                            const int a0 = cfg_a0[c], nc = cfg_a0[c + 1] - a0;
                            double b = 2.0 * bsum;
                            if (bzeroflag) b -= bzero[JO];
                            blist[static_cast<size_t>(a0) * IB + (atom - a0) + static_cast<size_t>(nc) * jjb] = b;
                        }
                    }
                }
            });
        });
    });
}

// ---- dB/dr ----------------------------------------------------------------------------------------

// Persistent: each wave takes whole atoms from a global counter; lanes are (pair, Cartesian
// component), 21 pairs x 3 per pass (lane 63 idle), all pairs of the pass belonging to the wave's atom.
// Each lane runs the U and dU/dr_c row-chain recursion of its pair in registers and, as every
// half-domain row (l, mb) of d(sfac U)/dr_c = dsfac u_c U + sfac dU/dr_c is finished, contracts it
// with the Z block of every bispectrum term on layer l (cpuSnap.cpp:2101-2240: middle column of even
// layers 1/2, term weights 2, 2(j+1)/(j1+1), 2(j+1)/(j2+1), or 2 under bnorm).  The terms of layer l
// are entries term_lo[l]..term_lo[l+1]-1 of the term table (bispectrum index, Z block offset,
// weight), walked by a uniform loop; the atom is wave-uniform, so Z values, table entries and
// weights are scalar loads shared by all 63 lanes.  Each lane's 55 accumulators live in LDS.
// dblist is written in the per-configuration layout dblist[c][lp + P_c*k + 3 P_c*jjb].
// PF: the term loop is software-pipelined, the table entry and Z row of term k + 1 being loaded while
// term k is contracted (a term otherwise waits for two dependent scalar loads).
template <int J, int WPE = 1, bool PF = false>
__global__ __launch_bounds__(64, WPE) void k_fit_dbdr(const double* __restrict__ rootpq, const double* __restrict__ zt,
                                                 const int* __restrict__ term_lo, const int* __restrict__ term_b,
                                                 const int* __restrict__ term_z, const double* __restrict__ term_w,
                                                 const double* __restrict__ radelem, const double* __restrict__ wjelem,
                                                 double rcutfac, double rmin0, double rfac0, int switchflag,
                                                 const double* __restrict__ rij, const int* __restrict__ pti,
                                                 const int* __restrict__ ptj, const int* __restrict__ poff,
                                                 const int* __restrict__ cfg_of, const int* __restrict__ cfg_p0,
                                                 double* __restrict__ dblist, int* __restrict__ counter, int N) {
    constexpr int IZ = n_z(J), IB = n_b(J);
    __shared__ double acc[IB * 64];
    const int lane = threadIdx.x;
    const int comp = lane % 3, slot = lane / 3;
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(counter, 1);
        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
        if (t >= N) break;
        const int atom = t;
        const int q0 = __builtin_amdgcn_readfirstlane(poff[atom]);
        const int q1 = __builtin_amdgcn_readfirstlane(poff[atom + 1]);
        const int c = __builtin_amdgcn_readfirstlane(cfg_of[atom]);
        const int p0 = __builtin_amdgcn_readfirstlane(cfg_p0[c]);
        const long Pc = __builtin_amdgcn_readfirstlane(cfg_p0[c + 1]) - p0;
        const double* __restrict__ z = zt + static_cast<size_t>(atom) * IZ * 2;
        double* __restrict__ db = dblist + static_cast<size_t>(p0) * 3 * IB;
        for (int qb = q0; qb < q1; qb += 21) {
            const int q = qb + slot;
            const bool active = lane < 63 && q < q1;
            const int p = active ? q : q0;   // idle lanes recompute a valid pair and store nothing
            const int tjp = ptj[p];
            const double rc = (radelem[pti[p]] + radelem[tjp]) * rcutfac;
            PairGeom g;
            pair_geometry<true>(rij[3 * p], rij[3 * p + 1], rij[3 * p + 2], rc, rmin0, rfac0, switchflag, wjelem[tjp], g);
            const double uk = comp == 0 ? g.u[0] : (comp == 1 ? g.u[1] : g.u[2]);
            g.dar[0] = comp == 0 ? g.dar[0] : (comp == 1 ? g.dar[1] : g.dar[2]);
            g.dai[0] = comp == 0 ? g.dai[0] : (comp == 1 ? g.dai[1] : g.dai[2]);
            g.dbr[0] = comp == 0 ? g.dbr[0] : (comp == 1 ? g.dbr[1] : g.dbr[2]);
            g.dbi[0] = comp == 0 ? g.dbi[0] : (comp == 1 ? g.dbi[1] : g.dbi[2]);
            const double dsu = g.dsfac * uk, sf = g.sfac;
            for (int k = 0; k < IB; ++k) acc[k * 64 + lane] = 0.0;
            row_chain<J, 1, 0, J / 2>(g, rootpq, [&](auto Jc, auto MBc, auto& pr, auto& pim, auto& dpr, auto& dpi) {
                constexpr int L = decltype(Jc)::value, MB = decltype(MBc)::value;
                constexpr int LAST = (L == 2 * MB) ? MB : L;
                double dr[LAST + 1], di[LAST + 1];
#pragma unroll
                for (int ma = 0; ma <= LAST; ++ma) {
                    dr[ma] = dsu * pr[ma] + sf * dpr[0][ma];
                    di[ma] = dsu * pim[ma] + sf * dpi[0][ma];
                }
                if constexpr (L == 2 * MB) {   // middle column of an even layer counts 1/2
                    dr[MB] *= 0.5;
                    di[MB] *= 0.5;
                }
                const int k0 = term_lo[L], k1 = term_lo[L + 1];
                if constexpr (!PF) {
                    for (int k = k0; k < k1; ++k) {
                        const double* __restrict__ zr = z + 2 * (term_z[k] + (L + 1) * MB);
                        double s = 0.0;
#pragma unroll
                        for (int ma = 0; ma <= LAST; ++ma) s += dr[ma] * zr[2 * ma] + di[ma] * zr[2 * ma + 1];
                        acc[term_b[k] * 64 + lane] += term_w[k] * s;
                    }
                } else if (k0 < k1) {
                    double zn[2 * (LAST + 1)], wn;
                    int bn;
                    auto fetch = [&](int k) {
                        const double* __restrict__ zr = z + 2 * (term_z[k] + (L + 1) * MB);
#pragma unroll
                        for (int m = 0; m < 2 * (LAST + 1); ++m) zn[m] = zr[m];
                        bn = term_b[k];
                        wn = term_w[k];
                    };
                    fetch(k0);
                    for (int k = k0; k < k1; ++k) {
                        double zc[2 * (LAST + 1)];
#pragma unroll
                        for (int m = 0; m < 2 * (LAST + 1); ++m) zc[m] = zn[m];
                        const int bc = bn;
                        const double wc = wn;
                        if (k + 1 < k1) fetch(k + 1);
                        double s = 0.0;
#pragma unroll
                        for (int ma = 0; ma <= LAST; ++ma) s += dr[ma] * zc[2 * ma] + di[ma] * zc[2 * ma + 1];
                        acc[bc * 64 + lane] += wc * s;
                    }
                }
            });
            if (active) {
                const size_t lp = static_cast<size_t>(q - p0);
                for (int k = 0; k < IB; ++k) db[lp + Pc * comp + 3 * Pc * k] = acc[k * 64 + lane];
            }
        }
    }
}

// dB/dr with dU/dr_c built from U alone through the SU(2) generators (as the MD dE/dr):
// dU/dx_c = sum_g (dq/dx_c . V_g) L_g U, so each lane runs only the U recursion.  The bispectrum indices
// are split into two halves (FitTerms::hb); workgroup b (one wave) serves half b % 2 and takes atoms from
// that half's counter, so its accumulators need half the LDS and more waves fit on a CU.  Lanes, passes,
// term walk (pipelined) and output layout as k_fit_dbdr.
template <int J>
__global__ __launch_bounds__(64) void k_fit_dbdr_g(const double* __restrict__ rootpq, const double* __restrict__ zt,
                                                   const int* __restrict__ h_lo, const int* __restrict__ h_b,
                                                   const int* __restrict__ h_z, const double* __restrict__ h_w, int hb,
                                                   const double* __restrict__ radelem,
                                                   const double* __restrict__ wjelem, double rcutfac, double rmin0,
                                                   double rfac0, int switchflag, const double* __restrict__ rij,
                                                   const int* __restrict__ pti, const int* __restrict__ ptj,
                                                   const int* __restrict__ poff, const int* __restrict__ cfg_of,
                                                   const int* __restrict__ cfg_p0, double* __restrict__ dblist,
                                                   int* __restrict__ counter, int N) {
    constexpr int IZ = n_z(J), IB = n_b(J);
    extern __shared__ double acc[];   // max(hb, IB - hb) * 64 doubles, sized at launch
    const int lane = threadIdx.x;
    const int comp = lane % 3, slot = lane / 3;
    const int half = static_cast<int>(blockIdx.x & 1);
    const int b0 = half ? hb : 0, nbh = half ? IB - hb : hb;
    const int* __restrict__ lo = h_lo + half * (J + 2);
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(counter + half, 1);
        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
        if (t >= N) break;
        const int atom = t;
        const int q0 = __builtin_amdgcn_readfirstlane(poff[atom]);
        const int q1 = __builtin_amdgcn_readfirstlane(poff[atom + 1]);
        const int c = __builtin_amdgcn_readfirstlane(cfg_of[atom]);
        const int p0 = __builtin_amdgcn_readfirstlane(cfg_p0[c]);
        const long Pc = __builtin_amdgcn_readfirstlane(cfg_p0[c + 1]) - p0;
        const double* __restrict__ z = zt + static_cast<size_t>(atom) * IZ * 2;
        double* __restrict__ db = dblist + static_cast<size_t>(p0) * 3 * IB;
        for (int qb = q0; qb < q1; qb += 21) {
            const int q = qb + slot;
            const bool active = lane < 63 && q < q1;
            const int p = active ? q : q0;   // idle lanes recompute a valid pair and store nothing
            const int tjp = ptj[p];
            const double rc = (radelem[pti[p]] + radelem[tjp]) * rcutfac;
            PairGeom g;
            pair_geometry<true>(rij[3 * p], rij[3 * p + 1], rij[3 * p + 2], rc, rmin0, rfac0, switchflag, wjelem[tjp], g);
            // projections of dq/dx_comp on V_0 = q*i, V_1 = q*j, V_2 = q*k
            const double dar = comp == 0 ? g.dar[0] : (comp == 1 ? g.dar[1] : g.dar[2]);
            const double dai = comp == 0 ? g.dai[0] : (comp == 1 ? g.dai[1] : g.dai[2]);
            const double dbr = comp == 0 ? g.dbr[0] : (comp == 1 ? g.dbr[1] : g.dbr[2]);
            const double dbi = comp == 0 ? g.dbi[0] : (comp == 1 ? g.dbi[1] : g.dbi[2]);
            const double uk = comp == 0 ? g.u[0] : (comp == 1 ? g.u[1] : g.u[2]);
            const double sf = g.sfac, dsu = g.dsfac * uk;
            const double k0 = sf * (-g.ai * dar + g.ar * dai + g.bi * dbr - g.br * dbi);
            const double k1 = sf * (-g.br * dar - g.bi * dai + g.ar * dbr + g.ai * dbi);
            const double k2 = sf * (-g.bi * dar + g.br * dai - g.ai * dbr + g.ar * dbi);
            for (int k = 0; k < nbh; ++k) acc[k * 64 + lane] = 0.0;
            row_chain<J, 0, 0, J / 2>(g, rootpq, [&](auto Jc, auto MBc, auto& pr, auto& pim, auto&, auto&) {
                constexpr int L = decltype(Jc)::value, MB = decltype(MBc)::value;
                constexpr int LAST = (L == 2 * MB) ? MB : L;
                double dr[LAST + 1], di[LAST + 1];
                static_for<0, LAST + 1>([&](auto MAc) {
                    constexpr int MA = decltype(MAc)::value;
                    // (L_0 U) = i (2ma - l) U; (L_1 U) = c(ma) U+ - c(ma-1) U-; (L_2 U) = i (c(ma) U+ + c(ma-1) U-)
                    constexpr double m0 = static_cast<double>(2 * MA - L);
                    constexpr double cu = (MA < L) ? cx_sqrt(static_cast<double>((MA + 1) * (L - MA))) : 0.0;
                    constexpr double cd = (MA > 0) ? cx_sqrt(static_cast<double>(MA * (L - MA + 1))) : 0.0;
                    const double upr = (MA < L) ? pr[MA < L ? MA + 1 : MA] : 0.0;
                    const double upi = (MA < L) ? pim[MA < L ? MA + 1 : MA] : 0.0;
                    const double dnr = (MA > 0) ? pr[MA > 0 ? MA - 1 : 0] : 0.0;
                    const double dni = (MA > 0) ? pim[MA > 0 ? MA - 1 : 0] : 0.0;
                    const double l1r = cu * upr - cd * dnr, l1i = cu * upi - cd * dni;
                    const double l2r = -(cu * upi + cd * dni), l2i = cu * upr + cd * dnr;
                    dr[MA] = dsu * pr[MA] + k0 * (-m0 * pim[MA]) + k1 * l1r + k2 * l2r;
                    di[MA] = dsu * pim[MA] + k0 * (m0 * pr[MA]) + k1 * l1i + k2 * l2i;
                });
                if constexpr (L == 2 * MB) {   // middle column of an even layer counts 1/2
                    dr[MB] *= 0.5;
                    di[MB] *= 0.5;
                }
                const int ka = lo[L], kb = lo[L + 1];
                if (ka < kb) {
                    double zn[2 * (LAST + 1)], wn;
                    int bn;
                    auto fetch = [&](int k) {
                        const double* __restrict__ zr = z + 2 * (h_z[k] + (L + 1) * MB);
#pragma unroll
                        for (int m = 0; m < 2 * (LAST + 1); ++m) zn[m] = zr[m];
                        bn = h_b[k] - b0;
                        wn = h_w[k];
                    };
                    fetch(ka);
                    for (int k = ka; k < kb; ++k) {
                        double zc[2 * (LAST + 1)];
#pragma unroll
                        for (int m = 0; m < 2 * (LAST + 1); ++m) zc[m] = zn[m];
                        const int bc = bn;
                        const double wc = wn;
                        if (k + 1 < kb) fetch(k + 1);
                        double sacc = 0.0;
#pragma unroll
                        for (int ma = 0; ma <= LAST; ++ma) sacc += dr[ma] * zc[2 * ma] + di[ma] * zc[2 * ma + 1];
                        acc[bc * 64 + lane] += wc * sacc;
                    }
                }
            });
            if (active) {
                const size_t lp = static_cast<size_t>(q - p0);
                for (int k = 0; k < nbh; ++k) db[lp + Pc * comp + 3 * Pc * (b0 + k)] = acc[k * 64 + lane];
            }
        }
    }
}

// As k_fit_dbdr_g with the atom's Z block staged in LDS.  A workgroup is two waves, one per half of
// the bispectrum indices, working on the same atom: they copy its Z block (2 IZ doubles) from HBM with
// wide vector loads once, and every term then reads its Z row from LDS instead of through dependent
// scalar loads from L2.  LDS: the Z block, then the accumulators of half 0 (hb * 64) and of half 1;
// two workgroups fit on a CU.
template <int J>
__global__ __launch_bounds__(128) void k_fit_dbdr_l(const double* __restrict__ rootpq, const double* __restrict__ zt,
                                                    const int* __restrict__ h_lo, const int* __restrict__ h_b,
                                                    const int* __restrict__ h_z, const double* __restrict__ h_w, int hb,
                                                    const double* __restrict__ radelem,
                                                    const double* __restrict__ wjelem, double rcutfac, double rmin0,
                                                    double rfac0, int switchflag, const double* __restrict__ rij,
                                                    const int* __restrict__ pti, const int* __restrict__ ptj,
                                                    const int* __restrict__ poff, const int* __restrict__ cfg_of,
                                                    const int* __restrict__ cfg_p0, double* __restrict__ dblist,
                                                    int* __restrict__ counter, int N) {
    constexpr int IZ = n_z(J), IB = n_b(J);
    __shared__ __align__(16) double sm[2 * IZ + IB * 64];   // 66 KB at twojmax = 8
    __shared__ int s_atom;
    double* __restrict__ zs = sm;
    const int lane = threadIdx.x & 63;
    const int half = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int comp = lane % 3, slot = lane / 3;
    const int b0 = half ? hb : 0, nbh = half ? IB - hb : hb;
    double* __restrict__ acc = sm + 2 * IZ + b0 * 64;
    const int* __restrict__ lo = h_lo + half * (J + 2);
    for (;;) {
        if (threadIdx.x == 0) s_atom = atomicAdd(counter, 1);
        __syncthreads();
        const int atom = __builtin_amdgcn_readfirstlane(s_atom);
        if (atom >= N) break;   // the same value in both waves
        {
            const double2* __restrict__ src = reinterpret_cast<const double2*>(zt + static_cast<size_t>(atom) * IZ * 2);
            double2* dst = reinterpret_cast<double2*>(zs);
            for (int e = threadIdx.x; e < IZ; e += 128) dst[e] = src[e];
        }
        __syncthreads();   // Z staged; s_atom read by both waves
        const int q0 = __builtin_amdgcn_readfirstlane(poff[atom]);
        const int q1 = __builtin_amdgcn_readfirstlane(poff[atom + 1]);
        const int c = __builtin_amdgcn_readfirstlane(cfg_of[atom]);
        const int p0 = __builtin_amdgcn_readfirstlane(cfg_p0[c]);
        const long Pc = __builtin_amdgcn_readfirstlane(cfg_p0[c + 1]) - p0;
        double* __restrict__ db = dblist + static_cast<size_t>(p0) * 3 * IB;
        for (int qb = q0; qb < q1; qb += 21) {
            const int q = qb + slot;
            const bool active = lane < 63 && q < q1;
            const int p = active ? q : q0;   // idle lanes recompute a valid pair and store nothing
            const int tjp = ptj[p];
            const double rc = (radelem[pti[p]] + radelem[tjp]) * rcutfac;
            PairGeom g;
            pair_geometry<true>(rij[3 * p], rij[3 * p + 1], rij[3 * p + 2], rc, rmin0, rfac0, switchflag, wjelem[tjp], g);
            const double dar = comp == 0 ? g.dar[0] : (comp == 1 ? g.dar[1] : g.dar[2]);
            const double dai = comp == 0 ? g.dai[0] : (comp == 1 ? g.dai[1] : g.dai[2]);
            const double dbr = comp == 0 ? g.dbr[0] : (comp == 1 ? g.dbr[1] : g.dbr[2]);
            const double dbi = comp == 0 ? g.dbi[0] : (comp == 1 ? g.dbi[1] : g.dbi[2]);
            const double uk = comp == 0 ? g.u[0] : (comp == 1 ? g.u[1] : g.u[2]);
            const double sf = g.sfac, dsu = g.dsfac * uk;
            const double k0 = sf * (-g.ai * dar + g.ar * dai + g.bi * dbr - g.br * dbi);
            const double k1 = sf * (-g.br * dar - g.bi * dai + g.ar * dbr + g.ai * dbi);
            const double k2 = sf * (-g.bi * dar + g.br * dai - g.ai * dbr + g.ar * dbi);
            for (int k = 0; k < nbh; ++k) acc[k * 64 + lane] = 0.0;
            row_chain<J, 0, 0, J / 2>(g, rootpq, [&](auto Jc, auto MBc, auto& pr, auto& pim, auto&, auto&) {
                constexpr int L = decltype(Jc)::value, MB = decltype(MBc)::value;
                constexpr int LAST = (L == 2 * MB) ? MB : L;
                double dr[LAST + 1], di[LAST + 1];
                static_for<0, LAST + 1>([&](auto MAc) {
                    constexpr int MA = decltype(MAc)::value;
                    constexpr double m0 = static_cast<double>(2 * MA - L);
                    constexpr double cu = (MA < L) ? cx_sqrt(static_cast<double>((MA + 1) * (L - MA))) : 0.0;
                    constexpr double cd = (MA > 0) ? cx_sqrt(static_cast<double>(MA * (L - MA + 1))) : 0.0;
                    const double upr = (MA < L) ? pr[MA < L ? MA + 1 : MA] : 0.0;
                    const double upi = (MA < L) ? pim[MA < L ? MA + 1 : MA] : 0.0;
                    const double dnr = (MA > 0) ? pr[MA > 0 ? MA - 1 : 0] : 0.0;
                    const double dni = (MA > 0) ? pim[MA > 0 ? MA - 1 : 0] : 0.0;
                    const double l1r = cu * upr - cd * dnr, l1i = cu * upi - cd * dni;
                    const double l2r = -(cu * upi + cd * dni), l2i = cu * upr + cd * dnr;
                    dr[MA] = dsu * pr[MA] + k0 * (-m0 * pim[MA]) + k1 * l1r + k2 * l2r;
                    di[MA] = dsu * pim[MA] + k0 * (m0 * pr[MA]) + k1 * l1i + k2 * l2i;
                });
                if constexpr (L == 2 * MB) {   // middle column of an even layer counts 1/2
                    dr[MB] *= 0.5;
                    di[MB] *= 0.5;
                }
                // Two terms per step (two independent chains), each summed as two FMAs per element.  The
                // next pair's Z rows and table entries are read from LDS one step ahead, and both
                // accumulators are read before the chains; when the two terms share a bispectrum index their
                // updates are merged, in the sequential order.
                const int ka = lo[L], kb = lo[L + 1];
                const int zo = 2 * (L + 1) * MB;
                int k = ka;
                if (kb - ka >= 2) {
                    double na[2 * (LAST + 1)], nc[2 * (LAST + 1)], nwa, nwc;
                    int nba, nbc;
                    auto load = [&](int kk) {
                        const double* za = zs + 2 * h_z[kk] + zo;
                        const double* zc = zs + 2 * h_z[kk + 1] + zo;
#pragma unroll
                        for (int m = 0; m < 2 * (LAST + 1); ++m) {
                            na[m] = za[m];
                            nc[m] = zc[m];
                        }
                        nba = h_b[kk] - b0;
                        nbc = h_b[kk + 1] - b0;
                        nwa = h_w[kk];
                        nwc = h_w[kk + 1];
                    };
                    load(k);
                    for (; k + 2 <= kb; k += 2) {
                        double za[2 * (LAST + 1)], zc[2 * (LAST + 1)];
#pragma unroll
                        for (int m = 0; m < 2 * (LAST + 1); ++m) {
                            za[m] = na[m];
                            zc[m] = nc[m];
                        }
                        const int ba = nba, bc = nbc;
                        const double wa = nwa, wc = nwc;
                        if (k + 4 <= kb) load(k + 2);
                        double* pa = acc + ba * 64 + lane;
                        double* pc = acc + bc * 64 + lane;
                        const double ra = *pa, rc = *pc;
                        double sa = 0.0, sc = 0.0;
#pragma unroll
                        for (int ma = 0; ma <= LAST; ++ma) {
                            sa = fma(dr[ma], za[2 * ma], sa);
                            sc = fma(dr[ma], zc[2 * ma], sc);
                            sa = fma(di[ma], za[2 * ma + 1], sa);
                            sc = fma(di[ma], zc[2 * ma + 1], sc);
                        }
                        const double va = ra + wa * sa;
                        if (ba == bc) {
                            *pa = va + wc * sc;
                        } else {
                            *pa = va;
                            *pc = rc + wc * sc;
                        }
                    }
                }
                if (k < kb) {
                    const double* za = zs + 2 * h_z[k] + zo;
                    double sa = 0.0;
#pragma unroll
                    for (int ma = 0; ma <= LAST; ++ma) sa = fma(di[ma], za[2 * ma + 1], fma(dr[ma], za[2 * ma], sa));
                    acc[(h_b[k] - b0) * 64 + lane] += h_w[k] * sa;
                }
            });
            if (active) {
                const size_t lp = static_cast<size_t>(q - p0);
                for (int k = 0; k < nbh; ++k) db[lp + Pc * comp + 3 * Pc * (b0 + k)] = acc[k * 64 + lane];
            }
        }
        __syncthreads();   // both waves done with this Z block before the next atom's copy
    }
}

// As k_fit_dbdr with the terms of every row enumerated at compile time: each term's Z row is a
// contiguous run of the atom's Z block at a constant offset, so it arrives in wide scalar loads, and
// the accumulator of each term sits at a constant LDS offset.  wtab[3*t + w] = weight of term w of t.
template <int J>
__global__ __launch_bounds__(64) void k_fit_dbdr_u(const double* __restrict__ rootpq, const double* __restrict__ zt,
                                                   const double* __restrict__ wtab, const double* __restrict__ radelem,
                                                   const double* __restrict__ wjelem, double rcutfac, double rmin0,
                                                   double rfac0, int switchflag, const double* __restrict__ rij,
                                                   const int* __restrict__ pti, const int* __restrict__ ptj,
                                                   const int* __restrict__ poff, const int* __restrict__ cfg_of,
                                                   const int* __restrict__ cfg_p0, double* __restrict__ dblist,
                                                   int* __restrict__ counter, int N) {
    constexpr int IZ = n_z(J), IB = n_b(J);
    __shared__ double acc[IB * 64];
    const int lane = threadIdx.x;
    const int comp = lane % 3, slot = lane / 3;
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(counter, 1);
        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
        if (t >= N) break;
        const int atom = t;
        const int q0 = __builtin_amdgcn_readfirstlane(poff[atom]);
        const int q1 = __builtin_amdgcn_readfirstlane(poff[atom + 1]);
        const int c = __builtin_amdgcn_readfirstlane(cfg_of[atom]);
        const int p0 = __builtin_amdgcn_readfirstlane(cfg_p0[c]);
        const long Pc = __builtin_amdgcn_readfirstlane(cfg_p0[c + 1]) - p0;
        const double* zbase = zt + static_cast<size_t>(atom) * IZ * 2;
        double* __restrict__ db = dblist + static_cast<size_t>(p0) * 3 * IB;
        for (int qb = q0; qb < q1; qb += 21) {
            // an opaque copy of the Z pointer per pass keeps its thousands of loads from being hoisted
            // out of the pass loop (where they would all stay live at once)
            const double* zp = zbase;
            asm volatile("" : "+s"(zp));
            const double* __restrict__ z = zp;
            const int q = qb + slot;
            const bool active = lane < 63 && q < q1;
            const int p = active ? q : q0;
            const int tjp = ptj[p];
            const double rc = (radelem[pti[p]] + radelem[tjp]) * rcutfac;
            PairGeom g;
            pair_geometry<true>(rij[3 * p], rij[3 * p + 1], rij[3 * p + 2], rc, rmin0, rfac0, switchflag, wjelem[tjp], g);
            const double uk = comp == 0 ? g.u[0] : (comp == 1 ? g.u[1] : g.u[2]);
            g.dar[0] = comp == 0 ? g.dar[0] : (comp == 1 ? g.dar[1] : g.dar[2]);
            g.dai[0] = comp == 0 ? g.dai[0] : (comp == 1 ? g.dai[1] : g.dai[2]);
            g.dbr[0] = comp == 0 ? g.dbr[0] : (comp == 1 ? g.dbr[1] : g.dbr[2]);
            g.dbi[0] = comp == 0 ? g.dbi[0] : (comp == 1 ? g.dbi[1] : g.dbi[2]);
            const double dsu = g.dsfac * uk, sf = g.sfac;
            for (int k = 0; k < IB; ++k) acc[k * 64 + lane] = 0.0;
            row_chain<J, 1, 0, J / 2>(g, rootpq, [&](auto Jc, auto MBc, auto& pr, auto& pim, auto& dpr, auto& dpi) {
                constexpr int L = decltype(Jc)::value, MB = decltype(MBc)::value;
                constexpr int LAST = (L == 2 * MB) ? MB : L;
                double dr[LAST + 1], di[LAST + 1];
#pragma unroll
                for (int ma = 0; ma <= LAST; ++ma) {
                    dr[ma] = dsu * pr[ma] + sf * dpr[0][ma];
                    di[ma] = dsu * pim[ma] + sf * dpi[0][ma];
                }
                if constexpr (L == 2 * MB) {
                    dr[MB] *= 0.5;
                    di[MB] *= 0.5;
                }
                static_for<0, IB>([&](auto Tc) {
                    constexpr int T = decltype(Tc)::value;
                    static_for<0, 3>([&](auto Wc) {
                        constexpr int WT = decltype(Wc)::value;
                        if constexpr (term_layer(J, T, WT) == L) {
                            constexpr int ZO = 2 * (term_zblock(J, T, WT) + (L + 1) * MB);
                            double sacc = 0.0;
#pragma unroll
                            for (int ma = 0; ma <= LAST; ++ma) sacc += dr[ma] * z[ZO + 2 * ma] + di[ma] * z[ZO + 2 * ma + 1];
                            acc[T * 64 + lane] += wtab[3 * T + WT] * sacc;
                        }
                    });
                    __builtin_amdgcn_sched_barrier(0);
                });
            });
            if (active) {
                const size_t lp = static_cast<size_t>(q - p0);
                for (int k = 0; k < IB; ++k) db[lp + Pc * comp + 3 * Pc * k] = acc[k * 64 + lane];
            }
        }
    }
}

// Term table of the dB/dr contraction grouped by layer: for every layer l the terms (bispectrum
// index, Z block offset, weight) that contract dU of layer l; term_lo has J+2 entries.
void fit_term_table(int J, int bnorm, std::vector<int>& lo, std::vector<int>& tb, std::vector<int>& tz,
                    std::vector<double>& tw) {
    struct Term { int l, b, z; double w; };
    std::vector<Term> all;
    auto zb = [&](int j1, int j2, int j) {
        int n = 0;
        for (int a = 0; a <= J; ++a)
            for (int b = 0; b <= a; ++b)
                for (int c = a - b; c <= std::min(J, a + b); c += 2) {
                    if (a == j1 && b == j2 && c == j) return n;
                    n += (c / 2 + 1) * (c + 1);
                }
        return -1;
    };
    int nb = 0;
    for (int j1 = 0; j1 <= J; ++j1)
        for (int j2 = 0; j2 <= j1; ++j2)
            for (int j = j1 - j2; j <= std::min(J, j1 + j2); j += 2) {
                if (j < j1) continue;
                const double f1 = bnorm ? 2.0 : 2.0 * (j + 1) / (j1 + 1.0);
                const double f2 = bnorm ? 2.0 : 2.0 * (j + 1) / (j2 + 1.0);
                all.push_back({j, nb, zb(j1, j2, j), 2.0});
                all.push_back({j1, nb, zb(j, j2, j1), f1});
                all.push_back({j2, nb, zb(j, j1, j2), f2});
                ++nb;
            }
    std::stable_sort(all.begin(), all.end(), [](const Term& a, const Term& b) { return a.l < b.l; });
    lo.assign(J + 2, 0);
    tb.clear(); tz.clear(); tw.clear();
    for (const Term& t : all) {
        tb.push_back(t.b);
        tz.push_back(t.z);
        tw.push_back(t.w);
        lo[t.l + 1] += 1;
    }
    for (int l = 0; l <= J; ++l) lo[l + 1] += lo[l];
}

int fit_term_table_halves(int J, int bnorm, std::vector<int>& lo, std::vector<int>& tb, std::vector<int>& tz,
                          std::vector<double>& tw) {
    std::vector<int> lo1, b1, z1;
    std::vector<double> w1;
    fit_term_table(J, bnorm, lo1, b1, z1, w1);
    // contraction work of each bispectrum index: the half-domain size of each of its terms' layers
    int nb = 0;
    for (int v : b1) nb = std::max(nb, v + 1);
    std::vector<long> work(nb, 0);
    for (int l = 0; l <= J; ++l)
        for (int k = lo1[l]; k < lo1[l + 1]; ++k) work[b1[k]] += static_cast<long>(l / 2 + 1) * (l + 1);
    long tot = 0;
    for (long v : work) tot += v;
    int hb = 0;
    long acc = 0;
    while (hb < nb && 2 * (acc + work[hb]) <= tot) acc += work[hb++];
    if (hb == 0) hb = 1;
    lo.assign(2 * (J + 2), 0);
    tb.clear(); tz.clear(); tw.clear();
    for (int h = 0; h < 2; ++h) {
        lo[h * (J + 2)] = static_cast<int>(tb.size());
        for (int l = 0; l <= J; ++l) {
            for (int k = lo1[l]; k < lo1[l + 1]; ++k)
                if ((b1[k] >= hb) == (h == 1)) {
                    tb.push_back(b1[k]);
                    tz.push_back(z1[k]);
                    tw.push_back(w1[k]);
                }
            lo[h * (J + 2) + l + 1] = static_cast<int>(tb.size());
        }
    }
    return hb;
}

// ---- launchers ----------------------------------------------------------------------------------

void launch_fit_pairs(bool fill, const double* x, const int* type, const int* cfg_of, const int* cfg_a0,
                      const double* cfg_L, const double* cfg_T, const double* rc2, int ntypes, int N,
                      const int* chunk_order, int* pcnt, const int* poff, long cap, double* rij, int* pi, int* pj,
                      int* pti, int* ptj, int* flags, hipStream_t st) {
    if (N <= 0) return;
    const dim3 grid(((N + 63) / 64) * FIT_PAIR_SPLIT), block(64);
    if (fill) {
        hipLaunchKernelGGL((k_fit_pairs<true>), grid, block, 0, st, x, type, cfg_of, cfg_a0, cfg_L, cfg_T, rc2, ntypes, N,
                           chunk_order, pcnt, poff, cap, rij, pi, pj, pti, ptj, flags);
    } else {
        hipLaunchKernelGGL((k_fit_pairs<false>), grid, block, 0, st, x, type, cfg_of, cfg_a0, cfg_L, cfg_T, rc2, ntypes,
                           N, chunk_order, pcnt, poff, cap, rij, pi, pj, pti, ptj, flags);
        hipLaunchKernelGGL(k_fit_pair_total, dim3((N + 255) / 256), dim3(256), 0, st, pcnt, N);
    }
}

void launch_fit_guard(const int* flags, int* poff, int N, hipStream_t st) {
    hipLaunchKernelGGL(k_fit_guard, dim3((N + 1 + 255) / 256), dim3(256), 0, st, flags, poff, N);
}

void launch_fit_cfg_p0(const int* poff, const int* cfg_a0, int* cfg_p0, int C, hipStream_t st) {
    hipLaunchKernelGGL(k_fit_cfg_p0, dim3((C + 1 + 255) / 256), dim3(256), 0, st, poff, cfg_a0, cfg_p0, C);
}

void launch_fill_zero_int(int* p, int n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_fill_zero_int, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
}

void launch_fit_zb(const SnapDev& s, const double* Ur, const double* Ui, double* zt, double* blist,
                   const double* bzero, const int* cfg_of, const int* cfg_a0, int N, int bnorm, int bzeroflag,
                   hipStream_t st) {
    if (N <= 0) return;
#define ZB_CASE(JV)                                                                                            \
    case JV:                                                                                                   \
        hipLaunchKernelGGL((k_fit_zb<JV>), dim3((N + 63) / 64, n_triples(JV)), dim3(64), 0, st, s.cg, Ur, Ui, zt, \
                           blist, bzero, cfg_of, cfg_a0, N, bnorm, bzeroflag);                                 \
        break;
    switch (s.J) {
        ZB_CASE(2)
        ZB_CASE(4)
        ZB_CASE(6)
        ZB_CASE(8)
        default: break;
    }
#undef ZB_CASE
}

void launch_fit_dbdr(const SnapDev& s, const double* zt, const FitTerms& tt, const double* rij, const int* pti,
                     const int* ptj, const int* poff, const int* cfg_of, const int* cfg_p0, double* dblist,
                     int* counter, int N, int groups, hipStream_t st, int variant) {
    if (N <= 0) return;
    const int G = groups > 0 ? groups : 1280;
    if (variant == 2) {   // the table-walking kernel limited to 256 registers (two waves per SIMD)
        switch (s.J) {
#define D2_CASE(JV)                                                                                             \
    case JV:                                                                                                    \
        hipLaunchKernelGGL((k_fit_dbdr<JV, 2>), dim3(G), dim3(64), 0, st, s.rootpq, zt, tt.lo, tt.b, tt.z, tt.w,   \
                           s.radelem, s.wjelem, s.rcutfac, s.rmin0, s.rfac0, s.switchflag, rij, pti, ptj, poff,    \
                           cfg_of, cfg_p0, dblist, counter, N);                                                  \
        return;
            D2_CASE(6)
            D2_CASE(8)
#undef D2_CASE
            default: break;
        }
    }
    if (variant == 4 && tt.h_lo) {   // generator form, bispectrum indices split over two wave halves
        const int G2 = (G + 1) / 2 * 2;
        const size_t lds = sizeof(double) * 64 * static_cast<size_t>(std::max(tt.hb, s.idxb_max - tt.hb));
        switch (s.J) {
#define D4_CASE(JV)                                                                                             \
    case JV:                                                                                                    \
        hipLaunchKernelGGL((k_fit_dbdr_g<JV>), dim3(G2), dim3(64), lds, st, s.rootpq, zt, tt.h_lo, tt.h_b, tt.h_z, \
                           tt.h_w, tt.hb, s.radelem, s.wjelem, s.rcutfac, s.rmin0, s.rfac0, s.switchflag, rij, pti, \
                           ptj, poff, cfg_of, cfg_p0, dblist, counter, N);                                       \
        return;
            D4_CASE(4)
            D4_CASE(6)
            D4_CASE(8)
#undef D4_CASE
            default: break;
        }
    }
    if (variant == 5 && tt.h_lo) {   // variant 4 with the atom's Z block staged in LDS, two waves per atom
        switch (s.J) {
#define D5_CASE(JV)                                                                                              \
    case JV:                                                                                                     \
        hipLaunchKernelGGL((k_fit_dbdr_l<JV>), dim3(G), dim3(128), 0, st, s.rootpq, zt, tt.h_lo, tt.h_b, tt.h_z,    \
                           tt.h_w, tt.hb, s.radelem, s.wjelem, s.rcutfac, s.rmin0, s.rfac0, s.switchflag, rij, pti,  \
                           ptj, poff, cfg_of, cfg_p0, dblist, counter, N);                                        \
        return;
            D5_CASE(4)
            D5_CASE(6)
            D5_CASE(8)
#undef D5_CASE
            default: break;
        }
    }
    if (variant == 3) {   // variant 0 with the pipelined term loop
        switch (s.J) {
#define D3_CASE(JV)                                                                                             \
    case JV:                                                                                                    \
        hipLaunchKernelGGL((k_fit_dbdr<JV, 1, true>), dim3(G), dim3(64), 0, st, s.rootpq, zt, tt.lo, tt.b, tt.z,   \
                           tt.w, s.radelem, s.wjelem, s.rcutfac, s.rmin0, s.rfac0, s.switchflag, rij, pti, ptj, poff, \
                           cfg_of, cfg_p0, dblist, counter, N);                                                  \
        return;
            D3_CASE(4)
            D3_CASE(6)
            D3_CASE(8)
#undef D3_CASE
            default: break;
        }
    }
    if (variant == 1 && tt.wtab) {
        switch (s.J) {
#define DU_CASE(JV)                                                                                             \
    case JV:                                                                                                    \
        hipLaunchKernelGGL((k_fit_dbdr_u<JV>), dim3(G), dim3(64), 0, st, s.rootpq, zt, tt.wtab, s.radelem, s.wjelem, \
                           s.rcutfac, s.rmin0, s.rfac0, s.switchflag, rij, pti, ptj, poff, cfg_of, cfg_p0, dblist,  \
                           counter, N);                                                                        \
        return;
            DU_CASE(4)
            DU_CASE(6)
            DU_CASE(8)
#undef DU_CASE
            default: break;
        }
    }
#define DB_CASE(JV)                                                                                            \
    case JV:                                                                                                   \
        hipLaunchKernelGGL((k_fit_dbdr<JV>), dim3(G), dim3(64), 0, st, s.rootpq, zt, tt.lo, tt.b, tt.z, tt.w,    \
                           s.radelem, s.wjelem, s.rcutfac, s.rmin0, s.rfac0, s.switchflag, rij, pti, ptj, poff,  \
                           cfg_of, cfg_p0, dblist, counter, N);                                                \
        break;
    switch (s.J) {
        DB_CASE(2)
        DB_CASE(4)
        DB_CASE(6)
        DB_CASE(8)
        default: break;
    }
#undef DB_CASE
}

}  // namespace mdp
