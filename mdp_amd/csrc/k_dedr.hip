// compute_duidrj + compute_deidrj for gfx950 (cpuSnap.cpp:991-1376 semantics, without the tally).
//
// Lane = pair.  Each lane runs the row-chain recursion of U and dU/dx,y,z of its pair entirely in
// registers (compile-time J) and contracts every finished row with Y of the pair's centre atom on
// the fly:  dE/dr_ij = 2 sum' (dsfac U u + sfac dU) . Y_i, with the j=0 term dsfac u Y_0 / 2 and the
// middle column of even layers weighted 1/2.  The result is written per pair; the force/virial
// tally is a separate deterministic gather.
#include <hip/hip_runtime.h>

#include "snap_device.hpp"
#include "snap_launch.hpp"

namespace mdp {

template <int J, bool HALF>
__global__ __launch_bounds__(64) void k_dedr(SnapDev s, const double* __restrict__ Yr, const double* __restrict__ Yi,
                                             const double* __restrict__ rij, const int* __restrict__ yatom,
                                             const int* __restrict__ ti, const int* __restrict__ tj,
                                             const long* __restrict__ ybase, double* __restrict__ dedr, int N,
                                             int npairs_host, const int* __restrict__ npairs_dev) {
    const int p = blockIdx.x * 64 + threadIdx.x;
    const int np = npairs_dev ? *npairs_dev : npairs_host;
    if (p >= np) return;
    const int tjp = tj[p];
    const double rc = (s.radelem[ti[p]] + s.radelem[tjp]) * s.rcutfac;
    PairGeom g;
    pair_geometry<true>(rij[3 * p], rij[3 * p + 1], rij[3 * p + 2], rc, s.rmin0, s.rfac0, s.switchflag, s.wjelem[tjp], g);
    const size_t n = static_cast<size_t>(N);
    const size_t yb = static_cast<size_t>(yatom[p]) + (ybase ? static_cast<size_t>(ybase[p]) : 0);
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
    row_chain<J, 3, 0, J / 2>(g, s.rootpq, [&](auto Jc, auto MBc, auto& pr, auto& pi, auto& dpr, auto& dpi) {
        constexpr int JJ = decltype(Jc)::value, MB = decltype(MBc)::value;
        if constexpr (JJ == 0) {
            const double y0 = 0.5 * Yr[yb];
            d0 = g.dsfac * g.u[0] * y0;
            d1 = g.dsfac * g.u[1] * y0;
            d2 = g.dsfac * g.u[2] * y0;
        } else {
            constexpr int LAST = (JJ == 2 * MB) ? MB : JJ;
            constexpr int row = (HALF ? h_block(JJ) : u_block(JJ)) + (JJ + 1) * MB;
            const size_t o = static_cast<size_t>(row) * n + yb;
#pragma unroll
            for (int ma = 0; ma <= LAST; ++ma) {
                double yr = Yr[o + ma * n], yi = Yi[o + ma * n];
                if (JJ == 2 * MB && ma == MB) { yr *= 0.5; yi *= 0.5; }
                const double ar = g.dsfac * pr[ma], ai = g.dsfac * pi[ma];
                d0 += (ar * g.u[0] + g.sfac * dpr[0][ma]) * yr + (ai * g.u[0] + g.sfac * dpi[0][ma]) * yi;
                d1 += (ar * g.u[1] + g.sfac * dpr[1][ma]) * yr + (ai * g.u[1] + g.sfac * dpi[1][ma]) * yi;
                d2 += (ar * g.u[2] + g.sfac * dpr[2][ma]) * yr + (ai * g.u[2] + g.sfac * dpi[2][ma]) * yi;
            }
        }
    });
    dedr[3 * static_cast<size_t>(p)] = 2.0 * d0;
    dedr[3 * static_cast<size_t>(p) + 1] = 2.0 * d1;
    dedr[3 * static_cast<size_t>(p) + 2] = 2.0 * d2;
}


// Same contraction with lane = (pair, Cartesian component): a wave holds 21 pairs x 3 components
// (lane 63 idle).  Each lane carries U and one derivative recursion, about half the register state of
// k_dedr, which lets several waves share a SIMD; U is recomputed by the three lanes of a pair.
// NW waves per workgroup take NW consecutive 21-pair tiles: the pairs of a few neighbouring centre
// atoms then share one CU, so their Y rows are reused from that CU's L1.
template <int J, bool HALF, int WPE, int NW = 1>
__global__ __launch_bounds__(64 * NW, WPE) void k_dedr_split(SnapDev s, const double* __restrict__ Yr,
                                                   const double* __restrict__ Yi, const double* __restrict__ rij,
                                                   const int* __restrict__ yatom, const int* __restrict__ ti,
                                                   const int* __restrict__ tj, const long* __restrict__ ybase,
                                                   double* __restrict__ dedr, int N, int npairs_host,
                                                   const int* __restrict__ npairs_dev) {
    const int lane = threadIdx.x & 63;
    const int p = (blockIdx.x * NW + static_cast<int>(threadIdx.x >> 6)) * 21 + lane / 3, comp = lane % 3;
    const int np = npairs_dev ? *npairs_dev : npairs_host;
    if (lane >= 63 || p >= np) return;
    const int tjp = tj[p];
    const double rc = (s.radelem[ti[p]] + s.radelem[tjp]) * s.rcutfac;
    PairGeom g;
    pair_geometry<true>(rij[3 * p], rij[3 * p + 1], rij[3 * p + 2], rc, s.rmin0, s.rfac0, s.switchflag, s.wjelem[tjp], g);
    // move this lane's component into slot 0
    const double uk = comp == 0 ? g.u[0] : (comp == 1 ? g.u[1] : g.u[2]);
    g.dar[0] = comp == 0 ? g.dar[0] : (comp == 1 ? g.dar[1] : g.dar[2]);
    g.dai[0] = comp == 0 ? g.dai[0] : (comp == 1 ? g.dai[1] : g.dai[2]);
    g.dbr[0] = comp == 0 ? g.dbr[0] : (comp == 1 ? g.dbr[1] : g.dbr[2]);
    g.dbi[0] = comp == 0 ? g.dbi[0] : (comp == 1 ? g.dbi[1] : g.dbi[2]);
    const size_t n = static_cast<size_t>(N);
    const size_t yb = static_cast<size_t>(yatom[p]) + (ybase ? static_cast<size_t>(ybase[p]) : 0);
    const double dsu = g.dsfac * uk;
    double d = 0.0;
    row_chain<J, 1, 0, J / 2>(g, s.rootpq, [&](auto Jc, auto MBc, auto& pr, auto& pi, auto& dpr, auto& dpi) {
        constexpr int JJ = decltype(Jc)::value, MB = decltype(MBc)::value;
        if constexpr (JJ == 0) {
            d = dsu * (0.5 * Yr[yb]);
        } else {
            constexpr int LAST = (JJ == 2 * MB) ? MB : JJ;
            constexpr int row = (HALF ? h_block(JJ) : u_block(JJ)) + (JJ + 1) * MB;
            const size_t o = static_cast<size_t>(row) * n + yb;
#pragma unroll
            for (int ma = 0; ma <= LAST; ++ma) {
                double yr = Yr[o + ma * n], yi = Yi[o + ma * n];
                if (JJ == 2 * MB && ma == MB) { yr *= 0.5; yi *= 0.5; }
                d += (dsu * pr[ma] + g.sfac * dpr[0][ma]) * yr + (dsu * pi[ma] + g.sfac * dpi[0][ma]) * yi;
            }
        }
    });
    dedr[3 * static_cast<size_t>(p) + comp] = 2.0 * d;
}

// dE/dr_ij from U alone.  U is a representation of SU(2) on the unit quaternion q = (a_r, a_i, b_r, b_i):
// along the fields V_0 = q*i, V_1 = q*j, V_2 = q*k its derivative is L_g U, with constant tridiagonal
// generators that mix ma within a row,
//   (L_0 U)(ma) = i (2ma - j) U(ma),
//   (L_1 U)(ma) = c(ma) U(ma+1) - c(ma-1) U(ma-1),
//   (L_2 U)(ma) = i (c(ma) U(ma+1) + c(ma-1) U(ma-1)),      c(m) = sqrt((m+1)(j-m)).
// |q| = 1 for every r_ij, so dq/dx_k = sum_g (dq/dx_k . V_g) V_g and
//   dF/dx_k = sum_g (dq/dx_k . V_g) T_g,   T_g = sum' Re(conj(Y) L_g U),   F = sum' Re(conj(Y) U)
// (half domain, middle column weighted 1/2, j = 0 term Y_0/2).  The result equals the contraction of
// compute_duidrj's dU/dr with Y (cpuSnap.cpp:1091-1376) up to rounding: one U recursion per pair and four
// running sums replace its three derivative recursions.
// With XCD, workgroup b runs tile (b % 8) * C + b / 8, C = ceil(tiles / 8) for the pairs present:
// workgroups are dealt to the 8 XCDs round-robin, so each XCD then walks one contiguous eighth of the
// pair list and the Y lines of its atoms (one 128-byte line holds an element of 16 consecutive atoms)
// are fetched into one L2 only.  The grid is a multiple of 8 covering the pair capacity; workgroups with
// b / 8 >= C and tiles past the end return.
template <int J, bool HALF, int NW, bool PREFETCH, bool XCD = false>
__global__ __launch_bounds__(64 * NW) void k_dedr_gen(SnapDev s, const double* __restrict__ Yr,
                                                      const double* __restrict__ Yi, const double* __restrict__ rij,
                                                      const int* __restrict__ yatom, const int* __restrict__ ti,
                                                      const int* __restrict__ tj, const long* __restrict__ ybase,
                                                      double* __restrict__ dedr, int N, int npairs_host,
                                                      const int* __restrict__ npairs_dev) {
    const int np = npairs_dev ? *npairs_dev : npairs_host;
    int tile = static_cast<int>(blockIdx.x);
    if constexpr (XCD) {
        const int chunk = ((np + 64 * NW - 1) / (64 * NW) + 7) / 8;
        const int k = static_cast<int>(blockIdx.x / 8);
        if (k >= chunk) return;
        tile = static_cast<int>(blockIdx.x % 8) * chunk + k;
    }
    const int p = (tile * NW + static_cast<int>(threadIdx.x >> 6)) * 64 + static_cast<int>(threadIdx.x & 63);
    if (p >= np) return;
    const int tjp = tj[p];
    const double rc = (s.radelem[ti[p]] + s.radelem[tjp]) * s.rcutfac;
    PairGeom g;
    pair_geometry<true>(rij[3 * p], rij[3 * p + 1], rij[3 * p + 2], rc, s.rmin0, s.rfac0, s.switchflag, s.wjelem[tjp], g);
    const size_t n = static_cast<size_t>(N);
    const size_t yb = static_cast<size_t>(yatom[p]) + (ybase ? static_cast<size_t>(ybase[p]) : 0);
    double f = 0.5 * Yr[yb], t0 = 0.0, t1 = 0.0, t2 = 0.0;
    // With PREFETCH, Y rows are loaded one emitted row ahead (rows come in the order (1,0) .. (J,0),
    // (2,1) .. (J,1), ...) so that their latency overlaps the recursion of the current row.
    double ynr[J + 1], yni[J + 1];
    auto load_row = [&](auto Jc, auto MBc) {
        constexpr int JJ = decltype(Jc)::value, MB = decltype(MBc)::value;
        constexpr int LAST = (JJ == 2 * MB) ? MB : JJ;
        constexpr int row = (HALF ? h_block(JJ) : u_block(JJ)) + (JJ + 1) * MB;
        const size_t o = static_cast<size_t>(row) * n + yb;
#pragma unroll
        for (int ma = 0; ma <= LAST; ++ma) {
            ynr[ma] = Yr[o + ma * n];
            yni[ma] = Yi[o + ma * n];
        }
    };
    if constexpr (PREFETCH) load_row(ic<1>{}, ic<0>{});
    row_chain<J, 0, 0, J / 2>(g, s.rootpq, [&](auto Jc, auto MBc, auto& pr, auto& pi, auto&, auto&) {
        constexpr int JJ = decltype(Jc)::value, MB = decltype(MBc)::value;
        if constexpr (JJ > 0) {
            constexpr bool MID = (JJ == 2 * MB);
            constexpr int LAST = MID ? MB : JJ;
            double ycr[LAST + 1], yci[LAST + 1];
            if constexpr (!PREFETCH) load_row(Jc, MBc);
#pragma unroll
            for (int ma = 0; ma <= LAST; ++ma) { ycr[ma] = ynr[ma]; yci[ma] = yni[ma]; }
            if constexpr (PREFETCH) {
                if constexpr (JJ < J) load_row(ic<JJ + 1>{}, ic<MB>{});
                else if constexpr (2 * (MB + 1) <= J) load_row(ic<2 * (MB + 1)>{}, ic<MB + 1>{});
            }
            static_for<0, LAST + 1>([&](auto MAc) {
                constexpr int MA = decltype(MAc)::value;
                constexpr double w = (MID && MA == MB) ? 0.5 : 1.0;
                const double yr = w * ycr[MA], yi = w * yci[MA];
                f += yr * pr[MA] + yi * pi[MA];
                t0 += static_cast<double>(2 * MA - JJ) * (yi * pr[MA] - yr * pi[MA]);
                if constexpr (MA < JJ) {
                    constexpr double c = cx_sqrt(static_cast<double>((MA + 1) * (JJ - MA)));
                    t1 += c * (yr * pr[MA + 1] + yi * pi[MA + 1]);
                    t2 += c * (yi * pr[MA + 1] - yr * pi[MA + 1]);
                }
                if constexpr (MA > 0) {
                    constexpr double c = cx_sqrt(static_cast<double>(MA * (JJ - MA + 1)));
                    t1 -= c * (yr * pr[MA - 1] + yi * pi[MA - 1]);
                    t2 += c * (yi * pr[MA - 1] - yr * pi[MA - 1]);
                }
            });
        }
    });
    // V_0 = q*i = (-a_i, a_r, b_i, -b_r), V_1 = q*j = (-b_r, -b_i, a_r, a_i), V_2 = q*k = (-b_i, b_r, -a_i, a_r)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double c0 = -g.ai * g.dar[k] + g.ar * g.dai[k] + g.bi * g.dbr[k] - g.br * g.dbi[k];
        const double c1 = -g.br * g.dar[k] - g.bi * g.dai[k] + g.ar * g.dbr[k] + g.ai * g.dbi[k];
        const double c2 = -g.bi * g.dar[k] + g.br * g.dai[k] - g.ai * g.dbr[k] + g.ar * g.dbi[k];
        dedr[3 * static_cast<size_t>(p) + k] = 2.0 * (g.dsfac * g.u[k] * f + g.sfac * (c0 * t0 + c1 * t1 + c2 * t2));
    }
}

// Y of the half layout ([element][atom]) copied atom-major and complex-interleaved,
// yt[(atom*HE + e)*2 + {0,1}], so that one atom's Y is a contiguous block.
__global__ void k_y_atom_major(const double* __restrict__ Yr, const double* __restrict__ Yi, int HE, int N,
                               double* __restrict__ yt) {
    const long k = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= static_cast<long>(HE) * N) return;
    const int e = static_cast<int>(k / N), atom = static_cast<int>(k % N);
    const size_t o = (static_cast<size_t>(atom) * HE + e) * 2;
    yt[o] = Yr[k];
    yt[o + 1] = Yi[k];
}

// One wave per centre atom (grid = atoms), lanes (pair, Cartesian component) as in k_dedr_split for up
// to 21 pairs per pass.  All pairs of the wave share the atom, so its Y block (atom-major copy) is
// wave-uniform and every element comes in through scalar loads at compile-time offsets.
template <int J>
__global__ __launch_bounds__(64, 2) void k_dedr_atom(const double* __restrict__ rootpq, const double* __restrict__ yt,
                                                    const double* __restrict__ radelem,
                                                    const double* __restrict__ wjelem, double rcutfac, double rmin0,
                                                    double rfac0, int switchflag, const double* __restrict__ rij,
                                                    const int* __restrict__ ti, const int* __restrict__ tj,
                                                    const int* __restrict__ poff, double* __restrict__ dedr) {
    constexpr int HE = n_half_elems(J);
    const int atom = blockIdx.x;
    const int lane = threadIdx.x, comp = lane % 3, slot = lane / 3;
    const int q0 = poff[atom], q1 = poff[atom + 1];
    const double* __restrict__ y = yt + static_cast<size_t>(atom) * HE * 2;
    for (int qb = q0; qb < q1; qb += 21) {
        const int q = qb + slot;
        const bool active = lane < 63 && q < q1;
        const int p = active ? q : q0;   // idle lanes recompute a valid pair and store nothing
        const int tjp = tj[p];
        const double rc = (radelem[ti[p]] + radelem[tjp]) * rcutfac;
        PairGeom g;
        pair_geometry<true>(rij[3 * p], rij[3 * p + 1], rij[3 * p + 2], rc, rmin0, rfac0, switchflag, wjelem[tjp], g);
        const double uk = comp == 0 ? g.u[0] : (comp == 1 ? g.u[1] : g.u[2]);
        g.dar[0] = comp == 0 ? g.dar[0] : (comp == 1 ? g.dar[1] : g.dar[2]);
        g.dai[0] = comp == 0 ? g.dai[0] : (comp == 1 ? g.dai[1] : g.dai[2]);
        g.dbr[0] = comp == 0 ? g.dbr[0] : (comp == 1 ? g.dbr[1] : g.dbr[2]);
        g.dbi[0] = comp == 0 ? g.dbi[0] : (comp == 1 ? g.dbi[1] : g.dbi[2]);
        const double dsu = g.dsfac * uk;
        double d = 0.0;
        row_chain<J, 1, 0, J / 2>(g, rootpq, [&](auto Jc, auto MBc, auto& pr, auto& pi, auto& dpr, auto& dpi) {
            constexpr int JJ = decltype(Jc)::value, MB = decltype(MBc)::value;
            if constexpr (JJ == 0) {
                d = dsu * (0.5 * y[0]);
            } else {
                constexpr int LAST = (JJ == 2 * MB) ? MB : JJ;
                constexpr int row = h_block(JJ) + (JJ + 1) * MB;
#pragma unroll
                for (int ma = 0; ma <= LAST; ++ma) {
                    double a = y[2 * (row + ma)], b = y[2 * (row + ma) + 1];
                    if (JJ == 2 * MB && ma == MB) { a *= 0.5; b *= 0.5; }
                    d += (dsu * pr[ma] + g.sfac * dpr[0][ma]) * a + (dsu * pi[ma] + g.sfac * dpi[0][ma]) * b;
                }
            }
        });
        if (active) dedr[3 * static_cast<size_t>(q) + comp] = 2.0 * d;
    }
}

void launch_dedr_atom(const SnapDev& s, const double* Yr, const double* Yi, double* yt, const double* rij,
                      const int* ti, const int* tj, const int* pair_off, double* dedr, int N, hipStream_t st) {
    if (N <= 0) return;
    int HE = 0;
    switch (s.J) {
        case 2: HE = n_half_elems(2); break;
        case 4: HE = n_half_elems(4); break;
        case 6: HE = n_half_elems(6); break;
        case 8: HE = n_half_elems(8); break;
        default: return;
    }
    const long tot = static_cast<long>(HE) * N;
    hipLaunchKernelGGL(k_y_atom_major, dim3((tot + 255) / 256), dim3(256), 0, st, Yr, Yi, HE, N, yt);
    switch (s.J) {
#define DA_CASE(JV)                                                                                               \
    case JV:                                                                                                      \
        hipLaunchKernelGGL((k_dedr_atom<JV>), dim3(N), dim3(64), 0, st, s.rootpq, yt, s.radelem, s.wjelem, s.rcutfac, \
                           s.rmin0, s.rfac0, s.switchflag, rij, ti, tj, pair_off, dedr);                          \
        return;
        DA_CASE(2)
        DA_CASE(4)
        DA_CASE(6)
        DA_CASE(8)
#undef DA_CASE
        default: return;
    }
}

void launch_dedr(const SnapDev& s, const double* Yr, const double* Yi, const double* rij, const int* yatom,
                 const int* ti, const int* tj, const long* ybase, double* dedr, int N, int npairs_host,
                 const int* npairs_dev, hipStream_t st, bool half, int variant) {
    if (npairs_host <= 0) return;
    const dim3 g3((npairs_host + 63) / 64), g1((npairs_host + 20) / 21), g2((npairs_host + 41) / 42),
        g4((npairs_host + 83) / 84), g256((npairs_host + 255) / 256),
        g3x(((npairs_host + 63) / 64 + 7) / 8 * 8), g256x(((npairs_host + 255) / 256 + 7) / 8 * 8);
#define DE_ARGS Yr, Yi, rij, yatom, ti, tj, ybase, dedr, N, npairs_host, npairs_dev
#define DE_CASE(JV)                                                                                            \
    case JV:                                                                                                   \
        if (!half)                                                                                             \
            hipLaunchKernelGGL((k_dedr_gen<JV, false, 1, true>), g3, dim3(64), 0, st, s, DE_ARGS);             \
        else if (variant == 0)                                                                                 \
            hipLaunchKernelGGL((k_dedr<JV, true>), g3, dim3(64), 0, st, s, DE_ARGS);                           \
        else if (variant == 2)                                                                                 \
            hipLaunchKernelGGL((k_dedr_split<JV, true, 2>), g1, dim3(64), 0, st, s, DE_ARGS);                  \
        else if (variant == 3)                                                                                 \
            hipLaunchKernelGGL((k_dedr_split<JV, true, 3>), g1, dim3(64), 0, st, s, DE_ARGS);                  \
        else if (variant == 6)                                                                                 \
            hipLaunchKernelGGL((k_dedr_split<JV, true, 2, 2>), g2, dim3(128), 0, st, s, DE_ARGS);              \
        else if (variant == 7)                                                                                 \
            hipLaunchKernelGGL((k_dedr_split<JV, true, 2, 4>), g4, dim3(256), 0, st, s, DE_ARGS);              \
        else if (variant == 8)                                                                                 \
            hipLaunchKernelGGL((k_dedr_gen<JV, true, 1, false>), g3, dim3(64), 0, st, s, DE_ARGS);             \
        else if (variant == 9)                                                                                 \
            hipLaunchKernelGGL((k_dedr_gen<JV, true, 4, false>), g256, dim3(256), 0, st, s, DE_ARGS);          \
        else if (variant == 10)                                                                                \
            hipLaunchKernelGGL((k_dedr_gen<JV, true, 1, true>), g3, dim3(64), 0, st, s, DE_ARGS);              \
        else if (variant == 11)                                                                                \
            hipLaunchKernelGGL((k_dedr_gen<JV, true, 4, true>), g256, dim3(256), 0, st, s, DE_ARGS);           \
        else if (variant == 12)                                                                                \
            hipLaunchKernelGGL((k_dedr_gen<JV, true, 1, true, true>), g3x, dim3(64), 0, st, s, DE_ARGS);       \
        else if (variant == 13)                                                                                \
            hipLaunchKernelGGL((k_dedr_gen<JV, true, 4, true, true>), g256x, dim3(256), 0, st, s, DE_ARGS);    \
        else                                                                                                   \
            hipLaunchKernelGGL((k_dedr_split<JV, true, 4>), g1, dim3(64), 0, st, s, DE_ARGS);                  \
        break;
    switch (s.J) {
        DE_CASE(2)
        DE_CASE(4)
        DE_CASE(6)
        DE_CASE(8)
        default: break;
    }
#undef DE_CASE
#undef DE_ARGS
}

int dedr_variants(int* ids) {
    static const int v[] = {0, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13};
    for (int k = 0; k < 13; ++k) ids[k] = v[k];
    return 13;   // 5: launch_dedr_atom; 6, 7: variant 2 with 2 and 4 waves per workgroup; 8-13: k_dedr_gen
}

}  // namespace mdp
