// compute_ui for gfx950 (cpuSnap.cpp:312-542 semantics).
//
// Lane = atom: a wavefront covers 64 consecutive atoms and each lane walks its own pairs.  The U
// rows of one "row group" (rows RLO..RHI of the half domain) are accumulated in registers over all
// pairs, so there are no atomics and the result is independent of scheduling.  Rows below RLO are
// advanced only up to the seed of the next row.  At twojmax = 8 the 155 half-U elements are split
// into four groups ({0}, {1}, {2}, {3,4}: 45/42/35/33 elements) run by separate workgroups of one
// launch.  U is stored atom-fastest, so every store of a wave is one contiguous 512-byte segment.
#include <hip/hip_runtime.h>

#include "snap_device.hpp"
#include "snap_launch.hpp"

namespace mdp {

bool templated_j(int J) { return J == 2 || J == 4 || J == 6 || J == 8; }

// GEO: the pair's Cayley-Klein parameters and weighted switching factor come precomputed from
// pgeo[5q..5q+4] = (a_r, a_i, b_r, b_i, sfac * w_j) (the MD pair list writes them once per step with the
// same pair_geometry), instead of being derived here from r_ij in every row group.
template <int J, int RLO, int RHI, int MODE, bool GEO = false>
__device__ __forceinline__ void ui_rows(const SnapDev& s, double* __restrict__ Ur, double* __restrict__ Ui,
                                        const double* __restrict__ rij, const int* __restrict__ pair_off,
                                        const int* __restrict__ perm, const int* __restrict__ ti,
                                        const int* __restrict__ tj, int N, double wself, int elem_sel, int atom,
                                        const double* __restrict__ pgeo = nullptr) {
    constexpr int NE = rows_elems(J, RLO, RHI);
    double accr[NE], acci[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) { accr[k] = 0.0; acci[k] = 0.0; }
    auto accumulate = [&](const PairGeom& g, double sf) {
        row_chain<J, 0, RLO, RHI>(g, s.rootpq, [&](auto Jc, auto MBc, auto& pr, auto& pi, auto&, auto&) {
            constexpr int JJ = decltype(Jc)::value, MB = decltype(MBc)::value;
            constexpr int base = rows_slot(J, RLO, JJ, MB, 0);
#pragma unroll
            for (int ma = 0; ma <= JJ; ++ma) {
                accr[base + ma] += sf * pr[ma];
                acci[base + ma] += sf * pi[ma];
            }
        });
    };

    // the next pair's data are loaded while the current pair's recursion runs
    const int q0 = pair_off[atom], q1 = pair_off[atom + 1];
    if constexpr (GEO) {
        double n[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (q0 < q1)
#pragma unroll
            for (int k = 0; k < 5; ++k) n[k] = pgeo[5 * static_cast<size_t>(q0) + k];
        for (int q = q0; q < q1; ++q) {
            PairGeom g;
            g.ar = n[0];
            g.ai = n[1];
            g.br = n[2];
            g.bi = n[3];
            const double sf = n[4];
            if (q + 1 < q1)
#pragma unroll
                for (int k = 0; k < 5; ++k) n[k] = pgeo[5 * static_cast<size_t>(q + 1) + k];
            accumulate(g, sf);
        }
    } else {
        double nx = 0.0, ny = 0.0, nz = 0.0;
        int nti = 0, ntj = 0;
        if (q0 < q1) {
            const int p = perm ? perm[q0] : q0;
            nx = rij[3 * p]; ny = rij[3 * p + 1]; nz = rij[3 * p + 2];
            nti = ti[p]; ntj = tj[p];
        }
        for (int q = q0; q < q1; ++q) {
            const double x = nx, y = ny, z = nz;
            const int tip = nti, tjp = ntj;
            if (q + 1 < q1) {
                const int p = perm ? perm[q + 1] : q + 1;
                nx = rij[3 * p]; ny = rij[3 * p + 1]; nz = rij[3 * p + 2];
                nti = ti[p]; ntj = tj[p];
            }
            if (elem_sel >= 0 && s.map[tjp] != elem_sel) continue;
            const double rc = (s.radelem[tip] + s.radelem[tjp]) * s.rcutfac;
            PairGeom g;
            pair_geometry<false>(x, y, z, rc, s.rmin0, s.rfac0, s.switchflag, s.wjelem[tjp], g);
            accumulate(g, g.sfac);
        }
    }

    const size_t n = static_cast<size_t>(N);
    static_for<RLO, RHI + 1>([&](auto MBc) {
        constexpr int MB = decltype(MBc)::value;
        static_for<2 * MB, J + 1>([&](auto Jc) {
            constexpr int JJ = decltype(Jc)::value;
            constexpr int base = rows_slot(J, RLO, JJ, MB, 0);
            constexpr int row = u_block(JJ) + (JJ + 1) * MB;
#pragma unroll
            for (int ma = 0; ma <= JJ; ++ma) {
                const size_t o = static_cast<size_t>(row + ma) * n + atom;
                if constexpr (MODE == UI_ACCUMULATE) {
                    Ur[o] += accr[base + ma];
                    Ui[o] += acci[base + ma];
                } else if constexpr (MODE == UI_HALF_WSELF) {
                    const size_t oh = static_cast<size_t>(h_block(JJ) + (JJ + 1) * MB + ma) * n + atom;
                    Ur[oh] = accr[base + ma] + (ma == MB ? wself : 0.0);
                    Ui[oh] = acci[base + ma];
                } else {
                    const double vr = accr[base + ma] + (ma == MB ? wself : 0.0);
                    const double vi = acci[base + ma];
                    Ur[o] = vr;
                    Ui[o] = vi;
                    if (2 * MB < JJ) {   // mirror into row j-mb: U(j-mb, j-ma) = (-1)^(ma+mb) conj U(mb, ma)
                        const size_t om =
                            static_cast<size_t>(u_block(JJ) + (JJ + 1) * (JJ - MB) + (JJ - ma)) * n + atom;
                        const bool even = ((ma + MB) % 2) == 0;
                        Ur[om] = even ? vr : -vr;
                        Ui[om] = even ? -vi : vi;
                    }
                }
            }
        });
    });
}

// The row groups of twojmax = 8 ({0}, {1}, {2}, {3,4}) and 6 ({0}, {1}, {2,3}) as one launch: workgroup
// b runs group (b / 8) % NG on atom block (b / 8 / NG) * 8 + b % 8, so the groups of one atom block go to
// the same XCD (workgroups are dealt to the 8 XCDs round-robin) and share its L2 for r_ij.
template <int J>
constexpr int ui_ngroups() { return J == 8 ? 4 : (J == 6 ? 3 : 1); }

template <int J, int MODE, bool GEO = false>
__global__ __launch_bounds__(64) void k_ui(SnapDev s, double* __restrict__ Ur, double* __restrict__ Ui,
                                           const double* __restrict__ rij, const int* __restrict__ pair_off,
                                           const int* __restrict__ perm, const int* __restrict__ ti,
                                           const int* __restrict__ tj, int N, double wself, int elem_sel,
                                           const double* __restrict__ pgeo) {
    constexpr int NG = ui_ngroups<J>();
    const int b = blockIdx.x, k = b / 8;
    const int group = k % NG, ablock = (k / NG) * 8 + b % 8;
    const int atom = ablock * 64 + threadIdx.x;
    if (atom >= N) return;
    if constexpr (J == 8) {
        if (group == 0) ui_rows<8, 0, 0, MODE, GEO>(s, Ur, Ui, rij, pair_off, perm, ti, tj, N, wself, elem_sel, atom, pgeo);
        else if (group == 1) ui_rows<8, 1, 1, MODE, GEO>(s, Ur, Ui, rij, pair_off, perm, ti, tj, N, wself, elem_sel, atom, pgeo);
        else if (group == 2) ui_rows<8, 2, 2, MODE, GEO>(s, Ur, Ui, rij, pair_off, perm, ti, tj, N, wself, elem_sel, atom, pgeo);
        else ui_rows<8, 3, 4, MODE, GEO>(s, Ur, Ui, rij, pair_off, perm, ti, tj, N, wself, elem_sel, atom, pgeo);
    } else if constexpr (J == 6) {
        if (group == 0) ui_rows<6, 0, 0, MODE, GEO>(s, Ur, Ui, rij, pair_off, perm, ti, tj, N, wself, elem_sel, atom, pgeo);
        else if (group == 1) ui_rows<6, 1, 1, MODE, GEO>(s, Ur, Ui, rij, pair_off, perm, ti, tj, N, wself, elem_sel, atom, pgeo);
        else ui_rows<6, 2, 3, MODE, GEO>(s, Ur, Ui, rij, pair_off, perm, ti, tj, N, wself, elem_sel, atom, pgeo);
    } else {
        ui_rows<J, 0, J / 2, MODE, GEO>(s, Ur, Ui, rij, pair_off, perm, ti, tj, N, wself, elem_sel, atom, pgeo);
    }
}

template <int J, int MODE, bool GEO = false>
static void ui_groups(const SnapDev& s, double* Ur, double* Ui, const double* rij, const int* off, const int* perm,
                      const int* ti, const int* tj, int N, double wself, int esel, hipStream_t st,
                      const double* pgeo = nullptr) {
    constexpr int NG = ui_ngroups<J>();
    const int ablocks = ((N + 63) / 64 + 7) / 8 * 8;   // padded to whole XCD rounds; excess blocks return
    hipLaunchKernelGGL((k_ui<J, MODE, GEO>), dim3(ablocks * NG), dim3(64), 0, st, s, Ur, Ui, rij, off, perm, ti, tj,
                       N, wself, esel, pgeo);
}

void launch_ui(const SnapDev& s, UiMode mode, double* Ur, double* Ui, const double* rij, const int* pair_off,
               const int* perm, const int* ti, const int* tj, int N, double wself, int esel, hipStream_t st) {
    if (N <= 0) return;
#define UI_CASE(JV)                                                                                   \
    case JV:                                                                                          \
        if (mode == UI_ACCUMULATE)                                                                    \
            ui_groups<JV, UI_ACCUMULATE>(s, Ur, Ui, rij, pair_off, perm, ti, tj, N, wself, esel, st); \
        else if (mode == UI_HALF_WSELF)                                                               \
            ui_groups<JV, UI_HALF_WSELF>(s, Ur, Ui, rij, pair_off, perm, ti, tj, N, wself, esel, st); \
        else                                                                                          \
            ui_groups<JV, UI_FULL_WSELF>(s, Ur, Ui, rij, pair_off, perm, ti, tj, N, wself, esel, st); \
        break;
    switch (s.J) {
        UI_CASE(2)
        UI_CASE(4)
        UI_CASE(6)
        UI_CASE(8)
        default: break;
    }
#undef UI_CASE
}

// The per-pair geometry that launch_ui_geo reads, pgeo[5q..5q+4] = (a_r, a_i, b_r, b_i, sfac w_j), for the
// pairs of a list whose length is on the device (np_dev): one lane per pair.
__global__ __launch_bounds__(256) void k_pair_geo(SnapDev s, const double* __restrict__ rij,
                                                  const int* __restrict__ ti, const int* __restrict__ tj,
                                                  const int* __restrict__ np_dev, double* __restrict__ pgeo) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= *np_dev) return;
    const int tjq = tj[q];
    PairGeom g;
    pair_geometry<false>(rij[3 * static_cast<size_t>(q)], rij[3 * static_cast<size_t>(q) + 1],
                         rij[3 * static_cast<size_t>(q) + 2], (s.radelem[ti[q]] + s.radelem[tjq]) * s.rcutfac, s.rmin0,
                         s.rfac0, s.switchflag, s.wjelem[tjq], g);
    double* __restrict__ o = pgeo + 5 * static_cast<size_t>(q);
    o[0] = g.ar;
    o[1] = g.ai;
    o[2] = g.br;
    o[3] = g.bi;
    o[4] = g.sfac;
}

void launch_pair_geo(const SnapDev& s, const double* rij, const int* ti, const int* tj, const int* np_dev, int cap,
                     double* pgeo, hipStream_t st) {
    if (cap <= 0) return;
    hipLaunchKernelGGL(k_pair_geo, dim3((cap + 255) / 256), dim3(256), 0, st, s, rij, ti, tj, np_dev, pgeo);
}

// compute_ui with the per-atom sums in LDS (half layout plus wself, from the pair geometry pgeo).  Lanes
// are pairs: workgroup block b holds the atoms whose first pair lies in [b*UI_PB, (b+1)*UI_PB), so with at
// most UI_MAXP pairs per atom a block has fewer than UI_PB + UI_MAXP <= 512 pairs, one per lane.  Each
// lane runs the row-chain recursion of its pair; every finished half row, scaled by sfac w_j, goes to LDS;
// then thread (ma, atom) adds the row entry of the atom's pairs in pair order and writes U, with wself added
// to the diagonal last (upstream adds it first, cpuSnap.cpp:312-542; the sums agree to rounding, as those of
// the per-atom kernel do).  The workgroups are persistent over the blocks; all lanes share the control flow
// of the recursion.
constexpr int UI_PB = 416, UI_MAXP = 96, UI_NL = 512;

template <int J>
__global__ __launch_bounds__(UI_NL) void k_ui_lds(const double* __restrict__ rootpq, double* __restrict__ Ur,
                                                 double* __restrict__ Ui, const double* __restrict__ pgeo,
                                                 const int* __restrict__ poff, int N, double wself) {
    constexpr int RW = J + 1;
    __shared__ double sr[UI_NL * RW], si[UI_NL * RW];
    __shared__ int s_off[UI_NL + 1];   // pair offsets (relative to the block) of the block's atoms
    const int lane = threadIdx.x;
    const int ptot = poff[N];
    const size_t n = static_cast<size_t>(N);
    for (int b = blockIdx.x; static_cast<long>(b) * UI_PB <= ptot; b += gridDim.x) {
        // atoms a0..a1-1: a0 (a1) is the first atom with poff[a] >= b*UI_PB (>= (b+1)*UI_PB), by binary search
        auto lower = [&](long v) {
            int lo = 0, hi = N;
            while (lo < hi) {
                const int mid = lo + (hi - lo) / 2;
                if (poff[mid] < v) lo = mid + 1;
                else hi = mid;
            }
            return lo;
        };
        const int a0 = lower(static_cast<long>(b) * UI_PB), a1 = lower(static_cast<long>(b + 1) * UI_PB);
        const int na = a1 - a0;
        if (na <= 0) continue;
        const int q0 = poff[a0], np = poff[a1] - q0;   // np < UI_PB + UI_MAXP <= UI_NL
        __syncthreads();   // the previous block's readers are done with s_off and the row stage
        for (int k = lane; k <= na; k += UI_NL) s_off[k] = poff[a0 + k] - q0;
        PairGeom g;   // lanes past the block's pairs run the recursion on a zero pair; nothing reads them
        g.ar = g.ai = g.br = g.bi = 0.0;
        double sf = 0.0;
        if (lane < np) {
            const size_t q = static_cast<size_t>(q0 + lane);
            g.ar = pgeo[5 * q];
            g.ai = pgeo[5 * q + 1];
            g.br = pgeo[5 * q + 2];
            g.bi = pgeo[5 * q + 3];
            sf = pgeo[5 * q + 4];
        }
        row_chain<J, 0, 0, J / 2>(g, rootpq, [&](auto Jc, auto MBc, auto& pr, auto& pim, auto&, auto&) {
            constexpr int L = decltype(Jc)::value, MB = decltype(MBc)::value;
            __syncthreads();   // the previous row's sums have read the stage
#pragma unroll
            for (int ma = 0; ma <= L; ++ma) {
                sr[lane * RW + ma] = sf * pr[ma];
                si[lane * RW + ma] = sf * pim[ma];
            }
            __syncthreads();
            constexpr int row = h_block(L) + (L + 1) * MB;
            for (int t = lane; t < na * (L + 1); t += UI_NL) {
                const int ma = t / na, al = t - ma * na;   // consecutive threads: consecutive atoms
                const int p1 = min(s_off[al + 1], UI_NL);   // the bound holds with <= UI_MAXP pairs per atom
                double ar = 0.0, ai = 0.0;
                for (int p = s_off[al]; p < p1; ++p) {
                    ar += sr[p * RW + ma];
                    ai += si[p * RW + ma];
                }
                const size_t o = static_cast<size_t>(row + ma) * n + a0 + al;
                Ur[o] = ar + (ma == MB ? wself : 0.0);
                Ui[o] = ai;
            }
        });
    }
}

void launch_ui_geo_lds(const SnapDev& s, double* Ur, double* Ui, const double* pgeo, const int* pair_off, int N,
                       double wself, int groups, hipStream_t st) {
    if (N <= 0) return;
    const int G = groups > 0 ? groups : 512;
    switch (s.J) {
        case 2: hipLaunchKernelGGL((k_ui_lds<2>), dim3(G), dim3(UI_NL), 0, st, s.rootpq, Ur, Ui, pgeo, pair_off, N, wself); break;
        case 4: hipLaunchKernelGGL((k_ui_lds<4>), dim3(G), dim3(UI_NL), 0, st, s.rootpq, Ur, Ui, pgeo, pair_off, N, wself); break;
        case 6: hipLaunchKernelGGL((k_ui_lds<6>), dim3(G), dim3(UI_NL), 0, st, s.rootpq, Ur, Ui, pgeo, pair_off, N, wself); break;
        case 8: hipLaunchKernelGGL((k_ui_lds<8>), dim3(G), dim3(UI_NL), 0, st, s.rootpq, Ur, Ui, pgeo, pair_off, N, wself); break;
        default: break;
    }
}

void launch_ui_geo(const SnapDev& s, double* Ur, double* Ui, const double* pgeo, const int* pair_off, int N,
                   double wself, hipStream_t st) {
    if (N <= 0) return;
    switch (s.J) {
        case 2: ui_groups<2, UI_HALF_WSELF, true>(s, Ur, Ui, nullptr, pair_off, nullptr, nullptr, nullptr, N, wself, -1, st, pgeo); break;
        case 4: ui_groups<4, UI_HALF_WSELF, true>(s, Ur, Ui, nullptr, pair_off, nullptr, nullptr, nullptr, N, wself, -1, st, pgeo); break;
        case 6: ui_groups<6, UI_HALF_WSELF, true>(s, Ur, Ui, nullptr, pair_off, nullptr, nullptr, nullptr, N, wself, -1, st, pgeo); break;
        case 8: ui_groups<8, UI_HALF_WSELF, true>(s, Ur, Ui, nullptr, pair_off, nullptr, nullptr, nullptr, N, wself, -1, st, pgeo); break;
        default: break;
    }
}

}  // namespace mdp
