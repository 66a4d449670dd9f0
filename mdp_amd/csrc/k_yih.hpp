// compute_yi of the MD path on gfx950: half-layout U staged in LDS (kernel templates; the
// instantiations live in k_yih_*.hip so that they compile in parallel).
//
// One workgroup owns 64 consecutive atoms.  It first copies their half-domain U (rows 2mb <= j, 155
// complex elements per atom at twojmax = 8: 2 x 155 x 64 doubles = 158,720 bytes of the 160 KB LDS)
// from HBM with coalesced loads, then its W wavefronts take output rows (j, mb) from a shared counter,
// heaviest first.  Every lane is one atom; the U rows a coupling needs are read from LDS (rows with
// 2mb > j through the mirror symmetry U(mb,ma) = (-1)^(mb+ma) conj U(j-mb, j-ma)), so each U element
// crosses HBM once per step instead of once per output row.  Y (only the entries the force uses) and
// the per-row energy partials sum' U.Y go straight to HBM.
//
// Options (template flags of k_yi_half): SA orders a row's couplings (j1, mb1, j2) so that each U row of
// the larger layer is read from LDS once per output row; CX stores (re, im) of an element next to each
// other in LDS, one 16-byte read per element; SYM takes each unordered pair of U rows of a (j1, j1)
// coupling once (its sum is symmetric under the swap of the two factors).  All three together (the MD
// default at every twojmax) read 13,446 LDS elements and execute 84% of the upstream flops per atom at
// twojmax 8.
#pragma once
#include <hip/hip_runtime.h>

#include "snap_device.hpp"
#include "snap_launch.hpp"

namespace mdp {

// one U row of layer J1 (row mb1, any mb1 in 0..J1) of this lane's atom from the LDS half layout.
// CX: the real and imaginary parts of an element are adjacent (lr holds (re, im) pairs, li is unused), so
// one 16-byte LDS read fetches an element.
template <int J1, bool CX = false>
__device__ __forceinline__ void lds_row(int mb1, const double* __restrict__ lr, const double* __restrict__ li,
                                        int lane, double (&ar)[J1 + 1], double (&ai)[J1 + 1]) {
    const double2* __restrict__ lc = reinterpret_cast<const double2*>(lr);
    if (2 * mb1 <= J1) {
        const int b = (h_block(J1) + (J1 + 1) * mb1) * 64 + lane;
#pragma unroll
        for (int m = 0; m <= J1; ++m) {
            if constexpr (CX) {
                const double2 v = lc[b + m * 64];
                ar[m] = v.x;
                ai[m] = v.y;
            } else {
                ar[m] = lr[b + m * 64];
                ai[m] = li[b + m * 64];
            }
        }
    } else {
        const int b = (h_block(J1) + (J1 + 1) * (J1 - mb1)) * 64 + lane;
        const double sm = (mb1 & 1) ? -1.0 : 1.0;
#pragma unroll
        for (int m = 0; m <= J1; ++m) {
            const double s = (m & 1) ? -sm : sm;
            if constexpr (CX) {
                const double2 v = lc[b + (J1 - m) * 64];
                ar[m] = s * v.x;
                ai[m] = -s * v.y;
            } else {
                ar[m] = s * lr[b + (J1 - m) * 64];
                ai[m] = -s * li[b + (J1 - m) * 64];
            }
        }
    }
}

// The Z contributions of one coupling (J1, J2) -> JO to the output row for one pair of U rows a (layer
// J1) and b (layer J2), scaled by the mb-direction coefficient cb (the terms of z_row_lds).
template <int J1, int J2, int JO>
__device__ __forceinline__ void z_terms(int malim, const double (&ar)[J1 + 1], const double (&ai)[J1 + 1],
                                        const double (&br)[J2 + 1], const double (&bi)[J2 + 1], double cb,
                                        double (&zr)[JO + 1], double (&zi)[JO + 1]) {
    static_for<0, JO + 1>([&](auto MAc) {
        constexpr int MA = decltype(MAc)::value;
        constexpr int LO = cg_lo(MA, J1, J2, JO), HI2 = cg_hi2(MA, J1, J2, JO), NA = cg_n(MA, J1, J2, JO);
        if (MA <= malim) {
            double sr = 0.0, si = 0.0;
            static_for<0, NA>([&](auto IAc) {
                constexpr int IA = decltype(IAc)::value;
                constexpr int M1 = LO + IA, M2 = HI2 - IA;
                constexpr double cga = cg_const(J1, J2, JO, M1, M2);
                sr += cga * (ar[M1] * br[M2] - ai[M1] * bi[M2]);
                si += cga * (ar[M1] * bi[M2] + ai[M1] * br[M2]);
            });
            zr[MA] += cb * sr;
            zi[MA] += cb * si;
        }
    });
}

// The terms of a coupling (J1, J1) -> JO whose two U rows are the same row (mb1 = mb2): the terms (m1, m2)
// and (m2, m1) are equal (the two CG factors change sign together under the swap), so each unordered pair
// is taken once, the off-diagonal ones twice.
template <int J1, int JO>
__device__ __forceinline__ void z_terms_same_row(int malim, const double (&ar)[J1 + 1], const double (&ai)[J1 + 1],
                                                 double cb, double (&zr)[JO + 1], double (&zi)[JO + 1]) {
    static_for<0, JO + 1>([&](auto MAc) {
        constexpr int MA = decltype(MAc)::value;
        constexpr int LO = cg_lo(MA, J1, J1, JO), HI2 = cg_hi2(MA, J1, J1, JO), NA = cg_n(MA, J1, J1, JO);
        if (MA <= malim) {
            double sr = 0.0, si = 0.0;
            static_for<0, NA>([&](auto IAc) {
                constexpr int IA = decltype(IAc)::value;
                constexpr int M1 = LO + IA, M2 = HI2 - IA;
                if constexpr (M1 <= M2) {
                    constexpr double cga = (M1 < M2 ? 2.0 : 1.0) * cg_const(J1, J1, JO, M1, M2);
                    sr += cga * (ar[M1] * ar[M2] - ai[M1] * ai[M2]);
                    si += cga * (ar[M1] * ai[M2] + ai[M1] * ar[M2]);
                }
            });
            zr[MA] += cb * sr;
            zi[MA] += cb * si;
        }
    });
}

// SYM: a coupling (J1, J1) takes each unordered pair of rows {mb1, mb2} once, as y_couplings_shared_a does.
template <int J, int J1, int J2, int JO, bool CX = false, bool SYM = false>
__device__ __forceinline__ void z_row_lds(int mb, int malim, const double* __restrict__ lr,
                                          const double* __restrict__ li, int lane, const double* __restrict__ cg,
                                          double scale, double (&zr)[JO + 1], double (&zi)[JO + 1]) {
    constexpr int CGB = cg_block(J, J1, J2, JO);
    constexpr bool SAME = SYM && (J1 == J2);
    const double* __restrict__ c = cg + CGB;
    const int mb1min = max(0, (2 * mb - JO - J2 + J1) / 2);
    const int mb2max = (2 * mb - JO - (2 * mb1min - J1) + J2) / 2;
    const int nb = min(J1, (2 * mb - JO + J2 + J1) / 2) - mb1min + 1;
    for (int ib = 0; ib < nb; ++ib) {
        const int mb1 = mb1min + ib, mb2 = mb2max - ib;
        if (SAME && mb1 > mb2) break;   // mb1 rises and mb2 falls with ib
        double cb = c[mb1 * (J2 + 1) + mb2] * scale;
        double ar[J1 + 1], ai[J1 + 1];
        lds_row<J1, CX>(mb1, lr, li, lane, ar, ai);
        if (SAME && mb1 == mb2) {
            z_terms_same_row<J1, JO>(malim, ar, ai, cb, zr, zi);
        } else {
            if (SAME) cb *= 2.0;
            double br[J2 + 1], bi[J2 + 1];
            lds_row<J2, CX>(mb2, lr, li, lane, br, bi);
            z_terms<J1, J2, JO>(malim, ar, ai, br, bi, cb, zr, zi);
        }
    }
}


// All couplings of output row (JO, mb) with the U row of the larger layer shared: for each J1 the rows mb1
// of layer J1 are read from LDS once and used by every J2 whose partner row mb2 = mb + (J1+J2-JO)/2 - mb1
// lies in layer J2, instead of once per coupling (30% fewer LDS reads at twojmax 8).  The same terms as
// z_row_lds, summed in the order (J1, mb1, J2) instead of (J1, J2, mb1).
// SYM: a coupling (J1, J1) takes each unordered pair of rows {mb1, mb2} once (the sum is symmetric under
// the swap of its two U factors): mb1 > mb2 is skipped, mb1 < mb2 counts twice, mb1 = mb2 through
// z_terms_same_row -- 17% fewer terms at twojmax 8.
template <int J, int JO, bool CX = false, bool SYM = false, int UNI = 0>
__device__ __forceinline__ void y_couplings_shared_a(int mb, int malim, const double* __restrict__ lr,
                                                     const double* __restrict__ li, int lane,
                                                     const double* __restrict__ cg,
                                                     const double* __restrict__ beta_t, double (&yr)[JO + 1],
                                                     double (&yi)[JO + 1]) {
    static_for<0, J + 1>([&](auto J1c) {
        constexpr int J1 = decltype(J1c)::value;
        int lo = J1 + 1, hi = -1;   // mb1 range over this layer's couplings
        static_for<0, J1 + 1>([&](auto J2c) {
            constexpr int J2 = decltype(J2c)::value;
            if constexpr (coupled(J1, J2, JO)) {
                lo = min(lo, max(0, (2 * mb - JO - J2 + J1) / 2));
                hi = max(hi, min(J1, (2 * mb - JO + J2 + J1) / 2));
            }
        });
        for (int mb1v = lo; mb1v <= hi; ++mb1v) {
            // UNI: the row index made wave-uniform explicitly (scalar branches for the row tests below)
            const int mb1 = UNI ? __builtin_amdgcn_readfirstlane(mb1v) : mb1v;
            double ar[J1 + 1], ai[J1 + 1];
            lds_row<J1, CX>(mb1, lr, li, lane, ar, ai);
            static_for<0, J1 + 1>([&](auto J2c) {
                constexpr int J2 = decltype(J2c)::value;
                if constexpr (coupled(J1, J2, JO)) {
                    const int mb2 = mb + (J1 + J2 - JO) / 2 - mb1;
                    constexpr bool SAME = SYM && (J1 == J2);
                    if (mb2 >= 0 && mb2 <= J2 && (!SAME || mb2 >= mb1)) {
                        constexpr int T = triple_id(J, J1, J2, JO);
                        double cb = cg[cg_block(J, J1, J2, JO) + mb1 * (J2 + 1) + mb2] * beta_t[T];
                        if (SAME && mb2 == mb1) {
                            z_terms_same_row<J1, JO>(malim, ar, ai, cb, yr, yi);
                        } else {
                            if (SAME) cb *= 2.0;
                            double br[J2 + 1], bi[J2 + 1];
                            lds_row<J2, CX>(mb2, lr, li, lane, br, bi);
                            z_terms<J1, J2, JO>(malim, ar, ai, br, bi, cb, yr, yi);
                        }
                    }
                }
            });
        }
    });
}

// PA: per-atom weights beta_t[triple * N + atom] (the learned surrogate's dE/dB) and no energy partials;
// otherwise wave-uniform weights beta_t[triple].  SA: couplings through y_couplings_shared_a; CX: the
// complex-interleaved LDS layout; SYM: the swap symmetry of the (J1, J1) couplings; UNI: 1 makes the
// shared-row loop's row index wave-uniform explicitly (140 instead of 166 VGPRs at twojmax 8, the row tests
// become scalar branches), 2 also the output row's (j, mb), which costs spills.
template <int J, int JO, bool PA = false, bool SA = false, bool CX = false, bool SYM = false, int UNI = 0>
__device__ __forceinline__ void y_row_lds(const double* __restrict__ cg, int row, int mb, int atom, bool live, int N,
                                          const double* __restrict__ lr, const double* __restrict__ li, int lane,
                                          const double* __restrict__ beta_t, double* __restrict__ Yr,
                                          double* __restrict__ Yi, double* __restrict__ epart) {
    const int malim = (2 * mb == JO) ? mb : JO;
    const size_t n = static_cast<size_t>(N);
    double yr[JO + 1], yi[JO + 1];
#pragma unroll
    for (int m = 0; m <= JO; ++m) { yr[m] = 0.0; yi[m] = 0.0; }
    if constexpr (SA && !PA) {
        y_couplings_shared_a<J, JO, CX, SYM, UNI>(mb, malim, lr, li, lane, cg, beta_t, yr, yi);
    } else {
        static_for<0, J + 1>([&](auto J1c) {
            constexpr int J1 = decltype(J1c)::value;
            static_for<0, J1 + 1>([&](auto J2c) {
                constexpr int J2 = decltype(J2c)::value;
                if constexpr (coupled(J1, J2, JO)) {
                    constexpr int T = triple_id(J, J1, J2, JO);
                    const double beta =
                        PA ? (live ? beta_t[static_cast<size_t>(T) * n + atom] : 0.0) : beta_t[T];
                    z_row_lds<J, J1, J2, JO, CX, SYM>(mb, malim, lr, li, lane, cg, beta, yr, yi);
                }
            });
        });
    }
    if (!live) return;
    const int hrow = h_block(JO) + (JO + 1) * mb;
    const size_t o = static_cast<size_t>(hrow) * n + atom;
    double e = 0.0;
#pragma unroll
    for (int ma = 0; ma <= JO; ++ma) {
        if (ma <= malim) {
            Yr[o + ma * n] = yr[ma];
            Yi[o + ma * n] = yi[ma];
            const double w = (2 * mb < JO || ma < mb) ? 1.0 : 0.5;
            if constexpr (CX) {
                const double2 u = reinterpret_cast<const double2*>(lr)[(hrow + ma) * 64 + lane];
                e += w * (u.x * yr[ma] + u.y * yi[ma]);
            } else {
                e += w * (lr[(hrow + ma) * 64 + lane] * yr[ma] + li[(hrow + ma) * 64 + lane] * yi[ma]);
            }
        }
    }
    if constexpr (!PA) epart[static_cast<size_t>(row) * n + atom] = e;
}

// The coefficient tables (cg, beta_t, order) are separate read-only, non-aliased kernel arguments so
// that their wave-uniform loads go through the scalar cache into SGPRs.
template <int J, int W, bool PA = false, bool SA = false, bool CX = false, bool SYM = false, int UNI = 0>
__global__ __launch_bounds__(64 * W) void k_yi_half(const double* __restrict__ cg, const double* __restrict__ Ur,
                                                    const double* __restrict__ Ui, const double* __restrict__ beta_t,
                                                    double* __restrict__ Yr, double* __restrict__ Yi,
                                                    double* __restrict__ epart, const int* __restrict__ order, int N) {
    constexpr int HE = n_half_elems(J);
    constexpr int NR = n_half_rows(J);
    __shared__ double lbuf[2 * HE * 64];   // lr | li, or (re, im) pairs under CX
    double* const lr = lbuf;
    double* const li = CX ? nullptr : lbuf + HE * 64;
    __shared__ int next_row;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int atom = blockIdx.x * 64 + lane;
    const bool live = atom < N;
    const size_t n = static_cast<size_t>(N);
    if (threadIdx.x == 0) next_row = 0;
    for (int e = wave; e < HE; e += W) {
        const double ur = live ? Ur[static_cast<size_t>(e) * n + atom] : 0.0;
        const double ui = live ? Ui[static_cast<size_t>(e) * n + atom] : 0.0;
        if constexpr (CX) {
            reinterpret_cast<double2*>(lbuf)[e * 64 + lane] = make_double2(ur, ui);
        } else {
            lr[e * 64 + lane] = ur;
            li[e * 64 + lane] = ui;
        }
    }
    __syncthreads();
    for (;;) {
        int r = 0;
        if (lane == 0) r = atomicAdd(&next_row, 1);
        // scalar (wave-uniform) row id: every coefficient address below becomes a scalar load
        r = __builtin_amdgcn_readfirstlane(__shfl(r, 0));
        if (r >= NR) break;
        const int row = __builtin_amdgcn_readfirstlane(order[r]);
        const int jo = UNI > 1 ? __builtin_amdgcn_readfirstlane(half_row_j(row)) : half_row_j(row);
        const int mb = UNI > 1 ? __builtin_amdgcn_readfirstlane(half_row_mb(row)) : half_row_mb(row);
        static_for<0, J + 1>([&](auto JOc) {
            constexpr int JO = decltype(JOc)::value;
            if (jo == JO)
                y_row_lds<J, JO, PA, SA, CX, SYM, UNI>(cg, row, mb, atom, live, N, lr, li, lane, beta_t, Yr, Yi,
                                                      epart);
        });
    }
}

// B of every bispectrum triple (j >= j1) from the half U of 64 atoms staged in LDS, as k_yi_half stages
// it: B = 2 sum' U*(j).Z(j1,j2,j) (cpuSnap.cpp:1899-2072 semantics, no bzero subtraction), Z rows from
// z_row_lds, the middle column of even layers weighted 1/2.  Waves take whole triples from a shared
// counter; blist[jjb * N + atom].
template <int J, int W, bool CX = true>
__global__ __launch_bounds__(64 * W) void k_b_half(const double* __restrict__ cg, const double* __restrict__ Ur,
                                                   const double* __restrict__ Ui, double* __restrict__ blist, int N,
                                                   int bnorm) {
    constexpr int HE = n_half_elems(J), NB = n_bisp(J);
    __shared__ double lbuf[2 * HE * 64];   // lr | li, or (re, im) pairs under CX
    double* const lr = lbuf;
    double* const li = CX ? nullptr : lbuf + HE * 64;
    __shared__ int next_task;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int atom = blockIdx.x * 64 + lane;
    const bool live = atom < N;
    const size_t n = static_cast<size_t>(N);
    if (threadIdx.x == 0) next_task = 0;
    for (int e = wave; e < HE; e += W) {
        const double ur = live ? Ur[static_cast<size_t>(e) * n + atom] : 0.0;
        const double ui = live ? Ui[static_cast<size_t>(e) * n + atom] : 0.0;
        if constexpr (CX) {
            reinterpret_cast<double2*>(lbuf)[e * 64 + lane] = make_double2(ur, ui);
        } else {
            lr[e * 64 + lane] = ur;
            li[e * 64 + lane] = ui;
        }
    }
    __syncthreads();
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(&next_task, 1);
        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
        if (t >= NB) break;
        static_for<0, J + 1>([&](auto J1c) {
            constexpr int J1 = decltype(J1c)::value;
            static_for<0, J1 + 1>([&](auto J2c) {
                constexpr int J2 = decltype(J2c)::value;
                static_for<J1, J + 1>([&](auto JOc) {
                    constexpr int JO = decltype(JOc)::value;
                    if constexpr (coupled(J1, J2, JO)) {
                        constexpr int JJB = bisp_index(J, J1, J2, JO);
                        if (t == JJB) {
                            double sum = 0.0;
                            for (int mb = 0; 2 * mb <= JO; ++mb) {
                                const int malim = (2 * mb == JO) ? mb : JO;
                                double zr[JO + 1], zi[JO + 1];
#pragma unroll
                                for (int m = 0; m <= JO; ++m) { zr[m] = 0.0; zi[m] = 0.0; }
                                z_row_lds<J, J1, J2, JO, CX, true>(mb, malim, lr, li, lane, cg,
                                                                   bnorm ? 1.0 / (JO + 1) : 1.0, zr, zi);
                                const int hrow = h_block(JO) + (JO + 1) * mb;
#pragma unroll
                                for (int ma = 0; ma <= JO; ++ma)
                                    if (ma <= malim) {
                                        const double w = (2 * mb < JO || ma < mb) ? 1.0 : 0.5;
                                        if constexpr (CX) {
                                            const double2 u =
                                                reinterpret_cast<const double2*>(lr)[(hrow + ma) * 64 + lane];
                                            sum += w * (u.x * zr[ma] + u.y * zi[ma]);
                                        } else {
                                            sum += w * (lr[(hrow + ma) * 64 + lane] * zr[ma] +
                                                        li[(hrow + ma) * 64 + lane] * zi[ma]);
                                        }
                                    }
                            }
                            if (live) blist[static_cast<size_t>(JJB) * n + atom] = 2.0 * sum;
                        }
                    }
                });
            });
        });
    }
}

}  // namespace mdp
