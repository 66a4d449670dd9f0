// Clebsch-Gordan product of two U layers for one output row, shared by compute_yi, compute_bi and
// compute_zi.
#pragma once
#include "snap_device.hpp"

namespace mdp {

// z[ma] += scale * Z(J1,J2,JO)(ma, mb) for ma = 0..min(JO, malim) of this lane's atom.
//   Z(ma,mb) = sum_{mb1} cg(mb1,mb2) sum_{ma1} cg(ma1,ma2) U_J1(mb1,ma1) U_J2(mb2,ma2)
// (cpuSnap.cpp:1899-1969).  For each row pair (mb1, mb2) both U rows are loaded once into registers
// and reused by every ma; the CG coefficients sit at compile-time offsets of the uniform table
// pointer, so they come in through scalar loads.  U element jju of the atom lives at jju*n + atom.
template <int J, int J1, int J2, int JO>
__device__ __forceinline__ void z_row(int mb, int malim, const double* __restrict__ U1r, const double* __restrict__ U1i,
                                      const double* __restrict__ U2r, const double* __restrict__ U2i, size_t n, int atom,
                                      const double* __restrict__ cg, double scale, double (&zr)[JO + 1],
                                      double (&zi)[JO + 1]) {
    constexpr int CGB = cg_block(J, J1, J2, JO);
    const double* __restrict__ c = cg + CGB;
    const int mb1min = max(0, (2 * mb - JO - J2 + J1) / 2);
    const int mb2max = (2 * mb - JO - (2 * mb1min - J1) + J2) / 2;
    const int nb = min(J1, (2 * mb - JO + J2 + J1) / 2) - mb1min + 1;
    for (int ib = 0; ib < nb; ++ib) {
        const int mb1 = mb1min + ib, mb2 = mb2max - ib;
        const double cb = c[mb1 * (J2 + 1) + mb2] * scale;
        const size_t o1 = static_cast<size_t>(u_block(J1) + (J1 + 1) * mb1) * n + atom;
        const size_t o2 = static_cast<size_t>(u_block(J2) + (J2 + 1) * mb2) * n + atom;
        double ar[J1 + 1], ai[J1 + 1], br[J2 + 1], bi[J2 + 1];
#pragma unroll
        for (int m = 0; m <= J1; ++m) { ar[m] = U1r[o1 + m * n]; ai[m] = U1i[o1 + m * n]; }
#pragma unroll
        for (int m = 0; m <= J2; ++m) { br[m] = U2r[o2 + m * n]; bi[m] = U2i[o2 + m * n]; }
        static_for<0, JO + 1>([&](auto MAc) {
            constexpr int MA = decltype(MAc)::value;
            constexpr int LO = cg_lo(MA, J1, J2, JO), HI2 = cg_hi2(MA, J1, J2, JO), NA = cg_n(MA, J1, J2, JO);
            if (MA <= malim) {
                double sr = 0.0, si = 0.0;
                static_for<0, NA>([&](auto IAc) {
                    constexpr int IA = decltype(IAc)::value;
                    constexpr int M1 = LO + IA, M2 = HI2 - IA;
                    const double cga = c[M1 * (J2 + 1) + M2];
                    sr += cga * (ar[M1] * br[M2] - ai[M1] * bi[M2]);
                    si += cga * (ar[M1] * bi[M2] + ai[M1] * br[M2]);
                });
                zr[MA] += cb * sr;
                zi[MA] += cb * si;
            }
        });
    }
}

}  // namespace mdp
