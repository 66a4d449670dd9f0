// One output row (j, mb) of Y for this lane's atom from a full-layout U: the building block of the
// drop-in compute_yi kernels (k_yi.hip, k_yia.hip).  With WITH_E the wave also emits
// sum_ma w(ma) (U . Y) of its row, from which the per-atom energy follows by the Euler identity of
// the cubic bispectrum energy: E_i = beta0 - sum_k beta_k bzero_k + (2/3) sum' U_i . Y_i.
#pragma once
#include "snap_device.hpp"
#include "snap_launch.hpp"
#include "zrow.hpp"

namespace mdp {

// PER_ATOM: beta_t holds one weight per (coupling triple, atom), beta_t[triple*N + atom], instead of one
// per triple (the learned surrogate's dE/dB differs from atom to atom).
template <int J, int JO, bool FULLMID, bool WITH_E, bool ACCUM, bool PER_ATOM = false>
__device__ __forceinline__ void y_row(const SnapDev& s, int row, int mb, int atom, int N, const double* __restrict__ U1r,
                                      const double* __restrict__ U1i, const double* __restrict__ U2r,
                                      const double* __restrict__ U2i, const double* __restrict__ beta_t,
                                      double* __restrict__ Yr, double* __restrict__ Yi, double* __restrict__ epart) {
    const size_t n = static_cast<size_t>(N);
    const int malim = (!FULLMID && 2 * mb == JO) ? mb : JO;
    double yr[JO + 1], yi[JO + 1];
#pragma unroll
    for (int m = 0; m <= JO; ++m) { yr[m] = 0.0; yi[m] = 0.0; }
    static_for<0, J + 1>([&](auto J1c) {
        constexpr int J1 = decltype(J1c)::value;
        static_for<0, J1 + 1>([&](auto J2c) {
            constexpr int J2 = decltype(J2c)::value;
            if constexpr (coupled(J1, J2, JO)) {
                const double beta = PER_ATOM ? beta_t[static_cast<size_t>(triple_id(J, J1, J2, JO)) * n + atom]
                                             : beta_t[triple_id(J, J1, J2, JO)];
                z_row<J, J1, J2, JO>(mb, malim, U1r, U1i, U2r, U2i, n, atom, s.cg, beta, yr, yi);
            }
        });
    });
    const size_t o = static_cast<size_t>(u_block(JO) + (JO + 1) * mb) * n + atom;
    double e = 0.0;
#pragma unroll
    for (int ma = 0; ma <= JO; ++ma) {
        if (ma <= malim) {
            if constexpr (ACCUM) {
                Yr[o + ma * n] += yr[ma];
                Yi[o + ma * n] += yi[ma];
            } else {
                Yr[o + ma * n] = yr[ma];
                Yi[o + ma * n] = yi[ma];
            }
            if constexpr (WITH_E) {
                const double w = (2 * mb < JO || ma < mb) ? 1.0 : (ma == mb ? 0.5 : 0.0);
                e += w * (U1r[o + ma * n] * yr[ma] + U1i[o + ma * n] * yi[ma]);
            }
        }
    }
    if constexpr (WITH_E) epart[static_cast<size_t>(row) * n + atom] = e;
}

}  // namespace mdp
