// compute_yi for gfx950 (cpuSnap.cpp:891-989 semantics) with the energy partials of the MD path.
//
// Y(j)(ma,mb) = sum over the couplings (j1,j2)->j of beta_eff(j1,j2,j) Z(j1,j2,j)(ma,mb).
// One wavefront = 64 consecutive atoms (lane = atom) x one output row (j, mb) of the half domain
// (blockIdx.y).  The row's j+1 complex accumulators stay in registers for the whole sum, every
// output element is written exactly once (no atomics), and the inner products run out of registers
// (see z_row).  With epart, the wave also emits sum_ma w(ma) (U . Y) of its row, from which the
// per-atom energy follows by the Euler identity of the cubic bispectrum energy:
//   E_i = beta0 - sum_k beta_k bzero_k + (2/3) sum' U_i . Y_i.
#include <hip/hip_runtime.h>

#include "snap_device.hpp"
#include "snap_launch.hpp"
#include "zrow.hpp"

namespace mdp {

template <int J, int JO, bool FULLMID, bool WITH_E, bool ACCUM>
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
                const double beta = beta_t[triple_id(J, J1, J2, JO)];
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

template <int J, bool FULLMID, bool WITH_E>
__global__ __launch_bounds__(64) void k_yi(SnapDev s, const double* __restrict__ Ur, const double* __restrict__ Ui,
                                           const double* __restrict__ beta_t, double* __restrict__ Yr,
                                           double* __restrict__ Yi, double* __restrict__ epart, int N) {
    const int atom = blockIdx.x * 64 + threadIdx.x;
    if (atom >= N) return;
    const int row = blockIdx.y;
    const int jo = half_row_j(row), mb = half_row_mb(row);
    static_for<0, J + 1>([&](auto JOc) {
        constexpr int JO = decltype(JOc)::value;
        if (jo == JO) y_row<J, JO, FULLMID, WITH_E, false>(s, row, mb, atom, N, Ur, Ui, Ur, Ui, beta_t, Yr, Yi, epart);
    });
}

// element-block variant: Y[off3] += sum beta Z(U[off1], U[off2]) for atoms of element isel
template <int J>
__global__ __launch_bounds__(64) void k_yi_acc(SnapDev s, const double* __restrict__ Ur, const double* __restrict__ Ui,
                                               long off1, long off2, long off3, const double* __restrict__ beta_t,
                                               double* __restrict__ Yr, double* __restrict__ Yi,
                                               const int* __restrict__ atom_elem, int isel, int N) {
    const int atom = blockIdx.x * 64 + threadIdx.x;
    if (atom >= N) return;
    if (isel >= 0 && atom_elem[atom] != isel) return;
    const int row = blockIdx.y;
    const int jo = half_row_j(row), mb = half_row_mb(row);
    static_for<0, J + 1>([&](auto JOc) {
        constexpr int JO = decltype(JOc)::value;
        if (jo == JO)
            y_row<J, JO, true, false, true>(s, row, mb, atom, N, Ur + off1, Ui + off1, Ur + off2, Ui + off2, beta_t,
                                           Yr + off3, Yi + off3, nullptr);
    });
}

void launch_yi(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_t, double* Yr, double* Yi,
               double* epart, int N, bool full_mid, hipStream_t st) {
    if (N <= 0) return;
#define YI_CASE(JV)                                                                                             \
    case JV: {                                                                                                  \
        dim3 grid((N + 63) / 64, n_half_rows(JV));                                                              \
        if (full_mid)                                                                                           \
            hipLaunchKernelGGL((k_yi<JV, true, false>), grid, dim3(64), 0, st, s, Ur, Ui, beta_t, Yr, Yi, epart, N); \
        else if (epart)                                                                                         \
            hipLaunchKernelGGL((k_yi<JV, false, true>), grid, dim3(64), 0, st, s, Ur, Ui, beta_t, Yr, Yi, epart, N); \
        else                                                                                                    \
            hipLaunchKernelGGL((k_yi<JV, false, false>), grid, dim3(64), 0, st, s, Ur, Ui, beta_t, Yr, Yi, epart, N); \
        break;                                                                                                  \
    }
    switch (s.J) {
        YI_CASE(2)
        YI_CASE(4)
        YI_CASE(6)
        YI_CASE(8)
        default: break;
    }
#undef YI_CASE
}

void launch_yi_acc(const SnapDev& s, const double* Ur, const double* Ui, long off1, long off2, long off3,
                   const double* beta_t, double* Yr, double* Yi, const int* atom_elem, int isel, int N, hipStream_t st) {
    if (N <= 0) return;
#define YA_CASE(JV)                                                                                           \
    case JV:                                                                                                  \
        hipLaunchKernelGGL((k_yi_acc<JV>), dim3((N + 63) / 64, n_half_rows(JV)), dim3(64), 0, st, s, Ur, Ui, off1, \
                           off2, off3, beta_t, Yr, Yi, atom_elem, isel, N);                                   \
        break;
    switch (s.J) {
        YA_CASE(2)
        YA_CASE(4)
        YA_CASE(6)
        YA_CASE(8)
        default: break;
    }
#undef YA_CASE
}

}  // namespace mdp
