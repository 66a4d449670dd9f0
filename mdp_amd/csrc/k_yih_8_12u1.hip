// compute_yi of the MD path, twojmax = 8 with 12 waves per workgroup, the larger layer's U rows shared across couplings, complex-interleaved LDS, the swap symmetry of the (j1, j1) couplings, the row index made wave-uniform explicitly (see k_yih.hpp).
#include "k_yih.hpp"

namespace mdp {

void launch_yi_half_8_12u1(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_t, double* Yr,
                          double* Yi, double* epart, const int* order, int N, hipStream_t st) {
    hipLaunchKernelGGL((k_yi_half<8, 12, false, true, true, true, 1>), dim3((N + 63) / 64), dim3(64 * 12), 0, st, s.cg, Ur, Ui, beta_t, Yr, Yi,
                       epart, order, N);
}

}  // namespace mdp
