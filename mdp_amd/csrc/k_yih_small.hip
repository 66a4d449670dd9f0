// compute_yi of the MD path, twojmax = 2, 4, 6 with 8 waves per workgroup (see k_yih.hpp).
#include "k_yih.hpp"

namespace mdp {

void launch_yi_half_2_8(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_t, double* Yr,
                         double* Yi, double* epart, const int* order, int N, hipStream_t st) {
    hipLaunchKernelGGL((k_yi_half<2, 8>), dim3((N + 63) / 64), dim3(64 * 8), 0, st, s.cg, Ur, Ui, beta_t, Yr, Yi,
                       epart, order, N);
}

void launch_yi_half_4_8(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_t, double* Yr,
                         double* Yi, double* epart, const int* order, int N, hipStream_t st) {
    hipLaunchKernelGGL((k_yi_half<4, 8>), dim3((N + 63) / 64), dim3(64 * 8), 0, st, s.cg, Ur, Ui, beta_t, Yr, Yi,
                       epart, order, N);
}

void launch_yi_half_6_8(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_t, double* Yr,
                         double* Yi, double* epart, const int* order, int N, hipStream_t st) {
    hipLaunchKernelGGL((k_yi_half<6, 8>), dim3((N + 63) / 64), dim3(64 * 8), 0, st, s.cg, Ur, Ui, beta_t, Yr, Yi,
                       epart, order, N);
}

}  // namespace mdp
