// compute_yi of the MD path, twojmax = 2, 4, 6 with 8 waves per workgroup, the larger layer's U rows shared
// across couplings, (re, im) adjacent in LDS and the swap symmetry of the (j1, j1) couplings (see k_yih.hpp),
// and the learned surrogate's twojmax 2 / 4 Y (per-atom weights) and B kernels (likewise, without the shared
// rows).
#include "k_yih.hpp"

namespace mdp {

void launch_yi_half_2_8(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_t, double* Yr,
                         double* Yi, double* epart, const int* order, int N, hipStream_t st) {
    hipLaunchKernelGGL((k_yi_half<2, 8, false, true, true, true, 1>), dim3((N + 63) / 64), dim3(64 * 8), 0, st, s.cg, Ur, Ui, beta_t, Yr, Yi,
                       epart, order, N);
}

void launch_yi_half_4_8(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_t, double* Yr,
                         double* Yi, double* epart, const int* order, int N, hipStream_t st) {
    hipLaunchKernelGGL((k_yi_half<4, 8, false, true, true, true, 1>), dim3((N + 63) / 64), dim3(64 * 8), 0, st, s.cg, Ur, Ui, beta_t, Yr, Yi,
                       epart, order, N);
}

void launch_yi_half_6_8(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_t, double* Yr,
                         double* Yi, double* epart, const int* order, int N, hipStream_t st) {
    hipLaunchKernelGGL((k_yi_half<6, 8, false, true, true, true, 1>), dim3((N + 63) / 64), dim3(64 * 8), 0, st, s.cg, Ur, Ui, beta_t, Yr, Yi,
                       epart, order, N);
}

// the learned surrogate at twojmax 2 and 4: Y with per-atom weights and B, from LDS-staged half U
bool launch_yi_half_pa(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_pa, double* Yr,
                       double* Yi, const int* order, int N, hipStream_t st) {
    if (N <= 0) return true;
    const dim3 grid((N + 63) / 64), block(64 * 4);
    switch (s.J) {
        case 2: hipLaunchKernelGGL((k_yi_half<2, 4, true, false, true, true>), grid, block, 0, st, s.cg, Ur, Ui, beta_pa, Yr, Yi, nullptr, order, N); return true;
        case 4: hipLaunchKernelGGL((k_yi_half<4, 4, true, false, true, true>), grid, block, 0, st, s.cg, Ur, Ui, beta_pa, Yr, Yi, nullptr, order, N); return true;
        default: return false;
    }
}

bool launch_b_half(const SnapDev& s, const double* Ur, const double* Ui, double* blist, int N, int bnorm,
                   hipStream_t st) {
    if (N <= 0) return true;
    const dim3 grid((N + 63) / 64), block(64 * 4);
    switch (s.J) {
        case 2: hipLaunchKernelGGL((k_b_half<2, 4>), grid, block, 0, st, s.cg, Ur, Ui, blist, N, bnorm); return true;
        case 4: hipLaunchKernelGGL((k_b_half<4, 4>), grid, block, 0, st, s.cg, Ur, Ui, blist, N, bnorm); return true;
        default: return false;
    }
}

}  // namespace mdp
