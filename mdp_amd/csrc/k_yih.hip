// compute_yi of the MD path: row order and launch dispatch over the (twojmax, waves) instantiations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "snap_device.hpp"
#include "snap_launch.hpp"

namespace mdp {

#define YH_DECL(JV, WV)                                                                                   \
    void launch_yi_half_##JV##_##WV(const SnapDev&, const double*, const double*, const double*, double*, \
                                    double*, double*, const int*, int, hipStream_t);
YH_DECL(2, 8)
YH_DECL(4, 8)
YH_DECL(6, 8)
YH_DECL(8, 8)
YH_DECL(8, 12)
YH_DECL(8, 4)
YH_DECL(8, 12s)
YH_DECL(8, 12c)
YH_DECL(8, 12y)
YH_DECL(8, 12u1)
#undef YH_DECL

void yi_row_order(int J, int* out) {
    auto cgn = [](int m, int j1, int j2, int j) {
        const int lo = std::max(0, (2 * m - j - j2 + j1) / 2);
        return std::min(j1, (2 * m - j + j2 + j1) / 2) - lo + 1;
    };
    std::vector<std::pair<long, int>> cost;
    int r = 0;
    for (int j = 0; j <= J; ++j)
        for (int mb = 0; 2 * mb <= j; ++mb, ++r) {
            long t = 0;
            for (int j1 = 0; j1 <= J; ++j1)
                for (int j2 = 0; j2 <= j1; ++j2)
                    if (j >= j1 - j2 && j <= j1 + j2 && (j1 + j2 + j) % 2 == 0) {
                        const int nb = cgn(mb, j1, j2, j);
                        for (int ma = 0; ma <= j; ++ma)
                            if (2 * mb < j || ma <= mb) t += static_cast<long>(cgn(ma, j1, j2, j)) * nb;
                        t += 4L * nb * (j1 + j2 + 2);   // row loads
                    }
            cost.emplace_back(-t, r);
        }
    std::stable_sort(cost.begin(), cost.end());
    for (size_t k = 0; k < cost.size(); ++k) out[k] = cost[k].second;
}

int yi_half_variants(int J, int* ids) {
    if (J == 8) {   // waves per workgroup; 112: 12 waves, the larger layer's U rows shared across couplings;
                    // 113: the same with (re, im) adjacent in LDS; 115: 113 with the swap symmetry of the
                    // (j1, j1) couplings; 116: 115 with the row index made wave-uniform explicitly
        ids[0] = 8; ids[1] = 12; ids[2] = 4; ids[3] = 112; ids[4] = 113; ids[5] = 115; ids[6] = 116;
        return 7;
    }
    ids[0] = 8;
    return 1;
}

void launch_yi_half(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_t, double* Yr, double* Yi,
                    double* epart, const int* order, int N, hipStream_t st, int waves) {
    if (N <= 0) return;
    switch (s.J) {
        case 2: launch_yi_half_2_8(s, Ur, Ui, beta_t, Yr, Yi, epart, order, N, st); break;
        case 4: launch_yi_half_4_8(s, Ur, Ui, beta_t, Yr, Yi, epart, order, N, st); break;
        case 6: launch_yi_half_6_8(s, Ur, Ui, beta_t, Yr, Yi, epart, order, N, st); break;
        case 8:
            if (waves == 12)
                launch_yi_half_8_12(s, Ur, Ui, beta_t, Yr, Yi, epart, order, N, st);
            else if (waves == 112)
                launch_yi_half_8_12s(s, Ur, Ui, beta_t, Yr, Yi, epart, order, N, st);
            else if (waves == 113)
                launch_yi_half_8_12c(s, Ur, Ui, beta_t, Yr, Yi, epart, order, N, st);
            else if (waves == 115)
                launch_yi_half_8_12y(s, Ur, Ui, beta_t, Yr, Yi, epart, order, N, st);
            else if (waves == 116)
                launch_yi_half_8_12u1(s, Ur, Ui, beta_t, Yr, Yi, epart, order, N, st);
            else if (waves == 4)
                launch_yi_half_8_4(s, Ur, Ui, beta_t, Yr, Yi, epart, order, N, st);
            else
                launch_yi_half_8_8(s, Ur, Ui, beta_t, Yr, Yi, epart, order, N, st);
            break;
        default: break;
    }
}

}  // namespace mdp
