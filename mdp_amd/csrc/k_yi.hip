// compute_yi of the drop-in for gfx950 (cpuSnap.cpp:891-989 semantics).
//
// Y(j)(ma,mb) = sum over the couplings (j1,j2)->j of beta_eff(j1,j2,j) Z(j1,j2,j)(ma,mb).
// One wavefront = 64 consecutive atoms (lane = atom) x one output row (j, mb) of the half domain
// (blockIdx.y).  The row's j+1 complex accumulators stay in registers for the whole sum, every
// output element is written exactly once (no atomics), and the inner products run out of registers
// (see z_row).  U is read in the caller's full layout, exactly as upstream reads it.
#include <hip/hip_runtime.h>

#include "yrow.hpp"

namespace mdp {

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

// Per-atom weights (beta_t[triple*N + atom]), every idxz entry written: the Y of the learned surrogate.
template <int J>
__global__ __launch_bounds__(64) void k_yi_pa(SnapDev s, const double* __restrict__ Ur, const double* __restrict__ Ui,
                                              const double* __restrict__ beta_t, double* __restrict__ Yr,
                                              double* __restrict__ Yi, int N) {
    const int atom = blockIdx.x * 64 + threadIdx.x;
    if (atom >= N) return;
    const int row = blockIdx.y;
    const int jo = half_row_j(row), mb = half_row_mb(row);
    static_for<0, J + 1>([&](auto JOc) {
        constexpr int JO = decltype(JOc)::value;
        if (jo == JO)
            y_row<J, JO, true, false, false, true>(s, row, mb, atom, N, Ur, Ui, Ur, Ui, beta_t, Yr, Yi, nullptr);
    });
}

__global__ __launch_bounds__(256) void k_beta_peratom(const double* __restrict__ g, const int* __restrict__ jjb,
                                                      const double* __restrict__ fac, int ntrip, int N,
                                                      double* __restrict__ beta) {
    const long k = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= static_cast<long>(ntrip) * N) return;
    const int t = static_cast<int>(k / N), a = static_cast<int>(k % N);
    beta[k] = fac[t] * g[static_cast<size_t>(jjb[t]) * N + a];
}

void launch_beta_peratom(const double* g, const int* jjb, const double* fac, int ntrip, int N, double* beta_t,
                         hipStream_t st) {
    const long n = static_cast<long>(ntrip) * N;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_beta_peratom, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, g, jjb, fac,
                       ntrip, N, beta_t);
}

bool launch_yi_peratom(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_t, double* Yr,
                       double* Yi, int N, hipStream_t st) {
    if (N <= 0) return true;
#define YI_PA_CASE(JV)                                                                                            \
    case JV:                                                                                                      \
        hipLaunchKernelGGL((k_yi_pa<JV>), dim3((N + 63) / 64, n_half_rows(JV)), dim3(64), 0, st, s, Ur, Ui, beta_t, \
                           Yr, Yi, N);                                                                            \
        return true;
    switch (s.J) {
        YI_PA_CASE(2)
        YI_PA_CASE(4)
        YI_PA_CASE(6)
        YI_PA_CASE(8)
        default: return false;
    }
#undef YI_PA_CASE
}

void launch_yi(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_t, double* Yr, double* Yi,
               double* epart, int N, bool full_mid, hipStream_t st) {
    (void)epart;
    (void)full_mid;   // the drop-in produces every idxz entry (upstream output); the MD path uses k_yi_half
    if (N <= 0) return;
#define YI_CASE(JV)                                                                                               \
    case JV:                                                                                                      \
        hipLaunchKernelGGL((k_yi<JV, true, false>), dim3((N + 63) / 64, n_half_rows(JV)), dim3(64), 0, st, s, Ur, Ui, \
                           beta_t, Yr, Yi, epart, N);                                                             \
        break;
    switch (s.J) {
        YI_CASE(2)
        YI_CASE(4)
        YI_CASE(6)
        YI_CASE(8)
        default: break;
    }
#undef YI_CASE
}

}  // namespace mdp
