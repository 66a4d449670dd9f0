// compute_yi of the drop-in for several element blocks (cpuSnap.cpp:891-989 with chemflag = 1).
//
// Y(j)(ma,mb) = sum over the couplings (j1,j2)->j of beta_eff(j1,j2,j) Z(j1,j2,j)(ma,mb).
// One wavefront = 64 consecutive atoms (lane = atom) x one output row (j, mb) of the half domain
// (blockIdx.y); each launch adds the contribution of one (element1, element2) -> element3 block for
// the atoms of one centre element, with that combination's coefficient table.
#include <hip/hip_runtime.h>

#include "yrow.hpp"

namespace mdp {

// element-block variant: Y[off3] += sum beta Z(U[off1], U[off2]) for atoms of element isel; PA: one weight
// per (coupling triple, atom), beta_t[triple*N + atom]
template <int J, bool PA>
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
            y_row<J, JO, true, false, true, PA>(s, row, mb, atom, N, Ur + off1, Ui + off1, Ur + off2, Ui + off2, beta_t,
                                           Yr + off3, Yi + off3, nullptr);
    });
}

void launch_yi_acc(const SnapDev& s, const double* Ur, const double* Ui, long off1, long off2, long off3,
                   const double* beta_t, double* Yr, double* Yi, const int* atom_elem, int isel, int N, hipStream_t st,
                   bool per_atom) {
    if (N <= 0) return;
#define YA_CASE(JV)                                                                                           \
    case JV:                                                                                                  \
        if (per_atom)                                                                                         \
            hipLaunchKernelGGL((k_yi_acc<JV, true>), dim3((N + 63) / 64, n_half_rows(JV)), dim3(64), 0, st, s, Ur, \
                               Ui, off1, off2, off3, beta_t, Yr, Yi, atom_elem, isel, N);                     \
        else                                                                                                  \
            hipLaunchKernelGGL((k_yi_acc<JV, false>), dim3((N + 63) / 64, n_half_rows(JV)), dim3(64), 0, st, s, Ur, \
                               Ui, off1, off2, off3, beta_t, Yr, Yi, atom_elem, isel, N);                     \
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
