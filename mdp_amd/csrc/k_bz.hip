// compute_bi and compute_zi for gfx950 (cpuSnap.cpp:721-889 and :1899-1969 semantics).
// Lane = atom; one wavefront per (64 atoms, coupling triple), dispatched to the compile-time
// specialisation of that triple.
#include <hip/hip_runtime.h>

#include "snap_device.hpp"
#include "snap_launch.hpp"
#include "zrow.hpp"

namespace mdp {

// ordinal of (j1, j2, j) among the bispectrum triples (j >= j1), canonical order
__host__ __device__ constexpr int b_id(int J, int j1, int j2, int j) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) {
                if (c < a) continue;
                if (a == j1 && b == j2 && c == j) return n;
                ++n;
            }
    return -1;
}

// blist[jjb*N + atom] = 2 sum' U3*(j) . Z(U1, U2)(j1,j2,j), middle column of even j weighted 1/2
template <int J>
__global__ __launch_bounds__(64) void k_bi(SnapDev s, const double* __restrict__ Ur, const double* __restrict__ Ui,
                                           long off1, long off2, long off3, double* __restrict__ blist, int N,
                                           int bnorm) {
    const int atom = blockIdx.x * 64 + threadIdx.x;
    if (atom >= N) return;
    const int jjb = blockIdx.y;
    const size_t n = static_cast<size_t>(N);
    static_for<0, J + 1>([&](auto J1c) {
        constexpr int J1 = decltype(J1c)::value;
        static_for<0, J1 + 1>([&](auto J2c) {
            constexpr int J2 = decltype(J2c)::value;
            static_for<J1, J + 1>([&](auto JOc) {
                constexpr int JO = decltype(JOc)::value;
                if constexpr (coupled(J1, J2, JO)) {
                    if (jjb == b_id(J, J1, J2, JO)) {
                        double sum = 0.0;
                        for (int mb = 0; 2 * mb <= JO; ++mb) {
                            double zr[JO + 1], zi[JO + 1];
#pragma unroll
                            for (int m = 0; m <= JO; ++m) { zr[m] = 0.0; zi[m] = 0.0; }
                            const int malim = (2 * mb == JO) ? mb : JO;
                            z_row<J, J1, J2, JO>(mb, malim, Ur + off1, Ui + off1, Ur + off2, Ui + off2, n, atom, s.cg,
                                                 bnorm ? 1.0 / (JO + 1) : 1.0, zr, zi);
                            const size_t o = static_cast<size_t>(u_block(JO) + (JO + 1) * mb) * n + atom + off3;
#pragma unroll
                            for (int ma = 0; ma <= JO; ++ma) {
                                if (ma <= malim) {
                                    const double w = (2 * mb < JO || ma < mb) ? 1.0 : 0.5;
                                    sum += w * (Ur[o + ma * n] * zr[ma] + Ui[o + ma * n] * zi[ma]);
                                }
                            }
                        }
                        blist[static_cast<size_t>(jjb) * n + atom] = 2.0 * sum;
                    }
                }
            });
        });
    });
}

void launch_bi(const SnapDev& s, const double* Ur, const double* Ui, long off1, long off2, long off3, double* blist,
               int N, int bnorm, hipStream_t st) {
    if (N <= 0) return;
    dim3 grid((N + 63) / 64, s.idxb_max);
#define BI_CASE(JV) \
    case JV: hipLaunchKernelGGL((k_bi<JV>), grid, dim3(64), 0, st, s, Ur, Ui, off1, off2, off3, blist, N, bnorm); break;
    switch (s.J) {
        BI_CASE(2)
        BI_CASE(4)
        BI_CASE(6)
        BI_CASE(8)
        default: break;
    }
#undef BI_CASE
}

// Z of one coupling triple (blockIdx.y), all rows mb and columns ma, into Z[jjz*N + atom]
template <int J>
__global__ __launch_bounds__(64) void k_zi(SnapDev s, const double* __restrict__ Ur, const double* __restrict__ Ui,
                                           long off1, long off2, double* __restrict__ Zr, double* __restrict__ Zi,
                                           int N, int bnorm) {
    const int atom = blockIdx.x * 64 + threadIdx.x;
    if (atom >= N) return;
    const int t = blockIdx.y;
    const size_t n = static_cast<size_t>(N);
    static_for<0, J + 1>([&](auto J1c) {
        constexpr int J1 = decltype(J1c)::value;
        static_for<0, J1 + 1>([&](auto J2c) {
            constexpr int J2 = decltype(J2c)::value;
            static_for<0, J + 1>([&](auto JOc) {
                constexpr int JO = decltype(JOc)::value;
                if constexpr (coupled(J1, J2, JO)) {
                    if (t == triple_id(J, J1, J2, JO)) {
                        const int zb = s.idxz_block[JO + J2 * (J + 1) + J1 * (J + 1) * (J + 1)];
                        const double scale = bnorm ? 1.0 / (JO + 1) : 1.0;
                        for (int mb = 0; 2 * mb <= JO; ++mb) {
                            double zr[JO + 1], zi[JO + 1];
#pragma unroll
                            for (int m = 0; m <= JO; ++m) { zr[m] = 0.0; zi[m] = 0.0; }
                            z_row<J, J1, J2, JO>(mb, JO, Ur + off1, Ui + off1, Ur + off2, Ui + off2, n, atom, s.cg,
                                                 scale, zr, zi);
                            const size_t o = static_cast<size_t>(zb + (JO + 1) * mb) * n + atom;
#pragma unroll
                            for (int ma = 0; ma <= JO; ++ma) { Zr[o + ma * n] = zr[ma]; Zi[o + ma * n] = zi[ma]; }
                        }
                    }
                }
            });
        });
    });
}

void launch_zi(const SnapDev& s, const double* Ur, const double* Ui, long off1, long off2, double* Zr, double* Zi,
               int N, int bnorm, hipStream_t st) {
    if (N <= 0) return;
    int nt = 0;
    switch (s.J) {
        case 2: nt = n_triples(2); break;
        case 4: nt = n_triples(4); break;
        case 6: nt = n_triples(6); break;
        case 8: nt = n_triples(8); break;
        default: return;
    }
    dim3 grid((N + 63) / 64, nt);
#define ZI_CASE(JV) \
    case JV: hipLaunchKernelGGL((k_zi<JV>), grid, dim3(64), 0, st, s, Ur, Ui, off1, off2, Zr, Zi, N, bnorm); break;
    switch (s.J) {
        ZI_CASE(2)
        ZI_CASE(4)
        ZI_CASE(6)
        ZI_CASE(8)
        default: break;
    }
#undef ZI_CASE
}

}  // namespace mdp
