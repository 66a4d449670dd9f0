// Z + B of the batched fitting path with the half-domain U of 64 atoms staged in LDS (the structure of
// the MD path's compute_yi, k_yih.hpp): each workgroup copies its atoms' half U (158,720 B at
// twojmax = 8) from HBM once, then its waves take whole coupling triples, heaviest first, from a shared
// counter.  A triple's rows mb <= j/2 are produced with every U row read from LDS (mirrored rows through
// the symmetry of U), written atom-major to zt, and for bispectrum triples (j >= j1) contracted with
// U(j) into that row's share of B, which a second kernel sums per atom in a fixed order
// (cpuSnap.cpp:1899-2072 semantics, as k_fit_zb).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "fit_launch.hpp"
#include "k_yih.hpp"

namespace mdp {

namespace {

__host__ __device__ constexpr int zb_z_block(int J, int j1, int j2, int j) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) {
                if (a == j1 && b == j2 && c == j) return n;
                n += (c / 2 + 1) * (c + 1);
            }
    return -1;
}
__host__ __device__ constexpr int zb_n_z(int J) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) n += (c / 2 + 1) * (c + 1);
    return n;
}
__host__ __device__ constexpr int zb_b_index(int J, int j1, int j2, int j) {
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
__host__ __device__ constexpr int zb_n_b(int J) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) n += (c >= a);
    return n;
}
// partial sums of B, one per half row (mb = 0..j/2) of each bispectrum triple in idxb order: the first
// slot of (j1, j2, j), and the number of slots
__host__ __device__ constexpr int zb_bp_first(int J, int j1, int j2, int j) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) {
                if (c < a) continue;
                if (a == j1 && b == j2 && c == j) return n;
                n += c / 2 + 1;
            }
    return -1;
}
__host__ __device__ constexpr int zb_n_bp(int J) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) n += (c >= a) ? c / 2 + 1 : 0;
    return n;
}

}  // namespace

// Z rows (JO, mb) of every coupling (J1, J2) -> JO for this lane's atom, from the LDS half U, stored
// atom-major.  The middle row of an even layer stops at ma = mb: the dB/dr contraction and B read no
// entry beyond it.  For bispectrum triples (JO >= J1) the row's share of B, sum' U*(JO).Z over its
// entries (the middle entry of an even layer halved; cpuSnap.cpp:1971-2072), goes to its own slot of
// bpart ([slot][N]), summed per atom in fixed order by k_fit_bsum.
template <int J, int JO, bool CX>
__device__ __forceinline__ void z_rows_lds(int mb, int atom, bool live, const double* __restrict__ lr,
                                           const double* __restrict__ li, int lane, const double* __restrict__ cg,
                                           double scale, double* __restrict__ zt, double* __restrict__ bpart,
                                           size_t n) {
    constexpr int IZ = zb_n_z(J);
    const int malim = (2 * mb == JO) ? mb : JO;
    static_for<0, J + 1>([&](auto J1c) {
        constexpr int J1 = decltype(J1c)::value;
        static_for<0, J1 + 1>([&](auto J2c) {
            constexpr int J2 = decltype(J2c)::value;
            if constexpr (coupled(J1, J2, JO)) {
                constexpr int ZB = zb_z_block(J, J1, J2, JO);
                double zr[JO + 1], zi[JO + 1];
#pragma unroll
                for (int m = 0; m <= JO; ++m) { zr[m] = 0.0; zi[m] = 0.0; }
                z_row_lds<J, J1, J2, JO, CX, true>(mb, malim, lr, li, lane, cg, scale, zr, zi);
                if (live) {
                    // the empty asm keeps each coupling's store address here: hoisted out of the row loop, the
                    // addresses of all couplings took a VGPR pair each (256 VGPRs and 126 spilled at twojmax 8)
                    size_t aoff = static_cast<size_t>(atom) * IZ;
                    asm volatile("" : "+v"(aoff));
                    double* __restrict__ zo = zt + (aoff + ZB + (JO + 1) * mb) * 2;
#pragma unroll
                    for (int ma = 0; ma <= JO; ++ma)
                        if (ma <= malim) {
                            zo[2 * ma] = zr[ma];
                            zo[2 * ma + 1] = zi[ma];
                        }
                }
                if constexpr (JO >= J1) {
                    constexpr int HB = h_block(JO), BP = zb_bp_first(J, J1, J2, JO);
                    double sb = 0.0;
#pragma unroll
                    for (int ma = 0; ma <= JO; ++ma)
                        if (ma <= malim) {
                            const double w = (2 * mb < JO || ma < mb) ? 1.0 : 0.5;
                            const int e = HB + (JO + 1) * mb + ma;
                            if constexpr (CX) {
                                const double2 u = reinterpret_cast<const double2*>(lr)[e * 64 + lane];
                                sb += w * (u.x * zr[ma] + u.y * zi[ma]);
                            } else {
                                sb += w * (lr[e * 64 + lane] * zr[ma] + li[e * 64 + lane] * zi[ma]);
                            }
                        }
                    if (live) {
                        size_t bo = static_cast<size_t>(BP + mb) * n + atom;
                        asm volatile("" : "+v"(bo));
                        bpart[bo] = sb;
                    }
                }
                // couplings are independent: keep the scheduler from interleaving them (register pressure)
                __builtin_amdgcn_sched_barrier(0);
            }
        });
    });
}

// (re, im) of each U element adjacent in LDS: one 16-byte read per element
template <int J, int W, bool CX = true>
__global__ __launch_bounds__(64 * W) void k_fit_z_lds(const double* __restrict__ cg, const double* __restrict__ Ur,
                                                      const double* __restrict__ Ui, double* __restrict__ zt,
                                                      double* __restrict__ bpart, const int* __restrict__ order, int N,
                                                      int bnorm) {
    constexpr int HE = n_half_elems(J), NR = n_half_rows(J);
    __shared__ double lbuf[2 * HE * 64];   // lr | li, or (re, im) pairs under CX
    double* const lr = lbuf;
    double* const li = CX ? nullptr : lbuf + HE * 64;
    __shared__ int next_row;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int atom = blockIdx.x * 64 + lane;
    const bool live = atom < N;
    const size_t n = static_cast<size_t>(N);
    if (threadIdx.x == 0) next_row = 0;
    for (int e = wave; e < HE; e += W) {
        const double ur = live ? Ur[static_cast<size_t>(e) * n + atom] : 0.0;
        const double ui = live ? Ui[static_cast<size_t>(e) * n + atom] : 0.0;
        if constexpr (CX) {
            reinterpret_cast<double2*>(lbuf)[e * 64 + lane] = make_double2(ur, ui);
        } else {
            lr[e * 64 + lane] = ur;
            li[e * 64 + lane] = ui;
        }
    }
    __syncthreads();
    for (;;) {
        int r = 0;
        if (lane == 0) r = atomicAdd(&next_row, 1);
        r = __builtin_amdgcn_readfirstlane(__shfl(r, 0));
        if (r >= NR) break;
        const int row = __builtin_amdgcn_readfirstlane(order[r]);
        const int jo = half_row_j(row), mb = half_row_mb(row);
        static_for<0, J + 1>([&](auto JOc) {
            constexpr int JO = decltype(JOc)::value;
            if (jo == JO)
                z_rows_lds<J, JO, CX>(mb, atom, live, lr, li, lane, cg, bnorm ? 1.0 / (JO + 1) : 1.0, zt, bpart,
                                      n);
        });
    }
}

// B = 2 sum_mb (row shares in bpart) - bzero[j], rows in ascending mb; lane = atom
template <int J>
__global__ __launch_bounds__(256) void k_fit_bsum(const double* __restrict__ bpart, double* __restrict__ blist,
                                                  const double* __restrict__ bzero, const int* __restrict__ cfg_of,
                                                  const int* __restrict__ cfg_a0, int N, int bzeroflag) {
    constexpr int IB = zb_n_b(J);
    const int atom = blockIdx.x * blockDim.x + threadIdx.x;
    if (atom >= N) return;
    const size_t n = static_cast<size_t>(N);
    const int c = cfg_of[atom];
    const int a0 = cfg_a0[c], nc = cfg_a0[c + 1] - a0;
    double* __restrict__ bo = blist + static_cast<size_t>(a0) * IB + (atom - a0);
    static_for<0, J + 1>([&](auto J1c) {
        constexpr int J1 = decltype(J1c)::value;
        static_for<0, J1 + 1>([&](auto J2c) {
            constexpr int J2 = decltype(J2c)::value;
            static_for<J1, J + 1>([&](auto JOc) {
                constexpr int JO = decltype(JOc)::value;
                if constexpr (coupled(J1, J2, JO)) {
                    constexpr int BP = zb_bp_first(J, J1, J2, JO), jjb = zb_b_index(J, J1, J2, JO);
                    double sum = 0.0;
#pragma unroll
                    for (int mb = 0; 2 * mb <= JO; ++mb) sum += bpart[static_cast<size_t>(BP + mb) * n + atom];
                    double b = 2.0 * sum;
                    if (bzeroflag) b -= bzero[JO];
                    bo[static_cast<size_t>(nc) * jjb] = b;
                }
            });
        });
    });
}

int fit_zb_partials(int J) { return (J == 4 || J == 6 || J == 8) ? zb_n_bp(J) : 0; }

bool launch_fit_zb_lds(const SnapDev& s, const double* Uhr, const double* Uhi, double* zt, double* blist,
                       double* bpart, const double* bzero, const int* cfg_of, const int* cfg_a0,
                       const int* row_order, int N, int bnorm, int bzeroflag, hipStream_t st, int waves) {
    if (N <= 0) return true;
    const dim3 grid((N + 63) / 64), gb((N + 255) / 256), bb(256);
    switch (s.J) {
#define ZB_CASE(JV)                                                                                            \
    case JV:                                                                                                   \
        if (waves == 4)                                                                                        \
            hipLaunchKernelGGL((k_fit_z_lds<JV, 4>), grid, dim3(256), 0, st, s.cg, Uhr, Uhi, zt, bpart, row_order, \
                               N, bnorm);                                                                      \
        else if (waves == 12)                                                                                  \
            hipLaunchKernelGGL((k_fit_z_lds<JV, 12>), grid, dim3(768), 0, st, s.cg, Uhr, Uhi, zt, bpart,        \
                               row_order, N, bnorm);                                                           \
        else                                                                                                   \
            hipLaunchKernelGGL((k_fit_z_lds<JV, 8>), grid, dim3(512), 0, st, s.cg, Uhr, Uhi, zt, bpart, row_order, \
                               N, bnorm);                                                                      \
        hipLaunchKernelGGL((k_fit_bsum<JV>), gb, bb, 0, st, bpart, blist, bzero, cfg_of, cfg_a0, N, bzeroflag);   \
        return true;
        ZB_CASE(4)
        ZB_CASE(6)
        ZB_CASE(8)
#undef ZB_CASE
        default: return false;
    }
}

}  // namespace mdp
