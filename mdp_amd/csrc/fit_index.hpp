// Index spaces of the batched fitting kernels: Z blocks in the idxz ordering and the bispectrum
// triples in the idxb ordering (cpuSnap.cpp:112-171), as compile-time functions of twojmax.
#pragma once
#include "snap_device.hpp"

namespace mdp {

// offset of Z(j1,j2,j) in the idxz ordering (cpuSnap.cpp:147-171): triples in canonical order, each
// holding rows mb = 0..j/2 of j+1 columns
__host__ __device__ constexpr int z_block(int J, int j1, int j2, int j) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) {
                if (a == j1 && b == j2 && c == j) return n;
                n += (c / 2 + 1) * (c + 1);
            }
    return -1;
}
__host__ __device__ constexpr int n_z(int J) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) n += (c / 2 + 1) * (c + 1);
    return n;
}
// bispectrum triples (j >= j1) in idxb order (cpuSnap.cpp:112-120): component k of triple t
// (0: j1, 1: j2, 2: j) and their number
__host__ __device__ constexpr int b_comp(int J, int t, int k) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) {
                if (c < a) continue;
                if (n == t) return k == 0 ? a : (k == 1 ? b : c);
                ++n;
            }
    return -1;
}
__host__ __device__ constexpr int n_b(int J) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) n += (c >= a);
    return n;
}
__host__ __device__ constexpr int b_index(int J, int j1, int j2, int j) {
    for (int k = 0; k < n_b(J); ++k)
        if (b_comp(J, k, 0) == j1 && b_comp(J, k, 1) == j2 && b_comp(J, k, 2) == j) return k;
    return -1;
}
// The three terms of dB_t/dr (cpuSnap.cpp:2101-2240): term 0 contracts dU of layer j with Z(j1,j2,j),
// term 1 dU of layer j1 with Z(j,j2,j1), term 2 dU of layer j2 with Z(j,j1,j2).
__host__ __device__ constexpr int term_layer(int J, int t, int w) { return b_comp(J, t, w == 0 ? 2 : (w == 1 ? 0 : 1)); }
__host__ __device__ constexpr int term_zblock(int J, int t, int w) {
    return w == 0 ? z_block(J, b_comp(J, t, 0), b_comp(J, t, 1), b_comp(J, t, 2))
                  : (w == 1 ? z_block(J, b_comp(J, t, 2), b_comp(J, t, 1), b_comp(J, t, 0))
                            : z_block(J, b_comp(J, t, 2), b_comp(J, t, 0), b_comp(J, t, 1)));
}

}  // namespace mdp
