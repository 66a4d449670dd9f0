// Launchers of the batched fitting-descriptor kernels (k_fit.hip).  All pointers are device pointers.
// The batch concatenates the atoms of C configurations (config c owns atoms cfg_a0[c]..cfg_a0[c+1]-1,
// periodic cell a = (Lx, 0, 0), b = (bx, Ly, 0), c = (cx, cy, Lz) with cfg_L[3c..3c+2] = (Lx, Ly, Lz) and
// cfg_T[3c..3c+2] = (bx, cx, cy), cfg_T null for orthogonal boxes); pairs are one CSR list over all atoms.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "snap_launch.hpp"

namespace mdp {

// Pair build by per-configuration brute force over periodic images.  fill = false writes the pair
// count of every atom to pcnt; fill = true writes rij/pi/pj/pti/ptj at the CSR offsets poff (N+1
// entries); when poff[N] exceeds cap it leaves the outputs unwritten, sets bit 0 of flags[0] and
// stores poff[N] in flags[1].  Each atom's candidates are taken in FIT_PAIR_SPLIT slices by as many
// workgroups; pcnt holds (FIT_PAIR_SPLIT + 1) * N ints (the totals, then the slice counts that the fill
// pass reads).  chunk_order ((N+63)/64 entries, a permutation) gives the launch order of the 64-atom
// chunks.
constexpr int FIT_PAIR_SPLIT = 8;
void launch_fit_pairs(bool fill, const double* x, const int* type, const int* cfg_of, const int* cfg_a0,
                      const double* cfg_L, const double* cfg_T, const double* rc2, int ntypes, int N,
                      const int* chunk_order, int* pcnt,
                      const int* poff, long cap, double* rij, int* pi, int* pj, int* pti, int* ptj, int* flags,
                      hipStream_t st);
// after an overflow flagged by the fill pass: poff[0..N] = 0 (no pairs for the remaining stages)
void launch_fit_guard(const int* flags, int* poff, int N, hipStream_t st);
// cfg_p0[c] = poff[cfg_a0[c]] for c = 0..C (first pair of each configuration, total at C)
void launch_fit_cfg_p0(const int* poff, const int* cfg_a0, int* cfg_p0, int C, hipStream_t st);
// p[0..n) = 0 by a kernel (safe inside captured graphs)
void launch_fill_zero_int(int* p, int n, hipStream_t st);
// Z of every coupling triple (atom-major, complex-interleaved, zt[(g*idxz_max + jjz)*2 + {0,1}]) and B
// (per-configuration layout blist[a0_c*idxb_max + ii + N_c*jjb]) from the full-layout Utot.
void launch_fit_zb(const SnapDev& s, const double* Ur, const double* Ui, double* zt, double* blist,
                   const double* bzero, const int* cfg_of, const int* cfg_a0, int N, int bnorm, int bzeroflag,
                   hipStream_t st);
// Z + B with the half U (rows 2mb <= j, [e][N] layout) of 64 atoms staged in LDS per workgroup; row_order
// = the half rows (j, mb) by decreasing work (yi_row_order).  Compiled for twojmax 4, 6, 8 (returns
// false otherwise); same outputs as launch_fit_zb.  bpart: scratch of fit_zb_partials(J) * N doubles (the
// per-row shares of B).  waves: 8 (two per SIMD, 256 registers each) or 4 (one per SIMD, 512 registers).
bool launch_fit_zb_lds(const SnapDev& s, const double* Uhr, const double* Uhi, double* zt, double* blist,
                       double* bpart, const double* bzero, const int* cfg_of, const int* cfg_a0,
                       const int* row_order, int N, int bnorm, int bzeroflag, hipStream_t st, int waves = 8);
int fit_zb_partials(int J);
// Device copy of the dB/dr term table (see fit_term_table)
struct FitTerms {
    const int* lo = nullptr;      // [J+2]
    const int* b = nullptr;       // bispectrum index of each term
    const int* z = nullptr;       // Z block offset (idxz ordering)
    const double* w = nullptr;    // weight
    const double* wtab = nullptr; // weights by (bispectrum index, term): wtab[3*t + term]
    // the same terms split into two halves of the bispectrum index range (b < hb and b >= hb, of about
    // equal contraction work), each grouped by layer: half h's terms of layer l are h_lo[h*(J+2) + l] ..
    const int* h_lo = nullptr;
    const int* h_b = nullptr;
    const int* h_z = nullptr;
    const double* h_w = nullptr;
    int hb = 0;
    // the rows of the matrix-core variant (fit_mfma_tables), null where twojmax has no such kernel
    const int* m_z = nullptr;
    const double* m_w = nullptr;
    const int* m_comb = nullptr;
};
// Host term table split into two halves of the bispectrum index range (FitTerms::h_*): returns hb.
int fit_term_table_halves(int J, int bnorm, std::vector<int>& lo, std::vector<int>& b, std::vector<int>& z,
                          std::vector<double>& w);
// Host term table of the dB/dr contraction, grouped by the layer whose dU it contracts.
void fit_term_table(int J, int bnorm, std::vector<int>& lo, std::vector<int>& b, std::vector<int>& z,
                    std::vector<double>& w);
// dB/dr of every pair (per-configuration layout dblist[p0_c*3*idxb_max + lp + P_c*k + 3*P_c*jjb]) by a
// persistent kernel of `groups` workgroups; counter[0] (and counter[1] for variant 4) must be 0 at
// launch.  variant 0 walks the layer-grouped term table; variant 1 (twojmax 4, 6, 8) has every term
// unrolled at compile time; variant 2 (twojmax 6, 8) is variant 0 limited to 256 registers (two waves
// per SIMD); variant 3 is variant 0 with the term loop pipelined; variant 4 forms dU/dr through the
// SU(2) generators with the bispectrum indices split over two single-wave workgroups; variant 5 is
// variant 4 with both halves in one two-wave workgroup that stages the atom's Z block in LDS.
void launch_fit_dbdr(const SnapDev& s, const double* zt, const FitTerms& tt, const double* rij, const int* pti,
                     const int* ptj, const int* poff, const int* cfg_of, const int* cfg_p0, double* dblist,
                     int* counter, int N, int groups, hipStream_t st, int variant = 0);

// dB/dr on the FP64 matrix cores (k_fit_mfma.hip): twojmax 4, 6 and 8.  fit_mfma_tables writes the
// packed term rows (per tile T, layer l and tile row m: Z block offset mz and weight mw at
// (T*(J+1) + l)*16 + m; mw has one more entry, 0.0) and, for every bispectrum index t, its tile rows
// comb[3t..3t+2] (-1 padded);
// launch_fit_dbdr_mfma runs the persistent kernel (counter[0] must be 0) with the same output layout as
// launch_fit_dbdr.
bool fit_mfma_supported(int J);
bool fit_mfma_tables(int J, int bnorm, std::vector<int>& mz, std::vector<double>& mw, std::vector<int>& comb);
bool launch_fit_dbdr_mfma(const SnapDev& s, const double* zt, const int* mz, const double* mw, const int* comb,
                          const double* rij, const int* pti, const int* ptj, const int* poff, const int* cfg_of,
                          const int* cfg_p0, double* dblist, int* counter, int N, int groups, hipStream_t st);

// ---- linear fit over the batched descriptors (k_fitlsq.hip) ----
// one chunk of rows of a weighted least-squares group: a(r, k) = A[abase + r*rs + k*cs], target
// T[tbase + r], weight W[wbase + r / wdiv], r < nrows <= 64
struct LsqChunk {
    long abase, tbase;
    int rs, cs, nrows, wbase, wdiv, pad;
};
// incoming pairs of every atom, ascending pair index: in_idx[in_off[a] .. in_off[a+1]) (in_cnt, cursor:
// N ints of scratch; scan_tmp: scan_tmp_ints(N) ints)
void launch_lsq_incoming(const int* pj, long P, int N, int* in_cnt, int* in_off, int* cursor, int* in_idx,
                         int* scan_tmp, hipStream_t st);
// per configuration c: bi[c*Kt + col], bd at 3*a0_c*Kt ([comp + 3 ii + 3 N_c col]), bv[c*6*Kt + 6 col + m]
// (upstream tally layouts, Kt = ntypes*IB, col = (type-1)*IB + k) and the stress rows as[(6c + m')*Kt + col]
void launch_lsq_tallies(const double* blist, const double* dblist, const double* rij, const int* poff,
                        const int* in_off, const int* in_idx, const int* pti, const int* type, const int* cfg_of,
                        const int* cfg_a0, const int* cfg_p0, int C, int N, int IB, int ntypes, int max_natoms,
                        double* bi, double* bd, double* bv, double* as, hipStream_t st);
// out[0 .. K(K+1)/2) = upper triangle of sum w^2 a a^T (row-major), out[K(K+1)/2 + k] = sum w^2 a_k b;
// ekl[e] = (k, l) of output e (l = -1 for h); part: G * (K(K+1)/2 + K) doubles
size_t lsq_gram_lds(int K);
void launch_lsq_gram(const LsqChunk* chunks, int nchunks, const double* A, const double* T, const double* W, int K,
                     const int2* ekl, double* part, int G, double* out, hipStream_t st);

}  // namespace mdp
