// Launchers of the surrogate training kernels (k_sur.hip).  All pointers are device pointers.
#pragma once
#include <hip/hip_runtime.h>

namespace mdp {

// s[c][p] = sum_k g[k][pi[p]] dbdr[k][c][p]
void launch_sur_pair_s(const double* g, const double* dbdr, const int* pi, int K, int n, long P, double* s,
                       hipStream_t st);
// F[a] = sum of s over the pairs of a minus sum over the pairs pointing at a; R = F - fref (either
// output may be null)
void launch_sur_atom_f(const double* s, const int* own_off, const int* in_off, const int* in_idx, long P, int n,
                       const double* fref, double* F, double* R, hipStream_t st);
// U[k][a] = coef sum_{p of a} (R_a - R_pj) . dbdr[k][:][p]   (K <= 16)
void launch_sur_u(const double* R, const double* dbdr, const int* own_off, const int* pj, int K, int n, long P,
                  double coef, double* U, hipStream_t st);
// Ec[c] = sum of e over configuration c; alpha (if not null) = coef 2 (Ec - eref) / N_c^2 on its atoms
void launch_sur_energy(const double* e, const int* cfg_off, const double* eref, int C, double coef, double* Ec,
                       double* alpha, hipStream_t st);
// part[b] = partial sums of v^2 over nblocks blocks
void launch_sumsq(const double* v, long n, double* part, int nblocks, hipStream_t st);

}  // namespace mdp
