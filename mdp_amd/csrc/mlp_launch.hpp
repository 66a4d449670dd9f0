// Launchers of the surrogate MLP kernels (k_mlp.hip).  All pointers are device pointers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

namespace mdp {

constexpr int MLP_NIN = 16;       // descriptor inputs (padded; the model uses nin <= 16 of them)
constexpr int MLP_NH = 64;        // hidden width of both layers
constexpr int MLP_NPARAM = 5313;  // W1[16][64] b1[64] W2[64][64] b2[64] w3[64] b3
constexpr int MLP_NPACK = 10240;  // BF16 operand tiles (40 x 64 lanes x 4)

// FP32 parameters -> swizzled BF16 MFMA operand tiles
void launch_mlp_pack(const float* p, unsigned short* q, hipStream_t st);
// Per-atom energy e[n] and (g != nullptr) its gradient g[k*n + a] with respect to the descriptors
// X[k*n + a]; aff = mu[16] | 1/sigma[16] | e_scale | e_shift (doubles)
void launch_mlp_eval(const unsigned short* q, const float* p, const double* aff, int nin, const double* X, int n,
                     double* e, double* g, hipStream_t st);
// grad[MLP_NPARAM] = d/dparams sum_a (alpha_a E_a + U_a . g_a); part holds nwaves * MLP_NPARAM floats
void launch_mlp_train(const unsigned short* q, const float* p, const double* aff, int nin, const double* X,
                      const double* U, const double* alpha, int n, float* part, int nwaves, float* grad,
                      hipStream_t st);
// one Adam step (step number t >= 1), decoupled weight decay wd on the weight matrices
void launch_mlp_adam(float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps, int t,
                     float wd, hipStream_t st);

}  // namespace mdp
