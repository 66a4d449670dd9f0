// C ABI of the learned-surrogate kernels (mdp_mlp_* and mdp_sur_* in include/mdpsnap.h): thin,
// allocation-free wrappers over the launchers; every pointer is a device pointer owned by the caller
// and every call is asynchronous on the given HIP stream (NULL = the library stream).
#include <hip/hip_runtime.h>

#include <string>

#include "mdpsnap.h"
#include "mlp_launch.hpp"
#include "runtime.hpp"
#include "sur_launch.hpp"

using namespace mdp;

namespace {
template <class F>
int api(const char* what, F&& f) {
    clear_error();
    try {
        f();
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    }
}
hipStream_t S(void* s) { return s ? static_cast<hipStream_t>(s) : library_stream(); }
}  // namespace

extern "C" {

// Device buffers for callers without a GPU array library of their own (every copy is ordered with the
// kernels of the library stream and waits for them).
int mdp_dev_malloc(void** p, size_t bytes) {
    return api("mdp_dev_malloc", [&] {
        if (!p) throw Error("mdp_dev_malloc: null pointer");
        *p = nullptr;
        MDP_HIP(hipMalloc(p, bytes ? bytes : 8));
    });
}

int mdp_dev_free(void* p) {
    return api("mdp_dev_free", [&] {
        if (p) {
            MDP_HIP(hipStreamSynchronize(library_stream()));
            MDP_HIP(hipFree(p));
        }
    });
}

int mdp_dev_h2d(void* dst, const void* src, size_t bytes) {
    return api("mdp_dev_h2d", [&] {
        hipStream_t st = library_stream();
        MDP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
        MDP_HIP(hipStreamSynchronize(st));
    });
}

int mdp_dev_d2h(void* dst, const void* src, size_t bytes) {
    return api("mdp_dev_d2h", [&] {
        hipStream_t st = library_stream();
        MDP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        MDP_HIP(hipStreamSynchronize(st));
    });
}

int mdp_dev_zero(void* p, size_t bytes) {
    return api("mdp_dev_zero", [&] { MDP_HIP(hipMemsetAsync(p, 0, bytes, library_stream())); });
}

int mdp_dev_sync(void) {
    return api("mdp_dev_sync", [&] { MDP_HIP(hipStreamSynchronize(library_stream())); });
}

int mdp_mlp_sizes(int* nparam, int* npack, int* nin_max, int* nhidden) {
    return api("mdp_mlp_sizes", [&] {
        if (nparam) *nparam = MLP_NPARAM;
        if (npack) *npack = MLP_NPACK;
        if (nin_max) *nin_max = MLP_NIN;
        if (nhidden) *nhidden = MLP_NH;
    });
}

int mdp_mlp_pack(const float* params, unsigned short* packed, void* stream) {
    return api("mdp_mlp_pack", [&] { launch_mlp_pack(params, packed, S(stream)); });
}

int mdp_mlp_eval(const unsigned short* packed, const float* params, const double* aff, int nin, const double* x, int n,
                 double* e, double* g, void* stream) {
    return api("mdp_mlp_eval", [&] {
        if (nin < 1 || nin > MLP_NIN) throw Error("mdp_mlp_eval: nin must be 1..16");
        launch_mlp_eval(packed, params, aff, nin, x, n, e, g, S(stream));
    });
}

int mdp_mlp_train(const unsigned short* packed, const float* params, const double* aff, int nin, const double* x,
                  const double* u, const double* alpha, int n, float* partials, int nwaves, float* grad, void* stream) {
    return api("mdp_mlp_train", [&] {
        if (nin < 1 || nin > MLP_NIN) throw Error("mdp_mlp_train: nin must be 1..16");
        if (nwaves < 1) throw Error("mdp_mlp_train: nwaves must be positive");
        launch_mlp_train(packed, params, aff, nin, x, u, alpha, n, partials, nwaves, grad, S(stream));
    });
}

int mdp_mlp_adam(float* params, const float* grad, float* m, float* v, float lr, float beta1, float beta2, float eps,
                 int step, float weight_decay, void* stream) {
    return api("mdp_mlp_adam", [&] {
        if (step < 1) throw Error("mdp_mlp_adam: step counts from 1");
        launch_mlp_adam(params, grad, m, v, lr, beta1, beta2, eps, step, weight_decay, S(stream));
    });
}

int mdp_sur_pair_s(const double* g, const double* dbdr, const int* pi, int K, int n, long P, double* s, void* stream) {
    return api("mdp_sur_pair_s", [&] { launch_sur_pair_s(g, dbdr, pi, K, n, P, s, S(stream)); });
}

int mdp_sur_atom_f(const double* s, const int* own_off, const int* in_off, const int* in_idx, long P, int n,
                   const double* fref, double* F, double* R, void* stream) {
    return api("mdp_sur_atom_f", [&] {
        if (R && !fref) throw Error("mdp_sur_atom_f: a residual needs reference forces");
        launch_sur_atom_f(s, own_off, in_off, in_idx, P, n, fref, F, R, S(stream));
    });
}

int mdp_sur_u(const double* R, const double* dbdr, const int* own_off, const int* pj, int K, int n, long P, double coef,
              double* U, void* stream) {
    return api("mdp_sur_u", [&] {
        if (K < 1 || K > 16) throw Error("mdp_sur_u: K must be 1..16");
        launch_sur_u(R, dbdr, own_off, pj, K, n, P, coef, U, S(stream));
    });
}

int mdp_sur_energy(const double* e, const int* cfg_off, const double* eref, int C, double coef, double* Ec,
                   double* alpha, void* stream) {
    return api("mdp_sur_energy", [&] { launch_sur_energy(e, cfg_off, eref, C, coef, Ec, alpha, S(stream)); });
}

int mdp_sumsq(const double* v, long n, double* partials, int nblocks, void* stream) {
    return api("mdp_sumsq", [&] {
        if (nblocks < 1) throw Error("mdp_sumsq: nblocks must be positive");
        launch_sumsq(v, n, partials, nblocks, S(stream));
    });
}

}  // extern "C"
