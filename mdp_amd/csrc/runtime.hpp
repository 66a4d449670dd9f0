// Host runtime shared by the drop-in entry points and the MD engine: error reporting, the library
// stream, a grow-only device buffer, and the device copy of the SNAP tables of one twojmax.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "snap_launch.hpp"
#include "snap_tables.hpp"

namespace mdp {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void set_error(const std::string& msg);
void clear_error();
const char* get_error();

#define MDP_HIP(call)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw ::mdp::Error(std::string(#call) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + ":" + \
                               std::to_string(__LINE__) + ")");                                         \
    } while (0)

// grow-only device allocation
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void* ensure(size_t bytes) {
        if (bytes == 0) bytes = 8;
        if (bytes > cap) {
            release();
            MDP_HIP(hipMalloc(&p, bytes));
            cap = bytes;
        }
        return p;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Device copy of the SNAP tables for one parameter set.
struct DevTables {
    SnapIndex idx;
    DevBuf rootpq, cg, radelem, wjelem, map, idxz, idxb, idxu_block, idxz_block, idxb_block, idxcg_block;
    SnapDev view;   // pointers into the buffers above
    void upload(const SnapIndex& ix, const double* rootpq_h, const double* cg_h, const double* radelem_h,
                const double* wjelem_h, const int* map_h, int ntypes, hipStream_t st);
};

hipStream_t library_stream();
std::mutex& library_mutex();

// Effective weight of each coupling triple in Y for element combination (e1,e2,e3) of an atom of
// element ielem (cpuSnap.cpp:959-981), with the 1/(j+1) of bnorm folded in; indexed by triple ordinal.
std::vector<double> y_beta_table(const SnapIndex& ix, const double* beta, int e1, int e2, int e3, int nelem,
                                 int ielem, int ncoeffall, int bnorm);

// The same weights for one element as (bispectrum index, factor) per coupling triple: the weight of
// triple t is fac[t] * beta[jjb[t]] (multiplicity, the (j1+1)/(j+1) of unnormalised B, or bnorm's 1/(j+1)).
void y_triple_weights(const SnapIndex& ix, int bnorm, std::vector<int>& jjb, std::vector<double>& fac);

}  // namespace mdp
