// Atoms sorted by cell for the cell-list neighbour search (Cell2AtomList, src/C++/Core/cpuNeighborList2D.cpp:
// 298-307: a stable merge sort of the cell ids with the atom indices alongside, then each cell's first
// entry).  Here a stable LSD radix sort of the (cell id, atom index) pairs over the bits that ncell needs
// (rocPRIM), then the first entry of every cell by binary search over the sorted ids.  Work is
// O(natom * bits) however the atoms fall into cells.
//
// mdp_hipCell2AtomListNoAlloc: the same sort for callers that cannot allocate device memory.  Its radix
// sort's temporary storage and a range-check flag live in a file-scope device array of fixed size
// (kScratchBytes, held by every GPU the library runs on until the process ends), whose device address
// is looked up once per device.  Calls that overlap on one GPU share that array and race.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <rocprim/rocprim.hpp>

#include "mdpsnap.h"
#include "runtime.hpp"

using namespace mdp;

namespace {

// first[c] = first p with sorted[p] >= c, for c = 0..ncell (first[ncell] = n)
__global__ __launch_bounds__(256) void k_cell_starts(const int* __restrict__ sorted, int n, int ncell,
                                                     int* __restrict__ first) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncell) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (sorted[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    first[c] = lo;
}

__global__ void k_fill_flag(int* __restrict__ flag) {
    if (threadIdx.x == 0) flag[0] = 0;
}

// flag[0] = 1 when some cell id lies outside [0, ncell)
__global__ __launch_bounds__(256) void k_check_cells(const int* __restrict__ clist, int n, int ncell,
                                                     int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (clist[i] < 0 || clist[i] >= ncell)) flag[0] = 1;
}

int key_bits(int ncell) {
    int b = 1;
    while (b < 31 && (1 << b) < ncell) ++b;
    return b;
}

// rocPRIM's temporary storage for n pairs (no allocation: the query only)
size_t sort_temp_bytes(int n, int bits) {
    size_t bytes = 0;
    const rocprim::counting_iterator<int> idx(0);
    if (rocprim::radix_sort_pairs(nullptr, bytes, static_cast<const int*>(nullptr), static_cast<int*>(nullptr), idx,
                                  static_cast<int*>(nullptr), static_cast<unsigned>(n), 0, bits) != hipSuccess)
        throw Error("Cell2AtomList: radix sort size query failed");
    return bytes;
}

// c2anum = cell ids in ascending order, c2alist = the atoms alongside (ascending within a cell: the sort is
// stable and the atoms enter in index order), c2anumsum[c] = first entry of cell c (natom at ncell)
void sort_cells(int* c2alist, int* c2anumsum, int* c2anum, const int* clist, int natom, int ncell, void* temp,
                size_t temp_bytes, hipStream_t st) {
    const int bits = key_bits(ncell);
    if (natom > 0) {
        const rocprim::counting_iterator<int> idx(0);
        MDP_HIP(rocprim::radix_sort_pairs(temp, temp_bytes, clist, c2anum, idx, c2alist, static_cast<unsigned>(natom),
                                          0, bits, st));
    }
    hipLaunchKernelGGL(k_cell_starts, dim3((ncell + 1 + 255) / 256), dim3(256), 0, st, c2anum, natom, ncell,
                       c2anumsum);
}

// ---- the allocation-free entry point's scratch ------------------------------------------------------

constexpr size_t kScratchBytes = size_t(64) << 20;   // 64 MiB: the radix sort of up to ~4M atoms
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
void* g_scratch_ptr[kMaxDevices] = {};               // device address of `scratch` on each device
std::mutex g_scratch_mutex;

// the device address of this device's `scratch` (looked up once per device, with the device current)
unsigned char* scratch_ptr() {
    int dev = 0;
    MDP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error("Cell2AtomListNoAlloc: device index beyond the supported range");
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(std::string("Cell2AtomListNoAlloc: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return static_cast<unsigned char*>(g_scratch_ptr[dev]);
}

template <class F>
void guarded_call(F&& f) {
    std::lock_guard<std::mutex> lk(library_mutex());
    clear_error();
    try {
        f();
    } catch (const std::exception& e) {
        set_error(e.what());
    }
}

}  // namespace

namespace mdp {

void launch_cell2atom_sorted(int* c2alist, int* c2anumsum, int* c2anum, const int* clist, int natom, int ncell,
                             void* (*temp_alloc)(size_t), hipStream_t st) {
    if (ncell <= 0) return;
    const size_t bytes = natom > 0 ? sort_temp_bytes(natom, key_bits(ncell)) : 0;
    void* temp = natom > 0 ? temp_alloc(bytes) : nullptr;
    sort_cells(c2alist, c2anumsum, c2anum, clist, natom, ncell, temp, bytes, st);
}

}  // namespace mdp

extern "C" {

// Cell2AtomList without device allocation (no upstream counterpart; see include/mdpsnap.h)
void mdp_hipCell2AtomListNoAlloc(int* c2alist, int* c2anumsum, int* c2anum, int* clist, int* dtemp, int natom,
                                 int ncell) {
    (void)dtemp;
    guarded_call([&] {
        if (ncell <= 0 || natom < 0) throw Error("Cell2AtomListNoAlloc: need ncell > 0 and natom >= 0");
        const int bits = key_bits(ncell);
        const size_t need = natom > 0 ? sort_temp_bytes(natom, bits) : 0;
        // [0, 256): the range-check flag; the sort's storage after it
        if (need + 256 > kScratchBytes)
            throw Error("Cell2AtomListNoAlloc: natom = " + std::to_string(natom) +
                        " needs more scratch than the fixed " + std::to_string(kScratchBytes >> 20) + " MiB");
        unsigned char* s = scratch_ptr();
        int* flag = reinterpret_cast<int*>(s);
        hipStream_t st = library_stream();
        hipLaunchKernelGGL(k_fill_flag, dim3(1), dim3(64), 0, st, flag);
        if (natom > 0)
            hipLaunchKernelGGL(k_check_cells, dim3((natom + 255) / 256), dim3(256), 0, st, clist, natom, ncell, flag);
        sort_cells(c2alist, c2anumsum, c2anum, clist, natom, ncell, s + 256, need, st);
        int bad = 0;
        MDP_HIP(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, st));
        MDP_HIP(hipStreamSynchronize(st));
        MDP_HIP(hipGetLastError());
        if (bad) throw Error("Cell2AtomListNoAlloc: a cell id is outside [0, ncell)");
    });
}

}  // extern "C"
