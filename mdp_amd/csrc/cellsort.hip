// Atoms sorted by cell for the cell-list neighbour search (Cell2AtomList, src/C++/Core/cpuNeighborList2D.cpp:
// 298-307: a stable merge sort of the cell ids with the atom indices alongside, then each cell's first
// entry).  Here a stable LSD radix sort of the (cell id, atom index) pairs over the bits that ncell needs
// (rocPRIM), then the first entry of every cell by binary search over the sorted ids.  Work is
// O(natom * bits) however the atoms fall into cells.
#include <hip/hip_runtime.h>

#include <cstring>
#include <rocprim/rocprim.hpp>

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
