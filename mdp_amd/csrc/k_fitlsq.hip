// Linear-fit tallies and normal equations over the batched descriptors.
//
// From the per-configuration B and dB/dr of the batch, the kernels below form what the upstream
// fitting driver consumes: per configuration the type-resolved descriptor sum bi
// (cpuSnapTallyBispectrum, cpuSnap.cpp:2265-2291), the per-atom derivative tally bd
// (cpuSnapTallyBispectrumDeriv, :2293-2365) and the virial tally bv (cpuSnapTallyBispectrumVirial,
// :2367-2470), and then the weighted normal equations L = sum_r w_r^2 a_r a_r^T, h = sum_r w_r^2 a_r b_r
// of the energy, force and stress rows (linearfit.jl:168-200, rows as in emldescriptors.jl:80-140).
// Every sum runs in a fixed order, so the results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include "fit_launch.hpp"
#include "runtime.hpp"
#include "snap_launch.hpp"

namespace mdp {

namespace {
inline int nblocks(long n, int b) { return static_cast<int>((n + b - 1) / b); }
}  // namespace

// ---- incoming pairs of every atom (pairs whose neighbour is that atom), ascending pair index ----

__global__ void k_lsq_zero(int* __restrict__ a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ void k_lsq_in_count(const int* __restrict__ pj, long P, int* __restrict__ in_cnt) {
    const long p = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < P) atomicAdd(in_cnt + pj[p], 1);
}

__global__ void k_lsq_in_fill(const int* __restrict__ pj, long P, const int* __restrict__ in_off,
                              int* __restrict__ cursor, int* __restrict__ in_idx) {
    const long p = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int a = pj[p];
    in_idx[in_off[a] + atomicAdd(cursor + a, 1)] = static_cast<int>(p);
}

// the fill order depends on scheduling: sort each atom's short list
__global__ void k_lsq_in_sort(const int* __restrict__ in_off, int* __restrict__ in_idx, int N) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    const int lo = in_off[a], hi = in_off[a + 1];
    for (int i = lo + 1; i < hi; ++i) {
        const int v = in_idx[i];
        int k = i - 1;
        while (k >= lo && in_idx[k] > v) {
            in_idx[k + 1] = in_idx[k];
            --k;
        }
        in_idx[k + 1] = v;
    }
}

void launch_lsq_incoming(const int* pj, long P, int N, int* in_cnt, int* in_off, int* cursor, int* in_idx,
                         int* scan_tmp, hipStream_t st) {
    if (N <= 0) return;
    hipLaunchKernelGGL(k_lsq_zero, dim3(nblocks(N, 256)), dim3(256), 0, st, in_cnt, N);
    hipLaunchKernelGGL(k_lsq_zero, dim3(nblocks(N, 256)), dim3(256), 0, st, cursor, N);
    if (P > 0) hipLaunchKernelGGL(k_lsq_in_count, dim3(nblocks(P, 256)), dim3(256), 0, st, pj, P, in_cnt);
    exclusive_scan(in_cnt, in_off, N, scan_tmp, st);
    if (P > 0) hipLaunchKernelGGL(k_lsq_in_fill, dim3(nblocks(P, 256)), dim3(256), 0, st, pj, P, in_off, cursor, in_idx);
    hipLaunchKernelGGL(k_lsq_in_sort, dim3(nblocks(N, 128)), dim3(128), 0, st, in_off, in_idx, N);
}

// ---- tallies ------------------------------------------------------------------------------------

// bd of configuration c at bd + 3*a0_c*Kt, element (comp + 3*ii, col) at [comp + 3 ii + 3 N_c col],
// col = (t-1)*IB + k with t the type of the pair's owner: + dB/dr of the atom's own pairs, - dB/dr of the
// pairs whose neighbour it is.  Thread = (atom, k), atoms fastest.
__global__ void k_lsq_bd(const double* __restrict__ dblist, const int* __restrict__ poff,
                         const int* __restrict__ in_off, const int* __restrict__ in_idx, const int* __restrict__ pti,
                         const int* __restrict__ type, const int* __restrict__ cfg_of, const int* __restrict__ cfg_a0,
                         const int* __restrict__ cfg_p0, int N, int IB, int ntypes, double* __restrict__ bd) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long>(N) * IB) return;
    const int g = static_cast<int>(t % N), k = static_cast<int>(t / N);
    const int c = cfg_of[g], a0 = cfg_a0[c], nc = cfg_a0[c + 1] - a0, p0 = cfg_p0[c];
    const long pc = cfg_p0[c + 1] - p0;
    const double* __restrict__ db = dblist + static_cast<size_t>(p0) * 3 * IB;
    const int kt = ntypes * IB, tg = type[g];
    double* __restrict__ out = bd + 3 * static_cast<size_t>(a0) * kt + 3 * (g - a0);
    const int q0 = poff[g], q1 = poff[g + 1], i0 = in_off[g], i1 = in_off[g + 1];
    for (int ty = 1; ty <= ntypes; ++ty) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        if (tg == ty)
            for (int q = q0; q < q1; ++q) {
                const size_t e = static_cast<size_t>(q - p0) + 3 * pc * k;
                s0 += db[e];
                s1 += db[e + pc];
                s2 += db[e + 2 * pc];
            }
        for (int m = i0; m < i1; ++m) {
            const int q = in_idx[m];
            if (pti[q] != ty) continue;
            const size_t e = static_cast<size_t>(q - p0) + 3 * pc * k;
            s0 -= db[e];
            s1 -= db[e + pc];
            s2 -= db[e + 2 * pc];
        }
        const size_t col = static_cast<size_t>(ty - 1) * IB + k;
        out[3 * nc * col] = s0;
        out[3 * nc * col + 1] = s1;
        out[3 * nc * col + 2] = s2;
    }
}

// The same tally for one element type, one block per (configuration, comp + 3 k): the block streams
// that plane of dB/dr (pair-contiguous) through LDS in tiles of LSQ_TILE pairs with coalesced loads, and
// each atom adds its own pairs and subtracts its incoming ones (both ascending) into an LDS accumulator
// only it touches.  Every dB/dr value is read from HBM once.  LDS: LSQ_TILE + N_c doubles, N_c ints;
// 1024 threads per block keep the CU busy although only two such blocks fit in its LDS.
constexpr int LSQ_TILE = 4096;
__global__ __launch_bounds__(1024) void k_lsq_bd_tiled(const double* __restrict__ dblist, const int* __restrict__ poff,
                                                      const int* __restrict__ in_off, const int* __restrict__ in_idx,
                                                      const int* __restrict__ cfg_a0, const int* __restrict__ cfg_p0,
                                                      int IB, double* __restrict__ bd) {
    extern __shared__ double sm[];
    const int c = blockIdx.x / (3 * IB), plane = blockIdx.x % (3 * IB);
    const int comp = plane % 3, k = plane / 3;
    const int a0 = cfg_a0[c], nc = cfg_a0[c + 1] - a0, p0 = cfg_p0[c];
    const int pc = cfg_p0[c + 1] - p0;
    double* tile = sm;
    double* acc = sm + LSQ_TILE;
    int* cur = reinterpret_cast<int*>(acc + nc);
    const double* __restrict__ db = dblist + static_cast<size_t>(p0) * 3 * IB + static_cast<size_t>(pc) * (comp + 3 * k);
    for (int a = threadIdx.x; a < nc; a += 1024) {
        acc[a] = 0.0;
        cur[a] = in_off[a0 + a];
    }
    for (int t0 = 0; t0 < pc; t0 += LSQ_TILE) {
        const int t1 = min(pc, t0 + LSQ_TILE);
        __syncthreads();
        for (int i = threadIdx.x; i < t1 - t0; i += 1024) tile[i] = db[t0 + i];
        __syncthreads();
        for (int a = threadIdx.x; a < nc; a += 1024) {
            const int g = a0 + a;
            double s = acc[a];
            const int q0 = max(poff[g] - p0, t0), q1 = min(poff[g + 1] - p0, t1);
            for (int q = q0; q < q1; ++q) s += tile[q - t0];
            int m = cur[a];
            const int mend = in_off[g + 1];
            for (; m < mend; ++m) {
                const int q = in_idx[m] - p0;
                if (q >= t1) break;
                s -= tile[q - t0];
            }
            cur[a] = m;
            acc[a] = s;
        }
    }
    __syncthreads();
    for (int a = threadIdx.x; a < nc; a += 1024)
        bd[3 * static_cast<size_t>(a0) * IB + comp + 3 * a + 3 * static_cast<size_t>(nc) * k] = acc[a];
}

// bi[c*Kt + col] = sum of B_k over the atoms of type t of configuration c: one block per (c, col), its
// threads over the atoms (coalesced), fixed-order tree reduction
__global__ __launch_bounds__(256) void k_lsq_bi(const double* __restrict__ blist, const int* __restrict__ type,
                                                const int* __restrict__ cfg_a0, int IB, int ntypes,
                                                double* __restrict__ bi) {
    const int kt = ntypes * IB;
    const int c = blockIdx.x / kt, col = blockIdx.x % kt;
    const int ty = col / IB + 1, k = col % IB;
    const int a0 = cfg_a0[c], nc = cfg_a0[c + 1] - a0;
    const double* __restrict__ b = blist + static_cast<size_t>(a0) * IB + static_cast<size_t>(nc) * k;
    double s = 0.0;
    for (int ii = threadIdx.x; ii < nc; ii += 256)
        if (type[a0 + ii] == ty) s += b[ii];
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) bi[blockIdx.x] = red[0];
}

// virial tally of configuration c and column col (one block): v = -r (x) dB/dr summed over the pairs
// whose owner has the column's type, components (xx, yy, zz, xy, xz, yz) as upstream; written raw,
// bv[c*6*Kt + 6*col + m], and as the stress rows of the fit, as[(6c + m')*Kt + col] with the upstream
// row order m' -> m = (0, 1, 2, 5, 4, 3) (snapcpp.jl:660-662).  Fixed-order tree reduction.
__global__ __launch_bounds__(256) void k_lsq_bv(const double* __restrict__ dblist, const double* __restrict__ rij,
                                                const int* __restrict__ pti, const int* __restrict__ cfg_p0, int IB,
                                                int ntypes, double* __restrict__ bv, double* __restrict__ as) {
    const int kt = ntypes * IB;
    const int c = blockIdx.x / kt, col = blockIdx.x % kt;
    const int ty = col / IB + 1, k = col % IB;
    const int p0 = cfg_p0[c];
    const long pc = cfg_p0[c + 1] - p0;
    const double* __restrict__ db = dblist + static_cast<size_t>(p0) * 3 * IB + 3 * pc * k;
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (long lp = threadIdx.x; lp < pc; lp += 256) {
        const long q = p0 + lp;
        if (pti[q] != ty) continue;
        const double dx = -rij[3 * q], dy = -rij[3 * q + 1], dz = -rij[3 * q + 2];
        const double bx = db[lp], by = db[lp + pc], bz = db[lp + 2 * pc];
        v[0] += dx * bx;
        v[1] += dy * by;
        v[2] += dz * bz;
        v[3] += dx * by;
        v[4] += dx * bz;
        v[5] += dy * bz;
    }
    __shared__ double red[6][256];
#pragma unroll
    for (int m = 0; m < 6; ++m) red[m][threadIdx.x] = v[m];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w)
#pragma unroll
            for (int m = 0; m < 6; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) {
        const int m = threadIdx.x;
        bv[static_cast<size_t>(c) * 6 * kt + 6 * col + m] = red[m][0];
        constexpr int ord[6] = {0, 1, 2, 5, 4, 3};
        as[(static_cast<size_t>(c) * 6 + m) * kt + col] = red[ord[m]][0];
    }
}

void launch_lsq_tallies(const double* blist, const double* dblist, const double* rij, const int* poff,
                        const int* in_off, const int* in_idx, const int* pti, const int* type, const int* cfg_of,
                        const int* cfg_a0, const int* cfg_p0, int C, int N, int IB, int ntypes, int max_natoms,
                        double* bi, double* bd, double* bv, double* as, hipStream_t st) {
    const int kt = ntypes * IB;
    const size_t lds = sizeof(double) * (LSQ_TILE + static_cast<size_t>(max_natoms)) + sizeof(int) * max_natoms;
    if (ntypes == 1 && lds <= 65536)
        hipLaunchKernelGGL(k_lsq_bd_tiled, dim3(static_cast<unsigned>(C) * 3 * IB), dim3(1024), lds, st, dblist, poff,
                           in_off, in_idx, cfg_a0, cfg_p0, IB, bd);
    else
        hipLaunchKernelGGL(k_lsq_bd, dim3(nblocks(static_cast<long>(N) * IB, 256)), dim3(256), 0, st, dblist, poff,
                           in_off, in_idx, pti, type, cfg_of, cfg_a0, cfg_p0, N, IB, ntypes, bd);
    hipLaunchKernelGGL(k_lsq_bi, dim3(static_cast<unsigned>(C) * kt), dim3(256), 0, st, blist, type, cfg_a0, IB, ntypes,
                       bi);
    hipLaunchKernelGGL(k_lsq_bv, dim3(static_cast<unsigned>(C) * kt), dim3(256), 0, st, dblist, rij, pti, cfg_p0, IB,
                       ntypes, bv, as);
}

// ---- weighted normal equations --------------------------------------------------------------------

// Chunk ch covers rows r < nrows: a(r, k) = A[abase + r*rs + k*cs], target T[tbase + r], weight
// W[wbase + r / wdiv].  Outputs e < K(K+1)/2 are L(k, l), k <= l (row-major upper triangle), then
// K entries of h.  Block b takes chunks b, b + G, ... and writes its partial sums to part[b]; each
// thread owns a fixed set of outputs, accumulated in LDS.
__global__ __launch_bounds__(256) void k_lsq_gram(const LsqChunk* __restrict__ chunks, int nchunks,
                                                  const double* __restrict__ A, const double* __restrict__ T,
                                                  const double* __restrict__ W, int K,
                                                  const int2* __restrict__ ekl, double* __restrict__ part) {
    extern __shared__ double sm[];
    const int nout = K * (K + 1) / 2 + K;
    double* acc = sm;                        // nout
    double* tile = sm + nout;                // 64 * (K + 1): a(r, k) at r*(K+1) + k, w^2 b at r*(K+1) + K
    double* w2 = tile + 64 * (K + 1);        // 64
    for (int e = threadIdx.x; e < nout; e += 256) acc[e] = 0.0;
    for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const LsqChunk q = chunks[ch];
        __syncthreads();
        for (int i = threadIdx.x; i < q.nrows * K; i += 256) {
            const int k = i / q.nrows, r = i % q.nrows;   // rows fastest: contiguous reads when rs == 1
            tile[r * (K + 1) + k] = A[q.abase + static_cast<long>(r) * q.rs + static_cast<long>(k) * q.cs];
        }
        for (int r = threadIdx.x; r < q.nrows; r += 256) {
            const double w = W[q.wbase + r / q.wdiv];
            w2[r] = w * w;
            tile[r * (K + 1) + K] = w * w * T[q.tbase + r];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nout; e += 256) {
            const int2 kl = ekl[e];
            double s = 0.0;
            if (kl.y >= 0) {
                for (int r = 0; r < q.nrows; ++r) s += w2[r] * tile[r * (K + 1) + kl.x] * tile[r * (K + 1) + kl.y];
            } else {
                for (int r = 0; r < q.nrows; ++r) s += tile[r * (K + 1) + kl.x] * tile[r * (K + 1) + K];
            }
            acc[e] += s;
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nout; e += 256) part[static_cast<size_t>(blockIdx.x) * nout + e] = acc[e];
}

// out[e] = sum over blocks b = 0..G-1 of part[b][e], in block order
__global__ void k_lsq_gram_sum(const double* __restrict__ part, int G, int nout, double* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nout) return;
    double s = 0.0;
    for (int b = 0; b < G; ++b) s += part[static_cast<size_t>(b) * nout + e];
    out[e] = s;
}

size_t lsq_gram_lds(int K) { return sizeof(double) * (K * (K + 1) / 2 + K + 64 * (K + 1) + 64); }

void launch_lsq_gram(const LsqChunk* chunks, int nchunks, const double* A, const double* T, const double* W, int K,
                     const int2* ekl, double* part, int G, double* out, hipStream_t st) {
    const int nout = K * (K + 1) / 2 + K;
    hipLaunchKernelGGL(k_lsq_gram, dim3(G), dim3(256), lsq_gram_lds(K), st, chunks, nchunks, A, T, W, K, ekl, part);
    hipLaunchKernelGGL(k_lsq_gram_sum, dim3(nblocks(nout, 256)), dim3(256), 0, st, part, G, nout, out);
}

}  // namespace mdp
