// dB/dr of the batched fitting path on the FP64 matrix cores (v_mfma_f64_16x16x4f64).
//
// For one atom, dB_t/dr_c of a pair (cpuSnap.cpp:2074-2263) is, for every column n = (pair, Cartesian
// component c), a sum over layers l of
//     sum_e A_l[row][e] * dU_l[e][n],
// where e runs over the half-domain entries of layer l (rows 2mb <= l, the middle column of an even layer
// halved; real and imaginary parts as separate entries), dU_l[e][n] = d(sfac U_l)/dr_c of the column's
// pair, and each row of A_l is one term of the contraction: bispectrum index t = (j1, j2, j) contracts
// layer j with Z(j1,j2,j), layer j1 with Z(j,j2,j1) and layer j2 with Z(j,j1,j2), with weights 2,
// 2(j+1)/(j1+1) and 2(j+1)/(j2+1) (2 under bnorm).  Terms of one t on the same layer use the same Z block
// and are merged, so each layer holds one row per t that touches it.
//
// The rows of all layers are packed into at most 8 tiles of 16 rows (mfma_pack).  A wave takes one atom
// at a time and its columns 64 at a time (four tiles of 16 columns).  It holds the 16 x 16 accumulators of
// every (row tile, column tile) in registers while the row-chain recursion produces U row by row (mb
// outer, l inner) in each lane's registers: every finished row gives this lane's dU entries through the
// SU(2) generators, which go through LDS into the B operand; the A operand is read straight from the
// atom's Z block (weights applied); the tiles that hold the row's layer then take one MFMA per four
// entries and column tile.  At the end the accumulator rows of each t are summed in a fixed order and
// written in the per-configuration layout dblist[p0_c*3*nb + lp + P_c*c + 3*P_c*t].
#include <hip/hip_runtime.h>

#include <vector>

#include "fit_index.hpp"
#include "fit_launch.hpp"
#include "snap_device.hpp"

namespace mdp {

// ---- the row packing (shared by the kernel and the host tables) ---------------------------------

// MFMA K steps (4 entries each) of half row (l, mb): entries ma = 0..l, or 0..mb on the middle row,
// each with a real and an imaginary part
__host__ __device__ constexpr int mfma_row_steps(int l, int mb) { return (2 * ((l == 2 * mb ? mb : l) + 1) + 3) / 4; }
__host__ __device__ constexpr int mfma_layer_steps(int l) {
    int s = 0;
    for (int mb = 0; 2 * mb <= l; ++mb) s += mfma_row_steps(l, mb);
    return s;
}
__host__ __device__ constexpr bool b_touches(int J, int t, int l) {
    return b_comp(J, t, 0) == l || b_comp(J, t, 1) == l || b_comp(J, t, 2) == l;
}

struct MfmaPack {
    int ntiles = 0;
    int nrows[16] = {};        // rows of layer l: the bispectrum indices touching it, ascending
    int nseg[16] = {};         // the layer's rows lie in nseg segments, each in one tile
    int seg_tile[16][3] = {};
    int seg_row0[16][3] = {};  // first row (within the layer) of the segment
    int seg_cnt[16][3] = {};
    int seg_slot[16][3] = {};  // first tile row of the segment
    bool ok = true;
};

// Layers from the widest down: whole tiles of 16 rows first, then the remainder into the tile with the
// least free room that takes it; layers of at most two K steps may be split over tiles with free room.
// A tile costs one MFMA per K step of every layer it holds, so splitting the wide layers is avoided.
__host__ __device__ constexpr MfmaPack mfma_pack(int J) {
    MfmaPack p{};
    int fr[16] = {};
    for (int l = 0; l <= J; ++l) {
        int n = 0;
        for (int t = 0; t < n_b(J); ++t) n += b_touches(J, t, l) ? 1 : 0;
        p.nrows[l] = n;
    }
    for (int l = J; l >= 0; --l) {
        int r0 = 0;
        const int n = p.nrows[l];
        while (n - r0 > 0) {
            if (p.ntiles > 15 || p.nseg[l] > 2) {
                p.ok = false;
                return p;
            }
            const int rem = n - r0;
            int best = -1, most = -1;
            for (int t = 0; t < p.ntiles; ++t) {
                if (fr[t] >= rem && rem < 16 && (best < 0 || fr[t] < fr[best])) best = t;
                if (fr[t] > 0 && (most < 0 || fr[t] > fr[most])) most = t;
            }
            int tile = -1, take = 0;
            if (best >= 0) {
                tile = best;
                take = rem;
            } else if (rem < 16 && mfma_layer_steps(l) <= 2 && most >= 0 && p.nseg[l] < 2) {
                tile = most;
                take = fr[most] < rem ? fr[most] : rem;
            } else {
                tile = p.ntiles++;
                fr[tile] = 16;
                take = rem < 16 ? rem : 16;
            }
            const int k = p.nseg[l]++;
            p.seg_tile[l][k] = tile;
            p.seg_row0[l][k] = r0;
            p.seg_cnt[l][k] = take;
            p.seg_slot[l][k] = 16 - fr[tile];
            fr[tile] -= take;
            r0 += take;
        }
    }
    return p;
}

// position of half row (l, mb) in the order the row chain produces them: (0,0), then mb = 0 for
// l = 1..J, then mb = 1 for l = 2..J, ...
__host__ __device__ constexpr int chain_row_index(int J, int l, int mb) {
    if (mb == 0) return l;
    int r = J + 1;
    for (int b = 1; b < mb; ++b) r += J - 2 * b + 1;
    return r + l - 2 * mb;
}

// the half row after (l, mb) in chain order, and whether there is one
__host__ __device__ constexpr int chain_next_l(int J, int l, int mb) { return l < J ? l + 1 : 2 * (mb + 1); }
__host__ __device__ constexpr int chain_next_mb(int J, int l, int mb) { return l < J ? mb : mb + 1; }
__host__ __device__ constexpr bool chain_has_next(int J, int l, int mb) { return l < J || 2 * (mb + 1) <= J; }

// A operands one lane holds for a row: one per (segment, K step), at most this many over the rows
__host__ __device__ constexpr int mfma_a_max(int J) {
    const MfmaPack p = mfma_pack(J);
    int m = 1;
    for (int l = 0; l <= J; ++l)
        for (int mb = 0; 2 * mb <= l; ++mb)
            m = cx_max(m, p.nseg[l] * mfma_row_steps(l, mb));
    return m;
}

namespace {

typedef double d4_t __attribute__((ext_vector_type(4)));

// Raw Z values of row (L, MB) for this lane's A operands (entry e = g4*S + SS of the term row in tile slot
// m16), weights not yet applied; entries outside the row or slots of other layers read the zero at
// zero[0].  Issued one row ahead of their use.
template <int J, int L, int MB, int AMAX>
__device__ __forceinline__ void load_a_row(const double* __restrict__ z, const int* __restrict__ s_z,
                                           const double* __restrict__ zero, int m16, int g4, double (&dst)[AMAX]) {
    constexpr MfmaPack PK = mfma_pack(J);
    constexpr int LAST = (L == 2 * MB) ? MB : L;
    constexpr int KR = 2 * (LAST + 1), S = mfma_row_steps(L, MB), NS = PK.nseg[L];
    static_for<0, NS>([&](auto Kc) {
        constexpr int K = decltype(Kc)::value, T = PK.seg_tile[L][K];
        const int zo = s_z[(T * (J + 1) + L) * 16 + m16];
        static_for<0, S>([&](auto Sc) {
            constexpr int SS = decltype(Sc)::value;
            const int e = g4 * S + SS;
            const double* src = (zo >= 0 && e < KR) ? z + 2 * (zo + (L + 1) * MB) + e : zero;
            dst[K * S + SS] = *src;
        });
    });
}

template <int J>
__global__ __launch_bounds__(64, 1) void k_fit_dbdr_mfma(const double* __restrict__ rootpq, const double* __restrict__ zt,
                                                         const int* __restrict__ mz, const double* __restrict__ mw,
                                                         const int* __restrict__ comb,
                                                         const double* __restrict__ radelem,
                                                         const double* __restrict__ wjelem, double rcutfac,
                                                         double rmin0, double rfac0, int switchflag,
                                                         const double* __restrict__ rij, const int* __restrict__ pti,
                                                         const int* __restrict__ ptj, const int* __restrict__ poff,
                                                         const int* __restrict__ cfg_of, const int* __restrict__ cfg_p0,
                                                         double* __restrict__ dblist, int* __restrict__ counter, int N) {
    constexpr MfmaPack PK = mfma_pack(J);
    static_assert(PK.ok && PK.ntiles <= 8, "row packing needs more than 8 tiles");
    constexpr int NT = PK.ntiles;
    constexpr int IZ = n_z(J), IB = n_b(J);
    constexpr int STAGE = 4 * mfma_row_steps(J, 0) * 64;   // one row of B entries for 64 columns
    constexpr int TAB = NT * (J + 1) * 16;
    constexpr int AMAX = mfma_a_max(J);
    const double* __restrict__ zero = mw + TAB;   // the table ends with a 0.0
    __shared__ double s_buf[cx_max(2 * STAGE, NT * 256)];
    __shared__ double s_w[TAB];
    __shared__ int s_z[TAB];
    __shared__ int s_comb[3 * IB];
    const int lane = threadIdx.x, m16 = lane & 15, g4 = lane >> 4;
    for (int k = lane; k < TAB; k += 64) {
        s_z[k] = mz[k];
        s_w[k] = mw[k];
    }
    for (int k = lane; k < 3 * IB; k += 64) s_comb[k] = comb[k];
    __syncthreads();
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(counter, 1);
        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
        if (t >= N) break;
        const int atom = t;
        const int q0 = __builtin_amdgcn_readfirstlane(poff[atom]);
        const int P = __builtin_amdgcn_readfirstlane(poff[atom + 1]) - q0;
        if (P <= 0) continue;
        const int c = __builtin_amdgcn_readfirstlane(cfg_of[atom]);
        const int p0 = __builtin_amdgcn_readfirstlane(cfg_p0[c]);
        const long Pc = __builtin_amdgcn_readfirstlane(cfg_p0[c + 1]) - p0;
        const double* __restrict__ z = zt + static_cast<size_t>(atom) * IZ * 2;
        double* __restrict__ db = dblist + static_cast<size_t>(p0) * 3 * IB;
        const int ncol = 3 * P;
        // Bring the atom's Z block (IZ*16 bytes) into the caches with one load per 128-byte line, all in
        // flight together: the A operands below then hit in cache instead of each row waiting on HBM.
        unsigned touch = 0;
        {
            const unsigned* __restrict__ zw = reinterpret_cast<const unsigned*>(z);
            constexpr int LINES = (IZ * 16 + 127) / 128;
#pragma unroll
            for (int k = 0; k < (LINES + 63) / 64; ++k)
                if (k * 64 + lane < LINES) touch ^= zw[(k * 64 + lane) * 32];
        }
        for (int c0v = 0; c0v < ncol; c0v += 64) {
            const int c0 = __builtin_amdgcn_readfirstlane(c0v);   // wave-uniform explicitly: fewer VGPRs
            const int nt = min(4, (ncol - c0 + 15) >> 4);   // column tiles in use
            const int col = c0 + lane;
            const int comp = col < ncol ? col / P : 0;      // columns: component-major within the atom
            const int p = q0 + (col < ncol ? col - comp * P : 0);
            const int tjp = ptj[p];
            const double rc = (radelem[pti[p]] + radelem[tjp]) * rcutfac;
            PairGeom g;
            pair_geometry<true>(rij[3 * p], rij[3 * p + 1], rij[3 * p + 2], rc, rmin0, rfac0, switchflag, wjelem[tjp], g);
            // projections of dq/dx_comp on V_0 = q*i, V_1 = q*j, V_2 = q*k
            const double dar = comp == 0 ? g.dar[0] : (comp == 1 ? g.dar[1] : g.dar[2]);
            const double dai = comp == 0 ? g.dai[0] : (comp == 1 ? g.dai[1] : g.dai[2]);
            const double dbr = comp == 0 ? g.dbr[0] : (comp == 1 ? g.dbr[1] : g.dbr[2]);
            const double dbi = comp == 0 ? g.dbi[0] : (comp == 1 ? g.dbi[1] : g.dbi[2]);
            const double uk = comp == 0 ? g.u[0] : (comp == 1 ? g.u[1] : g.u[2]);
            const double sf = g.sfac, dsu = g.dsfac * uk;
            const double k0 = sf * (-g.ai * dar + g.ar * dai + g.bi * dbr - g.br * dbi);
            const double k1 = sf * (-g.br * dar - g.bi * dai + g.ar * dbr + g.ai * dbi);
            const double k2 = sf * (-g.bi * dar + g.br * dai - g.ai * dbr + g.ar * dbi);
            d4_t acc[NT][4];
            static_for<0, NT>([&](auto Tc) {
                static_for<0, 4>([&](auto Nc) {
                    acc[decltype(Tc)::value][decltype(Nc)::value] = d4_t{0.0, 0.0, 0.0, 0.0};
                });
            });
            double an[AMAX];   // raw A values of the next row, loaded one row ahead
            load_a_row<J, 0, 0, AMAX>(z, s_z, zero, m16, g4, an);
            const double* rq = rootpq;
            __asm__ volatile("" : "+s"(rq));   // keeps the rootpq loads inside this pass (not hoisted to SGPRs)
            row_chain<J, 0, 0, J / 2>(g, rq, [&](auto Jc, auto MBc, auto& pr, auto& pim, auto&, auto&) {
                constexpr int L = decltype(Jc)::value, MB = decltype(MBc)::value;
                constexpr int LAST = (L == 2 * MB) ? MB : L;
                constexpr int KR = 2 * (LAST + 1), S = mfma_row_steps(L, MB);
                constexpr int NS = PK.nseg[L];
                double* __restrict__ stg = s_buf + (chain_row_index(J, L, MB) & 1) * STAGE;
                // A operand of every tile holding layer L (loaded during the previous row), then the next
                // row's loads, which overlap this row's dU work and MFMAs
                double a[AMAX];
#pragma unroll
                for (int k = 0; k < AMAX; ++k) a[k] = an[k];
                if constexpr (chain_has_next(J, L, MB))
                    load_a_row<J, chain_next_l(J, L, MB), chain_next_mb(J, L, MB), AMAX>(z, s_z, zero, m16, g4, an);
                // d(sfac U)/dx_comp of this row: dsfac u_comp U + sfac sum_g (dq/dx_comp . V_g) L_g U
                double dr[LAST + 1], di[LAST + 1];
                static_for<0, LAST + 1>([&](auto MAc) {
                    constexpr int MA = decltype(MAc)::value;
                    constexpr double m0 = static_cast<double>(2 * MA - L);
                    constexpr double cu = (MA < L) ? cx_sqrt(static_cast<double>((MA + 1) * (L - MA))) : 0.0;
                    constexpr double cd = (MA > 0) ? cx_sqrt(static_cast<double>(MA * (L - MA + 1))) : 0.0;
                    const double upr = (MA < L) ? pr[MA < L ? MA + 1 : MA] : 0.0;
                    const double upi = (MA < L) ? pim[MA < L ? MA + 1 : MA] : 0.0;
                    const double dnr = (MA > 0) ? pr[MA > 0 ? MA - 1 : 0] : 0.0;
                    const double dni = (MA > 0) ? pim[MA > 0 ? MA - 1 : 0] : 0.0;
                    const double l1r = cu * upr - cd * dnr, l1i = cu * upi - cd * dni;
                    const double l2r = -(cu * upi + cd * dni), l2i = cu * upr + cd * dnr;
                    dr[MA] = dsu * pr[MA] + k0 * (-m0 * pim[MA]) + k1 * l1r + k2 * l2r;
                    di[MA] = dsu * pim[MA] + k0 * (m0 * pr[MA]) + k1 * l1i + k2 * l2i;
                });
                if constexpr (L == 2 * MB) {   // middle column of an even layer counts 1/2
                    dr[MB] *= 0.5;
                    di[MB] *= 0.5;
                }
                // B operand: entry e of this lane's column at stg[e*64 + lane], zero past the row
                static_for<0, 4 * S>([&](auto Ec) {
                    constexpr int E = decltype(Ec)::value;
                    double v = 0.0;
                    if constexpr (E < KR) v = (E % 2 == 0) ? dr[E / 2] : di[E / 2];
                    stg[E * 64 + lane] = v;
                });
                __syncthreads();
                double aw[AMAX];
                static_for<0, NS>([&](auto Kc) {
                    constexpr int K = decltype(Kc)::value, T = PK.seg_tile[L][K];
                    const double w = s_w[(T * (J + 1) + L) * 16 + m16];
#pragma unroll
                    for (int ss = 0; ss < S; ++ss) aw[K * S + ss] = w * a[K * S + ss];
                });
                // every column tile takes its MFMAs, in use or not: no accumulator is carried through a
                // branch (tiles past the atom's columns hold finite values that are never stored)
                static_for<0, S>([&](auto Sc) {
                    constexpr int SS = decltype(Sc)::value;
                    static_for<0, 4>([&](auto Nc) {
                        constexpr int NN = decltype(Nc)::value;
                        const double b = stg[(g4 * S + SS) * 64 + NN * 16 + m16];
                        static_for<0, NS>([&](auto Kc) {
                            constexpr int K = decltype(Kc)::value, T = PK.seg_tile[L][K];
                            acc[T][NN] = __builtin_amdgcn_mfma_f64_16x16x4f64(aw[K * S + SS], b, acc[T][NN], 0, 0, 0);
                        });
                    });
                });
            });
            // accumulator rows -> dB: register r of lane (m16, g4) holds row 4*r + g4 of column m16 of each tile
            static_for<0, 4>([&](auto Nc) {
                constexpr int NN = decltype(Nc)::value;
                if (NN < nt) {
                    __syncthreads();
                    static_for<0, NT>([&](auto Tc) {
                        constexpr int T = decltype(Tc)::value;
#pragma unroll
                        for (int r = 0; r < 4; ++r) s_buf[(T * 16 + 4 * r + g4) * 16 + m16] = acc[T][NN][r];
                    });
                    __syncthreads();
                    for (int k = lane; k < IB * 16; k += 64) {
                        const int tb = k >> 4, cc = k & 15;
                        const int colc = c0 + NN * 16 + cc;
                        if (colc < ncol) {
                            double v = 0.0;
#pragma unroll
                            for (int r = 0; r < 3; ++r) {
                                const int pos = s_comb[3 * tb + r];
                                if (pos >= 0) v += s_buf[pos * 16 + cc];
                            }
                            const int cp = colc / P;
                            const size_t lp = static_cast<size_t>(q0 + colc - cp * P - p0);
                            // streamed past the caches: dB/dr is written once and not read here again,
                            // and keeping it out of L2 leaves room for the Z blocks the A operands read
                            __builtin_nontemporal_store(v, &db[lp + Pc * cp + 3 * Pc * tb]);
                        }
                    }
                }
            });
            __syncthreads();
        }
        if (touch == 0x9e3779b9u && N < 0) dblist[0] = 0.0;   // never taken: keeps the cache-fill loads
    }
}

}  // namespace

bool fit_mfma_supported(int J) { return J == 4 || J == 6 || J == 8; }

// Host tables of the packing: for tile T, layer l and tile row m, mz = Z block offset (idxz ordering)
// of the row's term and mw its weight (-1 / 0 where the row holds another layer), mw with one more 0.0
// at the end; comb[3t + k] = tile rows (T*16 + m) of bispectrum index t, -1 padded.
bool fit_mfma_tables(int J, int bnorm, std::vector<int>& mz, std::vector<double>& mw, std::vector<int>& comb) {
    const MfmaPack pk = mfma_pack(J);
    if (!pk.ok || pk.ntiles > 8) return false;
    const int nb = n_b(J);
    mz.assign(static_cast<size_t>(pk.ntiles) * (J + 1) * 16, -1);
    mw.assign(mz.size() + 1, 0.0);   // the last entry stays 0.0: the kernel reads it for padded A entries
    comb.assign(3 * static_cast<size_t>(nb), -1);
    std::vector<int> ncomb(nb, 0);
    for (int l = 0; l <= J; ++l) {
        int row = 0;
        for (int t = 0; t < nb; ++t) {
            if (!b_touches(J, t, l)) continue;
            int zb = -1;
            double w = 0.0;
            const int j1 = b_comp(J, t, 0), j2 = b_comp(J, t, 1), j = b_comp(J, t, 2);
            for (int term = 0; term < 3; ++term) {
                if (term_layer(J, t, term) != l) continue;
                const int zt = term_zblock(J, t, term);
                if (zb >= 0 && zt != zb) return false;   // same layer, different Z block: not mergeable
                zb = zt;
                w += term == 0 ? 2.0 : (bnorm ? 2.0 : 2.0 * (j + 1) / ((term == 1 ? j1 : j2) + 1.0));
            }
            int seg = -1;
            for (int k = 0; k < pk.nseg[l]; ++k)
                if (row >= pk.seg_row0[l][k] && row < pk.seg_row0[l][k] + pk.seg_cnt[l][k]) seg = k;
            if (seg < 0 || ncomb[t] >= 3) return false;
            const int T = pk.seg_tile[l][seg], m = pk.seg_slot[l][seg] + row - pk.seg_row0[l][seg];
            mz[(static_cast<size_t>(T) * (J + 1) + l) * 16 + m] = zb;
            mw[(static_cast<size_t>(T) * (J + 1) + l) * 16 + m] = w;
            comb[3 * static_cast<size_t>(t) + ncomb[t]++] = T * 16 + m;
            ++row;
        }
    }
    return true;
}

bool launch_fit_dbdr_mfma(const SnapDev& s, const double* zt, const int* mz, const double* mw, const int* comb,
                          const double* rij, const int* pti, const int* ptj, const int* poff, const int* cfg_of,
                          const int* cfg_p0, double* dblist, int* counter, int N, int groups, hipStream_t st) {
    if (N <= 0) return true;
    const int G = groups > 0 ? groups : 1024;
    switch (s.J) {
#define DM_CASE(JV)                                                                                              \
    case JV:                                                                                                     \
        hipLaunchKernelGGL((k_fit_dbdr_mfma<JV>), dim3(G), dim3(64), 0, st, s.rootpq, zt, mz, mw, comb, s.radelem, \
                           s.wjelem, s.rcutfac, s.rmin0, s.rfac0, s.switchflag, rij, pti, ptj, poff, cfg_of, cfg_p0, \
                           dblist, counter, N);                                                                  \
        return true;
        DM_CASE(4)
        DM_CASE(6)
        DM_CASE(8)
#undef DM_CASE
        default:
            return false;
    }
}

}  // namespace mdp
