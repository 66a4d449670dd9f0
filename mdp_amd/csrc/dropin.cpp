// The 23 drop-in entry points of src/Julia/Potential/cpuSnap.h on the GPU.
//
// Every call: validate the request, stage the host inputs (and, for outputs upstream accumulates
// into, the current output contents) to device buffers, run the HIP kernels on the library stream,
// copy the results back and synchronise.  Array extents the upstream signature leaves implicit are
// derived from the index arrays the upstream code itself dereferences.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/mdpsnap.h"
#include "runtime.hpp"

using namespace mdp;

namespace {

// Per-argument device buffers, reused across calls (guarded by the library mutex).
DevBuf g_slot[40];
DevTables g_tab;

template <class T>
T* dev(int slot, size_t n) {
    return static_cast<T*>(g_slot[slot].ensure(sizeof(T) * std::max<size_t>(n, 1)));
}
template <class T>
T* put(int slot, const T* h, size_t n, hipStream_t st) {
    T* d = dev<T>(slot, n);
    if (n) MDP_HIP(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, st));
    return d;
}
template <class T>
void get(T* h, const T* d, size_t n, hipStream_t st) {
    if (n) MDP_HIP(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, st));
}

// Argument I/O of the shared bodies below.  The cpuSnap.h entry points pass host arrays, staged through
// the per-slot device buffers.  The backend-3 entry points (hip*, mdpcore.h) pass device arrays: their
// large arrays (U, Y, r_ij, forces, virials, energies) are then used in place, while the small index
// and table arrays arrive as host copies and are staged as usual.
thread_local bool g_devio = false;
// virial layout of the force tally: R1 AoS vatom[6i + k] (g_vsoa < 0) or, for the backend-3 SnapComputeFi,
// SoA vatom[k * anum + i] with -r(x)dE/2 added to both atoms of the pair (g_vsoa = anum)
thread_local int g_vsoa = -1;
template <class T>
T* put_arg(int slot, const T* p, size_t n, hipStream_t st) {
    return g_devio ? const_cast<T*>(p) : put(slot, p, n, st);
}
template <class T>
T* out_arg(int slot, T* p, size_t n) {
    return g_devio ? p : dev<T>(slot, n);
}
template <class T>
void get_arg(T* h, const T* d, size_t n, hipStream_t st) {
    if (!g_devio) get(h, d, n, st);
}
void finish(hipStream_t st) {
    MDP_HIP(hipGetLastError());
    MDP_HIP(hipStreamSynchronize(st));
}

int max_of(const int* a, long n, int floor_v = -1) {
    int m = floor_v;
    for (long i = 0; i < n; ++i) m = std::max(m, a[i]);
    return m;
}

void require_j(int J) {
    if (!templated_j(J))
        throw Error("twojmax=" + std::to_string(J) + " is not supported by this build (supported: 2, 4, 6, 8)");
}

const SnapIndex& index_for(int J) {
    static SnapIndex cache[9];
    static bool built[9] = {false};
    require_j(J);
    if (!built[J]) {
        cache[J].build(J, nullptr);
        built[J] = true;
    }
    return cache[J];
}

// Stage tables: canonical index lists for J plus the caller's numeric tables.
const SnapDev& tables(int J, const double* rootpq, const double* cg, const double* radelem, const double* wjelem,
                      const int* map, int ntypes, hipStream_t st) {
    g_tab.upload(index_for(J), rootpq, cg, radelem, wjelem, map, ntypes, st);
    return g_tab.view;
}

template <class F>
void guarded(F&& f) {
    std::lock_guard<std::mutex> lk(library_mutex());
    clear_error();
    try {
        f();
    } catch (const std::exception& e) {
        set_error(e.what());
    }
}

// CSR of pairs by centre atom (stable): off[inum+1]; perm empty when already sorted.
void pairs_by_atom(const int* aii, int ijnum, int inum, std::vector<int>& off, std::vector<int>& perm) {
    off.assign(inum + 1, 0);
    bool sorted = true;
    for (int p = 0; p < ijnum; ++p) {
        if (aii[p] < 0 || aii[p] >= inum) throw Error("pair centre index out of range [0, inum)");
        off[aii[p] + 1]++;
        if (p && aii[p] < aii[p - 1]) sorted = false;
    }
    std::partial_sum(off.begin(), off.end(), off.begin());
    perm.clear();
    if (!sorted) {
        perm.resize(ijnum);
        std::vector<int> pos(off.begin(), off.end() - 1);
        for (int p = 0; p < ijnum; ++p) perm[pos[aii[p]]++] = p;
    }
}

}  // namespace

extern "C" {

const char* mdp_snap_last_error(void) { return get_error(); }

int idxcgcount(int twojmax) { return count_cg(twojmax); }
int idxucount(int twojmax) { return count_u(twojmax); }
int idxbcount(int twojmax) { return count_b(twojmax); }
int idxzcount(int twojmax) { return count_z(twojmax); }

void cpuInitSna(double* rootpqarray, double* cglist, double* factorial, int* idx_max, int* idxz, int* idxz_block,
                int* idxb, int* idxb_block, int* idxu_block, int* idxcg_block, int twojmax) {
    guarded([&] {
        SnapIndex ix;
        ix.build(twojmax, factorial);
        const int jd = twojmax + 1;
        idx_max[0] = ix.idxcg_max;
        idx_max[1] = ix.idxu_max;
        idx_max[2] = ix.idxb_max;
        idx_max[3] = ix.idxz_max;
        std::copy(ix.idxz.begin(), ix.idxz.end(), idxz);
        std::copy(ix.idxb.begin(), ix.idxb.end(), idxb);
        std::copy(ix.idxu_block.begin(), ix.idxu_block.end(), idxu_block);
        for (int k = 0; k < jd * jd * jd; ++k) {
            idxz_block[k] = ix.idxz_block[k];
            idxb_block[k] = ix.idxb_block[k];
            idxcg_block[k] = ix.idxcg_block[k];
        }
        for (int p = 1; p <= twojmax; ++p)
            for (int q = 1; q <= twojmax; ++q) rootpqarray[p * jd + q] = ix.rootpq[p * jd + q];
        std::copy(ix.cglist.begin(), ix.cglist.end(), cglist);
    });
}

void cpuNeighPairList(int* pairnum, int* pairlist, double* x, double* rcutsq, int* ilist, int* neighlist, int* neighnum,
                      int* atomtype, int* alist, int inum, int dim, int ntypes) {
    guarded([&] {
        if (inum <= 0) return;
        hipStream_t st = library_stream();
        const int imax = max_of(ilist, inum);
        const int nn = neighnum[imax + 1];
        const int npts = std::max(max_of(neighlist, nn), imax) + 1;
        const int natoms = std::max(max_of(alist, npts), imax) + 1;
        int* d_il = put(0, ilist, inum, st);
        int* d_nl = put(1, neighlist, nn, st);
        int* d_nn = put(2, neighnum, imax + 2, st);
        int* d_at = put(3, atomtype, natoms, st);
        int* d_al = put(4, alist, npts, st);
        double* d_x = put(5, x, static_cast<size_t>(npts) * dim, st);
        double* d_rc = put(6, rcutsq, static_cast<size_t>(ntypes + 1) * (ntypes + 1), st);
        int* d_cnt = dev<int>(7, inum);
        int* d_off = dev<int>(8, inum + 1);
        int* d_tmp = dev<int>(9, scan_tmp_ints(inum));
        launch_pair_count(d_cnt, d_x, d_rc, d_il, d_nl, d_nn, d_at, d_al, inum, dim, ntypes, st);
        exclusive_scan(d_cnt, d_off, inum, d_tmp, st);
        int total = 0;
        get(&total, d_off + inum, 1, st);
        MDP_HIP(hipStreamSynchronize(st));
        int* d_pl = dev<int>(10, total);
        launch_pair_fill(d_pl, d_off, d_x, d_rc, d_il, d_nl, d_nn, d_at, d_al, inum, dim, ntypes, st);
        get(pairnum, d_cnt, inum, st);
        get(pairlist, d_pl, total, st);
        finish(st);
    });
}

void cpuNeighPairs(double* xij, double* x, int* aii, int* ai, int* aj, int* ti, int* tj, int* pairlist, int* pairnumsum,
                   int* ilist, int* atomtype, int* alist, int inum, int dim) {
    guarded([&] {
        if (inum <= 0) return;
        hipStream_t st = library_stream();
        const int total = pairnumsum[inum];
        const int imax = max_of(ilist, inum);
        const int npts = std::max(max_of(pairlist, total), imax) + 1;
        const int natoms = std::max(max_of(alist, npts), imax) + 1;
        int* d_pl = put(0, pairlist, total, st);
        int* d_ps = put(1, pairnumsum, inum + 1, st);
        int* d_il = put(2, ilist, inum, st);
        int* d_at = put(3, atomtype, natoms, st);
        int* d_al = put(4, alist, npts, st);
        double* d_x = put(5, x, static_cast<size_t>(npts) * dim, st);
        double* d_xij = dev<double>(6, static_cast<size_t>(total) * dim);
        int* d_aii = dev<int>(7, total);
        int* d_ai = dev<int>(8, total);
        int* d_aj = dev<int>(9, total);
        int* d_ti = dev<int>(10, total);
        int* d_tj = dev<int>(11, total);
        launch_pair_arrays(d_xij, d_x, d_aii, d_ai, d_aj, d_ti, d_tj, d_pl, d_ps, d_il, d_at, d_al, inum, dim, st);
        get(xij, d_xij, static_cast<size_t>(total) * dim, st);
        get(aii, d_aii, total, st);
        get(ai, d_ai, total, st);
        get(aj, d_aj, total, st);
        get(ti, d_ti, total, st);
        get(tj, d_tj, total, st);
        finish(st);
    });
}

void cpuSnapComputeUi(double* Utotr, double* Utoti, double* rootpqarray, double* rij, double* wjelem, double* radelem,
                      double rmin0, double rfac0, double rcutfac, int* map, int* aii, int* ti, int* tj, int twojmax,
                      int idxu_max, int inum, int ijnum, int switchflag, int chemflag) {
    guarded([&] {
        if (inum <= 0) return;
        hipStream_t st = library_stream();
        const int ntypes = std::max(max_of(ti, ijnum, 1), max_of(tj, ijnum, 1));
        const int nelem = chemflag ? max_of(map + 1, ntypes, 0) + 1 : 1;
        SnapDev s = tables(twojmax, rootpqarray, nullptr, radelem, wjelem, map, ntypes, st);
        if (idxu_max != s.idxu_max) throw Error("idxu_max does not match twojmax");
        s.rcutfac = rcutfac;
        s.rfac0 = rfac0;
        s.rmin0 = rmin0;
        s.switchflag = switchflag;
        std::vector<int> off, perm;
        pairs_by_atom(aii, ijnum, inum, off, perm);
        const size_t nu = static_cast<size_t>(inum) * idxu_max * nelem;
        double* d_ur = put_arg(0, Utotr, nu, st);
        double* d_ui = put_arg(1, Utoti, nu, st);
        double* d_rij = put_arg(2, rij, 3 * static_cast<size_t>(ijnum), st);
        int* d_ti = put(3, ti, ijnum, st);
        int* d_tj = put(4, tj, ijnum, st);
        int* d_off = put(5, off.data(), off.size(), st);
        int* d_perm = perm.empty() ? nullptr : put(6, perm.data(), perm.size(), st);
        if (!chemflag) {
            launch_ui(s, UI_ACCUMULATE, d_ur, d_ui, d_rij, d_off, d_perm, d_ti, d_tj, inum, 0.0, -1, st);
        } else {
            for (int e = 0; e < nelem; ++e) {
                const size_t o = static_cast<size_t>(inum) * idxu_max * e;
                launch_ui(s, UI_ACCUMULATE, d_ur + o, d_ui + o, d_rij, d_off, d_perm, d_ti, d_tj, inum, 0.0, e, st);
            }
        }
        get_arg(Utotr, d_ur, nu, st);
        get_arg(Utoti, d_ui, nu, st);
        finish(st);
    });
}

void cpuAddWself2Ui(double* Utotr, double* Utoti, double wself, int* idxu_block, int* type, int* map, int wselfall_flag,
                    int chemflag, int idxu_max, int nelements, int twojmax, int inum) {
    (void)idxu_block;
    guarded([&] {
        if (inum <= 0) return;
        hipStream_t st = library_stream();
        const int ntypes = max_of(type, inum, 1);
        SnapDev s = tables(twojmax, nullptr, nullptr, nullptr, nullptr, map, ntypes, st);
        if (idxu_max != s.idxu_max) throw Error("idxu_max does not match twojmax");
        const size_t nu = static_cast<size_t>(inum) * idxu_max * nelements;
        double* d_ur = put_arg(0, Utotr, nu, st);
        double* d_ui = put_arg(1, Utoti, nu, st);
        int* d_t = put(2, type, inum, st);
        launch_add_wself(s, d_ur, d_ui, wself, d_t, wselfall_flag, chemflag, nelements, inum, st);
        get_arg(Utotr, d_ur, nu, st);
        get_arg(Utoti, d_ui, nu, st);
        finish(st);
    });
}

namespace {
// blist for all element triples (layout ii + N*jjb + N*idxb_max*jelem), before bzero
void all_triples_b(const SnapDev& s, const double* d_ur, const double* d_ui, double* d_b, int nelem, int inum, int bnorm,
                   hipStream_t st) {
    const long nu = static_cast<long>(inum) * s.idxu_max;
    for (int e1 = 0; e1 < nelem; ++e1)
        for (int e2 = 0; e2 < nelem; ++e2)
            for (int e3 = 0; e3 < nelem; ++e3) {
                const long je = (e1 * nelem + e2) * nelem + e3;
                launch_bi(s, d_ur, d_ui, nu * e1, nu * e2, nu * e3, d_b + je * inum * s.idxb_max, inum, bnorm, st);
            }
}
}  // namespace

void cpuSnapComputeEi(double* eatom, double* Utotr, double* Utoti, double* cglist, double* bzero, double* coeffelem,
                      int* ilist, int* map, int* type, int* idxb, int* idxcg_block, int* idxu_block, int twojmax,
                      int idxb_max, int idxu_max, int nelements, int ncoeffall, int bnorm_flag, int bzero_flag,
                      int wselfall_flag, int chemflag, int inum) {
    (void)idxb; (void)idxcg_block; (void)idxu_block;
    guarded([&] {
        if (inum <= 0) return;
        hipStream_t st = library_stream();
        const int ntypes = max_of(type, inum, 1);
        SnapDev s = tables(twojmax, nullptr, cglist, nullptr, nullptr, map, ntypes, st);
        if (idxu_max != s.idxu_max || idxb_max != s.idxb_max) throw Error("table sizes do not match twojmax");
        const int ne = nelements;
        const size_t nu = static_cast<size_t>(inum) * idxu_max * ne;
        const int ncoef_rows = chemflag ? max_of(map + 1, ntypes, 0) + 1 : 1;
        const int eext = max_of(ilist, inum) + 1;
        double* d_ur = put_arg(0, Utotr, nu, st);
        double* d_ui = put_arg(1, Utoti, nu, st);
        double* d_bz = put(2, bzero, twojmax + 1, st);
        double* d_c = put(3, coeffelem, static_cast<size_t>(ncoeffall) * ncoef_rows, st);
        int* d_il = put(4, ilist, inum, st);
        int* d_t = put(5, type, inum, st);
        double* d_e = put_arg(6, eatom, eext, st);
        double* d_b = dev<double>(7, static_cast<size_t>(inum) * idxb_max * ne * ne * ne);
        all_triples_b(s, d_ur, d_ui, d_b, ne, inum, bnorm_flag, st);
        if (bzero_flag) launch_bzero(s, d_b, d_bz, d_il, d_t, wselfall_flag, chemflag, ne, inum, 0, st);
        launch_ei_from_b(s, d_e, d_b, d_c, d_il, d_t, chemflag, ne, ncoeffall, inum, st);
        get_arg(eatom, d_e, eext, st);
        finish(st);
    });
}

void cpuSnapComputeBi(double* blist, double* Utotr, double* Utoti, double* cglist, double* bzero, int* idxb,
                      int* idxcg_block, int* idxu_block, int* idxz_block, int twojmax, int idxb_max, int idxu_max,
                      int idxz_max, int nelements, int bnorm_flag, int bzero_flag, int wselfall_flag, int inum) {
    (void)idxb; (void)idxcg_block; (void)idxu_block; (void)idxz_block; (void)idxz_max;
    guarded([&] {
        if (inum <= 0) return;
        hipStream_t st = library_stream();
        SnapDev s = tables(twojmax, nullptr, cglist, nullptr, nullptr, nullptr, 1, st);
        if (idxu_max != s.idxu_max || idxb_max != s.idxb_max) throw Error("table sizes do not match twojmax");
        const int ne = nelements;
        const size_t nu = static_cast<size_t>(inum) * idxu_max * ne;
        const size_t nb = static_cast<size_t>(inum) * idxb_max * ne * ne * ne;
        double* d_ur = put(0, Utotr, nu, st);
        double* d_ui = put(1, Utoti, nu, st);
        double* d_bz = put(2, bzero, twojmax + 1, st);
        double* d_b = dev<double>(3, nb);
        all_triples_b(s, d_ur, d_ui, d_b, ne, inum, bnorm_flag, st);
        if (bzero_flag) launch_bzero(s, d_b, d_bz, nullptr, nullptr, wselfall_flag, 0, ne, inum, 0, st);
        get(blist, d_b, nb, st);
        finish(st);
    });
}

void cpuSnapComputeYi(double* ylist_r, double* ylist_i, double* Utotr, double* Utoti, double* cglist, double* beta,
                      int* map, int* type, int* idxz, int* idxb_block, int* idxu_block, int* idxcg_block, int twojmax,
                      int idxb_max, int idxu_max, int idxz_max, int nelements, int ncoeffall, int bnorm_flag,
                      int chemflag, int inum) {
    (void)idxz; (void)idxb_block; (void)idxu_block; (void)idxcg_block; (void)idxz_max;
    guarded([&] {
        if (inum <= 0) return;
        hipStream_t st = library_stream();
        const int ntypes = max_of(type, inum, 1);
        SnapDev s = tables(twojmax, nullptr, cglist, nullptr, nullptr, map, ntypes, st);
        if (idxu_max != s.idxu_max || idxb_max != s.idxb_max) throw Error("table sizes do not match twojmax");
        const SnapIndex& ix = index_for(twojmax);
        const int ne = nelements;
        const size_t nu = static_cast<size_t>(inum) * idxu_max * ne;
        double* d_ur = put_arg(0, Utotr, nu, st);
        double* d_ui = put_arg(1, Utoti, nu, st);
        double* d_yr = out_arg(2, ylist_r, nu);
        double* d_yi = out_arg(3, ylist_i, nu);
        MDP_HIP(hipMemsetAsync(d_yr, 0, sizeof(double) * nu, st));
        MDP_HIP(hipMemsetAsync(d_yi, 0, sizeof(double) * nu, st));
        if (ne == 1 && !chemflag) {
            std::vector<double> bt = y_beta_table(ix, beta, 0, 0, 0, 1, 0, ncoeffall, bnorm_flag);
            double* d_bt = put(4, bt.data(), bt.size(), st);
            launch_yi(s, d_ur, d_ui, d_bt, d_yr, d_yi, nullptr, inum, true, st);
            MDP_HIP(hipStreamSynchronize(st));
        } else {
            std::vector<int> ae(inum);
            for (int i = 0; i < inum; ++i) ae[i] = chemflag ? map[type[i]] : 0;
            int* d_ae = put(5, ae.data(), ae.size(), st);
            const int nie = chemflag ? max_of(ae.data(), inum, 0) + 1 : 1;
            const long nu1 = static_cast<long>(inum) * idxu_max;
            std::vector<std::vector<double>> keep;
            int slot = 6;
            for (int ie = 0; ie < nie; ++ie)
                for (int e1 = 0; e1 < ne; ++e1)
                    for (int e2 = 0; e2 < ne; ++e2)
                        for (int e3 = 0; e3 < ne; ++e3) {
                            keep.push_back(y_beta_table(ix, beta, e1, e2, e3, ne, ie, ncoeffall, bnorm_flag));
                            if (slot >= 40) throw Error("too many element combinations for the staging slots");
                            double* d_bt = put(slot++, keep.back().data(), keep.back().size(), st);
                            launch_yi_acc(s, d_ur, d_ui, nu1 * e1, nu1 * e2, nu1 * e3, d_bt, d_yr, d_yi, d_ae,
                                          chemflag ? ie : -1, inum, st);
                        }
            MDP_HIP(hipStreamSynchronize(st));
        }
        get_arg(ylist_r, d_yr, nu, st);
        get_arg(ylist_i, d_yi, nu, st);
        finish(st);
    });
}

void cpuSnapComputeFi(double* fatom, double* vatom, double* ylist_r, double* ylist_i, double* rootpqarray, double* rij,
                      double* wjelem, double* radelem, double rmin0, double rfac0, double rcutfac, int* map, int* aii,
                      int* ai, int* aj, int* ti, int* tj, int twojmax, int idxu_max, int inum, int anum, int ijnum,
                      int switchflag, int chemflag) {
    (void)anum;
    guarded([&] {
        if (ijnum <= 0) return;
        hipStream_t st = library_stream();
        const int ntypes = std::max(max_of(ti, ijnum, 1), max_of(tj, ijnum, 1));
        const int nelem = chemflag ? max_of(map + 1, ntypes, 0) + 1 : 1;
        SnapDev s = tables(twojmax, rootpqarray, nullptr, radelem, wjelem, map, ntypes, st);
        if (idxu_max != s.idxu_max) throw Error("idxu_max does not match twojmax");
        s.rcutfac = rcutfac;
        s.rfac0 = rfac0;
        s.rmin0 = rmin0;
        s.switchflag = switchflag;
        const size_t nu = static_cast<size_t>(inum) * idxu_max * nelem;
        const int fext = std::max(max_of(ai, ijnum), max_of(aj, ijnum)) + 1;
        const int vext = max_of(ai, ijnum) + 1;
        std::vector<long> yb;
        if (chemflag) {
            yb.resize(ijnum);
            for (int p = 0; p < ijnum; ++p) yb[p] = static_cast<long>(inum) * idxu_max * map[tj[p]];
        }
        double* d_yr = put_arg(0, ylist_r, nu, st);
        double* d_yi = put_arg(1, ylist_i, nu, st);
        double* d_rij = put_arg(2, rij, 3 * static_cast<size_t>(ijnum), st);
        int* d_aii = put(3, aii, ijnum, st);
        int* d_ai = put(4, ai, ijnum, st);
        int* d_aj = put(5, aj, ijnum, st);
        int* d_ti = put(6, ti, ijnum, st);
        int* d_tj = put(7, tj, ijnum, st);
        long* d_yb = chemflag ? put(8, yb.data(), yb.size(), st) : nullptr;
        double* d_de = dev<double>(9, 3 * static_cast<size_t>(ijnum));
        double* d_f = put_arg(10, fatom, 3 * static_cast<size_t>(fext), st);
        double* d_v = put_arg(11, vatom, 6 * static_cast<size_t>(vext), st);
        launch_dedr(s, d_yr, d_yi, d_rij, d_aii, d_ti, d_tj, d_yb, d_de, inum, ijnum, nullptr, st);
        if (g_vsoa >= 0)
            launch_tally_force_atomic_soa(d_f, d_v, d_de, d_rij, d_ai, d_aj, ijnum, g_vsoa, st);
        else
            launch_tally_force_atomic(d_f, d_v, d_de, d_rij, d_ai, d_aj, ijnum, st);
        get_arg(fatom, d_f, 3 * static_cast<size_t>(fext), st);
        get_arg(vatom, d_v, 6 * static_cast<size_t>(vext), st);
        finish(st);
    });
}

namespace {
void uij_common(double* Sr, double* Si, double* dr[3], double* di[3], double* rootpqarray, double* rij, double* wjelem,
                double* radelem, double rmin0, double rfac0, double rcutfac, int* ti, int* tj, int twojmax, int idxu_max,
                int ijnum, int switch_flag) {
    if (ijnum <= 0) return;
    hipStream_t st = library_stream();
    const int ntypes = std::max(max_of(ti, ijnum, 1), max_of(tj, ijnum, 1));
    SnapDev s = tables(twojmax, rootpqarray, nullptr, radelem, wjelem, nullptr, ntypes, st);
    if (idxu_max != s.idxu_max) throw Error("idxu_max does not match twojmax");
    s.rcutfac = rcutfac;
    s.rfac0 = rfac0;
    s.rmin0 = rmin0;
    s.switchflag = switch_flag;
    const size_t n = static_cast<size_t>(idxu_max) * ijnum;
    double* d_sr = out_arg(0, Sr, n);
    double* d_si = out_arg(1, Si, n);
    double* ddr[3];
    double* ddi[3];
    for (int k = 0; k < 3; ++k) {   // upstream leaves some derivative rows untouched: stage them
        ddr[k] = put_arg(2 + k, dr[k], n, st);
        ddi[k] = put_arg(5 + k, di[k], n, st);
    }
    double* d_rij = put_arg(8, rij, 3 * static_cast<size_t>(ijnum), st);
    int* d_ti = put(9, ti, ijnum, st);
    int* d_tj = put(10, tj, ijnum, st);
    launch_uij(s, d_sr, d_si, ddr, ddi, d_rij, d_ti, d_tj, ijnum, st);
    get_arg(Sr, d_sr, n, st);
    get_arg(Si, d_si, n, st);
    for (int k = 0; k < 3; ++k) {
        get_arg(dr[k], ddr[k], n, st);
        get_arg(di[k], ddi[k], n, st);
    }
    finish(st);
}
}  // namespace

void cpuComputeSij(double* Sr, double* Si, double* Srx, double* Six, double* Sry, double* Siy, double* Srz, double* Siz,
                   double* rootpqarray, double* rij, double* wjelem, double* radelem, double rmin0, double rfac0,
                   double rcutfac, int* idxu_block, int* ti, int* tj, int twojmax, int idxu_max, int ijnum,
                   int switch_flag) {
    (void)idxu_block;
    guarded([&] {
        double* dr[3] = {Srx, Sry, Srz};
        double* di[3] = {Six, Siy, Siz};
        uij_common(Sr, Si, dr, di, rootpqarray, rij, wjelem, radelem, rmin0, rfac0, rcutfac, ti, tj, twojmax, idxu_max,
                   ijnum, switch_flag);
    });
}

void cpuComputeUij(double* Sr, double* Si, double* dSr, double* dSi, double* rootpqarray, double* rij, double* wjelem,
                   double* radelem, double rmin0, double rfac0, double rcutfac, int* idxu_block, int* ti, int* tj,
                   int twojmax, int idxu_max, int ijnum, int switch_flag) {
    (void)idxu_block;
    guarded([&] {
        const size_t n = static_cast<size_t>(idxu_max) * ijnum;
        double* dr[3] = {dSr, dSr + n, dSr + 2 * n};
        double* di[3] = {dSi, dSi + n, dSi + 2 * n};
        uij_common(Sr, Si, dr, di, rootpqarray, rij, wjelem, radelem, rmin0, rfac0, rcutfac, ti, tj, twojmax, idxu_max,
                   ijnum, switch_flag);
    });
}

void cpuZeroUarraytot2(double* Stotr, double* Stoti, double wself, int* idxu_block, int* type, int* map, int* ai,
                       int wselfall_flag, int chemflag, int idxu_max, int nelements, int twojmax, int inum) {
    (void)idxu_block;
    guarded([&] {
        if (inum <= 0) return;
        hipStream_t st = library_stream();
        int* d_t = nullptr;
        int* d_ai = nullptr;
        int ntypes = 1;
        if (chemflag) {   // upstream reads ai[ii] and type[ai[ii]] only in this case
            const int text = max_of(ai, inum) + 1;
            ntypes = max_of(type, text, 1);
            d_ai = put(2, ai, inum, st);
            d_t = put(3, type, text, st);
        }
        SnapDev s = tables(twojmax, nullptr, nullptr, nullptr, nullptr, map, ntypes, st);
        if (idxu_max != s.idxu_max) throw Error("idxu_max does not match twojmax");
        const size_t nu = static_cast<size_t>(inum) * idxu_max * nelements;
        double* d_r = out_arg(0, Stotr, nu);
        double* d_i = out_arg(1, Stoti, nu);
        launch_zero_utot(s, d_r, d_i, wself, d_t, d_ai, wselfall_flag, chemflag, nelements, inum, st);
        get_arg(Stotr, d_r, nu, st);
        get_arg(Stoti, d_i, nu, st);
        finish(st);
    });
}

void cpuAddUarraytot(double* Stotr, double* Stoti, double* Sr, double* Si, int* map, int* ai, int* tj, int idxu_max,
                     int inum, int ijnum, int chemflag) {
    guarded([&] {
        if (inum <= 0 || ijnum <= 0) return;
        hipStream_t st = library_stream();
        const int ntypes = chemflag ? max_of(tj, ijnum, 1) : 1;
        const int nelem = chemflag ? max_of(map + 1, ntypes, 0) + 1 : 1;
        int J = 0;
        while (count_u(J) < idxu_max) ++J;
        if (count_u(J) != idxu_max) throw Error("idxu_max is not a U index-space size");
        SnapDev s = tables(J, nullptr, nullptr, nullptr, nullptr, chemflag ? map : nullptr, ntypes, st);
        const size_t nu = static_cast<size_t>(inum) * idxu_max * nelem, np = static_cast<size_t>(ijnum) * idxu_max;
        double* d_tr = put_arg(0, Stotr, nu, st);
        double* d_ti = put_arg(1, Stoti, nu, st);
        double* d_sr = put_arg(2, Sr, np, st);
        double* d_si = put_arg(3, Si, np, st);
        int* d_ai = put(4, ai, ijnum, st);
        int* d_tj = put(5, tj, ijnum, st);
        launch_add_utot(s, d_tr, d_ti, d_sr, d_si, d_ai, d_tj, chemflag, inum, ijnum, st);
        get_arg(Stotr, d_tr, nu, st);
        get_arg(Stoti, d_ti, nu, st);
        finish(st);
    });
}

void cpuComputeZi2(double* zlist_r, double* zlist_i, double* Stotr, double* Stoti, double* cglist, int* idxz,
                   int* idxu_block, int* idxcg_block, int twojmax, int idxu_max, int idxz_max, int nelements,
                   int bnorm_flag, int inum) {
    (void)idxz; (void)idxu_block; (void)idxcg_block;
    guarded([&] {
        if (inum <= 0) return;
        hipStream_t st = library_stream();
        SnapDev s = tables(twojmax, nullptr, cglist, nullptr, nullptr, nullptr, 1, st);
        if (idxu_max != s.idxu_max || idxz_max != s.idxz_max) throw Error("table sizes do not match twojmax");
        const int ne = nelements;
        const size_t nu = static_cast<size_t>(inum) * idxu_max * ne, nz = static_cast<size_t>(inum) * idxz_max * ne * ne;
        double* d_ur = put_arg(0, Stotr, nu, st);
        double* d_ui = put_arg(1, Stoti, nu, st);
        double* d_zr = out_arg(2, zlist_r, nz);
        double* d_zi = out_arg(3, zlist_i, nz);
        const long nu1 = static_cast<long>(inum) * idxu_max, nz1 = static_cast<long>(inum) * idxz_max;
        for (int e1 = 0; e1 < ne; ++e1)
            for (int e2 = 0; e2 < ne; ++e2) {
                const long o = nz1 * (e1 * ne + e2);
                launch_zi(s, d_ur, d_ui, nu1 * e1, nu1 * e2, d_zr + o, d_zi + o, inum, bnorm_flag, st);
            }
        get_arg(zlist_r, d_zr, nz, st);
        get_arg(zlist_i, d_zi, nz, st);
        finish(st);
    });
}

void cpuComputeBi2(double* blist, double* zlist_r, double* zlist_i, double* Stotr, double* Stoti, double* bzero,
                   int* ilist, int* type, int* map, int* idxb, int* idxu_block, int* idxz_block, int twojmax,
                   int idxb_max, int idxu_max, int idxz_max, int nelements, int bzero_flag, int wselfall_flag,
                   int chemflag, int inum) {
    (void)idxb; (void)idxu_block; (void)idxz_block;
    guarded([&] {
        if (inum <= 0) return;
        hipStream_t st = library_stream();
        const int text = max_of(ilist, inum) + 1;
        const int ntypes = max_of(type, text, 1);
        SnapDev s = tables(twojmax, nullptr, nullptr, nullptr, nullptr, map, ntypes, st);
        if (idxu_max != s.idxu_max || idxz_max != s.idxz_max || idxb_max != s.idxb_max)
            throw Error("table sizes do not match twojmax");
        const int ne = nelements;
        const size_t nu = static_cast<size_t>(inum) * idxu_max * ne, nz = static_cast<size_t>(inum) * idxz_max * ne * ne;
        const size_t nb = static_cast<size_t>(inum) * idxb_max * ne * ne * ne;
        double* d_zr = put_arg(0, zlist_r, nz, st);
        double* d_zi = put_arg(1, zlist_i, nz, st);
        double* d_ur = put_arg(2, Stotr, nu, st);
        double* d_ui = put_arg(3, Stoti, nu, st);
        double* d_bz = put(4, bzero, twojmax + 1, st);
        int* d_il = put(5, ilist, inum, st);
        int* d_t = put(6, type, text, st);
        double* d_b = out_arg(7, blist, nb);
        launch_bi2(s, d_b, d_zr, d_zi, d_ur, d_ui, d_bz, d_il, d_t, bzero_flag, wselfall_flag, chemflag, ne, inum, st);
        get_arg(blist, d_b, nb, st);
        finish(st);
    });
}

void cpuComputeDbidrj(double* dblist, double* zlist_r, double* zlist_i, double* dulist_r, double* dulist_i, int* idxb,
                      int* idxu_block, int* idxz_block, int* map, int* ai, int* tj, int twojmax, int idxb_max,
                      int idxu_max, int idxz_max, int nelements, int bnorm_flag, int chemflag, int inum, int ijnum) {
    (void)idxb; (void)idxu_block; (void)idxz_block;
    guarded([&] {
        if (inum <= 0 || ijnum <= 0) return;
        hipStream_t st = library_stream();
        const int ntypes = chemflag ? max_of(tj, ijnum, 1) : 1;
        SnapDev s = tables(twojmax, nullptr, nullptr, nullptr, nullptr, chemflag ? map : nullptr, ntypes, st);
        if (idxu_max != s.idxu_max || idxz_max != s.idxz_max || idxb_max != s.idxb_max)
            throw Error("table sizes do not match twojmax");
        const int ne = nelements;
        const size_t nz = static_cast<size_t>(inum) * idxz_max * ne * ne;
        const size_t ndu = 3 * static_cast<size_t>(ijnum) * idxu_max;
        const size_t ndb = static_cast<size_t>(ijnum) * idxb_max * 3 * ne * ne * ne;
        double* d_zr = put_arg(0, zlist_r, nz, st);
        double* d_zi = put_arg(1, zlist_i, nz, st);
        double* d_dur = put_arg(2, dulist_r, ndu, st);
        double* d_dui = put_arg(3, dulist_i, ndu, st);
        int* d_ai = put(4, ai, ijnum, st);
        int* d_tj = put(5, tj, ijnum, st);
        double* d_db = out_arg(6, dblist, ndb);
        MDP_HIP(hipMemsetAsync(d_db, 0, sizeof(double) * ndb, st));
        launch_dbidrj(s, d_db, d_zr, d_zi, d_dur, d_dui, d_ai, d_tj, bnorm_flag, chemflag, ne, inum, ijnum, st);
        get_arg(dblist, d_db, ndb, st);
        finish(st);
    });
}

void cpuSnapTallyBispectrum(double* bi, double* bispectrum, int* ilist, int* type, int inum, int ncoeff, int nperdim,
                            int ntype, int quadraticflag) {
    guarded([&] {
        hipStream_t st = library_stream();
        const size_t nbi = static_cast<size_t>(nperdim) * ntype;
        double* d_bi = out_arg(0, bi, nbi);
        if (nbi) MDP_HIP(hipMemsetAsync(d_bi, 0, sizeof(double) * nbi, st));
        if (inum > 0) {
            const int text = max_of(ilist, inum) + 1;
            double* d_b = put_arg(1, bispectrum, static_cast<size_t>(inum) * ncoeff, st);
            int* d_il = put(2, ilist, inum, st);
            int* d_t = put(3, type, text, st);
            launch_tally_bi(d_bi, d_b, d_il, d_t, inum, ncoeff, nperdim, quadraticflag, st);
        }
        get_arg(bi, d_bi, nbi, st);
        finish(st);
    });
}

void cpuSnapTallyBispectrumDeriv(double* db, double* bispectrum, double* dbdr, int* aii, int* ai, int* aj, int* ti,
                                 int inum, int ijnum, int ncoeff, int nperdim, int ntype, int quadraticflag) {
    guarded([&] {
        hipStream_t st = library_stream();
        const size_t ndb = static_cast<size_t>(inum) * 3 * nperdim * ntype;
        double* d_db = out_arg(0, db, ndb);
        if (ndb) MDP_HIP(hipMemsetAsync(d_db, 0, sizeof(double) * ndb, st));
        if (ijnum > 0) {
            double* d_b = put_arg(1, bispectrum, static_cast<size_t>(inum) * ncoeff, st);
            double* d_dbdr = put_arg(2, dbdr, static_cast<size_t>(ijnum) * 3 * ncoeff, st);
            int* d_aii = put(3, aii, ijnum, st);
            int* d_ai = put(4, ai, ijnum, st);
            int* d_aj = put(5, aj, ijnum, st);
            int* d_ti = put(6, ti, ijnum, st);
            launch_tally_bd(d_db, d_b, d_dbdr, d_aii, d_ai, d_aj, d_ti, inum, ijnum, ncoeff, nperdim, quadraticflag, st);
        }
        get_arg(db, d_db, ndb, st);
        finish(st);
    });
}

void cpuSnapTallyBispectrumVirial(double* bv, double* bispectrum, double* dbdr, double* rij, int* aii, int* ti, int inum,
                                  int ijnum, int ncoeff, int nperdim, int ntype, int quadraticflag) {
    guarded([&] {
        hipStream_t st = library_stream();
        const size_t nbv = 6 * static_cast<size_t>(nperdim) * ntype;
        double* d_bv = out_arg(0, bv, nbv);
        if (nbv) MDP_HIP(hipMemsetAsync(d_bv, 0, sizeof(double) * nbv, st));
        if (ijnum > 0) {
            double* d_b = put_arg(1, bispectrum, static_cast<size_t>(inum) * ncoeff, st);
            double* d_dbdr = put_arg(2, dbdr, static_cast<size_t>(ijnum) * 3 * ncoeff, st);
            double* d_rij = put_arg(3, rij, 3 * static_cast<size_t>(ijnum), st);
            int* d_aii = put(4, aii, ijnum, st);
            int* d_ti = put(5, ti, ijnum, st);
            launch_tally_bv(d_bv, d_b, d_dbdr, d_rij, d_aii, d_ti, inum, ijnum, ncoeff, nperdim, quadraticflag, st);
        }
        get_arg(bv, d_bv, nbv, st);
        finish(st);
    });
}


}  // extern "C"

// ---- backend-3 entry points of the C++ core dispatch (src/C++/Common/mdpcore.h) -----------------
// Device arrays in and out, same argument lists as the dispatch's backend branches (without backend).
// The large arrays are used in place; the small index and table arrays are read back once to drive the
// same host-side set-up as the cpuSnap.h entry points above.

namespace {
// device -> host copy of a small argument array (after the caller's device work has finished)
template <class T>
std::vector<T> fetch(const T* d, size_t n) {
    std::vector<T> h(std::max<size_t>(n, 1));
    if (n) MDP_HIP(hipMemcpy(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost));
    return h;
}
// device-I/O scope for one backend-3 call
struct DevIo {
    explicit DevIo(int vsoa = -1) { g_devio = true; g_vsoa = vsoa; }
    ~DevIo() { g_devio = false; g_vsoa = -1; }
};
// runs the read-back; returns false (with the library error set) when it fails
template <class F>
bool fetch_args(F&& f) {
    bool ok = false;
    guarded([&] {
        MDP_HIP(hipDeviceSynchronize());
        f();
        ok = true;
    });
    return ok;
}
int max_type(const std::vector<int>& a, size_t n) { return std::max(1, max_of(a.data(), static_cast<long>(n), 1)); }
}  // namespace

extern "C" {

// SnapComputeUi backend 3 (mdpcore.h:3093-3116; cpuSnapComputeUi src/C++/Core/cpuSnap4.cpp:11-191)
void mdp_hipSnapComputeUi(double* Stotr, double* Stoti, double* rootpqarray, double* rij, double* wjelem,
                          double* radelem, double rmin0, double rfac0, double rcutfac, int* idxu_block, int* map, int* ai,
                          int* ti, int* tj, int twojmax, int idxu_max, int inum, int ijnum, int switchflag, int chemflag) {
    (void)idxu_block;
    std::vector<int> h_ai, h_ti, h_tj, h_map;
    std::vector<double> h_rp, h_wj, h_rad;
    if (!fetch_args([&] {
            const size_t np = static_cast<size_t>(std::max(ijnum, 0));
            h_ai = fetch(ai, np);
            h_ti = fetch(ti, np);
            h_tj = fetch(tj, np);
            const int nt = std::max(max_type(h_ti, np), max_type(h_tj, np));
            h_map = fetch(map, nt + 1);
            h_rp = fetch(rootpqarray, static_cast<size_t>(twojmax + 2) * (twojmax + 2));
            h_wj = fetch(wjelem, nt + 1);
            h_rad = fetch(radelem, nt + 1);
        }))
        return;
    DevIo io;
    cpuSnapComputeUi(Stotr, Stoti, h_rp.data(), rij, h_wj.data(), h_rad.data(), rmin0, rfac0, rcutfac, h_map.data(),
                     h_ai.data(), h_ti.data(), h_tj.data(), twojmax, idxu_max, inum, ijnum, switchflag, chemflag);
}

// AddWself2Ui backend 3 (mdpcore.h:3118-3140; src/C++/Core/cpuSnap4.cpp:193-247; ai is unused there too)
void mdp_hipAddWself2Ui(double* ulisttot_r, double* ulisttot_i, double wself, int* idxu_block, int* type, int* map,
                        int* ai, int wselfall_flag, int chemflag, int idxu_max, int nelements, int twojmax, int inum) {
    (void)ai;
    std::vector<int> h_type, h_map;
    if (!fetch_args([&] {
            h_type = fetch(type, static_cast<size_t>(std::max(inum, 0)));
            h_map = fetch(map, max_type(h_type, std::max(inum, 0)) + 1);
        }))
        return;
    DevIo io;
    cpuAddWself2Ui(ulisttot_r, ulisttot_i, wself, idxu_block, h_type.data(), h_map.data(), wselfall_flag, chemflag,
                   idxu_max, nelements, twojmax, inum);
}

// SnapComputeEi backend 3 (mdpcore.h:3167-3194; src/C++/Core/cpuSnap4.cpp:250-424).  The element of atom i
// is map[type[i]] there whatever the flags (cpuSnap4.cpp:277-278), so the shared body runs with the
// element mapping on; quadraticflag is unused there.
void mdp_hipSnapComputeEi(double* eatom, double* ulisttot_r, double* ulisttot_i, double* cglist, double* bzero,
                          double* coeffelem, int* ilist, int* map, int* type, int* idxb, int* idxcg_block,
                          int* idxu_block, int twojmax, int idxb_max, int idxu_max, int nelements, int ncoeffall,
                          int bnorm_flag, int bzero_flag, int wselfall_flag, int quadraticflag, int inum) {
    (void)quadraticflag;
    std::vector<int> h_il, h_type, h_map;
    std::vector<double> h_cg, h_bz, h_c;
    if (!fetch_args([&] {
            const size_t n = static_cast<size_t>(std::max(inum, 0));
            h_il = fetch(ilist, n);
            h_type = fetch(type, n);
            const int nt = max_type(h_type, n);
            h_map = fetch(map, nt + 1);
            const int rows = max_of(h_map.data() + 1, nt, 0) + 1;
            h_cg = fetch(cglist, static_cast<size_t>(index_for(twojmax).idxcg_max));
            h_bz = fetch(bzero, twojmax + 1);
            h_c = fetch(coeffelem, static_cast<size_t>(ncoeffall) * rows);
        }))
        return;
    DevIo io;
    cpuSnapComputeEi(eatom, ulisttot_r, ulisttot_i, h_cg.data(), h_bz.data(), h_c.data(), h_il.data(), h_map.data(),
                     h_type.data(), idxb, idxcg_block, idxu_block, twojmax, idxb_max, idxu_max, nelements, ncoeffall,
                     bnorm_flag, bzero_flag, wselfall_flag, 1, inum);
}

// SnapComputeYi backend 3, coefficient form (mdpcore.h:3221-3244; src/C++/Core/cpuSnap4.cpp:697-796):
// element of atom i is map[type[i]]
void mdp_hipSnapComputeYi(double* ylist_r, double* ylist_i, double* ulisttot_r, double* ulisttot_i, double* cglist,
                          double* betalist, int* map, int* type, int* idxz, int* idxb_block, int* idxu_block,
                          int* idxcg_block, int twojmax, int idxb_max, int idxu_max, int idxz_max, int nelements,
                          int ncoeffall, int bnorm_flag, int inum) {
    std::vector<int> h_type, h_map;
    std::vector<double> h_cg, h_c;
    if (!fetch_args([&] {
            const size_t n = static_cast<size_t>(std::max(inum, 0));
            h_type = fetch(type, n);
            const int nt = max_type(h_type, n);
            h_map = fetch(map, nt + 1);
            const int rows = max_of(h_map.data() + 1, nt, 0) + 1;
            h_cg = fetch(cglist, static_cast<size_t>(index_for(twojmax).idxcg_max));
            h_c = fetch(betalist, static_cast<size_t>(ncoeffall) * rows);
        }))
        return;
    DevIo io;
    cpuSnapComputeYi(ylist_r, ylist_i, ulisttot_r, ulisttot_i, h_cg.data(), h_c.data(), h_map.data(), h_type.data(),
                     idxz, idxb_block, idxu_block, idxcg_block, twojmax, idxb_max, idxu_max, idxz_max, nelements,
                     ncoeffall, bnorm_flag, 1, inum);
}

// SnapComputeYi backend 3, per-atom form (mdpcore.h:3196-3219; src/C++/Core/cpuSnap4.cpp:601-690): every
// atom carries its own bispectrum weights, betalist[(elemblock * idxb_max + jjb) * inum + atom] (no constant term)
void mdp_hipSnapComputeYiPerAtom(double* ylist_r, double* ylist_i, double* ulisttot_r, double* ulisttot_i,
                                 double* cglist, double* betalist, int* idxz, int* idxb_block, int* idxu_block,
                                 int* idxcg_block, int twojmax, int idxb_max, int idxu_max, int idxz_max,
                                 int nelements, int bnorm_flag, int inum) {
    (void)idxz; (void)idxb_block; (void)idxu_block; (void)idxcg_block; (void)idxz_max;
    std::vector<double> h_cg;
    if (!fetch_args([&] { h_cg = fetch(cglist, static_cast<size_t>(index_for(twojmax).idxcg_max)); })) return;
    guarded([&] {
        if (inum <= 0) return;
        if (nelements < 1) throw Error("SnapComputeYi (per-atom weights): nelements must be positive");
        hipStream_t st = library_stream();
        const SnapDev& s = tables(twojmax, nullptr, h_cg.data(), nullptr, nullptr, nullptr, 1, st);
        if (idxu_max != s.idxu_max || idxb_max != s.idxb_max) throw Error("table sizes do not match twojmax");
        std::vector<int> jjb;
        std::vector<double> fac;
        y_triple_weights(index_for(twojmax), bnorm_flag, jjb, fac);
        const int ntrip = static_cast<int>(jjb.size());
        int* d_jjb = put(0, jjb.data(), jjb.size(), st);
        double* d_fac = put(1, fac.data(), fac.size(), st);
        double* d_bt = dev<double>(2, static_cast<size_t>(ntrip) * inum);
        const size_t nu = static_cast<size_t>(inum) * idxu_max;
        if (nelements == 1) {
            MDP_HIP(hipMemsetAsync(ylist_r, 0, sizeof(double) * nu, st));
            MDP_HIP(hipMemsetAsync(ylist_i, 0, sizeof(double) * nu, st));
            launch_beta_peratom(betalist, d_jjb, d_fac, ntrip, inum, d_bt, st);
            if (!launch_yi_peratom(s, ulisttot_r, ulisttot_i, d_bt, ylist_r, ylist_i, inum, st))
                throw Error("SnapComputeYi (per-atom weights): twojmax not compiled");
            finish(st);
            return;
        }
        // several elements: combination (e1, e2, e3) of coupling triple (j1, j2, j) reads the weights of element
        // block (e1, e2, e3) when j >= j1, (e3, e2, e1) when j1 > j >= j2, (e2, e3, e1) when j < j2
        const int ne = nelements, J = twojmax;
        std::vector<int> branch;
        for (int j1 = 0; j1 <= J; ++j1)
            for (int j2 = 0; j2 <= j1; ++j2)
                for (int j = j1 - j2; j <= std::min(J, j1 + j2); j += 2) branch.push_back(j >= j1 ? 0 : (j >= j2 ? 1 : 2));
        if (static_cast<int>(branch.size()) != ntrip) throw Error("coupling triple count mismatch");
        std::vector<int> idx(static_cast<size_t>(ne) * ne * ne * ntrip);
        for (int e1 = 0; e1 < ne; ++e1)
            for (int e2 = 0; e2 < ne; ++e2)
                for (int e3 = 0; e3 < ne; ++e3)
                    for (int t = 0; t < ntrip; ++t) {
                        const int blk = branch[t] == 0 ? (e1 * ne + e2) * ne + e3
                                        : branch[t] == 1 ? (e3 * ne + e2) * ne + e1
                                                         : (e2 * ne + e3) * ne + e1;
                        idx[((static_cast<size_t>(e1) * ne + e2) * ne + e3) * ntrip + t] = blk * idxb_max + jjb[t];
                    }
        int* d_idx = put(3, idx.data(), idx.size(), st);
        const long nu1 = static_cast<long>(inum) * idxu_max;
        const size_t nall = nu * ne;
        MDP_HIP(hipMemsetAsync(ylist_r, 0, sizeof(double) * nall, st));
        MDP_HIP(hipMemsetAsync(ylist_i, 0, sizeof(double) * nall, st));
        for (int e1 = 0; e1 < ne; ++e1)
            for (int e2 = 0; e2 < ne; ++e2)
                for (int e3 = 0; e3 < ne; ++e3) {
                    const int* di = d_idx + ((static_cast<size_t>(e1) * ne + e2) * ne + e3) * ntrip;
                    launch_beta_peratom(betalist, di, d_fac, ntrip, inum, d_bt, st);
                    launch_yi_acc(s, ulisttot_r, ulisttot_i, nu1 * e1, nu1 * e2, nu1 * e3, d_bt, ylist_r, ylist_i,
                                  nullptr, -1, inum, st, true);
                }
        finish(st);
    });
}

// SnapComputeFi backend 3 (mdpcore.h:3246-3269; src/C++/Core/cpuSnap4.cpp:798-1210): forces to ai/aj,
// virial in the SoA layout vatom[k * anum + atom] with -r(x)dE/2 to both atoms
void mdp_hipSnapComputeFi(double* fatom, double* vatom, double* ylist_r, double* ylist_i, double* rootpqarray,
                          double* rij, double* wjelem, double* radelem, double rmin0, double rfac0, double rcutfac,
                          int* idxu_block, int* map, int* aii, int* ai, int* aj, int* ti, int* tj, int twojmax,
                          int idxu_max, int inum, int anum, int ijnum, int switchflag, int chemflag) {
    (void)idxu_block;
    std::vector<int> h_aii, h_ai, h_aj, h_ti, h_tj, h_map;
    std::vector<double> h_rp, h_wj, h_rad;
    if (!fetch_args([&] {
            const size_t np = static_cast<size_t>(std::max(ijnum, 0));
            h_aii = fetch(aii, np);
            h_ai = fetch(ai, np);
            h_aj = fetch(aj, np);
            h_ti = fetch(ti, np);
            h_tj = fetch(tj, np);
            const int nt = std::max(max_type(h_ti, np), max_type(h_tj, np));
            h_map = fetch(map, nt + 1);
            h_rp = fetch(rootpqarray, static_cast<size_t>(twojmax + 2) * (twojmax + 2));
            h_wj = fetch(wjelem, nt + 1);
            h_rad = fetch(radelem, nt + 1);
            if (max_of(h_ai.data(), static_cast<long>(np)) >= anum || max_of(h_aj.data(), static_cast<long>(np)) >= anum)
                throw Error("SnapComputeFi: a pair atom index is not below anum");
        }))
        return;
    DevIo io(anum);
    cpuSnapComputeFi(fatom, vatom, ylist_r, ylist_i, h_rp.data(), rij, h_wj.data(), h_rad.data(), rmin0, rfac0,
                     rcutfac, h_map.data(), h_aii.data(), h_ai.data(), h_aj.data(), h_ti.data(), h_tj.data(), twojmax,
                     idxu_max, inum, anum, ijnum, switchflag, chemflag);
}

// ---- the descriptor chain of the upstream fitting driver (snappotential.cpp:762-910) --------------

// NeighborPairList backend 3, typed cutoffs (mdpcore.h:3830-3851; src/C++/Core/cpuSnap.cpp:1883-1905)
void mdp_hipNeighPairList(int* pairnum, int* pairlist, double* x, double* rcutsq, int* ilist, int* neighlist,
                          int* neighnum, int* atomtype, int* alist, int inum, int jnum, int dim, int ntypes) {
    guarded([&] {
        if (inum <= 0) return;
        if (jnum <= 0 || dim != 3 || ntypes < 1) throw Error("NeighPairList: needs jnum > 0, dim = 3, ntypes >= 1");
        hipStream_t st = library_stream();
        launch_pair_list_strided(pairnum, pairlist, x, rcutsq, ilist, neighlist, neighnum, atomtype, alist, inum, jnum,
                                 dim, ntypes, st);
        finish(st);
    });
}

// NeighborPairList backend 3, one cutoff (mdpcore.h:3807-3828; src/C++/Core/cpuSnap.cpp:1833-1853)
void mdp_hipNeighPairListCut(int* pairnum, int* pairlist, double* x, double rcutsq, int* ilist, int* neighlist,
                             int* neighnum, int inum, int jnum, int dim) {
    guarded([&] {
        if (inum <= 0) return;
        if (jnum <= 0 || dim != 3) throw Error("NeighPairList: needs jnum > 0 and dim = 3");
        hipStream_t st = library_stream();
        launch_pair_list_strided_cut(pairnum, pairlist, x, rcutsq, ilist, neighlist, neighnum, inum, jnum, dim, st);
        finish(st);
    });
}

// Cumsum backend 3 (mdpcore.h:1114-1131; cpuCumsum src/C++/Core/cpuArrayOperations.cpp:22-27)
void mdp_hipCumsum(int* d_out, int* d_in, int* d_sums, int* d_incr, int length) {
    (void)d_sums; (void)d_incr;
    guarded([&] {
        if (length <= 0) return;
        hipStream_t st = library_stream();
        const int n = length - 1;
        int* d_tmp = dev<int>(0, scan_tmp_ints(std::max(n, 1)));
        exclusive_scan(d_in, d_out, n, d_tmp, st);   // out[0..n], out[n] = total
        finish(st);
    });
}

// AddUarraytot backend 3, rij form of the per-pair route (mdpcore.h:2911-2938; src/C++/Core/cpuSnap.cpp:358-395):
// Utot[(ii ne + jelem) idxu_max + u] += sfac(r) w_j U[k idxu_max + u] over the pairs k of each listed atom
void mdp_hipAddUarraytotRij(double* ulisttot_r, double* ulisttot_i, double* ulist_r, double* ulist_i, double* rij,
                            double* wjelem, double* radelem, double rmin0, double rcutfac, int* idxu_block, int* ilist,
                            int* type, int* pairnum, int* pairnumsum, int* map, int* tj, int twojmax, int idxu_max,
                            int nelements, int inum, int switch_flag, int chemflag) {
    (void)idxu_block;
    std::vector<int> h_il, h_type, h_pn, h_ps, h_tj, h_map;
    if (!fetch_args([&] {
            const size_t n = static_cast<size_t>(std::max(inum, 0));
            h_il = fetch(ilist, n);
            h_type = fetch(type, static_cast<size_t>(max_of(h_il.data(), static_cast<long>(n)) + 1));
            h_pn = fetch(pairnum, n);
            h_ps = fetch(pairnumsum, n);
            long np = 0;
            for (size_t ii = 0; ii < n; ++ii) np = std::max<long>(np, static_cast<long>(h_ps[ii]) + h_pn[ii]);
            h_tj = fetch(tj, static_cast<size_t>(np));
            h_map = fetch(map, std::max(max_type(h_tj, h_tj.size()), max_type(h_type, h_type.size())) + 1);
        }))
        return;
    guarded([&] {
        if (inum <= 0) return;
        if (idxu_max != index_for(twojmax).idxu_max) throw Error("AddUarraytot: idxu_max does not match twojmax");
        for (int ii = 0; ii < inum; ++ii)
            if (h_il[ii] < 0 || h_pn[ii] < 0 || h_ps[ii] < 0) throw Error("AddUarraytot: negative list entry");
        for (int t : h_tj)
            if (t < 0 || (chemflag && (h_map[t] < 0 || h_map[t] >= nelements)))
                throw Error("AddUarraytot: pair type or element out of range");
        for (int t : h_type)
            if (t < 0) throw Error("AddUarraytot: negative atom type");
        hipStream_t st = library_stream();
        double* sfw = dev<double>(0, std::max<size_t>(h_tj.size(), 1));
        launch_add_utot_rij(ulisttot_r, ulisttot_i, ulist_r, ulist_i, rij, wjelem, radelem, rmin0, rcutfac, ilist, type,
                            pairnum, pairnumsum, map, tj, idxu_max, nelements, inum, switch_flag, chemflag, sfw, st);
        finish(st);
    });
}

// CellList3D backend 3 (mdpcore.h:1332-1349; src/C++/Core/cpuNeighborList3D.cpp:334-356): cell of every atom
void mdp_hipCellList3D(int* clist, double* x, double* eta1, double* eta2, double* eta3, double* s2rmap, int* nc,
                       int inum, int natom, int dim) {
    std::vector<int> h_nc;
    if (!fetch_args([&] { h_nc = fetch(nc, 3); })) return;
    guarded([&] {
        if (dim != 3) throw Error("CellList3D: dim must be 3");
        if (natom <= 0) return;
        if (h_nc[0] < 1 || h_nc[1] < 1 || h_nc[2] < 1) throw Error("CellList3D: cell counts must be positive");
        hipStream_t st = library_stream();
        launch_cell_list3d(clist, x, eta1, eta2, eta3, s2rmap, h_nc.data(), inum, natom, st);
        finish(st);
    });
}

// Cell2AtomList backend 3 (mdpcore.h:1351-1368; src/C++/Core/cpuNeighborList2D.cpp:298-308): atoms sorted by
// cell (stable), the first entry of each cell and the sorted cell ids.  dtemp is not used (library scratch).
void mdp_hipCell2AtomList(int* c2alist, int* c2anumsum, int* c2anum, int* clist, int* dtemp, int natom, int ncell) {
    (void)dtemp;
    std::vector<int> h_cl;
    if (!fetch_args([&] { h_cl = fetch(clist, static_cast<size_t>(std::max(natom, 0))); })) return;
    guarded([&] {
        if (ncell <= 0 || natom <= 0) return;
        for (int c : h_cl)
            if (c < 0 || c >= ncell) throw Error("Cell2AtomList: a cell id is outside [0, ncell)");
        hipStream_t st = library_stream();
        launch_cell2atom_sorted(c2alist, c2anumsum, c2anum, clist, natom, ncell,
                                [](size_t bytes) -> void* { return dev<unsigned char>(0, bytes); }, st);
        finish(st);
    });
}

// FullNeighborList3D backend 3, cell-list form (mdpcore.h:1432-1453; cpuNeighborList3D.cpp:361-412):
// neighbours of atom i = alist[ii] in the 27 cells around its own, neighlist[l + jnum * i], count in neighnum[ii]
void mdp_hipFullNeighborList3DCell(int* neighlist, int* neighnum, double* x, double* rcutsq, int* alist, int* clist,
                                   int* c2alist, int* c2anumsum, int* nc, int inum, int jnum, int dim) {
    std::vector<int> h_nc;
    std::vector<double> h_rc;
    if (!fetch_args([&] {
            h_nc = fetch(nc, 3);
            h_rc = fetch(rcutsq, 1);
        }))
        return;
    guarded([&] {
        if (dim != 3) throw Error("FullNeighborList3D: dim must be 3");
        if (inum <= 0) return;
        if (jnum < 1) throw Error("FullNeighborList3D: jnum must be positive");
        hipStream_t st = library_stream();
        launch_full_neigh_cell(neighlist, neighnum, x, h_rc[0], alist, clist, c2alist, c2anumsum, h_nc.data(), inum,
                               jnum, st);
        finish(st);
    });
}

// FullNeighborList3D backend 3, all-pairs form (mdpcore.h:1390-1408; cpuNeighborList3D.cpp:417-441)
void mdp_hipFullNeighborList3D(int* neighlist, int* neighnum, double* x, double* rcutsq, int anum, int inum, int jnum,
                               int dim) {
    std::vector<double> h_rc;
    if (!fetch_args([&] { h_rc = fetch(rcutsq, 1); })) return;
    guarded([&] {
        if (dim != 3) throw Error("FullNeighborList3D: dim must be 3");
        if (inum <= 0) return;
        if (jnum < 1) throw Error("FullNeighborList3D: jnum must be positive");
        if (anum < inum) throw Error("FullNeighborList3D: anum must be at least inum");
        hipStream_t st = library_stream();
        launch_full_neigh_all(neighlist, neighnum, x, h_rc[0], anum, inum, jnum, st);
        finish(st);
    });
}

// NeighborPairs backend 3 (mdpcore.h:3853-3875; src/C++/Core/cpuSnap.cpp:1912-1935)
void mdp_hipNeighPairs(double* xij, double* x, int* aii, int* ai, int* aj, int* ti, int* tj, int* pairnum,
                       int* pairlist, int* pairnumsum, int* ilist, int* atomtype, int* alist, int inum, int jnum,
                       int dim) {
    guarded([&] {
        if (inum <= 0) return;
        if (jnum <= 0 || dim < 1) throw Error("NeighPairs: needs jnum > 0 and dim >= 1");
        hipStream_t st = library_stream();
        launch_pair_arrays_strided(xij, x, aii, ai, aj, ti, tj, pairnum, pairlist, pairnumsum, ilist, atomtype, alist,
                                   inum, jnum, dim, st);
        finish(st);
    });
}

// ComputeSij backend 3 (mdpcore.h:2837-2861; src/C++/Core/cpuSnap2.cpp:23-233)
void mdp_hipComputeSij(double* Sr, double* Si, double* Srx, double* Six, double* Sry, double* Siy, double* Srz,
                       double* Siz, double* rootpqarray, double* rij, double* wjelem, double* radelem, double rmin0,
                       double rfac0, double rcutfac, int* idxu_block, int* ti, int* tj, int twojmax, int idxu_max,
                       int ijnum, int switchflag) {
    std::vector<int> h_ti, h_tj;
    std::vector<double> h_rp, h_wj, h_rad;
    if (!fetch_args([&] {
            const size_t np = static_cast<size_t>(std::max(ijnum, 0));
            h_ti = fetch(ti, np);
            h_tj = fetch(tj, np);
            const int nt = std::max(max_type(h_ti, np), max_type(h_tj, np));
            h_rp = fetch(rootpqarray, static_cast<size_t>(twojmax + 2) * (twojmax + 2));
            h_wj = fetch(wjelem, nt + 1);
            h_rad = fetch(radelem, nt + 1);
        }))
        return;
    DevIo io;
    cpuComputeSij(Sr, Si, Srx, Six, Sry, Siy, Srz, Siz, h_rp.data(), rij, h_wj.data(), h_rad.data(), rmin0, rfac0,
                  rcutfac, idxu_block, h_ti.data(), h_tj.data(), twojmax, idxu_max, ijnum, switchflag);
}

// ZeroUarraytot2 backend 3 (mdpcore.h:2887-2909; src/C++/Core/cpuSnap2.cpp:235-266)
void mdp_hipZeroUarraytot2(double* ulisttot_r, double* ulisttot_i, double wself, int* idxu_block, int* type, int* map,
                           int* ai, int wselfall_flag, int chemflag, int idxu_max, int nelements, int twojmax,
                           int inum) {
    std::vector<int> h_ai, h_type, h_map;
    if (!fetch_args([&] {
            int nt = 1;
            if (chemflag && inum > 0) {   // ai[ii] and type[ai[ii]] are read only in this case
                h_ai = fetch(ai, inum);
                h_type = fetch(type, max_of(h_ai.data(), inum) + 1);
                nt = max_type(h_type, h_type.size());
            }
            h_map = fetch(map, nt + 1);
        }))
        return;
    DevIo io;
    cpuZeroUarraytot2(ulisttot_r, ulisttot_i, wself, idxu_block, chemflag ? h_type.data() : type, h_map.data(),
                      chemflag ? h_ai.data() : ai, wselfall_flag, chemflag, idxu_max, nelements, twojmax, inum);
}

// AddUarraytot backend 3, (map, ai, tj) form (mdpcore.h:2964-2986; src/C++/Core/cpuSnap2.cpp:322-347)
void mdp_hipAddUarraytot(double* ulisttot_r, double* ulisttot_i, double* ulist_r, double* ulist_i, int* map, int* ai,
                         int* tj, int idxu_max, int inum, int ijnum, int chemflag) {
    std::vector<int> h_ai, h_tj, h_map;
    if (!fetch_args([&] {
            const size_t np = static_cast<size_t>(std::max(ijnum, 0));
            h_ai = fetch(ai, np);
            h_tj = fetch(tj, np);
            h_map = fetch(map, (chemflag ? max_type(h_tj, np) : 1) + 1);
        }))
        return;
    DevIo io;
    cpuAddUarraytot(ulisttot_r, ulisttot_i, ulist_r, ulist_i, h_map.data(), h_ai.data(), h_tj.data(), idxu_max, inum,
                    ijnum, chemflag);
}

// AddUarraytot backend 3, (pairnum, pairnumsum) form (mdpcore.h:2940-2962; src/C++/Core/cpuSnap2.cpp:269-318)
void mdp_hipAddUarraytotPairs(double* ulisttot_r, double* ulisttot_i, double* ulist_r, double* ulist_i, int* pairnum,
                              int* pairnumsum, int* map, int* tj, int idxu_max, int nelements, int inum, int ijnum,
                              int chemflag) {
    (void)nelements;
    guarded([&] {
        if (inum <= 0 || idxu_max <= 0) return;
        hipStream_t st = library_stream();
        launch_add_utot_pairs(ulisttot_r, ulisttot_i, ulist_r, ulist_i, pairnum, pairnumsum, map, tj, idxu_max, inum,
                              ijnum, chemflag, st);
        finish(st);
    });
}

// ComputeZi2 backend 3 (mdpcore.h:3011-3033; src/C++/Core/cpuSnap2.cpp:364-438)
void mdp_hipComputeZi2(double* zlist_r, double* zlist_i, double* ulisttot_r, double* ulisttot_i, double* cglist,
                       int* idxz, int* idxu_block, int* idxcg_block, int twojmax, int idxu_max, int idxz_max,
                       int nelements, int bnorm_flag, int inum) {
    std::vector<double> h_cg;
    if (!fetch_args([&] { h_cg = fetch(cglist, static_cast<size_t>(index_for(twojmax).idxcg_max)); })) return;
    DevIo io;
    cpuComputeZi2(zlist_r, zlist_i, ulisttot_r, ulisttot_i, h_cg.data(), idxz, idxu_block, idxcg_block, twojmax,
                  idxu_max, idxz_max, nelements, bnorm_flag, inum);
}

// ComputeBi2 backend 3 (mdpcore.h:3064-3090; src/C++/Core/cpuSnap2.cpp:440-545)
void mdp_hipComputeBi2(double* blist, double* zlist_r, double* zlist_i, double* ulisttot_r, double* ulisttot_i,
                       double* bzero, int* ilist, int* type, int* map, int* idxb, int* idxu_block, int* idxz_block,
                       int twojmax, int idxb_max, int idxu_max, int idxz_max, int nelements, int bzero_flag,
                       int wselfall_flag, int chemflag, int inum) {
    std::vector<int> h_il, h_type, h_map;
    std::vector<double> h_bz;
    if (!fetch_args([&] {
            const size_t n = static_cast<size_t>(std::max(inum, 0));
            h_il = fetch(ilist, n);
            h_type = fetch(type, static_cast<size_t>(max_of(h_il.data(), static_cast<long>(n)) + 1));
            h_map = fetch(map, max_type(h_type, h_type.size()) + 1);
            h_bz = fetch(bzero, twojmax + 1);
        }))
        return;
    DevIo io;
    cpuComputeBi2(blist, zlist_r, zlist_i, ulisttot_r, ulisttot_i, h_bz.data(), h_il.data(), h_type.data(),
                  h_map.data(), idxb, idxu_block, idxz_block, twojmax, idxb_max, idxu_max, idxz_max, nelements,
                  bzero_flag, wselfall_flag, chemflag, inum);
}

// ComputeDbidrj backend 3, (inum, ijnum) form (mdpcore.h:3395-3420; src/C++/Core/cpuSnap2.cpp:826-1050)
void mdp_hipComputeDbidrj(double* dblist, double* zlist_r, double* zlist_i, double* dulist_r, double* dulist_i,
                          int* idxb, int* idxu_block, int* idxz_block, int* map, int* ai, int* tj, int twojmax,
                          int idxb_max, int idxu_max, int idxz_max, int nelements, int bnorm_flag, int chemflag,
                          int inum, int ijnum) {
    std::vector<int> h_ai, h_tj, h_map;
    if (!fetch_args([&] {
            const size_t np = static_cast<size_t>(std::max(ijnum, 0));
            h_ai = fetch(ai, np);
            h_tj = fetch(tj, np);
            h_map = fetch(map, (chemflag ? max_type(h_tj, np) : 1) + 1);
        }))
        return;
    DevIo io;
    cpuComputeDbidrj(dblist, zlist_r, zlist_i, dulist_r, dulist_i, idxb, idxu_block, idxz_block, h_map.data(),
                     h_ai.data(), h_tj.data(), twojmax, idxb_max, idxu_max, idxz_max, nelements, bnorm_flag, chemflag,
                     inum, ijnum);
}

// ComputeDbidrj backend 3, (ai, aj, ti, tj) form (mdpcore.h:3366-3393; src/C++/Core/cpuSnap.cpp:1050-1266):
// Z atom-major (zlist[(iT + ne^2 i) idxz_max + jjz]), dU and dB pair-major with x, y, z innermost
void mdp_hipComputeDbidrjAtomMajor(double* dblist, double* zlist_r, double* zlist_i, double* dulist_r,
                                   double* dulist_i, int* idxb, int* idxu_block, int* idxz_block, int* map, int* ai,
                                   int* aj, int* ti, int* tj, int twojmax, int idxb_max, int idxu_max, int idxz_max,
                                   int nelements, int bnorm_flag, int chemflag, int ijnum) {
    (void)idxb; (void)idxu_block; (void)idxz_block; (void)aj; (void)ti;
    std::vector<int> h_ai, h_tj, h_map;
    if (!fetch_args([&] {
            const size_t np = static_cast<size_t>(std::max(ijnum, 0));
            h_ai = fetch(ai, np);
            h_tj = fetch(tj, np);
            h_map = fetch(map, (chemflag ? max_type(h_tj, np) : 1) + 1);
        }))
        return;
    guarded([&] {
        if (ijnum <= 0) return;
        if (nelements < 1) throw Error("ComputeDbidrj: nelements must be positive");
        for (int p = 0; p < ijnum; ++p)
            if (h_ai[p] < 0 || (chemflag && h_tj[p] < 0)) throw Error("ComputeDbidrj: negative pair index");
        if (chemflag)
            for (int t : h_tj)
                if (h_map[t] < 0 || h_map[t] >= nelements) throw Error("ComputeDbidrj: element map out of range");
        hipStream_t st = library_stream();
        const int ntypes = chemflag ? max_type(h_tj, h_tj.size()) : 1;
        const SnapDev& s = tables(twojmax, nullptr, nullptr, nullptr, nullptr, chemflag ? h_map.data() : nullptr,
                                  ntypes, st);
        if (idxu_max != s.idxu_max || idxz_max != s.idxz_max || idxb_max != s.idxb_max)
            throw Error("table sizes do not match twojmax");
        launch_dbidrj_atom_major(s, dblist, zlist_r, zlist_i, dulist_r, dulist_i, ai, tj, nelements, bnorm_flag,
                                 chemflag, ijnum, st);
        finish(st);
    });
}

// SnapTallyBispectrum backend 3 (mdpcore.h:3750-3767; src/C++/Core/cpuSnap2.cpp:1297-1322)
void mdp_hipSnapTallyBispectrum(double* bi, double* bispectrum, int* ilist, int* type, int inum, int ncoeff,
                                int nperdim, int ntype, int quadraticflag) {
    std::vector<int> h_il, h_type;
    if (!fetch_args([&] {
            const size_t n = static_cast<size_t>(std::max(inum, 0));
            h_il = fetch(ilist, n);
            h_type = fetch(type, n ? static_cast<size_t>(max_of(h_il.data(), static_cast<long>(n)) + 1) : 0);
        }))
        return;
    DevIo io;
    cpuSnapTallyBispectrum(bi, bispectrum, h_il.data(), h_type.data(), inum, ncoeff, nperdim, ntype, quadraticflag);
}

// SnapTallyBispectrumDeriv backend 3 (mdpcore.h:3769-3786; src/C++/Core/cpuSnap2.cpp:1324-1396)
void mdp_hipSnapTallyBispectrumDeriv(double* db, double* bispectrum, double* dbdr, int* aii, int* ai, int* aj, int* ti,
                                     int inum, int ijnum, int ncoeff, int nperdim, int ntype, int quadraticflag) {
    std::vector<int> h_aii, h_ai, h_aj, h_ti;
    if (!fetch_args([&] {
            const size_t np = static_cast<size_t>(std::max(ijnum, 0));
            h_aii = fetch(aii, np);
            h_ai = fetch(ai, np);
            h_aj = fetch(aj, np);
            h_ti = fetch(ti, np);
        }))
        return;
    DevIo io;
    cpuSnapTallyBispectrumDeriv(db, bispectrum, dbdr, h_aii.data(), h_ai.data(), h_aj.data(), h_ti.data(), inum, ijnum,
                                ncoeff, nperdim, ntype, quadraticflag);
}

// SnapTallyBispectrumVirial backend 3 (mdpcore.h:3788-3805; src/C++/Core/cpuSnap2.cpp:1398-1503; ai and aj
// are read but unused there)
void mdp_hipSnapTallyBispectrumVirial(double* bv, double* bispectrum, double* dbdr, double* rij, int* aii, int* ai,
                                      int* aj, int* ti, int inum, int ijnum, int ncoeff, int nperdim, int ntype,
                                      int quadraticflag) {
    (void)ai; (void)aj;
    std::vector<int> h_aii, h_ti;
    if (!fetch_args([&] {
            const size_t np = static_cast<size_t>(std::max(ijnum, 0));
            h_aii = fetch(aii, np);
            h_ti = fetch(ti, np);
        }))
        return;
    DevIo io;
    cpuSnapTallyBispectrumVirial(bv, bispectrum, dbdr, rij, h_aii.data(), h_ti.data(), inum, ijnum, ncoeff, nperdim,
                                 ntype, quadraticflag);
}

// ---- energy / force / virial tallies (the force path's split form, src/C++/Core/cpuSnap.cpp:1746-1831) --

// SnapTallyEnergyFull backend 3 (mdpcore.h:3630-3651; bispectrum [ii*ncoeff + k])
void mdp_hipSnapTallyEnergyFull(double* eatom, double* bispectrum, double* coeffelem, int* ilist, int* map, int* type,
                                int inum, int ncoeff, int ncoeffall, int quadraticflag) {
    guarded([&] {
        hipStream_t st = library_stream();
        launch_tally_energy_full(eatom, bispectrum, coeffelem, ilist, map, type, inum, ncoeff, ncoeffall,
                                 quadraticflag, 0, st);
        finish(st);
    });
}

// SnapTallyEnergyFull2 backend 3 (mdpcore.h:3653-3674; bispectrum [ii + inum*k], cpuSnap2.cpp:1265-1291)
void mdp_hipSnapTallyEnergyFull2(double* eatom, double* bispectrum, double* coeffelem, int* ilist, int* map,
                                 int* type, int inum, int ncoeff, int ncoeffall, int quadraticflag) {
    guarded([&] {
        hipStream_t st = library_stream();
        launch_tally_energy_full(eatom, bispectrum, coeffelem, ilist, map, type, inum, ncoeff, ncoeffall,
                                 quadraticflag, 1, st);
        finish(st);
    });
}

// SnapTallyForceFull backend 3 (mdpcore.h:3676-3691; alist is unused upstream too)
void mdp_hipSnapTallyForceFull(double* fatom, double* fij, int* ai, int* aj, int* alist, int ijnum) {
    (void)alist;
    guarded([&] {
        hipStream_t st = library_stream();
        launch_tally_force_full(fatom, fij, ai, aj, ijnum, st);
        finish(st);
    });
}

// SnapTallyVirialFull backend 3 (mdpcore.h:3712-3730; SoA vatom[c*inum + atom], half to each atom)
void mdp_hipSnapTallyVirialFull(double* vatom, double* fij, double* rij, int* ai, int* aj, int inum, int ijnum) {
    guarded([&] {
        hipStream_t st = library_stream();
        launch_tally_virial_full(vatom, fij, rij, ai, aj, inum, ijnum, st);
        finish(st);
    });
}

}  // extern "C"
