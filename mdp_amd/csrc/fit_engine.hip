// Batched fitting descriptors: host engine and C ABI (mdp_fit_* in include/mdpsnap.h).
//
// Plays the role of the upstream per-configuration descriptor loop (snapcpp.jl:539-665 for one
// configuration, repeated by Regression.cpp:68-95 / the Julia fitting drivers for every configuration
// of a training set): B and dB/dr of many independent periodic configurations, computed for the
// whole batch at once.  The stages (pair build, Utot, Z + B, dB/dr) are captured once into a
// hipGraph and replayed by every mdp_fit_compute call; the dB/dr stage is a persistent kernel whose
// waves take atoms from a work counter, so configurations of very different sizes share the GPU
// without per-configuration launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "fit_launch.hpp"
#include "mdpsnap.h"
#include "runtime.hpp"
#include "snap_launch.hpp"

using namespace mdp;

namespace {
int n_half_rows_host(int J) {
    int n = 0;
    for (int j = 0; j <= J; ++j) n += j / 2 + 1;
    return n;
}
}  // namespace

struct mdp_fit {
    int C = 0, N = 0, J = 0, ntypes = 0;
    int bnorm = 0, bzeroflag = 1;
    hipStream_t st = nullptr;
    DevTables tab;
    std::vector<int> a0_h;   // [C+1]
    std::vector<double> boxes_h, tilts_h;   // tilts_h empty: every cell orthogonal
    DevBuf x, type, cfg_of, cfg_a0, cfg_L, cfg_T, cfg_p0, rc2, pcnt, poff, scan_tmp, flags, counter;
    DevBuf chunk_ord;   // 64-atom chunks of the pair build, most candidate tests first
    DevBuf pi, pj, pti, ptj, rij, Ur, Ui, zt, blist, bpart, dblist, bzero, t_lo, t_b, t_z, t_w, zorder, t_wtab, t_hlo, t_hb, t_hz,
        t_hw, t_mz, t_mw, t_mcomb, pgeo;
    int dbdr_variant = 0;
    bool zb_lds = false;     // Z + B from the LDS-staged half U (twojmax 4, 6, 8)
    int zb_waves = 8;        // waves per Z workgroup (autotuned: 8, 4 or 12)
    long cap = 0;            // pair capacity of the pair and dB/dr buffers
    long npairs = 0;         // pairs of the last compute
    FitTerms terms;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    DevBuf l_cnt, l_off, l_cur, l_idx, l_bi, l_bd, l_bv, l_as, l_chunks, l_ekl, l_part, l_out, l_T, l_W;
    int l_ekl_k = 0;         // column count the (k, l) table was built for
    int dbdr_groups = 1024;
    bool tuned = false;
    std::string tuning_json = "{}";
    std::string zb_json;
    double stage_ms[4] = {0, 0, 0, 0};

    ~mdp_fit() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (st) (void)hipStreamDestroy(st);
    }

    // cells (may be null): per configuration a, b, c row vectors, a = (Lx, 0, 0), b = (bx, Ly, 0),
    // c = (cx, cy, Lz); boxes then unused
    void init(const mdp_fit_params* p, int nconfig, const int* natoms, const double* boxes, const double* x0,
              const int* types, const double* cells = nullptr) {
        if (!p) throw Error("mdp_fit_create: null parameters");
        if (!templated_j(p->twojmax)) throw Error("mdp_fit_create: twojmax must be 2, 4, 6 or 8");
        if (nconfig <= 0) throw Error("mdp_fit_create: nconfig must be positive");
        if (p->ntypes <= 0) throw Error("mdp_fit_create: ntypes must be positive");
        J = p->twojmax;
        C = nconfig;
        ntypes = p->ntypes;
        bnorm = p->bnormflag;
        bzeroflag = p->bzeroflag;
        a0_h.assign(C + 1, 0);
        std::vector<double> bx_h;
        if (cells) {
            bx_h.resize(3 * C);
            tilts_h.resize(3 * C);
            bool any = false;
            for (int c = 0; c < C; ++c) {
                const double* m = cells + 9 * static_cast<size_t>(c);
                for (int k = 0; k < 9; ++k)
                    if (!std::isfinite(m[k])) throw Error("mdp_fit_create_cells: non-finite cell vector");
                if (m[1] != 0.0 || m[2] != 0.0 || m[5] != 0.0)
                    throw Error("mdp_fit_create_cells: cells must have a along x and b in the xy plane "
                                "(a = (Lx,0,0), b = (bx,Ly,0), c = (cx,cy,Lz))");
                bx_h[3 * c] = m[0];
                bx_h[3 * c + 1] = m[4];
                bx_h[3 * c + 2] = m[8];
                tilts_h[3 * c] = m[3];
                tilts_h[3 * c + 1] = m[6];
                tilts_h[3 * c + 2] = m[7];
                if (!(std::fabs(m[3]) < m[0]) || !(std::fabs(m[6]) < m[0]) || !(std::fabs(m[7]) < m[4]))
                    throw Error("mdp_fit_create_cells: tilts must be smaller than the box lengths");
                any = any || m[3] != 0.0 || m[6] != 0.0 || m[7] != 0.0;
            }
            if (!any) tilts_h.clear();
            boxes = bx_h.data();
        }
        for (int c = 0; c < C; ++c) {
            if (natoms[c] <= 0) throw Error("mdp_fit_create: every configuration needs at least one atom");
            for (int d = 0; d < 3; ++d)
                if (!(boxes[3 * c + d] > 0.0)) throw Error("mdp_fit_create: box lengths must be positive");
            a0_h[c + 1] = a0_h[c] + natoms[c];
        }
        N = a0_h[C];
        boxes_h.assign(boxes, boxes + 3 * C);
        for (int g = 0; g < N; ++g)
            if (types[g] < 1 || types[g] > ntypes) throw Error("mdp_fit_create: atom types must be 1..ntypes");
        check_positions(x0, "mdp_fit_create");
        MDP_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));

        SnapIndex ix;
        ix.build(J, nullptr);
        std::vector<double> radelem(ntypes + 1, 0.0), wjelem(ntypes + 1, 0.0);
        std::vector<int> map(ntypes + 1, 0);
        for (int t = 1; t <= ntypes; ++t) {
            radelem[t] = p->elemradius[t - 1];
            wjelem[t] = p->elemweight[t - 1];
        }
        tab.upload(ix, nullptr, nullptr, radelem.data(), wjelem.data(), map.data(), ntypes, st);
        tab.view.rcutfac = p->rcutfac;
        tab.view.rfac0 = p->rfac0;
        tab.view.rmin0 = p->rmin0;
        tab.view.switchflag = p->switchflag;

        std::vector<double> rc2h((ntypes + 1) * (ntypes + 1), 0.0);
        for (int a = 1; a <= ntypes; ++a)
            for (int b = 1; b <= ntypes; ++b) {
                const double rc = (radelem[a] + radelem[b]) * p->rcutfac;
                rc2h[a * (ntypes + 1) + b] = rc * rc;
            }
        // bzero[j] = wself^3 (j+1), wself = 1 (snapcpp.jl:290-298); 1 under bnorm
        std::vector<double> bz(J + 1);
        for (int j = 0; j <= J; ++j) bz[j] = bnorm ? 1.0 : j + 1.0;
        std::vector<int> cof(N);
        for (int c = 0; c < C; ++c)
            for (int g = a0_h[c]; g < a0_h[c + 1]; ++g) cof[g] = c;
        std::vector<int> lo, tb, tz;
        std::vector<double> tw;
        fit_term_table(J, bnorm, lo, tb, tz, tw);

        auto put = [&](DevBuf& b, const void* h, size_t bytes) {
            b.ensure(bytes);
            MDP_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, st));
        };
        const size_t n = N;
        put(x, x0, sizeof(double) * 3 * n);
        put(type, types, sizeof(int) * n);
        put(cfg_of, cof.data(), sizeof(int) * n);
        put(cfg_a0, a0_h.data(), sizeof(int) * (C + 1));
        put(cfg_L, boxes, sizeof(double) * 3 * C);
        if (!tilts_h.empty()) put(cfg_T, tilts_h.data(), sizeof(double) * 3 * C);
        put(rc2, rc2h.data(), sizeof(double) * rc2h.size());
        {   // the pair build runs one wave per 64 consecutive atoms, each lane testing every atom of its
            // configuration (every image of it in a box narrower than two cutoffs or tilted); launching the
            // chunks with the most tests first keeps the largest configurations from ending last
            double rcmax = 0.0;
            for (double r : rc2h) rcmax = std::max(rcmax, std::sqrt(r));
            std::vector<double> cost_c(C);
            for (int c = 0; c < C; ++c) {
                bool narrow = !tilts_h.empty() && (tilts_h[3 * c] != 0.0 || tilts_h[3 * c + 1] != 0.0 || tilts_h[3 * c + 2] != 0.0);
                for (int d = 0; d < 3; ++d) narrow = narrow || boxes[3 * c + d] <= 2.0 * rcmax;
                cost_c[c] = static_cast<double>(a0_h[c + 1] - a0_h[c]) * (narrow ? 27.0 : 1.0);
            }
            const int nch = (N + 63) / 64;
            std::vector<double> cost(nch, 0.0);
            for (int g = 0; g < N; ++g) cost[g / 64] += cost_c[cof[g]];
            std::vector<int> ord(nch);
            for (int k = 0; k < nch; ++k) ord[k] = k;
            std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return cost[a] > cost[b]; });
            put(chunk_ord, ord.data(), sizeof(int) * std::max(nch, 1));
        }
        put(bzero, bz.data(), sizeof(double) * bz.size());
        put(t_lo, lo.data(), sizeof(int) * lo.size());
        put(t_b, tb.data(), sizeof(int) * tb.size());
        put(t_z, tz.data(), sizeof(int) * tz.size());
        put(t_w, tw.data(), sizeof(double) * tw.size());
        zb_lds = (J == 4 || J == 6 || J == 8);
        if (zb_lds) {
            std::vector<int> order(n_half_rows_host(J));
            yi_row_order(J, order.data());
            put(zorder, order.data(), sizeof(int) * order.size());
        }
        {   // weights by (bispectrum triple, term) for the unrolled dB/dr variant (cpuSnap.cpp:2101-2240)
            std::vector<double> wt;
            for (int j1 = 0; j1 <= J; ++j1)
                for (int j2 = 0; j2 <= j1; ++j2)
                    for (int j = j1 - j2; j <= std::min(J, j1 + j2); j += 2) {
                        if (j < j1) continue;
                        wt.push_back(2.0);
                        wt.push_back(bnorm ? 2.0 : 2.0 * (j + 1) / (j1 + 1.0));
                        wt.push_back(bnorm ? 2.0 : 2.0 * (j + 1) / (j2 + 1.0));
                    }
            put(t_wtab, wt.data(), sizeof(double) * wt.size());
        }
        {   // the term table split into two halves of the bispectrum index range (dB/dr variant 4)
            std::vector<int> hl, hbv, hz;
            std::vector<double> hw;
            terms.hb = fit_term_table_halves(J, bnorm, hl, hbv, hz, hw);
            put(t_hlo, hl.data(), sizeof(int) * hl.size());
            put(t_hb, hbv.data(), sizeof(int) * hbv.size());
            put(t_hz, hz.data(), sizeof(int) * hz.size());
            put(t_hw, hw.data(), sizeof(double) * hw.size());
            terms.h_lo = t_hlo.as<int>();
            terms.h_b = t_hb.as<int>();
            terms.h_z = t_hz.as<int>();
            terms.h_w = t_hw.as<double>();
        }
        if (fit_mfma_supported(J)) {   // packed term rows of the matrix-core dB/dr (variant 6)
            std::vector<int> mz, mcomb;
            std::vector<double> mw;
            if (!fit_mfma_tables(J, bnorm, mz, mw, mcomb)) throw Error("fit: matrix-core dB/dr term packing failed");
            put(t_mz, mz.data(), sizeof(int) * mz.size());
            put(t_mw, mw.data(), sizeof(double) * mw.size());
            put(t_mcomb, mcomb.data(), sizeof(int) * mcomb.size());
            terms.m_z = t_mz.as<int>();
            terms.m_w = t_mw.as<double>();
            terms.m_comb = t_mcomb.as<int>();
        }
        terms.wtab = t_wtab.as<double>();
        terms.lo = t_lo.as<int>();
        terms.b = t_b.as<int>();
        terms.z = t_z.as<int>();
        terms.w = t_w.as<double>();
        cfg_p0.ensure(sizeof(int) * (C + 1));
        pcnt.ensure(sizeof(int) * n * (FIT_PAIR_SPLIT + 1));
        poff.ensure(sizeof(int) * (n + 1));
        scan_tmp.ensure(sizeof(int) * (scan_tmp_ints(N) + 8));
        flags.ensure(sizeof(int) * 4);
        counter.ensure(sizeof(int) * 4);
        const SnapIndex& I = tab.idx;
        Ur.ensure(sizeof(double) * n * I.idxu_max);
        Ui.ensure(sizeof(double) * n * I.idxu_max);
        zt.ensure(sizeof(double) * 2 * n * I.idxz_max);
        blist.ensure(sizeof(double) * n * I.idxb_max);
        if (zb_lds) bpart.ensure(sizeof(double) * n * static_cast<size_t>(fit_zb_partials(J)));
        MDP_HIP(hipStreamSynchronize(st));   // host staging vectors go out of scope
    }

    // the pair build derives image ranges from the coordinates: they must be finite and within one box
    // length of the box, which bounds the images each atom pair visits
    void check_positions(const double* xh, const char* where) const {
        for (int c = 0; c < C; ++c)
            for (long g = a0_h[c]; g < a0_h[c + 1]; ++g)
                for (int d = 0; d < 3; ++d) {
                    const double v = xh[3 * g + d], L = boxes_h[3 * c + d];
                    // a tilted cell reaches up to one more box length along x and y
                    const double lo = tilts_h.empty() ? -L : -2.0 * L, hi = tilts_h.empty() ? 2.0 * L : 3.0 * L;
                    if (!std::isfinite(v)) throw Error(std::string(where) + ": non-finite position");
                    if (v < lo || v >= hi) throw Error(std::string(where) + ": position outside the box bounds");
                }
    }

    void set_positions(const double* xh) {
        check_positions(xh, "mdp_fit_set_positions");
        MDP_HIP(hipMemcpyAsync(x.p, xh, sizeof(double) * 3 * N, hipMemcpyHostToDevice, st));
        MDP_HIP(hipStreamSynchronize(st));
    }

    void ensure_pairs(long need) {
        if (need <= cap) return;
        const long c = need + need / 8 + 64;
        pi.ensure(sizeof(int) * c);
        pj.ensure(sizeof(int) * c);
        pti.ensure(sizeof(int) * c);
        ptj.ensure(sizeof(int) * c);
        rij.ensure(sizeof(double) * 3 * c);
        dblist.ensure(sizeof(double) * 3 * c * tab.idx.idxb_max);
        pgeo.ensure(sizeof(double) * 5 * c);
        cap = c;
        if (gexec) {   // buffers moved: the captured graph refers to the old ones
            MDP_HIP(hipGraphExecDestroy(gexec));
            MDP_HIP(hipGraphDestroy(graph));
            gexec = nullptr;
            graph = nullptr;
        }
    }

    // stage k of the pipeline (0 pairs, 1 Utot, 2 Z+B, 3 dB/dr)
    void stage(int k) {
        const SnapDev& s = tab.view;
        switch (k) {
            case 0:
                launch_fill_zero_int(flags.as<int>(), 4, st);
                launch_fit_pairs(false, x.as<double>(), type.as<int>(), cfg_of.as<int>(), cfg_a0.as<int>(),
                                 cfg_L.as<double>(), tilts_h.empty() ? nullptr : cfg_T.as<double>(), rc2.as<double>(), ntypes, N, chunk_ord.as<int>(), pcnt.as<int>(), nullptr, 0, nullptr,
                                 nullptr, nullptr, nullptr, nullptr, flags.as<int>(), st);
                exclusive_scan(pcnt.as<int>(), poff.as<int>(), N, scan_tmp.as<int>(), st);
                launch_fit_cfg_p0(poff.as<int>(), cfg_a0.as<int>(), cfg_p0.as<int>(), C, st);
                launch_fit_pairs(true, x.as<double>(), type.as<int>(), cfg_of.as<int>(), cfg_a0.as<int>(),
                                 cfg_L.as<double>(), tilts_h.empty() ? nullptr : cfg_T.as<double>(), rc2.as<double>(), ntypes, N, chunk_ord.as<int>(), pcnt.as<int>(), poff.as<int>(), cap,
                                 rij.as<double>(), pi.as<int>(), pj.as<int>(), pti.as<int>(), ptj.as<int>(),
                                 flags.as<int>(), st);
                launch_fit_guard(flags.as<int>(), poff.as<int>(), N, st);
                break;
            case 1:
                if (zb_lds) {   // half U + wself from the per-pair geometry (one geometry per pair)
                    launch_pair_geo(s, rij.as<double>(), pti.as<int>(), ptj.as<int>(), poff.as<int>() + N,
                                    static_cast<int>(cap), pgeo.as<double>(), st);
                    launch_ui_geo(s, Ur.as<double>(), Ui.as<double>(), pgeo.as<double>(), poff.as<int>(), N, 1.0, st);
                } else {
                    launch_ui(s, UI_FULL_WSELF, Ur.as<double>(), Ui.as<double>(), rij.as<double>(), poff.as<int>(),
                              nullptr, pti.as<int>(), ptj.as<int>(), N, 1.0, -1, st);
                }
                break;
            case 2:
                if (zb_lds)
                    launch_fit_zb_lds(s, Ur.as<double>(), Ui.as<double>(), zt.as<double>(), blist.as<double>(),
                                      bpart.as<double>(), bzero.as<double>(), cfg_of.as<int>(), cfg_a0.as<int>(),
                                      zorder.as<int>(), N, bnorm, bzeroflag, st, zb_waves);
                else
                    launch_fit_zb(s, Ur.as<double>(), Ui.as<double>(), zt.as<double>(), blist.as<double>(),
                                  bzero.as<double>(), cfg_of.as<int>(), cfg_a0.as<int>(), N, bnorm, bzeroflag, st);
                break;
            default:
                launch_fill_zero_int(counter.as<int>(), 2, st);
                if (dbdr_variant == 6 && terms.m_z)
                    launch_fit_dbdr_mfma(s, zt.as<double>(), terms.m_z, terms.m_w, terms.m_comb, rij.as<double>(),
                                         pti.as<int>(), ptj.as<int>(), poff.as<int>(), cfg_of.as<int>(),
                                         cfg_p0.as<int>(), dblist.as<double>(), counter.as<int>(), N, dbdr_groups, st);
                else
                    launch_fit_dbdr(s, zt.as<double>(), terms, rij.as<double>(), pti.as<int>(), ptj.as<int>(),
                                    poff.as<int>(), cfg_of.as<int>(), cfg_p0.as<int>(), dblist.as<double>(),
                                    counter.as<int>(), N, dbdr_groups, st, dbdr_variant);
                break;
        }
    }

    long count_pairs() {
        launch_fit_pairs(false, x.as<double>(), type.as<int>(), cfg_of.as<int>(), cfg_a0.as<int>(), cfg_L.as<double>(),
                         tilts_h.empty() ? nullptr : cfg_T.as<double>(), rc2.as<double>(), ntypes, N, chunk_ord.as<int>(), pcnt.as<int>(), nullptr, 0, nullptr, nullptr, nullptr, nullptr,
                         nullptr, flags.as<int>(), st);
        exclusive_scan(pcnt.as<int>(), poff.as<int>(), N, scan_tmp.as<int>(), st);
        int tot = 0;
        MDP_HIP(hipMemcpyAsync(&tot, poff.as<int>() + N, sizeof(int), hipMemcpyDeviceToHost, st));
        MDP_HIP(hipStreamSynchronize(st));
        return tot;
    }

    void capture() {
        MDP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < 4; ++k) stage(k);
        MDP_HIP(hipStreamEndCapture(st, &graph));
        MDP_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    }

    float time_stage(int k, int reps) {
        hipEvent_t a, b;
        MDP_HIP(hipEventCreate(&a));
        MDP_HIP(hipEventCreate(&b));
        MDP_HIP(hipEventRecord(a, st));
        for (int r = 0; r < reps; ++r) stage(k);
        MDP_HIP(hipEventRecord(b, st));
        MDP_HIP(hipEventSynchronize(b));
        float ms = 0;
        MDP_HIP(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return ms / reps;
    }

    // Time the persistent dB/dr kernel at a few grid sizes on the live data and keep the fastest
    // (the kernel is a pure function of its inputs; its counter is reset inside the stage).
    void autotune() {
        const char* env = std::getenv("MDP_AUTOTUNE");
        tuned = true;
        const char* fv = std::getenv("MDP_FIT_DBDR_VARIANT");   // pin the dB/dr variant (tests)
        if (fv) {
            const std::string v(fv);
            const bool j468 = (J == 4 || J == 6 || J == 8);
            dbdr_variant = (v == "1" && j468) ? 1 : ((v == "2" && (J == 6 || J == 8)) ? 2 : ((v == "3" && j468) ? 3 : 0));
            if (v == "4" && j468) dbdr_variant = 4;
            if (v == "5" && j468) dbdr_variant = 5;
            if (v == "6" && j468) dbdr_variant = 6;
        }
        const char* zp = std::getenv("MDP_FIT_ZB_WAVES");   // pin the Z kernel's waves per workgroup
        if (zp) {
            const int w = std::atoi(zp);
            if (w != 4 && w != 8 && w != 12) throw Error("MDP_FIT_ZB_WAVES must be 4, 8 or 12");
            zb_waves = w;
        }
        if (!zp && !(env && std::string(env) == "0") && zb_lds) {   // Z kernel: 8, 4 or 12 waves per workgroup
            const int zw[3] = {8, 4, 12};
            double tz[3];
            zb_json = "\"fit_zb\": {";
            int best = 0;
            for (int k = 0; k < 3; ++k) {
                zb_waves = zw[k];
                stage(2);   // warm-up
                tz[k] = time_stage(2, 2);
                if (tz[k] < tz[best]) best = k;
                zb_json += (k ? ", \"" : "\"") + std::to_string(zw[k]) + "\": " + std::to_string(tz[k]);
            }
            zb_waves = zw[best];
            zb_json += "}, ";
        }
        if ((env && std::string(env) == "0") || fv) {
            tuning_json = "{" + zb_json + "\"chosen\": {\"fit_zb_waves\": " + std::to_string(zb_waves) + "}}";
            return;
        }
        int props_cu = 256;
        hipDeviceProp_t prop;
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            props_cu = prop.multiProcessorCount;
        const int cand[] = {props_cu * 2, props_cu * 4, props_cu * 8, props_cu * 12};
        // dB/dr variants: 0 table walk, 1 unrolled terms, 2 table walk at two waves per SIMD, 3 variant 0
        // with the term loop pipelined, 4 generator form with the bispectrum indices split over two waves,
        // 5 variant 4 with the atom's Z block staged in LDS, 6 the contraction on the FP64 matrix cores
        std::vector<int> vars = {0};
        if (J == 4) vars = {0, 1, 3, 4, 5, 6};
        if (J == 6 || J == 8) vars = {0, 1, 2, 3, 4, 5, 6};
        double best = 1e30;
        int pick = dbdr_groups, pickv = 0;
        std::string js = "{" + zb_json + "\"fit_dbdr\": {";
        for (int v : vars)
            for (int k = 0; k < 4; ++k) {
                dbdr_variant = v;
                dbdr_groups = cand[k];
                stage(3);   // warm-up
                const double t = time_stage(3, 2);
                js += (v || k ? ", " : "") + std::string("\"") + std::to_string(v) + "/" + std::to_string(cand[k]) +
                      "\": " + std::to_string(t);
                if (t < best) {
                    best = t;
                    pick = cand[k];
                    pickv = v;
                }
            }
        dbdr_groups = pick;
        dbdr_variant = pickv;
        js += "}, \"chosen\": {\"fit_dbdr_variant\": " + std::to_string(pickv) + ", \"fit_dbdr_groups\": " +
              std::to_string(pick) + ", \"fit_zb_waves\": " + std::to_string(zb_waves) + "}}";
        tuning_json = js;
    }

    void compute() {
        if (cap == 0) ensure_pairs(count_pairs());
        if (!tuned) {
            for (int k = 0; k < 4; ++k) stage(k);
            MDP_HIP(hipGetLastError());
            autotune();
        }
        for (int attempt = 0; attempt < 2; ++attempt) {
            if (!gexec) capture();
            MDP_HIP(hipGraphLaunch(gexec, st));
            int fl[2] = {0, 0}, tot = 0;
            MDP_HIP(hipMemcpyAsync(fl, flags.p, sizeof(fl), hipMemcpyDeviceToHost, st));
            MDP_HIP(hipMemcpyAsync(&tot, poff.as<int>() + N, sizeof(int), hipMemcpyDeviceToHost, st));
            MDP_HIP(hipStreamSynchronize(st));
            npairs = tot;
            if (!(fl[0] & 1)) return;
            // positions changed and the pair list outgrew its buffers: the guard emptied the pair list
            // for the rest of this run; resize to the reported need and run again
            npairs = 0;
            ensure_pairs(fl[1]);
        }
        throw Error("mdp_fit_compute: pair buffers could not be sized");
    }

    // replay without the readback (for timing loops); returns nothing, errors surface in sync()
    void replay() {
        if (!gexec) throw Error("mdp_fit_replay: call mdp_fit_compute first");
        MDP_HIP(hipGraphLaunch(gexec, st));
    }

    int max_natoms() const {
        int m = 0;
        for (int c = 0; c < C; ++c) m = std::max(m, a0_h[c + 1] - a0_h[c]);
        return m;
    }

    // Tallies of the last compute for the linear fit (upstream cpuSnapTallyBispectrum{,Deriv,Virial}
    // layouts per configuration) and the stress rows of the fit.
    void lsq_tallies() {
        if (!gexec) throw Error("mdp_fit_descriptors: call mdp_fit_compute first");
        const int IB = tab.idx.idxb_max, kt = ntypes * IB;
        l_cnt.ensure(sizeof(int) * N);
        l_off.ensure(sizeof(int) * (N + 1));
        l_cur.ensure(sizeof(int) * N);
        l_idx.ensure(sizeof(int) * std::max<long>(npairs, 1));
        l_bi.ensure(sizeof(double) * C * kt);
        l_bd.ensure(sizeof(double) * 3 * static_cast<size_t>(N) * kt);
        l_bv.ensure(sizeof(double) * 6 * static_cast<size_t>(C) * kt);
        l_as.ensure(sizeof(double) * 6 * static_cast<size_t>(C) * kt);
        launch_lsq_incoming(pj.as<int>(), npairs, N, l_cnt.as<int>(), l_off.as<int>(), l_cur.as<int>(), l_idx.as<int>(),
                            scan_tmp.as<int>(), st);
        launch_lsq_tallies(blist.as<double>(), dblist.as<double>(), rij.as<double>(), poff.as<int>(), l_off.as<int>(),
                           l_idx.as<int>(), pti.as<int>(), type.as<int>(), cfg_of.as<int>(), cfg_a0.as<int>(),
                           cfg_p0.as<int>(), C, N, IB, ntypes, max_natoms(), l_bi.as<double>(), l_bd.as<double>(),
                           l_bv.as<double>(), l_as.as<double>(), st);
        MDP_HIP(hipGetLastError());
    }

    // Weighted normal equations of the energy, force and stress rows (linearfit.jl:168-200):
    // L[g] = sum_r w_r^2 a_r a_r^T and h[g] = sum_r w_r^2 a_r b_r for g = energy, force, stress.  Energy
    // rows a = bi_c with target e_c (both / N_c under normalizeenergy, emldescriptors.jl:108-128), force
    // rows the 3 N_c rows of bd_c with targets f (AoS per atom), stress rows the 6 reordered virial rows
    // with targets s[6c + m'].  A null target array leaves its group zero; null weights are 1.
    void lsq_normal(const double* e, const double* f, const double* s, const double* we, const double* wf,
                    const double* ws, int normalizeenergy, double* L, double* hv) {
        lsq_tallies();
        const int K = ntypes * tab.idx.idxb_max;
        if (lsq_gram_lds(K) > 65536) throw Error("mdp_fit_normal_equations: too many descriptor columns");
        const int ntri = K * (K + 1) / 2, nout = ntri + K, G = 1024;
        if (l_ekl_k != K) {
            std::vector<int2> ekl(nout);
            int e0 = 0;
            for (int k = 0; k < K; ++k)
                for (int l = k; l < K; ++l) ekl[e0++] = make_int2(k, l);
            for (int k = 0; k < K; ++k) ekl[ntri + k] = make_int2(k, -1);
            l_ekl.ensure(sizeof(int2) * nout);
            MDP_HIP(hipMemcpyAsync(l_ekl.p, ekl.data(), sizeof(int2) * nout, hipMemcpyHostToDevice, st));
            MDP_HIP(hipStreamSynchronize(st));
            l_ekl_k = K;
        }
        l_part.ensure(sizeof(double) * static_cast<size_t>(G) * nout);
        l_out.ensure(sizeof(double) * nout);
        std::fill(L, L + 3 * static_cast<size_t>(K) * K, 0.0);
        std::fill(hv, hv + 3 * static_cast<size_t>(K), 0.0);
        auto run = [&](int grp, const double* A, const std::vector<LsqChunk>& ch, const double* T, size_t nT,
                       const std::vector<double>& W) {
            if (ch.empty()) return;
            l_chunks.ensure(sizeof(LsqChunk) * ch.size());
            l_T.ensure(sizeof(double) * nT);
            l_W.ensure(sizeof(double) * W.size());
            MDP_HIP(hipMemcpyAsync(l_chunks.p, ch.data(), sizeof(LsqChunk) * ch.size(), hipMemcpyHostToDevice, st));
            MDP_HIP(hipMemcpyAsync(l_T.p, T, sizeof(double) * nT, hipMemcpyHostToDevice, st));
            MDP_HIP(hipMemcpyAsync(l_W.p, W.data(), sizeof(double) * W.size(), hipMemcpyHostToDevice, st));
            launch_lsq_gram(l_chunks.as<LsqChunk>(), static_cast<int>(ch.size()), A, l_T.as<double>(), l_W.as<double>(),
                            K, l_ekl.as<int2>(), l_part.as<double>(), G, l_out.as<double>(), st);
            std::vector<double> o(nout);
            MDP_HIP(hipMemcpyAsync(o.data(), l_out.p, sizeof(double) * nout, hipMemcpyDeviceToHost, st));
            MDP_HIP(hipStreamSynchronize(st));
            double* Lg = L + static_cast<size_t>(grp) * K * K;
            int e0 = 0;
            for (int k = 0; k < K; ++k)
                for (int l = k; l < K; ++l, ++e0) Lg[k * K + l] = Lg[l * K + k] = o[e0];
            for (int k = 0; k < K; ++k) hv[static_cast<size_t>(grp) * K + k] = o[ntri + k];
        };
        if (e) {   // one row per configuration, 64 per chunk
            std::vector<LsqChunk> ch;
            std::vector<double> W(C);
            for (int c = 0; c < C; ++c)
                W[c] = (we ? we[c] : 1.0) / (normalizeenergy ? static_cast<double>(a0_h[c + 1] - a0_h[c]) : 1.0);
            for (int c0 = 0; c0 < C; c0 += 64)
                ch.push_back(LsqChunk{static_cast<long>(c0) * K, c0, K, 1, std::min(64, C - c0), c0, 1, 0});
            run(0, l_bi.as<double>(), ch, e, C, W);
        }
        if (f) {   // 3 N_c rows per configuration, column stride 3 N_c
            std::vector<LsqChunk> ch;
            std::vector<double> W(C);
            for (int c = 0; c < C; ++c) {
                W[c] = wf ? wf[c] : 1.0;
                const int a0 = a0_h[c], rows = 3 * (a0_h[c + 1] - a0);
                for (int r0 = 0; r0 < rows; r0 += 64)
                    ch.push_back(LsqChunk{3L * a0 * K + r0, 3L * a0 + r0, 1, rows, std::min(64, rows - r0), c,
                                          1 << 30, 0});
            }
            run(1, l_bd.as<double>(), ch, f, 3 * static_cast<size_t>(N), W);
        }
        if (s) {   // 6 rows per configuration, 60 (ten configurations) per chunk
            std::vector<LsqChunk> ch;
            std::vector<double> W(C);
            for (int c = 0; c < C; ++c) W[c] = ws ? ws[c] : 1.0;
            for (int r0 = 0; r0 < 6 * C; r0 += 60)
                ch.push_back(LsqChunk{static_cast<long>(r0) * K, r0, K, 1, std::min(60, 6 * C - r0), r0 / 6, 6, 0});
            run(2, l_as.as<double>(), ch, s, 6 * static_cast<size_t>(C), W);
        }
    }

    void profile() {
        if (!gexec) compute();
        for (int k = 0; k < 4; ++k) stage_ms[k] = time_stage(k, 1);
    }
};

namespace {
template <class F>
int api(F&& f) {
    clear_error();
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    }
}
}  // namespace

extern "C" {

int mdp_fit_create(mdp_fit** h, const mdp_fit_params* p, int nconfig, const int* natoms, const double* boxes,
                   const double* x, const int* types) {
    return api([&] {
        if (!h) throw Error("mdp_fit_create: null handle pointer");
        *h = nullptr;
        auto e = std::make_unique<mdp_fit>();
        e->init(p, nconfig, natoms, boxes, x, types);
        *h = e.release();
    });
}

int mdp_fit_create_cells(mdp_fit** h, const mdp_fit_params* p, int nconfig, const int* natoms, const double* cells,
                         const double* x, const int* types) {
    return api([&] {
        if (!h) throw Error("mdp_fit_create_cells: null handle pointer");
        if (!cells) throw Error("mdp_fit_create_cells: null cells");
        *h = nullptr;
        auto e = std::make_unique<mdp_fit>();
        e->init(p, nconfig, natoms, nullptr, x, types, cells);
        *h = e.release();
    });
}

int mdp_fit_destroy(mdp_fit* h) {
    return api([&] { delete h; });
}

int mdp_fit_set_positions(mdp_fit* h, const double* x) {
    return api([&] { h->set_positions(x); });
}

int mdp_fit_compute(mdp_fit* h) {
    return api([&] { h->compute(); });
}

int mdp_fit_replay(mdp_fit* h, int nreps) {
    return api([&] {
        for (int r = 0; r < nreps; ++r) h->replay();
    });
}

int mdp_fit_sync(mdp_fit* h) {
    return api([&] {
        MDP_HIP(hipStreamSynchronize(h->st));
        MDP_HIP(hipGetLastError());
    });
}

int mdp_fit_layout(mdp_fit* h, long* atom_off, long* pair_off) {
    return api([&] {
        if (atom_off)
            for (int c = 0; c <= h->C; ++c) atom_off[c] = h->a0_h[c];
        if (pair_off) {
            std::vector<int> p(h->C + 1);
            MDP_HIP(hipMemcpy(p.data(), h->cfg_p0.p, sizeof(int) * (h->C + 1), hipMemcpyDeviceToHost));
            for (int c = 0; c <= h->C; ++c) pair_off[c] = p[c];
        }
    });
}

int mdp_fit_get(mdp_fit* h, double* blist, double* dblist) {
    return api([&] {
        const int ib = h->tab.idx.idxb_max;
        if (blist)
            MDP_HIP(hipMemcpy(blist, h->blist.p, sizeof(double) * static_cast<size_t>(h->N) * ib,
                              hipMemcpyDeviceToHost));
        if (dblist)
            MDP_HIP(hipMemcpy(dblist, h->dblist.p, sizeof(double) * 3 * static_cast<size_t>(h->npairs) * ib,
                              hipMemcpyDeviceToHost));
    });
}

int mdp_fit_pairs(mdp_fit* h, int* ai, int* aj, double* rij) {
    return api([&] {
        const size_t P = h->npairs;
        std::vector<int> gi(P), gj(P);
        MDP_HIP(hipMemcpy(gi.data(), h->pi.p, sizeof(int) * P, hipMemcpyDeviceToHost));
        MDP_HIP(hipMemcpy(gj.data(), h->pj.p, sizeof(int) * P, hipMemcpyDeviceToHost));
        std::vector<int> cof(h->N);
        for (int c = 0; c < h->C; ++c)
            for (int g = h->a0_h[c]; g < h->a0_h[c + 1]; ++g) cof[g] = c;
        for (size_t q = 0; q < P; ++q) {
            const int a0 = h->a0_h[cof[gi[q]]];
            if (ai) ai[q] = gi[q] - a0;
            if (aj) aj[q] = gj[q] - a0;
        }
        if (rij) MDP_HIP(hipMemcpy(rij, h->rij.p, sizeof(double) * 3 * P, hipMemcpyDeviceToHost));
    });
}

int mdp_fit_tuning(mdp_fit* h, char* buf, int len) {
    return api([&] {
        if (!buf || len <= 0) return;
        std::strncpy(buf, h->tuning_json.c_str(), static_cast<size_t>(len) - 1);
        buf[len - 1] = 0;
    });
}

int mdp_fit_descriptors(mdp_fit* h, double* bi, double* bd, double* bv) {
    return api([&] {
        if (!h) throw Error("mdp_fit_descriptors: null handle");
        h->lsq_tallies();
        const size_t kt = static_cast<size_t>(h->ntypes) * h->tab.idx.idxb_max;
        if (bi) MDP_HIP(hipMemcpyAsync(bi, h->l_bi.p, sizeof(double) * h->C * kt, hipMemcpyDeviceToHost, h->st));
        if (bd)
            MDP_HIP(hipMemcpyAsync(bd, h->l_bd.p, sizeof(double) * 3 * h->N * kt, hipMemcpyDeviceToHost, h->st));
        if (bv) MDP_HIP(hipMemcpyAsync(bv, h->l_bv.p, sizeof(double) * 6 * h->C * kt, hipMemcpyDeviceToHost, h->st));
        MDP_HIP(hipStreamSynchronize(h->st));
    });
}

int mdp_fit_normal_equations(mdp_fit* h, const double* e, const double* f, const double* s, const double* we,
                             const double* wf, const double* ws, int normalizeenergy, double* L, double* hv) {
    return api([&] {
        if (!h || !L || !hv) throw Error("mdp_fit_normal_equations: null handle or output");
        h->lsq_normal(e, f, s, we, wf, ws, normalizeenergy, L, hv);
    });
}

int mdp_fit_profile(mdp_fit* h, double* ms4) {
    return api([&] {
        h->profile();
        for (int k = 0; k < 4; ++k) ms4[k] = h->stage_ms[k];
    });
}

// lsqsolve of linearfit.jl:29-75 on the assembled normal equations: each system is solved when its
// matrices and right-hand sides are not all zero (sum |.| > 1e-15), with the diagonal scaled by
// (1 + 1e-15) and, for the force-only system, a(0,0) raised to 1e-10 when smaller in magnitude.  The
// solve is LU with partial pivoting (what the upstream `\` does for a square dense matrix).
// solved[0..3] says which of ce, cf, cef, cefs were computed.
int mdp_lsqsolve(int K, const double* Le, const double* Lf, const double* Ls, const double* he, const double* hf,
                 const double* hs, double* ce, double* cf, double* cef, double* cefs, int* solved) {
    return api([&] {
        if (K <= 0) throw Error("mdp_lsqsolve: K must be positive");
        const size_t kk = static_cast<size_t>(K) * K;
        auto nonzero = [](const double* a, size_t n) {
            double t = 0.0;
            if (a)
                for (size_t i = 0; i < n; ++i) t += std::fabs(a[i]);
            return t > 1e-15;
        };
        const bool sAe = nonzero(Le, kk), sAf = nonzero(Lf, kk), sAs = nonzero(Ls, kk);
        const bool sbe = nonzero(he, K), sbf = nonzero(hf, K), sbs = nonzero(hs, K);
        auto solve = [&](std::initializer_list<const double*> As, std::initializer_list<const double*> bs, bool fguard,
                         double* x) {
            std::vector<double> A(kk, 0.0), b(K, 0.0);
            for (const double* a : As)
                for (size_t i = 0; i < kk; ++i) A[i] += a[i];
            for (const double* v : bs)
                for (int i = 0; i < K; ++i) b[i] += v[i];
            for (int i = 0; i < K; ++i) A[static_cast<size_t>(i) * K + i] *= 1.0 + 1e-15;
            if (fguard && std::fabs(A[0]) <= 1e-10) A[0] = 1e-10;
            std::vector<int> piv(K);
            for (int c = 0; c < K; ++c) {   // LU with partial pivoting, row-major
                int pr = c;
                for (int r = c + 1; r < K; ++r)
                    if (std::fabs(A[static_cast<size_t>(r) * K + c]) > std::fabs(A[static_cast<size_t>(pr) * K + c])) pr = r;
                if (A[static_cast<size_t>(pr) * K + c] == 0.0) throw Error("mdp_lsqsolve: singular matrix");
                if (pr != c) {
                    for (int j = 0; j < K; ++j) std::swap(A[static_cast<size_t>(c) * K + j], A[static_cast<size_t>(pr) * K + j]);
                    std::swap(b[c], b[pr]);
                }
                const double d = A[static_cast<size_t>(c) * K + c];
                for (int r = c + 1; r < K; ++r) {
                    const double m = A[static_cast<size_t>(r) * K + c] / d;
                    A[static_cast<size_t>(r) * K + c] = m;
                    for (int j = c + 1; j < K; ++j) A[static_cast<size_t>(r) * K + j] -= m * A[static_cast<size_t>(c) * K + j];
                    b[r] -= m * b[c];
                }
            }
            for (int r = K - 1; r >= 0; --r) {
                double t = b[r];
                for (int j = r + 1; j < K; ++j) t -= A[static_cast<size_t>(r) * K + j] * x[j];
                x[r] = t / A[static_cast<size_t>(r) * K + r];
            }
        };
        int ok[4] = {0, 0, 0, 0};
        if (sAe && sAf && sAs && sbe && sbf && sbs && cefs) {
            solve({Le, Lf, Ls}, {he, hf, hs}, false, cefs);
            ok[3] = 1;
        }
        if (sAe && sAf && sbe && sbf && cef) {
            solve({Le, Lf}, {he, hf}, false, cef);
            ok[2] = 1;
        }
        if (sAe && sbe && ce) {
            solve({Le}, {he}, false, ce);
            ok[0] = 1;
        }
        if (sAf && sbf && cf) {
            solve({Lf}, {hf}, true, cf);
            ok[1] = 1;
        }
        if (solved)
            for (int i = 0; i < 4; ++i) solved[i] = ok[i];
    });
}

}  // extern "C"
