// Host runtime shared by the drop-in entry points and the MD engine.
#include "runtime.hpp"

#include <algorithm>

namespace mdp {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& msg) { g_error = msg; }
void clear_error() { g_error.clear(); }
const char* get_error() { return g_error.c_str(); }

hipStream_t library_stream() {
    static hipStream_t s = [] {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = nullptr;
        return st;
    }();
    if (!s) throw Error("could not create the library HIP stream (no GPU?)");
    return s;
}

std::mutex& library_mutex() {
    static std::mutex m;
    return m;
}

void DevTables::upload(const SnapIndex& ix, const double* rootpq_h, const double* cg_h, const double* radelem_h,
                       const double* wjelem_h, const int* map_h, int ntypes, hipStream_t st) {
    idx = ix;
    auto put = [&](DevBuf& b, const void* h, size_t bytes) {
        b.ensure(bytes);
        if (h && bytes) MDP_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, st));
    };
    const int jd = ix.J + 1;
    put(rootpq, rootpq_h ? rootpq_h : ix.rootpq.data(), sizeof(double) * (ix.J + 2) * (ix.J + 2));
    put(cg, cg_h ? cg_h : ix.cglist.data(), sizeof(double) * ix.idxcg_max);
    std::vector<double> zeros(ntypes + 1, 0.0);
    std::vector<int> izeros(ntypes + 1, 0);
    put(radelem, radelem_h ? radelem_h : zeros.data(), sizeof(double) * (ntypes + 1));
    put(wjelem, wjelem_h ? wjelem_h : zeros.data(), sizeof(double) * (ntypes + 1));
    put(map, map_h ? map_h : izeros.data(), sizeof(int) * (ntypes + 1));
    put(idxz, ix.idxz.data(), sizeof(int) * ix.idxz.size());
    put(idxb, ix.idxb.data(), sizeof(int) * ix.idxb.size());
    put(idxu_block, ix.idxu_block.data(), sizeof(int) * jd);
    put(idxz_block, ix.idxz_block.data(), sizeof(int) * jd * jd * jd);
    put(idxb_block, ix.idxb_block.data(), sizeof(int) * jd * jd * jd);
    put(idxcg_block, ix.idxcg_block.data(), sizeof(int) * jd * jd * jd);
    view.J = ix.J;
    view.idxu_max = ix.idxu_max;
    view.idxz_max = ix.idxz_max;
    view.idxb_max = ix.idxb_max;
    view.idxcg_max = ix.idxcg_max;
    view.rootpq = rootpq.as<double>();
    view.cg = cg.as<double>();
    view.radelem = radelem.as<double>();
    view.wjelem = wjelem.as<double>();
    view.map = map.as<int>();
    view.idxz = idxz.as<int>();
    view.idxb = idxb.as<int>();
    view.idxu_block = idxu_block.as<int>();
    view.idxz_block = idxz_block.as<int>();
    view.idxb_block = idxb_block.as<int>();
    view.idxcg_block = idxcg_block.as<int>();
    MDP_HIP(hipStreamSynchronize(st));   // host staging vectors above go out of scope
}

std::vector<double> y_beta_table(const SnapIndex& ix, const double* beta, int e1, int e2, int e3, int ne, int ielem,
                                 int ncoeffall, int bnorm) {
    const int J = ix.J, jd = J + 1;
    std::vector<double> out;
    for (int j1 = 0; j1 <= J; ++j1)
        for (int j2 = 0; j2 <= j1; ++j2)
            for (int j = j1 - j2; j <= std::min(J, j1 + j2); j += 2) {
                double w;
                if (j >= j1) {
                    const int jjb = ix.idxb_block[j + j2 * jd + j1 * jd * jd];
                    const int it = ((e1 * ne + e2) * ne + e3) * ix.idxb_max + jjb + 1 + ielem * ncoeffall;
                    w = (j1 == j) ? ((j2 == j) ? 3 * beta[it] : 2 * beta[it]) : beta[it];
                } else if (j >= j2) {
                    const int jjb = ix.idxb_block[j1 + j2 * jd + j * jd * jd];
                    const int it = ((e3 * ne + e2) * ne + e1) * ix.idxb_max + jjb + 1 + ielem * ncoeffall;
                    w = (j2 == j) ? 2 * beta[it] : beta[it];
                } else {
                    const int jjb = ix.idxb_block[j1 + j * jd + j2 * jd * jd];
                    const int it = ((e2 * ne + e3) * ne + e1) * ix.idxb_max + jjb + 1 + ielem * ncoeffall;
                    w = beta[it];
                }
                if (!bnorm && j1 > j) w *= (j1 + 1) / (j + 1.0);
                if (bnorm) w /= (j + 1);
                out.push_back(w);
            }
    return out;
}

void y_triple_weights(const SnapIndex& ix, int bnorm, std::vector<int>& jjb, std::vector<double>& fac) {
    // read off the weight table with one-hot coefficient vectors
    const int nc = ix.idxb_max + 1;
    jjb.clear();
    fac.clear();
    for (int k = 0; k < ix.idxb_max; ++k) {
        std::vector<double> beta(nc, 0.0);
        beta[1 + k] = 1.0;
        const std::vector<double> t = y_beta_table(ix, beta.data(), 0, 0, 0, 1, 0, nc, bnorm);
        if (jjb.empty()) {
            jjb.assign(t.size(), -1);
            fac.assign(t.size(), 0.0);
        }
        for (size_t q = 0; q < t.size(); ++q)
            if (t[q] != 0.0) {
                if (jjb[q] >= 0) throw Error("coupling triple with two bispectrum coefficients");
                jjb[q] = k;
                fac[q] = t[q];
            }
    }
    for (int q : jjb)
        if (q < 0) throw Error("coupling triple without a bispectrum coefficient");
}

}  // namespace mdp
