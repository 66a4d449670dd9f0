// Host-side SNAP index spaces and coefficient tables (built once per parameter set, then staged
// to the device).  Layouts follow src/Julia/Potential/cpuSnap.cpp:71-246 so that the tables can be
// handed back through the drop-in cpuInitSna.
#pragma once
#include <vector>

namespace mdp {

struct SnapIndex {
    int J = 0;                      // twojmax
    int idxu_max = 0, idxz_max = 0, idxb_max = 0, idxcg_max = 0;
    std::vector<int> idxu_block;    // [J+1]
    std::vector<int> idxz;          // [10*idxz_max]: j1 j2 j ma1min ma2max na mb1min mb2max nb jju
    std::vector<int> idxz_block;    // [(J+1)^3] at j + j2*(J+1) + j1*(J+1)^2
    std::vector<int> idxb;          // [3*idxb_max]: j1 j2 j (j >= j1)
    std::vector<int> idxb_block;    // [(J+1)^3]
    std::vector<int> idxcg_block;   // [(J+1)^3]
    std::vector<double> cglist;     // [idxcg_max]
    std::vector<double> rootpq;     // [(J+2)^2], sqrt(p/q) at p*(J+1)+q

    void build(int twojmax, const double* factorial);   // factorial == nullptr: exact n!
    int half_count() const;         // number of U elements with 2*mb <= j
};

int count_u(int J);
int count_z(int J);
int count_b(int J);
int count_cg(int J);

}  // namespace mdp
