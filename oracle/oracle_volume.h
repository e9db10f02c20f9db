/* oracle_volume.h — TEST INFRASTRUCTURE ONLY (see xpbd_oracle.cpp). */
#pragma once
#include <vector>

#include "xpbd_oracle_internal.h"

namespace xo {

struct M3 { float m[3][3]; };   // column-major like glm: m[col][row]

struct VolumeState {
    bool built = false;
    int minMembers = 4;                 // precompute gate (forward.cu:7745 rejects < 4)
    int k = -1, cap = -1, r = -1;       // parameters the topology was built with
    bool explicitTopo = false;
    long long version = 0;              // bumped by every topology (re)build
    std::vector<char> isLeader;
    std::vector<int> memOff, memIdx, revOff, revIdx;
    std::vector<float> detRest, Vrest, restLen, lam, cS, coef;
    std::vector<int> matType;
    std::vector<V3> cC;
};

extern VolumeState vol;

float det3(const M3& m);
bool scale_shape_det(const M3& S, float& s, float& detHat);
bool inverse_safe(const M3& m, M3& inv, float epsDet);
void eig_sym3(const M3& S, float ev[3]);

void volume_build_reference(int k, int cap, int leaderR);
void volume_set_topology(const std::vector<char>& isLeader, const std::vector<std::vector<int>>& members,
                         int minMembers);
void volume_precompute();
void volume_iteration(std::vector<V3>*& pin, std::vector<V3>*& pout);
void volume_J(const std::vector<V3>& pos, std::vector<float>& J);
void volume_stats(const std::vector<V3>& pos, float& mean, float& stdv);

}  // namespace xo
