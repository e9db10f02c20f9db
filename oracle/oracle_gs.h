/* oracle_gs.h — TEST INFRASTRUCTURE ONLY (see xpbd_oracle.cpp).
 * Graph-coloured Gauss-Seidel passes: distance edges and volume clusters. */
#pragma once
#include <vector>

#include "xpbd_oracle_internal.h"

namespace xo {

// buildDistanceGSEdges (forward.cu:10407-10517): undirected edges a < b,
// colour-major, rest / stiffness averaged over the directed duplicates.
struct DistanceGS {
    bool built = false;
    std::vector<int> a, b, colorOff;
    std::vector<float> rest, stiff, lam;
    int oneWay = 0, maxDeg = 0;
};

// buildVolumeGSColors (forward.cu:9745-9812): leaders in colour-major order.
struct VolumeGS {
    bool built = false, skipped = false;
    long long version = -1;             // volume topology the colouring belongs to
    int lowerBound = 0;
    std::vector<int> order, colorOff;
};

extern DistanceGS dgs;
extern VolumeGS vgs;

void distance_gs_build();
void distance_gs_iteration(std::vector<V3>& pos, float dt);
void volume_gs_build();
void volume_gs_iteration(std::vector<V3>& pos, float dt);

}  // namespace xo
