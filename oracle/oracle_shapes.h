/* oracle_shapes.h — TEST INFRASTRUCTURE ONLY (see xpbd_oracle.cpp). */
#pragma once
#include <vector>

#include "oracle_volume.h"

namespace xo {

M3 mmul(const M3& A, const M3& B);
M3 mtrans(const M3& A);
void eig_sym3_vec(const M3& S, float ev[3], M3& V);
bool polar_double(const M3& A, M3& R);
void shape_matching(const std::vector<V3>& pin, std::vector<V3>& pout);

// Cluster shape matching (extension): clusters of nodes replace the 1-ring.
struct ShapeClusters {
    bool on = false;
    std::vector<int> off, mem, clusterOf;
    std::vector<V3> comRest;
};
extern ShapeClusters shc;
void shape_clusters_set(const int* clusterOf, int n);
void shape_cluster_matching(const std::vector<V3>& pin, std::vector<V3>& pout);
int deformed_shapes(float deformEps, std::vector<float>& outScale, std::vector<float>& outQuat);
void contact_pack(const float* sc, const float* qt, const float* op, float tau, float rcap, std::vector<float>& M,
                  std::vector<float>& rmax);

}  // namespace xo
