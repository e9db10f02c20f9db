// kernels.h — launcher declarations shared by the host orchestration.
#pragma once
#include <hip/hip_runtime.h>

#include "xpbd_device.h"

namespace xpbd {

// Node-space launch description shared by every per-node kernel.
struct NodeArgs {
    int nPad;                 // padded node count (multiple of kTile)
    const int* tileBody;      // [nPad / kTile]
    const BodyDev* bodies;    // [B]
};

// Buffers the per-node epilogue touches.
struct EpiArgs {
    float3* x;                // committed positions (friction reference, commit)
    float3* v;
    const ColliderDev* cols;
    int* colHits;
    float* colPush;
    const float* cM;          // contact ellipsoids k^2 Sigma [N x 6] or null
    const float* cRmax;       // [N] or null
    const float* rg;          // [N] ground support radius or null
};

void launch_predict(hipStream_t s, const NodeArgs& na, const float3* X, float3* V, const float* invMass, float4* P);
void launch_distance_jacobi(hipStream_t s, const NodeArgs& na, bool first, bool last, const float4* Pin, float4* Pout,
                            const int* rowOff, const int* nbr, const float* nrest, const float* nalpha, float* lam,
                            unsigned epi, const EpiArgs& ea);
void launch_epilogue(hipStream_t s, const NodeArgs& na, float4* P, unsigned epi, const EpiArgs& ea);
void launch_distance_alpha(hipStream_t s, const NodeArgs& na, const int* rowOff, const float* nstiff, float* nalpha);
void launch_attach(hipStream_t s, int count, const int* idx, const float* pos, int nodeBegin, int nodeEnd, float3* X,
                   float3* V, float* invMass, float* objW, float weight);
void launch_rigid_velocity(hipStream_t s, int nodeBegin, int nodeEnd, const float3* X, float3* V, const float* invMass,
                           float3 dv, float3 dw, float3 c);
void launch_fill_float(hipStream_t s, int n, float* a, float v);
void launch_fill_int(hipStream_t s, int n, int* a, int v);

}  // namespace xpbd
