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

// Volume-constraint buffers (node space; leaders own clusters).
struct VolumeArgs {
    int nActive;              // leaders with matType 0
    const int* active;        // [nActive]
    const int* volOff;        // [nPad + 1] members CSR by leader
    const int* volIdx;
    const int* revOff;        // [nPad + 1] clusters containing each node
    const int* revIdx;
    float* detRest;
    float* Vrest;
    float* restLen;
    int* matType;
    float* lam;               // [nPad] per leader
    float4* bbC;              // [nPad] centroid + coef
    float4* bbS;              // [nPad] S^-1 00 11 22 10
    float2* bbS2;             // [nPad] S^-1 20 21
};

void launch_vol_precompute(hipStream_t s, const NodeArgs& na, const float3* rest, const int* rowOff, const int* nbr,
                           const VolumeArgs& va, int minMembers, float aniso);
void launch_vol_solve(hipStream_t s, const NodeArgs& na, bool wave, bool first, bool last, const float4* P,
                      const VolumeArgs& va);
void launch_vol_apply(hipStream_t s, const NodeArgs& na, const float4* Pin, float4* Pout, const VolumeArgs& va,
                      unsigned epi, const EpiArgs& ea);
void launch_vol_J(hipStream_t s, const NodeArgs& na, const float3* X, const VolumeArgs& va, float* outJ);

void launch_shape_match(hipStream_t s, const NodeArgs& na, const float4* Pin, float4* Pout, const float3* rest,
                        const int* rowOff, const int* nbr, unsigned epi, const EpiArgs& ea);
// scales / rots / outScale / outQuat are the body's own arrays (local index);
// node arrays are indexed from nodeBegin.
void launch_deformed_shapes(hipStream_t s, int nodeBegin, int n, const float3* rest, const float3* cur,
                            const int* rowOff, const int* nbr, const float* scales, const float* rots, float deformEps,
                            float* outScale, float* outQuat, unsigned int* ctr);
void launch_contact_pack(hipStream_t s, int nodeBegin, int n, const float* sc, const float* qt, const float* op,
                         float tau, float rcap, float* M, float* rmax);
void launch_contact_ground(hipStream_t s, const NodeArgs& na, const float* M, float* rg);

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
