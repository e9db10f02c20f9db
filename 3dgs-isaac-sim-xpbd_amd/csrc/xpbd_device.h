// xpbd_device.h — device-side data layout of the MI355X XPBD solver.
//
// Node index space: every body (independent simulated instance) owns a node
// range that starts on a multiple of kTile (one wavefront).  A wave therefore
// belongs to exactly one body; kernels look the body up once per wave from
// tileBody[] and read its parameter block with scalar loads.  Padding nodes
// between bodies are never referenced by any CSR and are skipped by the
// `i >= body.nodeEnd` test.
//
// Per-node arrays (HBM, N_pad entries):
//   x        float3  committed positions (the ABI's float3[n] device buffer)
//   v        float3  velocities
//   p0, p1   float4  predicted positions, ping-pong; .w carries invMass so a
//                    neighbour gather is one 16-byte load
//   invMass  float
// Directed CSR (M = 2E entries, rows in node order, input edge order):
//   rowOff int[N_pad+1], nbr int[M], nrest float[M], nalpha float[M] (alpha
//   tilde = compliance / max(stiff * scale, 1e-6) / max(dt^2, 1e-8), rebuilt
//   when solver settings change), lam float[M] (per-substep multiplier).
#pragma once
#include <hip/hip_runtime.h>

namespace xpbd {

constexpr int kTile = 64;          // nodes per tile = one wave64
constexpr int kMaxColliders = 32;  // forward.cu:138 MAX_KIN_COLLIDERS
constexpr int kColAcc = 7;         // forward.cu:151 KINCOL_ACC

// Per-body parameter block (read-only inside a substep).
struct BodyDev {
    int nodeBegin, nodeEnd;
    int distanceOn, iters;
    float dt, underRelax, velDamping, invMassScale;
    float distCompliance, stiffnessScale;
    int distAlphaUniform;      // one stiffness on every edge: alpha tilde is distAlphaU
    float distAlphaU;
    int distGSOn, volGSOn;     // graph-coloured Gauss-Seidel instead of Jacobi / gather
    float grav[3];
    // ground (forward.cu:4430-4446); groundOn = enabled and no press active
    int groundOn;
    float gn[3];
    float planeD, slop, friction, restitution, restingSpeed;
    // kinematic colliders: [colBegin, colBegin + colCount) of the collider array
    int colBegin, colCount;
    float colMargin, colFriction;
    // press plates (forward.cu:9411-9424)
    int pressOn, pressAxis;
    float pressLo, pressHi;
    // contact ellipsoids (forward.cu:7358-7399)
    int contactColliders, contactGround;
    // umbrella angle constraint (forward.cu:4990-5130)
    int angleOn;
    float angleCompliance, angleBlend;
    // volume constraint (forward.cu:8531-9243)
    int volumeOn;
    float volCompliance;
    int volMinMembers;         // rest-precompute gate of the body's clusters (forward.cu:7745)
    // per-node shape matching (forward.cu:6088-6383)
    int shapeOn, shapeRobust;
    int shapeClusters;         // node clusters replace the 1-ring (extension)
    float shapeCompliance, shapeBlend;
    // object (connected component) shape matching (forward.cu:6801-6937)
    int objOn, objBase, objK, objWeighted, objBoxes, objActiveAllowed;
    float objStiffness;
    // self-collision (forward.cu:6986-7322); the hash table of the body is
    // [scTableBase, scTableBase + scTableMask + 1) of the shared cell arrays
    int scOn, scWithin, scTableBase, scTableMask;
    float scDc, scSearch, scExcl2, scFastRel, scCellOverride;
    // candidate search through the predicted-position grid (extension; same
    // lists): cells of edge >= search
    int scFine, scFineBase, scFineMask;
    float scFineInv;
};

// forward.cu:128-136 KinCollider, same field meaning.
struct ColliderDev {
    float R[9];
    float t[3];
    float h[3];
    int type;
    float Rp[9];
    float tp[3];
};

// Epilogue bits: per-node work appended to the last cross-node pass of an
// iteration, in the reference order press -> colliders -> ground -> velocity.
enum EpilogueBits : unsigned {
    kEpiPress = 1u,
    kEpiColliders = 2u,
    kEpiColliderHits = 4u,   // count hits (last iteration only)
    kEpiGround = 8u,
    kEpiVelocity = 16u,      // commit x, v (last iteration)
};

}  // namespace xpbd
