/*
 * xpbd_oracle_internal.h — TEST INFRASTRUCTURE ONLY (see xpbd_oracle.cpp).
 * Shared state of the serial CPU restatement of the reference XPBD path.
 */
#pragma once
#include <cmath>
#include <string>
#include <vector>

namespace xo {

// Threads of the order-independent per-node / per-cluster loops (every
// iteration writes only its own entries, so the result does not depend on the
// count); 1 unless the CPU-baseline timing asks for more.
extern int g_threads;
#define XO_PARALLEL_FOR _Pragma("omp parallel for num_threads(xo::g_threads) schedule(static)")

struct V3 { float x, y, z; };
inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len2(V3 a) { return dot(a, a); }
inline float len(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// forward.cu:128-136 KinCollider
struct Collider {
    float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float t[3] = {0, 0, 0};
    float h[3] = {0, 0, 0};
    int type = 0;
    float Rp[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float tp[3] = {0, 0, 0};
};

// Library-global settings with the reference defaults (forward.cu:41-612).
struct Settings {
    // XPBD (forward.cu:79-93)
    int iters = 20;
    float dt = 1.0f / 60.0f, underRelax = 0.6f, velDamping = 0.59f;
    float stiffnessScale = 1.0f, invMassScale = 1.0f, distCompliance = 1e-3f;
    float shapeCompliance = 0.0f, shapeBlend = 0.1f, angleCompliance = 0.0f, angleBlend = 0.2f;
    bool useDistance = true, useShape = false, useAngle = false, useVolume = false;
    bool shapeRobust = true;
    // ground (forward.cu:101-115)
    bool groundEnabled = false;
    float groundN[3] = {0, 0, 1};
    float groundHeight = 0, groundFriction = 0.5f, groundRestitution = 0.3f;
    float groundRadius = 0, groundSlop = 0, groundGravity = 0;
    // colliders (forward.cu:138-151)
    Collider col[32];
    int colCount = 0;
    float colMargin = 0, colFriction = 0;
    // contact shapes (forward.cu:160-166)
    bool contactColliders = false, contactGround = false;
    float contactTau = 0.2f, contactCap = 0.0f;
    // volume (forward.cu:313-345)
    float volCompliance = 1e-6f, volAniso = 0.05f;
    int volRingK = 1, volMaxMembers = 64, volLeaderMinHop = 1, volMemberSelect = 1;
    bool volJStats = true;
    // object shape (forward.cu:169-172, 195)
    float objShapeStiffness = 0.0f;
    int objMinComponent = 50;
    bool objShapeGPU = true;
    // self-collision (forward.cu:224-256)
    bool scEnabled = false, scWithinBody = false;
    float scRadiusScale = 1.5f, scExcludeScale = 2.0f, scSearchScale = 1.5f, scFastRel = 0.0f;
    float scCellOverride = 0.0f;   // extension: fixed cell size instead of the displacement rule
    int scTableOverride = 0;       // extension: fixed hash table size (power of two)
    // graph-coloured Gauss-Seidel passes (forward.cu:310-313, off by default)
    bool distGS = false, volGS = false;
    // squash / press (forward.cu:535-575)
    bool sqActive = false, sqPress = false, sqPressNext = false, sqPressFromLo = false, sqSlip = true;
    bool sqAutoLog = true;
    int sqAxis = 1, sqDwellFrames = 90, sqCkptIdx = 0, sqDwell = 0;
    float sqRamp = 0.15f, sqMaxDisp = 0.30f, sqCurDisp = 0.0f, sqLo = 0.0f, sqHi = 0.0f;
};

struct State {
    int n = 0;
    bool stepped = false, initialised = false, colStatsValid = false;
    std::string err;
    // graph (loadGraph CSR)
    std::vector<int> off, nbr;
    std::vector<float> ndist, nstiff;
    // per node
    std::vector<V3> rest, x, v, p, q;
    std::vector<float> invMass, invMassSaved, objW;
    std::vector<float> lamA, lamB;
    bool attActive = false;
    // inputs kept for shapes / volume
    std::vector<float> scales, rots, opacity;
    bool contactOn = false;
    // collider statistics
    std::vector<int> colHits = std::vector<int>(32, 0);
    std::vector<float> colPush = std::vector<float>(32 * 7, 0.0f);
    // press slabs
    std::vector<int> sqTop, sqBot;
    std::vector<V3> sqTopRest, sqBotRest;
};

extern State g;
extern Settings s;
int fail(const std::string& msg);

// Extension hooks implemented in xpbd_oracle_ext.cpp (volume, shape matching,
// self-collision, object shape, contact shapes, deformed shapes).
void extensions_on_init();
void extensions_free();
void extensions_on_reset();
void extensions_begin_step();
void extensions_after_predict();
void extensions_before_iterations();
bool extensions_distance(std::vector<V3>& pin, std::vector<V3>& pout, int it);
void extensions_iteration(std::vector<V3>*& pin, std::vector<V3>*& pout, int it);
void extensions_self_collision(std::vector<V3>& pin);
bool extensions_after_iterations(std::vector<V3>& pin, bool groundOn);   // true: re-project on the ground
float contact_rmax(int i);
bool contact_colliders_on();
float contact_support(int i, float nx, float ny, float nz);
float ground_contact_radius(int i);

}  // namespace xo
