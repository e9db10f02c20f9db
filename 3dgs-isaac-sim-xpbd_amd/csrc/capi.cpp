// capi.cpp — the drop-in C ABI (include/xpbd_capi.h, part 1 and 2).
//
// Mirrors native/xpbd_capi.cu: one library-global instance, 0 / -1 return
// codes with the message in xpbd_last_error and on stderr, the same call-order
// rules, and settings that persist across create / destroy like the
// reference's forward.cu globals.  The device work is done by an Engine with
// one body; the graph upload is deferred to the first xpbd_step as in the
// reference's ensureChainmailGPU (forward.cu:10519-10845).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "engine.h"
#include "xpbd_capi.h"

using namespace xpbd;

namespace {

struct Global {
    Settings set;                       // persists across create / destroy
    ColliderDev cols[kMaxColliders] = {};
    int colCount = 0;
    float colMargin = 0.0f, colFriction = 0.0f;
    bool pressAutoLog = true;
    bool volJStats = true;
    float jvw = 1.0f, jvwStd = 0.0f;
    std::unique_ptr<BodyInput> input;
    std::unique_ptr<Engine> engine;
    // explicit volume clusters, applied when the engine is built
    bool volExplicit = false;
    std::vector<int> volOff, volMembers;
    int volMinMembers = 4;
    // fixed self-collision grid (0 = the reference's displacement rule)
    float scCell = 0.0f;
    int scTable = 0;
    int nodeOrder = 1;   // device node order of the next engine (1 = Morton)
    int deviceBuilders = -1;   // where k-ring volume member lists are built (-1 by size)
    bool scFine = true;  // candidate search of one-object bodies on the predicted-position grid
    std::vector<int> shapeClusters;     // cluster of every node, empty = 1-ring
    int n = 0;
    bool stepped = false;
    bool contactOn = false;
    int contactEnabled = 0;
    // setSelfCollisionParams(false) zeroes the candidate and overflow
    // statistics (forward.cu:4539); the next build recounts them
    bool scStatsZero = false;
    std::string err;
};

Global& G()
{
    static Global g;
    return g;
}

int fail(const std::string& msg)
{
    G().err = msg;
    std::fprintf(stderr, "[xpbd_dll] %s\n", msg.c_str());
    return -1;
}

void free_all()
{
    Global& g = G();
    g.engine.reset();
    g.input.reset();
    g.volExplicit = false;
    g.volOff.clear();
    g.volMembers.clear();
    g.shapeClusters.clear();
    g.n = 0;
    g.stepped = false;
    g.contactOn = false;
    g.contactEnabled = 0;
    g.set.contactColliders = false;
    g.set.contactGround = false;
}

void push_colliders()
{
    Global& g = G();
    if (g.engine) g.engine->setColliderArray(0, g.cols, g.colCount, g.colMargin, g.colFriction);
}

}  // namespace

extern "C" {

const char* xpbd_last_error(void) { return G().err.c_str(); }
const char* xpbd_ext_backend(void) { return "hip-gfx950"; }

int xpbd_create(int n, const float* pos, const float* scales, const float* rots, const float* opacity,
                int e, const int* edges, const float* rest, const float* stiff)
{
    free_all();
    if (n <= 0 || !pos || !scales || !rots || !opacity || e < 0 || (e > 0 && (!edges || !rest || !stiff)))
        return fail("xpbd_create: invalid arguments");
    Global& g = G();
    g.input.reset(new BodyInput());
    build_input(*g.input, n, pos, scales, rots, opacity, e, edges, rest, stiff);
    g.n = n;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        free_all();
        return fail("xpbd_create: no HIP device");
    }
    return 0;
}

void xpbd_destroy(void)
{
    xpbd_set_colliders(0, nullptr, nullptr, nullptr, 0.0f, 0.0f);
    Global& g = G();
    if (g.engine) g.engine->attach(0, 0, nullptr, nullptr, 1.0f);
    free_all();
}

int xpbd_count(void) { return G().n; }

void xpbd_set_solver(int iters, float dt, float underRelax, float velDamping, float stiffnessScale,
                     float invMassScale, float distCompliance, float shapeCompliance, float shapeBlend,
                     float angleCompliance, float angleBlend)
{
    settings_solver(G().set, iters, dt, underRelax, velDamping, stiffnessScale, invMassScale, distCompliance,
                    shapeCompliance, shapeBlend, angleCompliance, angleBlend);
}

void xpbd_set_constraints(int distance, int shape, int angle, int volume)
{
    settings_constraints(G().set, distance, shape, angle, volume);
}

void xpbd_set_volume(float compliance, int ringK, int maxMembers, int leaderMinHop)
{
    settings_volume(G().set, compliance, ringK, maxMembers, leaderMinHop);
}

void xpbd_set_ground(int enabled, float ux, float uy, float uz, float height, float friction,
                     float restitution, float contactRadius, float contactSlop, float gravity)
{
    settings_ground(G().set, enabled, ux, uy, uz, height, friction, restitution, contactRadius, contactSlop, gravity);
}

void xpbd_set_object_shape(float stiffness)
{
    settings_object_shape(G().set, stiffness);
}

void xpbd_set_object_shape_gpu(int enabled) { G().set.objShapeGPU = enabled != 0; }

void xpbd_set_colliders(int count, const int* types, const float* poses, const float* dims, float margin,
                        float friction)
{
    Global& g = G();
    g.colCount = collider_list_update(g.cols, g.colCount, count, types, poses, dims);
    g.colMargin = positive_or_zero(margin);
    g.colFriction = positive_or_zero(friction);
    push_colliders();
}

int xpbd_collider_stats(int* hits, float* push, float* centroid, int maxCount)
{
    Global& g = G();
    if (!g.engine) return 0;
    return g.engine->colliderStats(0, hits, push, centroid, maxCount);
}

int xpbd_add_rigid_velocity(const float* dv, const float* dw, const float* center)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_add_rigid_velocity: no step yet");
    g.engine->addRigidVelocity(0, dv, dw, center);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(std::string("xpbd_add_rigid_velocity: ") + hipGetErrorString(e));
    return 0;
}

int xpbd_set_contact_shape(int enabled, float tau, float radiusCap)
{
    Global& g = G();
    if (!g.input) return fail("xpbd_set_contact_shape: call xpbd_create first");
    g.contactOn = enabled != 0;
    g.contactEnabled = enabled;
    g.set.contactTau = tau;
    g.set.contactCap = radiusCap;
    g.set.contactColliders = (enabled & 1) != 0;
    g.set.contactGround = (enabled & 2) != 0;
    if (g.engine) {
        std::string err;
        if (!g.engine->setContactShape(0, enabled, tau, radiusCap, err))
            return fail("xpbd_set_contact_shape: " + err);
    }
    return 0;
}

int xpbd_set_attached(int count, const int* idx, const float* pos, float weight)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_set_attached: no step yet");
    if (!g.engine->attach(0, count, idx, pos, weight)) return fail("xpbd_set_attached: device error");
    return 0;
}

void xpbd_set_self_collision(int enabled, float radiusScale, float excludeScale, int withinBody)
{
    settings_self_collision(G().set, enabled, radiusScale, excludeScale, withinBody);
    if (!enabled) G().scStatsZero = true;
}

void xpbd_self_collision_stats(int* out)
{
    if (!out) return;
    out[0] = 0; out[1] = 0; out[2] = 0;
    Global& g = G();
    if (!g.engine) return;
    std::string err;
    if (!g.engine->selfCollisionStats(0, out, err)) fail("xpbd_self_collision_stats: " + err);
    if (g.scStatsZero) {
        out[0] = 0;
        out[2] = 0;
    }
}

void xpbd_set_shape_robust(int enabled) { G().set.shapeRobust = enabled != 0; }

int xpbd_step(void)
{
    Global& g = G();
    if (!g.input) return fail("xpbd_step: call xpbd_create first");
    std::string err;
    if (!g.engine) {
        g.engine.reset(new Engine());
        g.engine->setNodeOrder(g.nodeOrder);
        g.engine->setDeviceBuilders(g.deviceBuilders);
        if (!g.engine->build({g.input.get()}, err)) {
            g.engine.reset();
            return fail("xpbd_step: " + err);
        }
        push_colliders();
        g.engine->setSelfCollisionGrid(0, g.scCell, g.scTable);
        g.engine->setSelfCollisionFine(g.scFine);
        if (!g.shapeClusters.empty() && !g.engine->setShapeClusters(0, g.shapeClusters.data(), err)) {
            g.engine.reset();
            return fail("xpbd_step: " + err);
        }
        if (g.volExplicit &&
            !g.engine->setVolumeClusters(0, g.volOff.data(), g.volMembers.data(), g.volMinMembers, err)) {
            g.engine.reset();
            return fail("xpbd_step: " + err);
        }
        if (g.contactEnabled &&
            !g.engine->setContactShape(0, g.contactEnabled, g.set.contactTau, g.set.contactCap, err)) {
            g.engine.reset();
            return fail("xpbd_step: " + err);
        }
    }
    Engine& E = *g.engine;
    E.settings(0) = g.set;
    E.press(0).autoLog = E.press(0).active ? E.press(0).autoLog : g.pressAutoLog;
    if (!E.step(1, err) || !E.sync(err)) return fail("xpbd_step: " + err);
    g.stepped = true;
    if (g.set.scEnabled && !E.press(0).active) g.scStatsZero = false;
    return 0;
}

unsigned long long xpbd_positions_device(void)
{
    Global& g = G();
    if (!g.stepped) return 0ull;
    return reinterpret_cast<unsigned long long>(g.engine->positionsDevice());
}

int xpbd_positions_host(float* out, int n)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_positions_host: no step yet");
    if (!out || n != g.n) return fail("xpbd_positions_host: size mismatch");
    if (!g.engine->getState(0, out, nullptr)) return fail("xpbd_positions_host: device copy failed");
    return 0;
}

int xpbd_compute_shapes(float deformEps, float* hostScales, float* hostQuats, int n)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_compute_shapes: no step yet");
    if (n != g.n) return fail("xpbd_compute_shapes: size mismatch");
    std::string err;
    int deformed = 0;
    if (!g.engine->computeShapes(0, deformEps, hostScales, hostQuats, deformed, err))
        return fail("xpbd_compute_shapes: " + err);
    return deformed;
}

int xpbd_ext_frame_outputs(float* posDev, float* scaleDev, float* quatDev, int quatXYZW, float deformEps,
                           float* prevDev, float moveEps, double* host4, int n)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_ext_frame_outputs: no step yet");
    if (n != g.n) return fail("xpbd_ext_frame_outputs: size mismatch");
    if (!host4) return fail("xpbd_ext_frame_outputs: host4 is null");
    std::string err;
    int deformed = 0;
    if (!g.engine->frameOutputs(0, posDev, scaleDev, quatDev, quatXYZW != 0, deformEps, prevDev, moveEps, host4,
                                deformed, err))
        return fail("xpbd_ext_frame_outputs: " + err);
    return deformed;
}

int xpbd_reset(void)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_reset: no step yet");
    g.engine->attach(0, 0, nullptr, nullptr, 1.0f);
    g.engine->resetToRest(0);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(std::string("xpbd_reset: ") + hipGetErrorString(e));
    return 0;
}

int xpbd_launch(const float* lin, const float* ang, const float* tilt)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_launch: no step yet");
    g.engine->launchFromRest(0, lin, ang, tilt);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(std::string("xpbd_launch: ") + hipGetErrorString(e));
    return 0;
}

// squashStart in press mode (forward.cu:4310-4387; native/xpbd_capi.cu:281-288)
int xpbd_press_start(int axis, float rampPerSec, float maxDispPct, int fromLo)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_press_start: step once first");
    PressState& ps = g.engine->press(0);
    ps.fromLo = fromLo != 0;
    if (axis < 0 || axis > 2) axis = 1;
    const float slabPct = 0.08f;
    maxDispPct = std::fmax(0.01f, std::fmin(1.0f, maxDispPct));
    ps.axis = axis;
    ps.ramp = rampPerSec;
    const BodyInput& in = *g.input;
    float lo = 3.402823466e+38f, hi = -3.402823466e+38f;
    for (int i = 0; i < in.n; ++i) {
        const float v = in.pos[3 * i + axis];
        lo = std::fmin(lo, v);
        hi = std::fmax(hi, v);
    }
    const float extent = hi - lo;
    ps.lo = lo;
    ps.hi = hi;
    ps.maxDisp = maxDispPct * extent;
    ps.curDisp = 0.0f;
    // the slabs must both be non-empty for the reference to start (4350-4355)
    const float topThr = hi - slabPct * extent, botThr = lo + slabPct * extent;
    int top = 0, bot = 0;
    for (int i = 0; i < in.n; ++i) {
        const float v = in.pos[3 * i + axis];
        if (v >= topThr) ++top;
        else if (v <= botThr) ++bot;
    }
    if (top == 0 || bot == 0) {
        ps.active = false;
        return fail("xpbd_press_start: squashStart failed");
    }
    ps.active = true;
    ps.autoLog = g.pressAutoLog;
    ps.ckptIdx = 0;
    ps.dwell = 0;
    return 0;
}

void xpbd_press_autolog(int enabled) { G().pressAutoLog = enabled != 0; }

void xpbd_press_stop(void)
{
    Global& g = G();
    if (!g.engine || !g.engine->press(0).active) return;
    g.engine->press(0).active = false;
    g.engine->releaseInvMass(0);
}

int xpbd_press_state(float* lo, float* hi, float* curDisp, float* maxDisp)
{
    Global& g = G();
    PressState ps;
    if (g.engine) ps = g.engine->press(0);
    const bool moving = ps.active;
    if (curDisp) *curDisp = ps.curDisp;
    if (maxDisp) *maxDisp = ps.maxDisp;
    if (lo) *lo = ps.lo + ((moving && ps.fromLo) ? ps.curDisp : 0.0f);
    if (hi) *hi = ps.hi - ((moving && !ps.fromLo) ? ps.curDisp : 0.0f);
    return ps.active ? 1 : 0;
}

// J statistics of the last committed positions, evaluated on request instead
// of after every step (forward.cu:12658-12736 computes them per step).
void xpbd_volume_stats(int enable, float* jvw, float* jvwStd)
{
    Global& g = G();
    g.volJStats = enable != 0;
    if (g.volJStats && g.stepped) {
        std::string err;
        g.engine->settings(0) = g.set;
        if (!g.engine->volumeStats(0, g.jvw, g.jvwStd, err)) fail("xpbd_volume_stats: " + err);
    }
    if (jvw) *jvw = g.jvw;
    if (jvwStd) *jvwStd = g.jvwStd;
}

int xpbd_ext_set_volume_clusters(const int* off, const int* members, int minMembers)
{
    Global& g = G();
    if (!g.input) return fail("xpbd_ext_set_volume_clusters: call xpbd_create first");
    const int n = g.n;
    if (!off) {
        g.volExplicit = false;
        g.volOff.clear();
        g.volMembers.clear();
    } else {
        if (off[0] != 0) return fail("xpbd_ext_set_volume_clusters: off[0] must be 0");
        for (int i = 0; i < n; ++i)
            if (off[i + 1] < off[i]) return fail("xpbd_ext_set_volume_clusters: offsets must be non-decreasing");
        for (int k = 0; k < off[n]; ++k)
            if (members[k] < 0 || members[k] >= n) return fail("xpbd_ext_set_volume_clusters: member out of range");
        g.volExplicit = true;
        g.volOff.assign(off, off + n + 1);
        g.volMembers.assign(members, members + off[n]);
        g.volMinMembers = minMembers;
    }
    if (g.engine) {
        std::string err;
        if (!g.engine->setVolumeClusters(0, g.volExplicit ? g.volOff.data() : nullptr, g.volMembers.data(),
                                         g.volMinMembers, err))
            return fail("xpbd_ext_set_volume_clusters: " + err);
    }
    return 0;
}

int xpbd_ext_volume_topology(int* leaders, int* memberCount, int* revCount, int* matType, float* detRest, int n)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_ext_volume_topology: no step yet");
    if (n != g.n) return fail("xpbd_ext_volume_topology: size mismatch");
    std::string err;
    g.engine->settings(0) = g.set;
    if (!g.engine->volumeTopology(0, leaders, memberCount, revCount, matType, detRest, err))
        return fail("xpbd_ext_volume_topology: " + err);
    return 0;
}

int xpbd_ext_volume_members(int* members, int len)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_ext_volume_members: no step yet");
    std::string err;
    int total = 0;
    g.engine->settings(0) = g.set;
    if (!g.engine->volumeMembers(0, members, len, total, err)) return fail("xpbd_ext_volume_members: " + err);
    return total;
}

int xpbd_ext_get_state(float* pos, float* vel, int n)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_ext_get_state: no step yet");
    if (n != g.n) return fail("xpbd_ext_get_state: size mismatch");
    if (!g.engine->getState(0, pos, vel)) return fail("xpbd_ext_get_state: device copy failed");
    return 0;
}

int xpbd_ext_set_state(const float* pos, const float* vel, int n)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_ext_set_state: no step yet");
    if (n != g.n) return fail("xpbd_ext_set_state: size mismatch");
    if (!g.engine->setState(0, pos, vel)) return fail("xpbd_ext_set_state: device copy failed");
    return 0;
}

int xpbd_ext_autotune(int substeps, int flags, char* json, int len)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_ext_autotune: no step yet");
    std::string out, err;
    Engine& E = *g.engine;
    E.settings(0) = g.set;
    if (!E.autotune(substeps, flags, out, err)) return fail("xpbd_ext_autotune: " + err);
    if (json && len > 0) {
        if ((int)out.size() + 1 > len) return fail("xpbd_ext_autotune: buffer too small");
        std::memcpy(json, out.c_str(), out.size() + 1);
    }
    return 0;
}

int xpbd_ext_set_group_lanes(int lanes)
{
    Global& g = G();
    if (!g.engine) return fail("xpbd_ext_set_group_lanes: no instance");
    if (lanes != -1 && lanes != 0 && lanes != 1)
        return fail("xpbd_ext_set_group_lanes: lanes must be -1, 0 or 1");
    g.engine->tuning().scGroupLanes = lanes;
    return 0;
}

int xpbd_ext_profile(int substeps, char* json, int len)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_ext_profile: no step yet");
    if (!json || len <= 0 || substeps <= 0) return fail("xpbd_ext_profile: invalid arguments");
    std::string out, err;
    Engine& E = *g.engine;
    E.settings(0) = g.set;
    if (!E.profile(substeps, out, err)) return fail("xpbd_ext_profile: " + err);
    if ((int)out.size() + 1 > len) return fail("xpbd_ext_profile: buffer too small");
    std::memcpy(json, out.c_str(), out.size() + 1);
    return 0;
}

int xpbd_ext_get_inv_mass(float* invMass, int n)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_ext_get_inv_mass: no step yet");
    if (n != g.n || !invMass) return fail("xpbd_ext_get_inv_mass: size mismatch");
    if (!g.engine->getInvMass(0, invMass)) return fail("xpbd_ext_get_inv_mass: device copy failed");
    return 0;
}

int xpbd_ext_set_node_order(int mode)
{
    if (mode != 0 && mode != 1) return fail("xpbd_ext_set_node_order: mode must be 0 (caller order) or 1 (Morton)");
    G().nodeOrder = mode;
    return 0;
}

int xpbd_ext_self_collision_grid_info(float* info)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_ext_self_collision_grid_info: no step yet");
    if (!info) return fail("xpbd_ext_self_collision_grid_info: null output");
    std::string err;
    if (!g.engine->selfCollisionGridInfo(0, info, err)) return fail("xpbd_ext_self_collision_grid_info: " + err);
    return 0;
}

int xpbd_ext_self_collision_paths(int* out)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_ext_self_collision_paths: no step yet");
    if (!out) return fail("xpbd_ext_self_collision_paths: null output");
    std::string err;
    if (!g.engine->selfCollisionPaths(0, out, err)) return fail("xpbd_ext_self_collision_paths: " + err);
    return 0;
}

int xpbd_ext_set_device_builders(int mode)
{
    if (mode != -1 && mode != 0 && mode != 1)
        return fail("xpbd_ext_set_device_builders: mode must be -1 (by size), 0 (host) or 1 (GPU)");
    Global& g = G();
    g.deviceBuilders = mode;
    if (g.engine) g.engine->setDeviceBuilders(mode);
    return 0;
}

int xpbd_ext_set_shape_clusters(const int* clusterOf, int n)
{
    Global& g = G();
    if (!g.input) return fail("xpbd_ext_set_shape_clusters: call xpbd_create first");
    if (clusterOf && n != g.n) return fail("xpbd_ext_set_shape_clusters: size mismatch");
    if (clusterOf)
        for (int i = 0; i < n; ++i)
            if (clusterOf[i] < -1 || clusterOf[i] >= n) return fail("xpbd_ext_set_shape_clusters: cluster id out of range");
    if (clusterOf) g.shapeClusters.assign(clusterOf, clusterOf + n);
    else g.shapeClusters.clear();
    if (g.engine) {
        std::string err;
        if (!g.engine->setShapeClusters(0, clusterOf, err)) return fail("xpbd_ext_set_shape_clusters: " + err);
    }
    return 0;
}

int xpbd_ext_set_gauss_seidel(int distance, int volume)
{
    Settings& s = G().set;
    s.distGS = distance != 0;
    s.volGS = volume != 0;
    return 0;
}

int xpbd_ext_gs_info(int* info)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_ext_gs_info: no step yet");
    if (!info) return fail("xpbd_ext_gs_info: null output");
    std::string err;
    g.engine->settings(0) = g.set;
    if (!g.engine->gsInfo(0, info, err)) return fail("xpbd_ext_gs_info: " + err);
    return 0;
}

int xpbd_ext_distance_gs_edges(int* ab, float* rest, float* stiff, int* colorOff)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_ext_distance_gs_edges: no step yet");
    std::string err;
    if (!g.engine->distanceGSEdges(0, ab, rest, stiff, colorOff, err)) return fail("xpbd_ext_distance_gs_edges: " + err);
    return 0;
}

int xpbd_ext_volume_gs_order(int* order, int* colorOff)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_ext_volume_gs_order: no step yet");
    std::string err;
    g.engine->settings(0) = g.set;
    if (!g.engine->volumeGSOrder(0, order, colorOff, err)) return fail("xpbd_ext_volume_gs_order: " + err);
    return 0;
}

int xpbd_ext_set_self_collision_grid(float cellSize, int tableSize)
{
    Global& g = G();
    if (tableSize != 0 && (tableSize < 1024 || (tableSize & (tableSize - 1))))
        return fail("xpbd_ext_set_self_collision_grid: table size must be a power of two >= 1024");
    g.scCell = (std::isfinite(cellSize) && cellSize > 0.0f) ? cellSize : 0.0f;
    g.scTable = tableSize;
    if (g.engine) g.engine->setSelfCollisionGrid(0, g.scCell, g.scTable);
    return 0;
}

int xpbd_ext_angle_pairs(int* pairNext, float* restCos, int m)
{
    Global& g = G();
    if (!g.input) return fail("xpbd_ext_angle_pairs: call xpbd_create first");
    if (m != g.input->rowOff[g.n]) return fail("xpbd_ext_angle_pairs: size mismatch");
    AngleTopo t;
    build_angle_pairs(*g.input, t);
    if (pairNext && m) std::memcpy(pairNext, t.pairNext.data(), sizeof(int) * m);
    if (restCos && m) std::memcpy(restCos, t.restCos.data(), sizeof(float) * m);
    return 0;
}

int xpbd_ext_get_settings(float* out, int len)
{
    const Global& g = G();
    const Settings& s = g.set;
    const float v[42] = {
        (float)s.iters, s.dt, s.underRelax, s.velDamping, s.stiffnessScale, s.invMassScale, s.distCompliance,
        s.shapeCompliance, s.shapeBlend, s.angleCompliance, s.angleBlend, (float)s.useDistance, (float)s.useShape,
        (float)s.useAngle, (float)s.useVolume, s.volCompliance, (float)s.volRingK, (float)s.volMaxMembers,
        (float)s.volLeaderMinHop, (float)s.groundEnabled, s.groundN[0], s.groundN[1], s.groundN[2], s.groundHeight,
        s.groundFriction, s.groundRestitution, s.groundRadius, s.groundSlop, s.groundGravity, s.objShapeStiffness,
        (float)s.objMinComponent, (float)s.objShapeGPU, (float)s.scEnabled, s.scRadiusScale, s.scExcludeScale,
        (float)s.scWithinBody, (float)s.shapeRobust, (float)g.colCount, g.colMargin, g.colFriction,
        (float)g.pressAutoLog, (float)g.volJStats};
    if (!out || len < 42) return fail("xpbd_ext_get_settings: need 42 floats");
    std::memcpy(out, v, sizeof(v));
    return 42;
}

int xpbd_ext_compute_shapes_noalloc(int n, const float* restDev, const float* posDev, const int* rowOffDev,
                                    const int* nbrDev, const float* scalesDev, const float* rotsDev, float deformEps,
                                    float* outScalesDev, float* outQuatsDev)
{
    if (n <= 0 || !restDev || !posDev || !rowOffDev || !nbrDev || !scalesDev || !rotsDev || !outScalesDev ||
        !outQuatsDev)
        return fail("xpbd_ext_compute_shapes_noalloc: invalid arguments");
    if (n > deformed_shape_scratch_max())
        return fail("xpbd_ext_compute_shapes_noalloc: n exceeds the limit of " +
                    std::to_string(deformed_shape_scratch_max()) + " Gaussians per call");
    unsigned int* counts = nullptr;
    std::string err;
    if (!deformed_shape_scratch(&counts, err)) return fail("xpbd_ext_compute_shapes_noalloc: " + err);
    launch_deformed_shapes(nullptr, 0, n, reinterpret_cast<const float3*>(restDev),
                           reinterpret_cast<const float3*>(posDev), rowOffDev, nbrDev, scalesDev, rotsDev, deformEps,
                           outScalesDev, outQuatsDev, counts);
    hipError_t e = hipGetLastError();
    unsigned int total = 0;
    if (e == hipSuccess)
        e = hipMemcpy(&total, counts + deformed_shape_blocks(n), sizeof(unsigned int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(std::string("xpbd_ext_compute_shapes_noalloc: ") + hipGetErrorString(e));
    return (int)total;
}

int xpbd_ext_set_self_collision_fine(int enabled)
{
    Global& g = G();
    g.scFine = enabled != 0;
    if (g.engine) g.engine->setSelfCollisionFine(g.scFine);
    return 0;
}

int xpbd_ext_self_collision_lists(unsigned int* keys, int* sorted, int* counts, int* contacts, int n)
{
    Global& g = G();
    if (!g.stepped) return fail("xpbd_ext_self_collision_lists: no step yet");
    if (n != g.n) return fail("xpbd_ext_self_collision_lists: size mismatch");
    std::string err;
    if (!g.engine->selfCollisionLists(0, keys, sorted, counts, contacts, err))
        return fail("xpbd_ext_self_collision_lists: " + err);
    return 0;
}

}  // extern "C"
