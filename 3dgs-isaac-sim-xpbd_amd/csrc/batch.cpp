// batch.cpp — the batched many-object C API (include/xpbd_capi.h, part 3).
//
// A batch is one Engine holding every object as a body.  Per-object settings
// go straight into the engine (they live in device parameter blocks, so
// changing them does not recapture the graph); the engine is built at create
// time, so every call is valid from then on.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "engine.h"
#include "xpbd_capi.h"

using namespace xpbd;

struct xpbd_batch {
    std::vector<std::unique_ptr<BodyInput>> inputs;
    std::vector<int> nodeOff;
    Engine engine;
    std::string err;
};

namespace {

thread_local std::string t_createErr;

int bfail(xpbd_batch* b, const std::string& msg)
{
    b->err = msg;
    std::fprintf(stderr, "[xpbd_dll] %s\n", msg.c_str());
    return -1;
}

bool valid_obj(const xpbd_batch* b, int obj) { return obj >= 0 && obj < (int)b->inputs.size(); }

// Runs f(body) for obj, or for every body when obj < 0.
template <typename F>
int for_objects(xpbd_batch* b, int obj, const char* what, F f)
{
    if (!b) return -1;
    if (obj >= (int)b->inputs.size()) return bfail(b, std::string(what) + ": object out of range");
    if (obj >= 0) {
        f(obj);
        return 0;
    }
    for (int k = 0; k < (int)b->inputs.size(); ++k) f(k);
    return 0;
}

}  // namespace

extern "C" {

xpbd_batch* xpbd_batch_create(int count, const int* nodeOff, const float* pos, const float* scales, const float* rots,
                              const float* opacity, const int* edgeOff, const int* edges, const float* rest,
                              const float* stiff)
{
    auto failc = [](const std::string& m) -> xpbd_batch* {
        t_createErr = m;
        std::fprintf(stderr, "[xpbd_dll] %s\n", m.c_str());
        return nullptr;
    };
    if (count <= 0 || !nodeOff || !edgeOff || !pos || !scales || !rots || !opacity)
        return failc("xpbd_batch_create: invalid arguments");
    if (nodeOff[0] != 0 || edgeOff[0] != 0) return failc("xpbd_batch_create: offsets must start at 0");
    for (int k = 0; k < count; ++k) {
        if (nodeOff[k + 1] <= nodeOff[k]) return failc("xpbd_batch_create: every object needs at least one node");
        if (edgeOff[k + 1] < edgeOff[k]) return failc("xpbd_batch_create: edge offsets must be non-decreasing");
    }
    if (edgeOff[count] > 0 && (!edges || !rest || !stiff)) return failc("xpbd_batch_create: missing edge arrays");
    std::unique_ptr<xpbd_batch> b(new xpbd_batch());
    b->nodeOff.assign(nodeOff, nodeOff + count + 1);
    std::vector<const BodyInput*> in;
    for (int k = 0; k < count; ++k) {
        const size_t n0 = nodeOff[k], e0 = edgeOff[k];
        const int n = nodeOff[k + 1] - nodeOff[k], e = edgeOff[k + 1] - edgeOff[k];
        b->inputs.emplace_back(new BodyInput());
        build_input(*b->inputs.back(), n, pos + 3 * n0, scales + 3 * n0, rots + 4 * n0, opacity + n0, e,
                    e ? edges + 2 * e0 : nullptr, e ? rest + e0 : nullptr, e ? stiff + e0 : nullptr);
        in.push_back(b->inputs.back().get());
    }
    std::string err;
    if (!b->engine.build(in, err)) return failc("xpbd_batch_create: " + err);
    return b.release();
}

void xpbd_batch_destroy(xpbd_batch* b) { delete b; }

const char* xpbd_batch_last_error(const xpbd_batch* b) { return b ? b->err.c_str() : t_createErr.c_str(); }

int xpbd_batch_objects(const xpbd_batch* b) { return b ? (int)b->inputs.size() : 0; }

int xpbd_batch_count(const xpbd_batch* b, int obj)
{
    if (!b) return 0;
    if (obj < 0) return b->nodeOff.back();
    return valid_obj(b, obj) ? b->inputs[obj]->n : 0;
}

int xpbd_batch_set_solver(xpbd_batch* b, int obj, int iters, float dt, float underRelax, float velDamping,
                          float stiffnessScale, float invMassScale, float distCompliance, float shapeCompliance,
                          float shapeBlend, float angleCompliance, float angleBlend)
{
    return for_objects(b, obj, "xpbd_batch_set_solver", [&](int k) {
        settings_solver(b->engine.settings(k), iters, dt, underRelax, velDamping, stiffnessScale, invMassScale,
                        distCompliance, shapeCompliance, shapeBlend, angleCompliance, angleBlend);
    });
}

int xpbd_batch_set_constraints(xpbd_batch* b, int obj, int distance, int shape, int angle, int volume)
{
    return for_objects(b, obj, "xpbd_batch_set_constraints",
                       [&](int k) { settings_constraints(b->engine.settings(k), distance, shape, angle, volume); });
}

int xpbd_batch_set_volume(xpbd_batch* b, int obj, float compliance, int ringK, int maxMembers, int leaderMinHop)
{
    return for_objects(b, obj, "xpbd_batch_set_volume", [&](int k) {
        settings_volume(b->engine.settings(k), compliance, ringK, maxMembers, leaderMinHop);
    });
}

int xpbd_batch_set_ground(xpbd_batch* b, int obj, int enabled, float ux, float uy, float uz, float height,
                          float friction, float restitution, float contactRadius, float contactSlop, float gravity)
{
    return for_objects(b, obj, "xpbd_batch_set_ground", [&](int k) {
        settings_ground(b->engine.settings(k), enabled, ux, uy, uz, height, friction, restitution, contactRadius,
                        contactSlop, gravity);
    });
}

int xpbd_batch_set_object_shape(xpbd_batch* b, int obj, float stiffness)
{
    return for_objects(b, obj, "xpbd_batch_set_object_shape",
                       [&](int k) { settings_object_shape(b->engine.settings(k), stiffness); });
}

int xpbd_batch_set_self_collision(xpbd_batch* b, int obj, int enabled, float radiusScale, float excludeScale,
                                  int withinBody)
{
    return for_objects(b, obj, "xpbd_batch_set_self_collision", [&](int k) {
        settings_self_collision(b->engine.settings(k), enabled, radiusScale, excludeScale, withinBody);
    });
}

int xpbd_batch_set_self_collision_grid(xpbd_batch* b, int obj, float cellSize, int tableSize)
{
    if (!b) return -1;
    if (tableSize != 0 && (tableSize < 1024 || (tableSize & (tableSize - 1))))
        return bfail(b, "xpbd_batch_set_self_collision_grid: table size must be a power of two >= 1024");
    return for_objects(b, obj, "xpbd_batch_set_self_collision_grid",
                       [&](int k) { b->engine.setSelfCollisionGrid(k, positive_or_zero(cellSize), tableSize); });
}

int xpbd_batch_set_gauss_seidel(xpbd_batch* b, int obj, int distance, int volume)
{
    return for_objects(b, obj, "xpbd_batch_set_gauss_seidel", [&](int k) {
        b->engine.settings(k).distGS = distance != 0;
        b->engine.settings(k).volGS = volume != 0;
    });
}

int xpbd_batch_set_shape_robust(xpbd_batch* b, int obj, int enabled)
{
    return for_objects(b, obj, "xpbd_batch_set_shape_robust",
                       [&](int k) { b->engine.settings(k).shapeRobust = enabled != 0; });
}

int xpbd_batch_set_volume_clusters(xpbd_batch* b, int obj, const int* off, const int* members, int minMembers)
{
    if (!b) return -1;
    if (!valid_obj(b, obj)) return bfail(b, "xpbd_batch_set_volume_clusters: object out of range");
    const int n = b->inputs[obj]->n;
    if (off) {
        if (off[0] != 0) return bfail(b, "xpbd_batch_set_volume_clusters: off[0] must be 0");
        for (int i = 0; i < n; ++i)
            if (off[i + 1] < off[i]) return bfail(b, "xpbd_batch_set_volume_clusters: offsets must be non-decreasing");
        for (int k = 0; k < off[n]; ++k)
            if (members[k] < 0 || members[k] >= n) return bfail(b, "xpbd_batch_set_volume_clusters: member out of range");
    }
    std::string err;
    if (!b->engine.setVolumeClusters(obj, off, members, minMembers, err))
        return bfail(b, "xpbd_batch_set_volume_clusters: " + err);
    return 0;
}

int xpbd_batch_set_shape_clusters(xpbd_batch* b, int obj, const int* clusterOf, int n)
{
    if (!b) return -1;
    if (!valid_obj(b, obj)) return bfail(b, "xpbd_batch_set_shape_clusters: object out of range");
    if (clusterOf && n != b->inputs[obj]->n) return bfail(b, "xpbd_batch_set_shape_clusters: size mismatch");
    std::string err;
    if (!b->engine.setShapeClusters(obj, clusterOf, err)) return bfail(b, "xpbd_batch_set_shape_clusters: " + err);
    return 0;
}

int xpbd_batch_set_colliders(xpbd_batch* b, int obj, int count, const int* types, const float* poses,
                             const float* dims, float margin, float friction)
{
    if (!b) return -1;
    if (count > 0 && (!types || !poses || !dims)) return bfail(b, "xpbd_batch_set_colliders: missing arrays");
    return for_objects(b, obj, "xpbd_batch_set_colliders",
                       [&](int k) { b->engine.setColliders(k, count, types, poses, dims, margin, friction); });
}

int xpbd_batch_set_attached(xpbd_batch* b, int requests, const int* objs, const int* off, const int* idx,
                            const float* pos, float weight)
{
    if (!b) return -1;
    if (requests <= 0) return 0;
    if (!objs || !off) return bfail(b, "xpbd_batch_set_attached: missing arrays");
    for (int r = 0; r < requests; ++r) {
        if (!valid_obj(b, objs[r])) return bfail(b, "xpbd_batch_set_attached: object out of range");
        if (off[r + 1] < off[r]) return bfail(b, "xpbd_batch_set_attached: offsets must be non-decreasing");
    }
    if (off[requests] > off[0] && (!idx || !pos)) return bfail(b, "xpbd_batch_set_attached: missing arrays");
    std::string err;
    if (!b->engine.attachMany(requests, objs, off, idx, pos, weight, err))
        return bfail(b, "xpbd_batch_set_attached: " + err);
    return 0;
}

int xpbd_batch_add_rigid_velocity(xpbd_batch* b, int obj, const float* dv, const float* dw, const float* center)
{
    if (!b) return -1;
    if (!dv || !dw || !center) return bfail(b, "xpbd_batch_add_rigid_velocity: null argument");
    for_objects(b, obj, "xpbd_batch_add_rigid_velocity", [&](int k) { b->engine.addRigidVelocity(k, dv, dw, center); });
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, std::string("xpbd_batch_add_rigid_velocity: ") + hipGetErrorString(e));
    return 0;
}

int xpbd_batch_reset(xpbd_batch* b, int obj)
{
    if (!b) return -1;
    for_objects(b, obj, "xpbd_batch_reset", [&](int k) {
        b->engine.attach(k, 0, nullptr, nullptr, 1.0f);
        b->engine.resetToRest(k);
    });
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, std::string("xpbd_batch_reset: ") + hipGetErrorString(e));
    return 0;
}

int xpbd_batch_prepare(xpbd_batch* b)
{
    if (!b) return -1;
    std::string err;
    if (!b->engine.warm(err)) return bfail(b, "xpbd_batch_prepare: " + err);
    return 0;
}

int xpbd_batch_step_async(xpbd_batch* b, int substeps)
{
    if (!b) return -1;
    if (substeps < 0) return bfail(b, "xpbd_batch_step: negative substep count");
    std::string err;
    if (!b->engine.step(substeps, err)) return bfail(b, "xpbd_batch_step: " + err);
    return 0;
}

int xpbd_batch_sync(xpbd_batch* b)
{
    if (!b) return -1;
    std::string err;
    if (!b->engine.sync(err)) return bfail(b, "xpbd_batch_sync: " + err);
    return 0;
}

int xpbd_batch_step(xpbd_batch* b, int substeps)
{
    if (xpbd_batch_step_async(b, substeps) != 0) return -1;
    return xpbd_batch_sync(b);
}

unsigned long long xpbd_batch_positions_device(xpbd_batch* b)
{
    return b ? reinterpret_cast<unsigned long long>(b->engine.positionsDevice()) : 0ull;
}

int xpbd_batch_positions_host(xpbd_batch* b, int obj, float* out, int n)
{
    if (!b) return -1;
    if (!valid_obj(b, obj) || !out || n != b->inputs[obj]->n) return bfail(b, "xpbd_batch_positions_host: size mismatch");
    if (!b->engine.getState(obj, out, nullptr)) return bfail(b, "xpbd_batch_positions_host: device copy failed");
    return 0;
}

int xpbd_batch_get_state(xpbd_batch* b, int obj, float* pos, float* vel, int n)
{
    if (!b) return -1;
    if (!valid_obj(b, obj) || n != b->inputs[obj]->n) return bfail(b, "xpbd_batch_get_state: size mismatch");
    if (!b->engine.getState(obj, pos, vel)) return bfail(b, "xpbd_batch_get_state: device copy failed");
    return 0;
}

int xpbd_batch_set_state(xpbd_batch* b, int obj, const float* pos, const float* vel, int n)
{
    if (!b) return -1;
    if (!valid_obj(b, obj) || n != b->inputs[obj]->n) return bfail(b, "xpbd_batch_set_state: size mismatch");
    if (!b->engine.setState(obj, pos, vel)) return bfail(b, "xpbd_batch_set_state: device copy failed");
    return 0;
}

int xpbd_batch_compute_shapes(xpbd_batch* b, int obj, float deformEps, float* hostScales, float* hostQuats, int n)
{
    if (!b) return -1;
    std::string err;
    if (obj < 0) {
        if (!b->engine.computeShapesAll(deformEps, err)) return bfail(b, "xpbd_batch_compute_shapes: " + err);
        return 0;
    }
    if (!valid_obj(b, obj) || n != b->inputs[obj]->n) return bfail(b, "xpbd_batch_compute_shapes: size mismatch");
    int deformed = 0;
    if (!b->engine.computeShapes(obj, deformEps, hostScales, hostQuats, deformed, err))
        return bfail(b, "xpbd_batch_compute_shapes: " + err);
    return deformed;
}

int xpbd_batch_shapes_device(xpbd_batch* b, unsigned long long* ptrs)
{
    if (!b) return -1;
    if (!ptrs) return bfail(b, "xpbd_batch_shapes_device: null output");
    if (!b->engine.shapeScalesDevice()) return bfail(b, "xpbd_batch_shapes_device: no shapes computed yet");
    ptrs[0] = reinterpret_cast<unsigned long long>(b->engine.shapeScalesDevice());
    ptrs[1] = reinterpret_cast<unsigned long long>(b->engine.shapeQuatsDevice());
    ptrs[2] = reinterpret_cast<unsigned long long>(b->engine.callerPosDevice());
    return b->engine.padNodes();
}

int xpbd_batch_self_collision_stats(xpbd_batch* b, int obj, int* out)
{
    if (!b) return -1;
    if (!valid_obj(b, obj) || !out) return bfail(b, "xpbd_batch_self_collision_stats: invalid arguments");
    std::string err;
    if (!b->engine.selfCollisionStats(obj, out, err)) return bfail(b, "xpbd_batch_self_collision_stats: " + err);
    return 0;
}

int xpbd_batch_collider_stats(xpbd_batch* b, int obj, int* hits, float* push, float* centroid, int maxCount)
{
    if (!b || !valid_obj(b, obj)) return 0;
    return b->engine.colliderStats(obj, hits, push, centroid, maxCount);
}

int xpbd_batch_copy_object(xpbd_batch* dst, int dstObj, xpbd_batch* src, int srcObj)
{
    if (!dst) return -1;
    if (!src || src == dst) return bfail(dst, "xpbd_batch_copy_object: needs two different batches");
    std::string err;
    if (!dst->engine.copyBodyFrom(dstObj, src->engine, srcObj, err)) return bfail(dst, "xpbd_batch_copy_object: " + err);
    return 0;
}

int xpbd_batch_autotune(xpbd_batch* b, int substeps, int flags, char* json, int len)
{
    if (!b) return -1;
    std::string rep, err;
    if (!b->engine.autotune(substeps, flags, rep, err)) return bfail(b, "xpbd_batch_autotune: " + err);
    if (json && len > 0) {
        if ((int)rep.size() + 1 > len) return bfail(b, "xpbd_batch_autotune: buffer too small");
        std::memcpy(json, rep.c_str(), rep.size() + 1);
    }
    return 0;
}

int xpbd_batch_get_tuning(xpbd_batch* b, int* t)
{
    if (!b || !t) return -1;
    const Tuning& u = b->engine.tuning();
    t[0] = u.jacobiTiles; t[1] = u.volMode; t[2] = u.shapeClusterWaves; t[3] = u.scBuildWaves; t[4] = u.scBuildNodes;
    t[5] = u.scBuildXcd; t[6] = u.scGroupLanes;
    return 0;
}

int xpbd_batch_set_tuning(xpbd_batch* b, const int* t)
{
    if (!b) return -1;
    if (!t) return bfail(b, "xpbd_batch_set_tuning: null argument");
    auto in = [](int v, std::initializer_list<int> ok) {
        for (int o : ok) if (v == o) return true;
        return false;
    };
    if (!in(t[0], {1, 2, 4}) || !in(t[1], {-1, 0, 1}) || !in(t[2], {1, 2, 4}) || !in(t[3], {1, 2, 4}) ||
        !in(t[4], {4, 8, 16, 32, 64}) || !(t[5] >= 0 && t[5] <= 256) || !in(t[6], {-1, 0, 1}))
        return bfail(b, "xpbd_batch_set_tuning: value outside the tuned range");
    Tuning& u = b->engine.tuning();
    u.jacobiTiles = t[0]; u.volMode = t[1]; u.shapeClusterWaves = t[2]; u.scBuildWaves = t[3]; u.scBuildNodes = t[4];
    u.scBuildXcd = t[5]; u.scGroupLanes = t[6];
    return 0;
}

int xpbd_batch_profile(xpbd_batch* b, int substeps, char* json, int len)
{
    if (!b) return -1;
    if (!json || len <= 0 || substeps <= 0) return bfail(b, "xpbd_batch_profile: invalid arguments");
    std::string out, err;
    if (!b->engine.profile(substeps, out, err)) return bfail(b, "xpbd_batch_profile: " + err);
    if ((int)out.size() + 1 > len) return bfail(b, "xpbd_batch_profile: buffer too small");
    std::memcpy(json, out.c_str(), out.size() + 1);
    return 0;
}

}  // extern "C"
