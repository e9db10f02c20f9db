/*
 * xpbd_oracle_ext.cpp — TEST INFRASTRUCTURE ONLY (see xpbd_oracle.cpp).
 * Setters and hooks for the constraint families beyond distance / colliders /
 * ground: volume, shape matching, self-collision, object shape, contact shapes
 * and deformed shapes.
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "xpbd_capi.h"
#include "xpbd_oracle_internal.h"

namespace xo {

void extensions_on_init() {}
void extensions_free() {}
void extensions_on_reset() {}
void extensions_begin_step() {}
void extensions_after_predict() {}
void extensions_before_iterations() {}
bool extensions_distance(std::vector<V3>&, std::vector<V3>&, int) { return false; }
void extensions_iteration(std::vector<V3>*&, std::vector<V3>*&, int) {}
void extensions_self_collision(std::vector<V3>&) {}
void extensions_after_iterations(std::vector<V3>&, bool) {}
float contact_rmax(int) { return 0.0f; }
bool contact_colliders_on() { return false; }
float contact_support(int, float, float, float) { return 0.0f; }
float ground_contact_radius(int) { return 0.0f; }

}  // namespace xo

using namespace xo;

extern "C" {

// native/xpbd_capi.cu:130-137 -> forward.cu:3849-3853, 3934-3944, 3953-3965
void xpbd_set_volume(float compliance, int ringK, int maxMembers, int leaderMinHop)
{
    s.volCompliance = compliance;   // setVolumeParams stores it unclamped (3849-3853)
    const int k = std::max(1, std::min(8, ringK));
    const int cap = std::max(8, std::min(8192, maxMembers));
    s.volRingK = k;
    s.volMaxMembers = cap;
    s.volLeaderMinHop = std::max(1, std::min(8, leaderMinHop));
    s.volMemberSelect = 1;
}

// forward.cu:4464-4476
void xpbd_set_object_shape(float stiffness)
{
    float v = std::fmin(std::fmax(stiffness, 0.0f), 1.0f);
    if (!std::isfinite(v)) v = 0.0f;
    s.objShapeStiffness = v;
    s.objMinComponent = 50;
}

void xpbd_set_object_shape_gpu(int enabled) { s.objShapeGPU = enabled != 0; }

void xpbd_set_shape_robust(int enabled) { s.shapeRobust = enabled != 0; }

// forward.cu:4532-4540
void xpbd_set_self_collision(int enabled, float radiusScale, float excludeScale, int withinBody)
{
    s.scEnabled = enabled != 0;
    s.scWithinBody = withinBody != 0;
    s.scRadiusScale = std::fmin(std::fmax(std::isfinite(radiusScale) ? radiusScale : 1.5f, 0.1f), 10.0f);
    s.scExcludeScale = std::fmin(std::fmax(std::isfinite(excludeScale) ? excludeScale : 2.0f, 1.05f), 20.0f);
}

void xpbd_self_collision_stats(int* out)
{
    if (!out) return;
    out[0] = 0; out[1] = 0; out[2] = 0;
}

int xpbd_set_contact_shape(int enabled, float tau, float radiusCap)
{
    if (g.n <= 0) return fail("xpbd_set_contact_shape: call xpbd_create first");
    g.contactOn = enabled != 0;
    s.contactTau = tau;
    s.contactCap = radiusCap;
    s.contactColliders = (enabled & 1) != 0;
    s.contactGround = (enabled & 2) != 0;
    return 0;
}

int xpbd_compute_shapes(float, float*, float*, int n)
{
    if (!g.stepped) return fail("xpbd_compute_shapes: no step yet");
    if (n != g.n) return fail("xpbd_compute_shapes: size mismatch");
    return fail("xpbd_compute_shapes: not restated in the oracle yet");
}

void xpbd_volume_stats(int enable, float* jvw, float* jvwStd)
{
    s.volJStats = enable != 0;
    if (jvw) *jvw = 1.0f;
    if (jvwStd) *jvwStd = 0.0f;
}

}  // extern "C"
