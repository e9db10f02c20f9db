// settings.cpp — the setter clamps of the reference's parameter setters,
// shared by the singleton C ABI and the batched API.
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "engine.h"

namespace xpbd {

// setXPBDParams (forward.cu:4896-4921) as called by native/xpbd_capi.cu:112-120
void settings_solver(Settings& s, int iters, float dt, float underRelax, float velDamping, float stiffnessScale,
                     float invMassScale, float distCompliance, float shapeCompliance, float shapeBlend,
                     float angleCompliance, float angleBlend)
{
    s.iters = iters < 1 ? 1 : iters;
    s.dt = std::fmax(1e-5f, dt);
    s.underRelax = std::fmax(0.0f, underRelax);
    s.velDamping = std::fmax(0.0f, velDamping);
    s.stiffnessScale = std::fmax(0.0f, stiffnessScale);
    s.invMassScale = std::fmax(0.0f, invMassScale);
    s.distCompliance = std::fmax(0.0f, distCompliance);
    s.shapeCompliance = std::fmax(0.0f, shapeCompliance);
    s.shapeBlend = std::fmin(std::fmax(0.0f, shapeBlend), 1.0f);
    s.angleCompliance = std::fmax(0.0f, angleCompliance);
    s.angleBlend = std::fmin(std::fmax(0.0f, angleBlend), 1.0f);
}

// constraint enables (forward.cu:3718-3766; native/xpbd_capi.cu:122-129)
void settings_constraints(Settings& s, int distance, int shape, int angle, int volume)
{
    s.useDistance = distance != 0;
    s.useShape = shape != 0;
    s.useAngle = angle != 0;
    s.useVolume = volume != 0;
}

// setVolumeParams / setVolumeRingK / setVolumeLeaderParams (forward.cu:3849-3965)
void settings_volume(Settings& s, float compliance, int ringK, int maxMembers, int leaderMinHop)
{
    s.volCompliance = compliance;
    s.volRingK = std::max(1, std::min(8, ringK));
    s.volMaxMembers = std::max(8, std::min(8192, maxMembers));
    s.volLeaderMinHop = std::max(1, std::min(8, leaderMinHop));
    s.volMemberSelect = 1;
}

// setGroundParams (native/xpbd_capi.cu:131-150)
void settings_ground(Settings& s, int enabled, float ux, float uy, float uz, float height, float friction,
                     float restitution, float contactRadius, float contactSlop, float gravity)
{
    float nx = ux, ny = uy, nz = uz;
    const float l = std::sqrt(nx * nx + ny * ny + nz * nz);
    if (!(l > 1e-8f) || !std::isfinite(l)) { nx = 0; ny = 0; nz = 1; }
    else { nx /= l; ny /= l; nz /= l; }
    s.groundN[0] = nx; s.groundN[1] = ny; s.groundN[2] = nz;
    s.groundEnabled = enabled != 0;
    s.groundHeight = std::isfinite(height) ? height : 0.0f;
    s.groundFriction = std::fmax(0.0f, friction);
    s.groundRestitution = std::fmin(std::fmax(restitution, 0.0f), 1.0f);
    s.groundRadius = std::fmax(0.0f, contactRadius);
    s.groundSlop = std::fmax(0.0f, contactSlop);
    s.groundGravity = std::fmax(0.0f, gravity);
}

// setObjectModeParams (forward.cu:4464-4476)
void settings_object_shape(Settings& s, float stiffness)
{
    float v = std::fmin(std::fmax(stiffness, 0.0f), 1.0f);
    if (!std::isfinite(v)) v = 0.0f;
    s.objShapeStiffness = v;
    s.objMinComponent = 50;
}

// setSelfCollisionParams (forward.cu:4532-4540)
void settings_self_collision(Settings& s, int enabled, float radiusScale, float excludeScale, int withinBody)
{
    s.scEnabled = enabled != 0;
    s.scWithinBody = withinBody != 0;
    s.scRadiusScale = std::fmin(std::fmax(std::isfinite(radiusScale) ? radiusScale : 1.5f, 0.1f), 10.0f);
    s.scExcludeScale = std::fmin(std::fmax(std::isfinite(excludeScale) ? excludeScale : 2.0f, 1.05f), 20.0f);
}

// setKinematicColliders (forward.cu:4478-4500): a list of the same length
// and types keeps each collider's previous pose for friction.
int collider_list_update(ColliderDev* cols, int prevCount, int count, const int* types, const float* poses,
                         const float* dims)
{
    if (count < 0) count = 0;
    if (count > kMaxColliders) {
        std::printf("[Collider] %d colliders requested, using the first %d\n", count, kMaxColliders);
        count = kMaxColliders;
    }
    const bool samePrev = (count == prevCount);
    for (int c = 0; c < count; ++c) {
        ColliderDev& K = cols[c];
        const bool keep = samePrev && K.type == types[c];
        for (int a = 0; a < 9; ++a) K.Rp[a] = keep ? K.R[a] : poses[12 * c + a];
        for (int a = 0; a < 3; ++a) K.tp[a] = keep ? K.t[a] : poses[12 * c + 9 + a];
        K.type = types[c];
        for (int a = 0; a < 9; ++a) K.R[a] = poses[12 * c + a];
        for (int a = 0; a < 3; ++a) { K.t[a] = poses[12 * c + 9 + a]; K.h[a] = std::fmax(dims[3 * c + a], 0.0f); }
    }
    return count;
}

float positive_or_zero(float v) { return (std::isfinite(v) && v > 0.0f) ? v : 0.0f; }

}  // namespace xpbd
