// node_ops.h — per-node device operations shared by the kernels that host the
// iteration epilogue: kinematic colliders (forward.cu:7425-7518), ground
// projection (7405-7421), press plates (9411-9424) and the velocity update /
// commit (6623-6661, 7528-7599).
#pragma once
#include "kernels.h"

namespace xpbd {

__device__ __forceinline__ float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 add3(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 mul3(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len3(float3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ float3 xyz(float4 a) { return f3(a.x, a.y, a.z); }

// Body of the wave that owns node i (wave-uniform by construction).
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  With a grid of 8 * per blocks, this maps hardware block b to logical
// block (b mod 8) * per + b / 8, so every XCD works on one contiguous range of
// tiles and the neighbour gathers of adjacent tiles hit the same L2.
__device__ __forceinline__ int xcd_block()
{
    const int per = gridDim.x >> 3;
    return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

__device__ __forceinline__ int wave_body(const int* __restrict__ tileBody, int i)
{
    return __builtin_amdgcn_readfirstlane(tileBody[i / kTile]);
}

// Support distance sqrt(n^T M n) of a contact ellipsoid (forward.cu:7387-7392).
__device__ __forceinline__ float contact_support(const float* m, float nx, float ny, float nz)
{
    const float q = m[0] * nx * nx + m[3] * ny * ny + m[5] * nz * nz
        + 2.0f * (m[1] * nx * ny + m[2] * nx * nz + m[4] * ny * nz);
    return sqrtf(fmaxf(q, 0.0f));
}

// ---------------------------------------------------------------------------
// Kinematic colliders (forward.cu:7425-7518).  Pushes p out of every collider
// of the body; statistics go to per-collider accumulators.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void collide_node(int i, float3& p, float3 x0, const BodyDev& B, const ColliderDev* __restrict__ cols,
                             int* __restrict__ hits, float* __restrict__ push, bool countHits,
                             const float* __restrict__ cM, const float* __restrict__ cRmax)
{
    const float margin = B.colMargin, mu = B.colFriction;
    const float rmax = cRmax ? cRmax[i] : 0.0f;
    const float reach = margin + rmax;
    for (int c = 0; c < B.colCount; ++c) {
        const int gc = B.colBegin + c;
        const ColliderDev& C = cols[gc];
        const float dx = p.x - C.t[0], dy = p.y - C.t[1], dz = p.z - C.t[2];
        const float q0 = C.R[0] * dx + C.R[3] * dy + C.R[6] * dz;
        const float q1 = C.R[1] * dx + C.R[4] * dy + C.R[7] * dz;
        const float q2 = C.R[2] * dx + C.R[5] * dy + C.R[8] * dz;
        float n0 = 0.0f, n1 = 0.0f, n2 = 1.0f, dist;
        if (C.type == 1) {
            const float a0 = fabsf(q0) - C.h[0], a1 = fabsf(q1) - C.h[1], a2 = fabsf(q2) - C.h[2];
            if (a0 > 0.0f || a1 > 0.0f || a2 > 0.0f) {
                const float o0 = fmaxf(a0, 0.0f), o1 = fmaxf(a1, 0.0f), o2 = fmaxf(a2, 0.0f);
                dist = sqrtf(o0 * o0 + o1 * o1 + o2 * o2);
                if (dist >= reach) continue;
                const float inv = 1.0f / fmaxf(dist, 1e-20f);
                n0 = copysignf(o0 * inv, q0); n1 = copysignf(o1 * inv, q1); n2 = copysignf(o2 * inv, q2);
            } else {
                if (a0 >= a1 && a0 >= a2) { dist = a0; n0 = (q0 >= 0.0f) ? 1.0f : -1.0f; n2 = 0.0f; }
                else if (a1 >= a2)        { dist = a1; n1 = (q1 >= 0.0f) ? 1.0f : -1.0f; n2 = 0.0f; }
                else                      { dist = a2; n2 = (q2 >= 0.0f) ? 1.0f : -1.0f; }
            }
        } else {
            float e2 = q2;
            if (C.type == 2) e2 = q2 - fminf(fmaxf(q2, -C.h[1]), C.h[1]);
            const float r = sqrtf(q0 * q0 + q1 * q1 + e2 * e2);
            dist = r - C.h[0];
            if (dist >= reach) continue;
            if (r > 1e-20f) { n0 = q0 / r; n1 = q1 / r; n2 = e2 / r; }
        }
        const float wx = C.R[0] * n0 + C.R[1] * n1 + C.R[2] * n2;
        const float wy = C.R[3] * n0 + C.R[4] * n1 + C.R[5] * n2;
        const float wz = C.R[6] * n0 + C.R[7] * n1 + C.R[8] * n2;
        if (cM && rmax > 0.0f) {
            dist -= contact_support(&cM[6 * i], wx, wy, wz);
            if (dist >= margin) continue;
        }
        const float corr = margin - dist;
        const float3 pIn = p;
        p.x += wx * corr; p.y += wy * corr; p.z += wz * corr;
        float3 dp = f3(wx * corr, wy * corr, wz * corr);
        if (mu > 0.0f) {
            const float ox = C.Rp[0] * q0 + C.Rp[1] * q1 + C.Rp[2] * q2 + C.tp[0];
            const float oy = C.Rp[3] * q0 + C.Rp[4] * q1 + C.Rp[5] * q2 + C.tp[1];
            const float oz = C.Rp[6] * q0 + C.Rp[7] * q1 + C.Rp[8] * q2 + C.tp[2];
            const float rx = (p.x - x0.x) - (pIn.x - ox), ry = (p.y - x0.y) - (pIn.y - oy), rz = (p.z - x0.z) - (pIn.z - oz);
            const float rn = rx * wx + ry * wy + rz * wz;
            const float tx = rx - rn * wx, ty = ry - rn * wy, tz = rz - rn * wz;
            const float lt = sqrtf(tx * tx + ty * ty + tz * tz);
            if (lt > 1e-20f) {
                const float k = fminf(1.0f, mu * corr / lt);
                p.x -= k * tx; p.y -= k * ty; p.z -= k * tz;
                dp.x -= k * tx; dp.y -= k * ty; dp.z -= k * tz;
            }
        }
        float* acc = &push[kColAcc * gc];
        const float w = sqrtf(dp.x * dp.x + dp.y * dp.y + dp.z * dp.z);
        atomicAdd(&acc[0], dp.x);
        atomicAdd(&acc[1], dp.y);
        atomicAdd(&acc[2], dp.z);
        atomicAdd(&acc[3], w * p.x);
        atomicAdd(&acc[4], w * p.y);
        atomicAdd(&acc[5], w * p.z);
        atomicAdd(&acc[6], w);
        if (countHits) atomicAdd(&hits[gc], 1);
    }
}

// Velocity update and commit (forward.cu:6623-6661 and 7528-7599).
__device__ __forceinline__ void commit_xv(int i, float3 p, float w, const BodyDev& B, float3* __restrict__ X,
                                          float3* __restrict__ V, const float* __restrict__ rg)
{
    const float3 x = X[i];
    if (w <= 0.0f) {
        V[i] = f3(0.0f, 0.0f, 0.0f);
        X[i] = p;
        return;
    }
    const float dt = B.dt;
    const float3 dx = sub3(p, x);
    const float dx2 = dot3(dx, dx);
    const float damping = fminf(fmaxf(B.velDamping, 0.0f), 1.0f);
    if (!B.groundOn) {
        if (!isfinite(dx2) || dx2 < 1e-18f) {
            V[i] = f3(0.0f, 0.0f, 0.0f);
            X[i] = p;
            return;
        }
        float3 vs = mul3(dx, 1.0f / fmaxf(dt, 1e-8f));
        vs = mul3(vs, 1.0f - damping);
        if (!isfinite(vs.x) || !isfinite(vs.y) || !isfinite(vs.z) || dot3(vs, vs) < 1e-16f) vs = f3(0.0f, 0.0f, 0.0f);
        V[i] = vs;
        X[i] = p;
        return;
    }
    const float3 vPre = V[i];
    if (!isfinite(dx2)) {
        V[i] = f3(0.0f, 0.0f, 0.0f);
        X[i] = p;
        return;
    }
    float3 v = (dx2 < 1e-18f) ? f3(0.0f, 0.0f, 0.0f) : mul3(dx, 1.0f / fmaxf(dt, 1e-8f));
    const float nx = B.gn[0], ny = B.gn[1], nz = B.gn[2];
    const float gap = nx * p.x + ny * p.y + nz * p.z - B.planeD - (rg ? rg[i] : 0.0f);
    if (gap <= B.slop) {
        const float vn = nx * v.x + ny * v.y + nz * v.z;
        const float vnPre = nx * vPre.x + ny * vPre.y + nz * vPre.z;
        float vnNew = vn;
        float jn = 0.0f;
        if (vnPre < 0.0f) {
            const float e = (-vnPre <= B.restingSpeed) ? 0.0f : B.restitution;
            vnNew = fmaxf(vn, -e * vnPre);
            jn = -(1.0f + e) * vnPre;
        }
        float3 vt = f3(v.x - vn * nx, v.y - vn * ny, v.z - vn * nz);
        const float vtLen = len3(vt);
        const float dvt = B.friction * jn;
        if (vtLen <= dvt || vtLen < 1e-12f) vt = f3(0.0f, 0.0f, 0.0f);
        else vt = mul3(vt, (vtLen - dvt) / vtLen);
        v = f3(vt.x + vnNew * nx, vt.y + vnNew * ny, vt.z + vnNew * nz);
    }
    v = mul3(v, 1.0f - damping);
    if (!isfinite(v.x) || !isfinite(v.y) || !isfinite(v.z) || dot3(v, v) < 1e-16f) v = f3(0.0f, 0.0f, 0.0f);
    V[i] = v;
    X[i] = p;
}

// Commit plus the caller-order copy of the position when the device order
// differs from the caller's.
__device__ __forceinline__ void commit_node(int i, float3 p, float w, const BodyDev& B, const EpiArgs& a,
                                            const float* __restrict__ rg)
{
    commit_xv(i, p, w, B, a.x, a.v, rg);
    if (a.xCaller) a.xCaller[a.callerPos[i]] = p;
}

// Per-node tail of an iteration.  Returns true when the node was committed
// (velocity bit), in which case the predicted buffer need not be written.
__device__ __forceinline__ bool epilogue_node(unsigned bits, int i, float3& p, float w, const BodyDev& B, const EpiArgs& a)
{
    if ((bits & kEpiPress) && B.pressOn) {
        if (w > 0.0f) {
            const float v = (B.pressAxis == 0) ? p.x : (B.pressAxis == 1) ? p.y : p.z;
            if ((v < B.pressLo) || (v > B.pressHi)) {
                const float t = (v < B.pressLo) ? B.pressLo : B.pressHi;
                if (B.pressAxis == 0) p.x = t; else if (B.pressAxis == 1) p.y = t; else p.z = t;
            }
        }
    }
    if ((bits & kEpiColliders) && w > 0.0f && B.colCount > 0)
        collide_node(i, p, a.x[i], B, a.cols, a.colHits, a.colPush, (bits & kEpiColliderHits) != 0,
                     B.contactColliders ? a.cM : nullptr, B.contactColliders ? a.cRmax : nullptr);
    const float* rg = B.contactGround ? a.rg : nullptr;
    if ((bits & kEpiGround) && B.groundOn && w > 0.0f) {
        const float c = B.gn[0] * p.x + B.gn[1] * p.y + B.gn[2] * p.z - B.planeD - (rg ? rg[i] : 0.0f);
        if (c < 0.0f) p = f3(p.x - c * B.gn[0], p.y - c * B.gn[1], p.z - c * B.gn[2]);
    }
    if (bits & kEpiVelocity) {
        commit_node(i, p, w, B, a, rg);
        return true;
    }
    return false;
}

}  // namespace xpbd
