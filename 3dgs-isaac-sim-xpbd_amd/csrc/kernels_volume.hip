// kernels_volume.hip — covariance-J volume constraint over node clusters.
//
// A cluster is a leader node plus its member list.  Per iteration:
//   solve  (per active leader): centroid c, covariance S of {leader} + members,
//          J = sqrt(det S / det S_rest) clamped to [0.333, 3], dead band 1e-4,
//          S^-1 (isotropic lift when det(S/s) < 0.2), XPBD multiplier step,
//          and a blackboard entry (c, coef = dLambda * J / n, S^-1);
//   apply  (per node): mean over the clusters containing the node of
//          w * coef * S^-1 (x - c), clamped to 0.25 * restLen, under-relaxed.
// (forward.cu:8531-8664 and 9187-9243.)  Two solve kernels: one thread per
// cluster for small clusters (operand order of the reference, bit-exact), one
// wave64 per cluster for large ones (members spread over lanes, cached in
// registers, butterfly all-reduce; same maths, different summation order).
#include "kernels.h"
#include "mat3.h"
#include "node_ops.h"

namespace xpbd {

__device__ __forceinline__ const BodyDev& body_of(const int* __restrict__ tileBody, const BodyDev* __restrict__ bodies,
                                                  int i)
{
    return bodies[tileBody[i / kTile]];
}

// ---------------------------------------------------------------------------
// Rest precompute (forward.cu:7691-7844, mixture rest off): restLen for every
// node; rest determinant, rest volume and material type for leaders.
// ---------------------------------------------------------------------------
__global__ void vol_precompute_kernel(int nPad, const int* __restrict__ tileBody, const BodyDev* __restrict__ bodies,
                                      const float3* __restrict__ rest, const int* __restrict__ rowOff,
                                      const int* __restrict__ nbr, const int* __restrict__ volOff,
                                      const int* __restrict__ volIdx, float aniso,
                                      float* __restrict__ detRest, float* __restrict__ Vrest,
                                      float* __restrict__ restLen, int* __restrict__ matType)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nPad) return;
    const BodyDev& B = body_of(tileBody, bodies, idx);
    detRest[idx] = 0.0f;
    Vrest[idx] = 0.0f;
    restLen[idx] = 0.0f;
    matType[idx] = 3;
    if (idx >= B.nodeEnd) return;
    const float3 p0 = rest[idx];
    float dmin = 3.402823466e+38f;
    for (int e = rowOff[idx]; e < rowOff[idx + 1]; ++e) {
        const int j = nbr[e];
        if (j == idx) continue;
        const float3 pj = rest[j];
        const float dx = pj.x - p0.x, dy = pj.y - p0.y, dz = pj.z - p0.z;
        const float dl = sqrtf(dx * dx + dy * dy + dz * dz);
        if (dl > 1e-12f) dmin = fminf(dmin, dl);
    }
    restLen[idx] = (dmin < 3.402823466e+38f) ? dmin : 0.0f;
    const int off = volOff[idx], cnt = volOff[idx + 1] - off;
    if (cnt < B.volMinMembers) return;
    float cx = p0.x, cy = p0.y, cz = p0.z;
    for (int k = 0; k < cnt; ++k) {
        const float3 q = rest[volIdx[off + k]];
        cx = cx + q.x; cy = cy + q.y; cz = cz + q.z;
    }
    const float invN = 1.0f / (float)(cnt + 1);
    cx = cx * invN; cy = cy * invN; cz = cz * invN;
    M3 S = m3_zero();
    m3_add_outer(S, p0.x - cx, p0.y - cy, p0.z - cz);
    for (int k = 0; k < cnt; ++k) {
        const float3 q = rest[volIdx[off + k]];
        m3_add_outer(S, q.x - cx, q.y - cy, q.z - cz);
    }
    m3_scale(S, invN);
    float s, dh;
    if (!m3_scale_shape_det(S, s, dh)) { matType[idx] = 4; return; }
    float ev[3];
    m3_eig_sym_values(S, ev);
    const float lmax = fmaxf(ev[2], 1e-24f);
    const float a1 = ev[0] / lmax, a2 = ev[1] / lmax;
    matType[idx] = (a1 > aniso) ? 0 : (a2 > aniso) ? 1 : 2;
    const float d = s * s * s * dh;
    detRest[idx] = d;
    Vrest[idx] = (4.0f / 3.0f) * 3.14159265f * sqrtf(fmaxf(d, 0.0f));
}

// Multiplier step shared by both solve kernels; returns false when the
// cluster does not act this iteration (coef stays 0).
__device__ __forceinline__ bool vol_step(const M3& S, float invN, float detR, float& J, float& C, M3& Sinv)
{
    float sC, dhC;
    if (!m3_scale_shape_det(S, sC, dhC)) return false;
    const float detCur = sC * sC * sC * dhC;
    if (detR < 1e-30f) return false;
    J = sqrtf(detCur / detR);
    if (!isfinite(J) || J <= 0.0f) return false;
    if (J > 3.0f || J < 0.333f) J = fminf(3.0f, fmaxf(0.333f, J));
    C = J - 1.0f;
    if (fabsf(C) < 1e-4f) return false;
    M3 Sreg = S;
    if (dhC < 0.2f) {
        const float d = 0.06f * sC;
        Sreg.m[0][0] += d; Sreg.m[1][1] += d; Sreg.m[2][2] += d;
    }
    return m3_inverse_safe(Sreg, Sinv, 1e-20f);
}

__device__ __forceinline__ void vol_write(int idx, float cx, float cy, float cz, float coef, const M3& Si,
                                          float4* __restrict__ bbC, float4* __restrict__ bbS, float2* __restrict__ bbS2)
{
    bbC[idx] = make_float4(cx, cy, cz, coef);
    bbS[idx] = make_float4(Si.m[0][0], Si.m[1][1], Si.m[2][2], Si.m[1][0]);
    bbS2[idx] = make_float2(Si.m[2][0], Si.m[2][1]);
}

// One thread per active leader (operand order of forward.cu:8531-8664).
template <bool kFirst, bool kLast>
__global__ void __launch_bounds__(256) vol_solve_thread_kernel(
    int nActive, const int* __restrict__ active, const int* __restrict__ tileBody, const BodyDev* __restrict__ bodies,
    const float4* __restrict__ P, const int* __restrict__ volOff, const int* __restrict__ volIdx,
    const float* __restrict__ detRest, float* __restrict__ lamVol, float4* __restrict__ bbC,
    float4* __restrict__ bbS, float2* __restrict__ bbS2)
{
    const int t = xcd_block() * blockDim.x + threadIdx.x;   // leaders in node order: neighbours share an XCD
    if (t >= nActive) return;
    const int idx = active[t];
    const BodyDev& B = body_of(tileBody, bodies, idx);
    if (!B.volumeOn) return;
    const float4 pi = P[idx];
    const int off = volOff[idx], cnt = volOff[idx + 1] - off;
    // the first kMemCache members' positions, fetched together and kept for
    // the three passes; every sum still runs in member order
    constexpr int kMemCache = 4;
    float4 qc[kMemCache];
    {
        int jc[kMemCache];
#pragma unroll
        for (int u = 0; u < kMemCache; ++u) jc[u] = u < cnt ? volIdx[off + u] : idx;
#pragma unroll
        for (int u = 0; u < kMemCache; ++u) qc[u] = P[jc[u]];
    }
    auto member = [&](int k) { return k < kMemCache ? qc[k] : P[volIdx[off + k]]; };
    float cx = pi.x, cy = pi.y, cz = pi.z;
    for (int k = 0; k < cnt; ++k) {
        const float4 q = member(k);
        cx = cx + q.x; cy = cy + q.y; cz = cz + q.z;
    }
    const float invN = 1.0f / (float)(cnt + 1);
    cx = cx * invN; cy = cy * invN; cz = cz * invN;
    M3 S = m3_zero();
    m3_add_outer(S, pi.x - cx, pi.y - cy, pi.z - cz);
    for (int k = 0; k < cnt; ++k) {
        const float4 q = member(k);
        m3_add_outer(S, q.x - cx, q.y - cy, q.z - cz);
    }
    m3_scale(S, invN);
    float J, C;
    M3 Si;
    float coef = 0.0f;
    if (vol_step(S, invN, detRest[idx], J, C, Si)) {
        const float gs = J * invN;
        float denom = 0.0f;
        {
            const float3 gv = m3_mul_vec(Si, pi.x - cx, pi.y - cy, pi.z - cz);
            const float gx = gs * gv.x, gy = gs * gv.y, gz = gs * gv.z;
            denom += (pi.w * B.invMassScale) * (gx * gx + gy * gy + gz * gz);
        }
        for (int k = 0; k < cnt; ++k) {
            const float4 q = member(k);
            const float3 gv = m3_mul_vec(Si, q.x - cx, q.y - cy, q.z - cz);
            const float gx = gs * gv.x, gy = gs * gv.y, gz = gs * gv.z;
            denom += (q.w * B.invMassScale) * (gx * gx + gy * gy + gz * gz);
        }
        const float dt2 = fmaxf(B.dt * B.dt, 1e-8f);
        const float alphaT = B.volCompliance / dt2;
        const float den = denom + alphaT;
        if (den > 1e-12f) {
            const float lam = kFirst ? 0.0f : lamVol[idx];
            const float dLam = (-C - alphaT * lam) / den;
            if (isfinite(dLam)) {
                if (!kLast) lamVol[idx] = lam + dLam;
                coef = dLam * gs;
            }
        }
    }
    if (coef != 0.0f) vol_write(idx, cx, cy, cz, coef, Si, bbC, bbS, bbS2);
    else bbC[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (kFirst && !kLast && coef == 0.0f) lamVol[idx] = 0.0f;
}

__device__ __forceinline__ float lane_val(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// One wave64 per active leader.  Lanes gather members l, l+64, ... in
// parallel (the first 64 * kCache stay in registers between the passes); every
// sum is then accumulated in member order with values read from the owning
// lane, so the result equals the thread-per-cluster kernel bit for bit.
template <bool kFirst, bool kLast, int kCache>
__global__ void __launch_bounds__(256) vol_solve_wave_kernel(
    int nActive, const int* __restrict__ active, const int* __restrict__ tileBody, const BodyDev* __restrict__ bodies,
    const float4* __restrict__ P, const int* __restrict__ volOff, const int* __restrict__ volIdx,
    const float* __restrict__ detRest, float* __restrict__ lamVol, float4* __restrict__ bbC,
    float4* __restrict__ bbS, float2* __restrict__ bbS2)
{
    const int wave = __builtin_amdgcn_readfirstlane((int)((xcd_block() * blockDim.x + threadIdx.x) >> 6));
    const int lane = threadIdx.x & 63;
    if (wave >= nActive) return;
    const int idx = __builtin_amdgcn_readfirstlane(active[wave]);
    const BodyDev& B = bodies[__builtin_amdgcn_readfirstlane(tileBody[idx / kTile])];
    if (!B.volumeOn) return;
    const float4 pi = P[idx];
    const int off = volOff[idx], cnt = volOff[idx + 1] - off;
    float4 cache[kCache];
    auto member = [&](int chunk) -> float4 {
        if (chunk < kCache) return cache[chunk];
        const int k = chunk * 64 + lane;
        return (k < cnt) ? P[volIdx[off + k]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    };
    const int chunks = (cnt + 63) / 64;
    // pass 1: centroid, summed in member order
    float cx = pi.x, cy = pi.y, cz = pi.z;
    for (int ch = 0; ch < chunks; ++ch) {
        float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const int k = ch * 64 + lane;
        if (k < cnt) q = P[volIdx[off + k]];
#pragma unroll
        for (int r = 0; r < kCache; ++r)
            if (r == ch) cache[r] = q;
        const int m = min(64, cnt - ch * 64);
        for (int t = 0; t < m; ++t) {
            cx = cx + lane_val(q.x, t);
            cy = cy + lane_val(q.y, t);
            cz = cz + lane_val(q.z, t);
        }
    }
    const float invN = 1.0f / (float)(cnt + 1);
    cx = cx * invN; cy = cy * invN; cz = cz * invN;
    // pass 2: covariance (six distinct entries of the symmetric sum)
    float s00, s11, s22, s10, s20, s21;
    {
        const float dx = pi.x - cx, dy = pi.y - cy, dz = pi.z - cz;
        s00 = 0.0f + dx * dx; s11 = 0.0f + dy * dy; s22 = 0.0f + dz * dz;
        s10 = 0.0f + dy * dx; s20 = 0.0f + dz * dx; s21 = 0.0f + dz * dy;
    }
    for (int ch = 0; ch < chunks; ++ch) {
        const float4 q = member(ch);
        const float dx = q.x - cx, dy = q.y - cy, dz = q.z - cz;
        const float p00 = dx * dx, p11 = dy * dy, p22 = dz * dz, p10 = dy * dx, p20 = dz * dx, p21 = dz * dy;
        const int m = min(64, cnt - ch * 64);
        for (int t = 0; t < m; ++t) {
            s00 = s00 + lane_val(p00, t);
            s11 = s11 + lane_val(p11, t);
            s22 = s22 + lane_val(p22, t);
            s10 = s10 + lane_val(p10, t);
            s20 = s20 + lane_val(p20, t);
            s21 = s21 + lane_val(p21, t);
        }
    }
    M3 S = m3_load_sym(s00, s11, s22, s10, s20, s21);
    m3_scale(S, invN);
    float J, C;
    M3 Si;
    float coef = 0.0f;
    if (vol_step(S, invN, detRest[idx], J, C, Si)) {
        const float gs = J * invN;
        float denom = 0.0f;
        {
            const float3 gv = m3_mul_vec(Si, pi.x - cx, pi.y - cy, pi.z - cz);
            const float gx = gs * gv.x, gy = gs * gv.y, gz = gs * gv.z;
            denom += (pi.w * B.invMassScale) * (gx * gx + gy * gy + gz * gz);
        }
        for (int ch = 0; ch < chunks; ++ch) {
            const float4 q = member(ch);
            const float3 gv = m3_mul_vec(Si, q.x - cx, q.y - cy, q.z - cz);
            const float gx = gs * gv.x, gy = gs * gv.y, gz = gs * gv.z;
            const float term = (q.w * B.invMassScale) * (gx * gx + gy * gy + gz * gz);
            const int m = min(64, cnt - ch * 64);
            for (int t = 0; t < m; ++t) denom += lane_val(term, t);
        }
        const float dt2 = fmaxf(B.dt * B.dt, 1e-8f);
        const float alphaT = B.volCompliance / dt2;
        const float den = denom + alphaT;
        if (den > 1e-12f) {
            const float lam = kFirst ? 0.0f : lamVol[idx];
            const float dLam = (-C - alphaT * lam) / den;
            if (isfinite(dLam)) {
                if (!kLast && lane == 0) lamVol[idx] = lam + dLam;
                coef = dLam * gs;
            }
        }
    }
    if (lane == 0) {
        if (coef != 0.0f) vol_write(idx, cx, cy, cz, coef, Si, bbC, bbS, bbS2);
        else bbC[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (kFirst && !kLast && coef == 0.0f) lamVol[idx] = 0.0f;
    }
}

// Gather apply (forward.cu:9187-9243) plus the iteration epilogue.
__global__ void __launch_bounds__(256) vol_apply_kernel(
    int nPad, const int* __restrict__ tileBody, const BodyDev* __restrict__ bodies, const float4* __restrict__ Pin,
    float4* __restrict__ Pout, const int* __restrict__ revOff, const int* __restrict__ revIdx,
    const float4* __restrict__ bbC, const float4* __restrict__ bbS, const float2* __restrict__ bbS2,
    const float* __restrict__ restLen, unsigned epi, EpiArgs ea)
{
    const int idx = xcd_block() * blockDim.x + threadIdx.x;
    if (idx >= nPad) return;
    const BodyDev& B = bodies[__builtin_amdgcn_readfirstlane(tileBody[idx / kTile])];
    if (idx >= B.nodeEnd) return;
    const float4 p4 = Pin[idx];
    float3 out = make_float3(p4.x, p4.y, p4.z);
    const float wi = p4.w * B.invMassScale;
    if (B.volumeOn && wi > 0.0f) {
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        int received = 0;
        // the node's clusters four at a time: indices, then every cluster's
        // blackboard fetched together; summed in entry order as one at a time
        constexpr int U = 4;
        const int k0 = revOff[idx], k1 = revOff[idx + 1];
        for (int k = k0; k < k1; k += U) {
            int ci[U];
#pragma unroll
            for (int u = 0; u < U; ++u) ci[u] = revIdx[k + u < k1 ? k + u : k];
            float4 cc[U], sa[U];
            float2 sb[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                cc[u] = bbC[ci[u]];
                sa[u] = bbS[ci[u]];
                sb[u] = bbS2[ci[u]];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (k + u >= k1) break;
                const float4 c = cc[u];
                if (c.w == 0.0f) continue;
                const M3 Si = m3_load_sym(sa[u].x, sa[u].y, sa[u].z, sa[u].w, sb[u].x, sb[u].y);
                const float3 g = m3_mul_vec(Si, p4.x - c.x, p4.y - c.y, p4.z - c.z);
                const float dx = g.x * wi * c.w, dy = g.y * wi * c.w, dz = g.z * wi * c.w;
                if (!isfinite(dx) || !isfinite(dy) || !isfinite(dz)) continue;
                sx = sx + dx; sy = sy + dy; sz = sz + dz;
                ++received;
            }
        }
        if (received > 0) {
            const float inv = 1.0f / (float)received;
            float ax = sx * inv, ay = sy * inv, az = sz * inv;
            if (isfinite(ax) && isfinite(ay) && isfinite(az)) {
                const float rl = restLen[idx];
                if (rl > 1e-12f) {
                    const float maxStep = 0.25f * rl;
                    const float l = sqrtf(ax * ax + ay * ay + az * az);
                    if (l > maxStep && l > 1e-20f) {
                        const float sc = maxStep / l;
                        ax = ax * sc; ay = ay * sc; az = az * sc;
                    }
                }
                out = make_float3(p4.x + ax * B.underRelax, p4.y + ay * B.underRelax, p4.z + az * B.underRelax);
            }
        }
    }
    if (epi && epilogue_node(epi, idx, out, p4.w, B, ea)) return;
    Pout[idx] = make_float4(out.x, out.y, out.z, p4.w);
}

// J per cluster for statistics (forward.cu:9428-9481), thread per leader.
__global__ void vol_J_kernel(int nPad, const int* __restrict__ tileBody, const BodyDev* __restrict__ bodies,
                             const float3* __restrict__ X, const int* __restrict__ volOff,
                             const int* __restrict__ volIdx, const float* __restrict__ detRest,
                             const int* __restrict__ matType, float* __restrict__ outJ)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nPad) return;
    outJ[idx] = 1.0f;
    if (matType[idx] != 0) return;
    const float3 p0 = X[idx];
    const int off = volOff[idx], cnt = volOff[idx + 1] - off;
    float cx = p0.x, cy = p0.y, cz = p0.z;
    for (int k = 0; k < cnt; ++k) {
        const float3 q = X[volIdx[off + k]];
        cx = cx + q.x; cy = cy + q.y; cz = cz + q.z;
    }
    if (cnt < 3) return;
    const float invN = 1.0f / (float)(cnt + 1);
    cx = cx * invN; cy = cy * invN; cz = cz * invN;
    M3 S = m3_zero();
    m3_add_outer(S, p0.x - cx, p0.y - cy, p0.z - cz);
    for (int k = 0; k < cnt; ++k) {
        const float3 q = X[volIdx[off + k]];
        m3_add_outer(S, q.x - cx, q.y - cy, q.z - cz);
    }
    m3_scale(S, invN);
    float s, dh;
    if (!m3_scale_shape_det(S, s, dh)) return;
    const float dR = detRest[idx];
    if (dR < 1e-30f) return;
    const float J = sqrtf(s * s * s * dh / dR);
    if (isfinite(J) && J > 0.0f) outJ[idx] = J;
}

// ---------------------------------------------------------------------------
static inline int nb(int n, int bs) { return (n + bs - 1) / bs; }

void launch_vol_precompute(hipStream_t s, const NodeArgs& na, const float3* rest, const int* rowOff, const int* nbr,
                           const VolumeArgs& va, float aniso)
{
    vol_precompute_kernel<<<nb(na.nPad, 256), 256, 0, s>>>(na.nPad, na.tileBody, na.bodies, rest, rowOff, nbr,
                                                          va.volOff, va.volIdx, aniso, va.detRest,
                                                          va.Vrest, va.restLen, va.matType);
}

void launch_vol_solve(hipStream_t s, const NodeArgs& na, bool wave, bool first, bool last, const float4* P,
                      const VolumeArgs& va)
{
    if (va.nActive <= 0) return;
#define XPBD_VOL_T(F, L)                                                                                         \
    vol_solve_thread_kernel<F, L><<<xcd_grid(nb(va.nActive, 256)), 256, 0, s>>>(va.nActive, va.active, na.tileBody, na.bodies, \
                                                                      P, va.volOff, va.volIdx, va.detRest,        \
                                                                      va.lam, va.bbC, va.bbS, va.bbS2)
#define XPBD_VOL_W(F, L)                                                                                        \
    vol_solve_wave_kernel<F, L, 4><<<xcd_grid(nb(va.nActive, 4)), 256, 0, s>>>(va.nActive, va.active, na.tileBody, na.bodies, \
                                                                     P, va.volOff, va.volIdx, va.detRest,        \
                                                                     va.lam, va.bbC, va.bbS, va.bbS2)
    if (!wave) {
        if (first && last) XPBD_VOL_T(true, true);
        else if (first) XPBD_VOL_T(true, false);
        else if (last) XPBD_VOL_T(false, true);
        else XPBD_VOL_T(false, false);
    } else {
        if (first && last) XPBD_VOL_W(true, true);
        else if (first) XPBD_VOL_W(true, false);
        else if (last) XPBD_VOL_W(false, true);
        else XPBD_VOL_W(false, false);
    }
#undef XPBD_VOL_T
#undef XPBD_VOL_W
}

void launch_vol_apply(hipStream_t s, const NodeArgs& na, const float4* Pin, float4* Pout, const VolumeArgs& va,
                      unsigned epi, const EpiArgs& ea)
{
    vol_apply_kernel<<<xcd_grid(nb(na.nPad, 256)), 256, 0, s>>>(na.nPad, na.tileBody, na.bodies, Pin, Pout, va.revOff,
                                                     va.revIdx, va.bbC, va.bbS, va.bbS2, va.restLen, epi, ea);
}

void launch_vol_J(hipStream_t s, const NodeArgs& na, const float3* X, const VolumeArgs& va, float* outJ)
{
    vol_J_kernel<<<nb(na.nPad, 256), 256, 0, s>>>(na.nPad, na.tileBody, na.bodies, X, va.volOff, va.volIdx,
                                                 va.detRest, va.matType, outJ);
}

}  // namespace xpbd
