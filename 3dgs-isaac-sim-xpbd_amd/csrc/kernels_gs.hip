// kernels_gs.hip — graph-coloured Gauss-Seidel passes: distance edges
// (forward.cu:6576-6621) and volume clusters (9041-9183).
//
// One launch per colour, in place on the predicted positions.  Inside a colour
// no two edges (clusters) share a node, so the threads of a launch never touch
// the same position and the result equals the serial colour-by-colour order
// bit for bit.  Colours merge across bodies: colour c of every body runs in
// the same launch.  Multipliers live per edge / per leader in place: the first
// iteration of a substep does not read them (zero), the last does not write.
#include "kernels.h"
#include "mat3.h"
#include "node_ops.h"

namespace xpbd {

__device__ __forceinline__ const BodyDev& gs_body(const int* __restrict__ tileBody, const BodyDev* __restrict__ bodies,
                                                  int i)
{
    return bodies[tileBody[i / kTile]];
}

// alpha tilde per edge: (compliance / max(stiff * scale, 1e-6)) / max(dt^2, 1e-8)
__global__ void dist_gs_alpha_kernel(int m, const int2* __restrict__ ab, const float* __restrict__ stiff,
                                     const int* __restrict__ tileBody, const BodyDev* __restrict__ bodies,
                                     float* __restrict__ alpha)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const BodyDev& B = gs_body(tileBody, bodies, ab[e].x);
    const float dt2 = fmaxf(B.dt * B.dt, 1e-8f);
    const float s = fmaxf(stiff[e] * B.stiffnessScale, 1e-6f);
    alpha[e] = (B.distCompliance / s) / dt2;
}

template <bool kFirst, bool kLast>
__global__ void __launch_bounds__(256) dist_gs_kernel(int begin, int count, const int2* __restrict__ ab,
                                                      const float* __restrict__ rest,
                                                      const float* __restrict__ alpha, float* __restrict__ lam,
                                                      float4* __restrict__ P, const int* __restrict__ tileBody,
                                                      const BodyDev* __restrict__ bodies)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int e = begin + t;
    const int2 ij = ab[e];
    const BodyDev& B = gs_body(tileBody, bodies, ij.x);
    if (!B.distGSOn) return;
    const float4 pi4 = P[ij.x], pj4 = P[ij.y];
    const float wi = pi4.w * B.invMassScale;
    const float wj = pj4.w * B.invMassScale;
    if (wi + wj <= 0.0f) {
        if (kFirst && !kLast) lam[e] = 0.0f;
        return;
    }
    const float3 pi = xyz(pi4), pj = xyz(pj4);
    const float3 diff = sub3(pi, pj);
    const float d = len3(diff);
    const float at = alpha[e];
    const float denom = (wi + wj) + at;
    if (d < 1e-7f || denom < 1e-8f) {
        if (kFirst && !kLast) lam[e] = 0.0f;
        return;
    }
    const float C = d - rest[e];
    const float l = kFirst ? 0.0f : lam[e];
    const float dl = (-C - at * l) / denom;
    if (!kLast) lam[e] = l + dl;
    const float3 n = mul3(diff, 1.0f / d);
    if (wi > 0.0f) {
        const float3 q = add3(pi, mul3(n, wi * dl));
        P[ij.x] = make_float4(q.x, q.y, q.z, pi4.w);
    }
    if (wj > 0.0f) {
        const float3 q = sub3(pj, mul3(n, wj * dl));
        P[ij.y] = make_float4(q.x, q.y, q.z, pj4.w);
    }
}

// One thread per cluster of the colour: centroid, covariance, J, multiplier
// step, then every member moved in place (leader first) with the 0.25 restLen
// step clamp.
template <bool kFirst, bool kLast>
__global__ void __launch_bounds__(64) vol_gs_kernel(int begin, int count, const int* __restrict__ order,
                                                    float4* __restrict__ P, const int* __restrict__ volOff,
                                                    const int* __restrict__ volIdx, const float* __restrict__ detRest,
                                                    const int* __restrict__ matType, float* __restrict__ lamVol,
                                                    const float* __restrict__ restLen,
                                                    const int* __restrict__ tileBody,
                                                    const BodyDev* __restrict__ bodies)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int idx = order[begin + t];
    const BodyDev& B = gs_body(tileBody, bodies, idx);
    if (!B.volGSOn || matType[idx] != 0) return;
    float coef = 0.0f;
    bool acted = false;
    const float4 pi = P[idx];
    const int off = volOff[idx], cnt = volOff[idx + 1] - off;
    float cx = pi.x, cy = pi.y, cz = pi.z;
    for (int k = 0; k < cnt; ++k) {
        const float4 q = P[volIdx[off + k]];
        cx = cx + q.x; cy = cy + q.y; cz = cz + q.z;
    }
    M3 Si;
    if (cnt >= 3) {
        const float invN = 1.0f / (float)(cnt + 1);
        cx = cx * invN; cy = cy * invN; cz = cz * invN;
        M3 S = m3_zero();
        m3_add_outer(S, pi.x - cx, pi.y - cy, pi.z - cz);
        for (int k = 0; k < cnt; ++k) {
            const float4 q = P[volIdx[off + k]];
            m3_add_outer(S, q.x - cx, q.y - cy, q.z - cz);
        }
        m3_scale(S, invN);
        float sC, dhC;
        const float dRest = detRest[idx];
        if (m3_scale_shape_det(S, sC, dhC) && !(dRest < 1e-30f)) {
            const float detCur = sC * sC * sC * dhC;
            float J = sqrtf(detCur / dRest);
            if (isfinite(J) && J > 0.0f) {
                if (J > 3.0f || J < 0.333f) J = fminf(3.0f, fmaxf(0.333f, J));
                const float C = J - 1.0f;
                M3 Sreg = S;
                if (dhC < 0.2f) {
                    const float dd = 0.06f * sC;
                    Sreg.m[0][0] += dd; Sreg.m[1][1] += dd; Sreg.m[2][2] += dd;
                }
                if (!(fabsf(C) < 1e-4f) && m3_inverse_safe(Sreg, Si, 1e-20f)) {
                    const float gs = J * invN;
                    float denom = 0.0f;
                    {
                        const float3 gv = m3_mul_vec(Si, pi.x - cx, pi.y - cy, pi.z - cz);
                        const float gx = gs * gv.x, gy = gs * gv.y, gz = gs * gv.z;
                        denom += (pi.w * B.invMassScale) * (gx * gx + gy * gy + gz * gz);
                    }
                    for (int k = 0; k < cnt; ++k) {
                        const float4 q = P[volIdx[off + k]];
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
                            acted = true;
                        }
                    }
                }
            }
        }
    }
    if (!acted) {
        if (kFirst && !kLast) lamVol[idx] = 0.0f;
        return;
    }
    for (int k = -1; k < cnt; ++k) {
        const int j = (k < 0) ? idx : volIdx[off + k];
        const float4 xj = P[j];
        const float wj = xj.w * B.invMassScale;
        if (wj <= 0.0f) continue;
        const float3 g = m3_mul_vec(Si, xj.x - cx, xj.y - cy, xj.z - cz);
        float dpx = g.x * wj * coef, dpy = g.y * wj * coef, dpz = g.z * wj * coef;
        if (!isfinite(dpx) || !isfinite(dpy) || !isfinite(dpz)) continue;
        const float rl = restLen[j];
        if (rl > 1e-12f) {
            const float maxStep = 0.25f * rl;
            const float len = sqrtf(dpx * dpx + dpy * dpy + dpz * dpz);
            if (len > maxStep && len > 1e-20f) {
                const float sc = maxStep / len;
                dpx = dpx * sc; dpy = dpy * sc; dpz = dpz * sc;
            }
        }
        P[j] = make_float4(xj.x + dpx, xj.y + dpy, xj.z + dpz, xj.w);
    }
}

static inline int nbs(int n, int bs) { return (n + bs - 1) / bs; }

void launch_dist_gs_alpha(hipStream_t s, int m, const int2* ab, const float* stiff, const int* tileBody,
                          const BodyDev* bodies, float* alpha)
{
    if (m > 0) dist_gs_alpha_kernel<<<nbs(m, 256), 256, 0, s>>>(m, ab, stiff, tileBody, bodies, alpha);
}

void launch_dist_gs(hipStream_t s, bool first, bool last, int begin, int count, const int2* ab, const float* rest,
                    const float* alpha, float* lam, float4* P, const int* tileBody, const BodyDev* bodies)
{
    if (count <= 0) return;
    const dim3 g(nbs(count, 256)), b(256);
    if (first && last) dist_gs_kernel<true, true><<<g, b, 0, s>>>(begin, count, ab, rest, alpha, lam, P, tileBody, bodies);
    else if (first) dist_gs_kernel<true, false><<<g, b, 0, s>>>(begin, count, ab, rest, alpha, lam, P, tileBody, bodies);
    else if (last) dist_gs_kernel<false, true><<<g, b, 0, s>>>(begin, count, ab, rest, alpha, lam, P, tileBody, bodies);
    else dist_gs_kernel<false, false><<<g, b, 0, s>>>(begin, count, ab, rest, alpha, lam, P, tileBody, bodies);
}

void launch_vol_gs(hipStream_t s, bool first, bool last, int begin, int count, const int* order, float4* P,
                   const VolumeArgs& va, const int* tileBody, const BodyDev* bodies)
{
    if (count <= 0) return;
    const dim3 g(nbs(count, 64)), b(64);
#define XPBD_VOL_GS(F, L)                                                                                       \
    vol_gs_kernel<F, L><<<g, b, 0, s>>>(begin, count, order, P, va.volOff, va.volIdx, va.detRest, va.matType, \
                                        va.lam, va.restLen, tileBody, bodies)
    if (first && last) XPBD_VOL_GS(true, true);
    else if (first) XPBD_VOL_GS(true, false);
    else if (last) XPBD_VOL_GS(false, true);
    else XPBD_VOL_GS(false, false);
#undef XPBD_VOL_GS
}

}  // namespace xpbd
