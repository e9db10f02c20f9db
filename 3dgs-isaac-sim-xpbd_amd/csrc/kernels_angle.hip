// kernels_angle.hip — umbrella angle constraint (forward.cu:4990-5130).
//
// The reference accumulates each pair's three corrections with float
// atomics; here the pair kernel stores them per directed entry and the apply
// kernel gathers each node's corrections in one fixed order (centre in caller
// order, entry, role), the order of a serial loop, so the result does not
// depend on scheduling.  One thread per node in both kernels; the apply
// kernel hosts the iteration epilogue when the angle pass is the last one.
#include "kernels.h"
#include "node_ops.h"

namespace xpbd {

__device__ __forceinline__ float3 a_sub(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 a_add(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 a_mul(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float a_dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float a_len(float3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ bool a_finite(float3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

// xpbdAngleAccumulateKernel (4990-5096): corrections of every pair centred on i.
__global__ void __launch_bounds__(256) angle_pair_kernel(int nPad, const int* __restrict__ tileBody,
                                                        const BodyDev* __restrict__ bodies,
                                                        const float4* __restrict__ P, AngleArgs aa)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const BodyDev& B = bodies[wave_body(tileBody, i)];
    if (!B.angleOn || i >= B.nodeEnd) return;
    const int off = aa.rowOff[i], cnt = aa.rowOff[i + 1] - off;
    if (cnt < 2) return;
    const float4 p04 = P[i];
    const float3 p0 = xyz(p04);
    const float w0 = p04.w * B.invMassScale;
    const float dt2 = fmaxf(B.dt * B.dt, 1e-8f);
    const float alpha = fmaxf(B.angleCompliance, 0.0f) / dt2;
    for (int k = 0; k < cnt; ++k) {
        const int e0 = off + k;
        const int i2 = aa.pair[e0];
        if (i2 < 0) continue;   // unpaired: never listed as a contribution
        const int i1 = aa.nbr[e0];
        float4* out = &aa.dp[3 * (size_t)e0];
        out[0].w = 0.0f; out[1].w = 0.0f; out[2].w = 0.0f;
        if (i1 == i2) continue;
        const float rc = aa.restCos[e0];
        if (!isfinite(rc) || rc < -1.0f || rc > 1.0f) continue;
        const float4 p14 = P[i1], p24 = P[i2];
        const float3 v1 = a_sub(xyz(p14), p0), v2 = a_sub(xyz(p24), p0);
        const float l1 = a_len(v1), l2 = a_len(v2);
        if (l1 <= 1e-8f || l2 <= 1e-8f) continue;
        const float3 n1 = a_mul(v1, 1.0f / l1), n2 = a_mul(v2, 1.0f / l2);
        float cosCurr = a_dot(n1, n2);
        cosCurr = fminf(1.0f, fmaxf(-1.0f, cosCurr));
        const float C = cosCurr - rc;
        if (!isfinite(C)) continue;
        const float3 g1 = a_mul(a_sub(n2, a_mul(n1, a_dot(n2, n1))), 1.0f / l1);
        const float3 g2 = a_mul(a_sub(n1, a_mul(n2, a_dot(n1, n2))), 1.0f / l2);
        const float3 g0 = a_mul(a_add(g1, g2), -1.0f);
        const float w1 = p14.w * B.invMassScale, w2 = p24.w * B.invMassScale;
        const float denom = w0 * a_dot(g0, g0) + w1 * a_dot(g1, g1) + w2 * a_dot(g2, g2);
        if (denom < 1e-8f) continue;
        float dl = -C / (denom + alpha);
        if (!isfinite(dl)) continue;
        dl = fmaxf(fminf(dl, 0.1f), -0.1f);
        float3 dp0 = a_mul(g0, dl * w0), dp1 = a_mul(g1, dl * w1), dp2 = a_mul(g2, dl * w2);
        if (!a_finite(dp0) || !a_finite(dp1) || !a_finite(dp2)) continue;
        const float maxStep = fmaxf(1e-6f, 0.25f * fminf(l1, l2));
        const float s0 = a_len(dp0);
        if (s0 > maxStep && s0 > 1e-8f) dp0 = a_mul(dp0, maxStep / s0);
        const float s1 = a_len(dp1);
        if (s1 > maxStep && s1 > 1e-8f) dp1 = a_mul(dp1, maxStep / s1);
        const float s2 = a_len(dp2);
        if (s2 > maxStep && s2 > 1e-8f) dp2 = a_mul(dp2, maxStep / s2);
        out[0] = make_float4(dp0.x, dp0.y, dp0.z, w0 > 0.0f ? 1.0f : 0.0f);
        out[1] = make_float4(dp1.x, dp1.y, dp1.z, w1 > 0.0f ? 1.0f : 0.0f);
        out[2] = make_float4(dp2.x, dp2.y, dp2.z, w2 > 0.0f ? 1.0f : 0.0f);
    }
}

// xpbdAngleApplyKernel (5098-5130): mean of the node's counted corrections,
// blended, then the iteration epilogue.
__global__ void __launch_bounds__(256) angle_apply_kernel(int nPad, const int* __restrict__ tileBody,
                                                         const BodyDev* __restrict__ bodies,
                                                         const float4* __restrict__ Pin, float4* __restrict__ Pout,
                                                         AngleArgs aa, unsigned epi, EpiArgs ea)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const BodyDev& B = bodies[wave_body(tileBody, i)];
    const float4 p4 = Pin[i];
    float3 out = xyz(p4);
    if (B.angleOn && i < B.nodeEnd && p4.w > 0.0f) {
        float3 sum = f3(0.0f, 0.0f, 0.0f);
        int c = 0;
        for (int q = aa.revOff[i]; q < aa.revOff[i + 1]; ++q) {
            const float4 d = aa.dp[aa.revEnt[q]];
            if (d.w != 0.0f) {
                sum = a_add(sum, xyz(d));
                ++c;
            }
        }
        if (c > 0) {
            const float3 avg = a_mul(sum, 1.0f / float(c));
            if (a_finite(avg)) {
                const float blend = fminf(fmaxf(B.angleBlend, 0.0f), 1.0f);
                out = a_add(out, a_mul(avg, blend));
            }
        }
    }
    if (i < B.nodeEnd && epi && epilogue_node(epi, i, out, p4.w, B, ea)) return;
    Pout[i] = make_float4(out.x, out.y, out.z, p4.w);
}

void launch_angle(hipStream_t s, const NodeArgs& na, const float4* Pin, float4* Pout, const AngleArgs& aa,
                  unsigned epi, const EpiArgs& ea)
{
    const int blocks = (na.nPad + 255) / 256;
    angle_pair_kernel<<<blocks, 256, 0, s>>>(na.nPad, na.tileBody, na.bodies, Pin, aa);
    angle_apply_kernel<<<blocks, 256, 0, s>>>(na.nPad, na.tileBody, na.bodies, Pin, Pout, aa, epi, ea);
}

}  // namespace xpbd
