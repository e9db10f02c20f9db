// kernels_core.hip — per-substep node kernels of the MI355X XPBD solver:
// predict, distance (Jacobi over the directed graph, sliced ELL) with a fused per-node
// epilogue (press plates, kinematic colliders, ground, velocity commit), and
// the small state kernels behind the C ABI setters.
//
// Arithmetic follows the reference kernels operand for operand (the library
// is compiled with -ffp-contract=off), so every pass here reproduces the
// reference's float results bit for bit on the same inputs.
#include "kernels.h"
#include "node_ops.h"

namespace xpbd {

// ---------------------------------------------------------------------------
// Predict (forward.cu:5954-5979): p = x + (v + g dt) dt for free nodes.
// 52 B/node: x 12 + v 12 + invMass 4 read, p 16 + v 12 written (64 B moved).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) predict_kernel(int nPad, const int* __restrict__ tileBody,
                                                      const BodyDev* __restrict__ bodies,
                                                      const float3* __restrict__ X, float3* __restrict__ V,
                                                      const float* __restrict__ invMass, float4* __restrict__ P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const BodyDev& B = bodies[wave_body(tileBody, i)];
    if (i >= B.nodeEnd) return;
    const float3 x = X[i];
    const float w = invMass[i];
    if (w > 0.0f) {
        float3 v = V[i];
        v = add3(v, mul3(f3(B.grav[0], B.grav[1], B.grav[2]), B.dt));
        const float3 p = add3(x, mul3(v, B.dt));
        P[i] = make_float4(p.x, p.y, p.z, w);
        V[i] = v;
    } else {
        P[i] = make_float4(x.x, x.y, x.z, w);
        V[i] = f3(0.0f, 0.0f, 0.0f);
    }
}

// ---------------------------------------------------------------------------
// Distance constraint, Jacobi (forward.cu:6484-6571) over the directed graph
// in sliced-ELL form: one wave64 per slice of 64 nodes, entry k of the slice's
// lane l at sliceOff + 64 k + l, so every per-entry load of a wave is one
// contiguous 256 B (nbr, lambda) or 512 B (rest, alpha) request.  Rows keep
// the reference's entry order (padding only at the end, nbr = -1), so the
// per-node sums are the reference's bit for bit.  The multiplier is kept per
// directed entry in place: the first iteration of a substep does not read it
// (it is zero), the last does not write it (it is reset next substep).
// Entries are processed four at a time: the loads and neighbour gathers of a
// group are issued before its arithmetic, which stays in row order.
// Per entry 20 B (nbr 4, rest 4, alpha 4, lambda 4 + 4; alpha is a per-body
// scalar when the body has one stiffness) plus the 16 B neighbour gather
// (cache resident); per node 32 B (p in, p out).
// ---------------------------------------------------------------------------
constexpr int kDistGroup = 4;

template <bool kFirst, bool kLast>
__global__ void __launch_bounds__(256) distance_jacobi_kernel(
    int nTiles, const int* __restrict__ tileBody, const BodyDev* __restrict__ bodies,
    const float4* __restrict__ Pin, float4* __restrict__ Pout, SellArgs sa, unsigned epi, EpiArgs ea)
{
    const int tile = xcd_block() * (int)(blockDim.x >> 6) + (threadIdx.x >> 6);
    if (tile >= nTiles) return;
    const int lane = threadIdx.x & 63;
    const int i = tile * kTile + lane;
    const BodyDev& B = bodies[__builtin_amdgcn_readfirstlane(tileBody[tile])];
    if (i >= B.nodeEnd) return;
    const float4 pi4 = Pin[i];
    const float3 pi = xyz(pi4);
    if (!B.distanceOn) {
        float3 p = pi;
        if (epi && epilogue_node(epi, i, p, pi4.w, B, ea)) return;
        Pout[i] = make_float4(p.x, p.y, p.z, pi4.w);
        return;
    }
    const float wi = pi4.w * B.invMassScale;
    float3 acc = f3(0.0f, 0.0f, 0.0f);
    int valid = 0;
    const bool uniformAlpha = B.distAlphaUniform != 0;   // one stiffness for the whole body
    const int width = __builtin_amdgcn_readfirstlane(sa.sliceW[tile]);
    const int base = __builtin_amdgcn_readfirstlane(sa.sliceOff[tile]) + lane;
    for (int e0 = 0; e0 < width; e0 += kDistGroup) {
        int jj[kDistGroup];
        float rl[kDistGroup], al[kDistGroup];
        float ll[kDistGroup];
        float4 pj[kDistGroup];
#pragma unroll
        for (int k = 0; k < kDistGroup; ++k) jj[k] = (e0 + k < width) ? sa.nbr[base + (e0 + k) * kTile] : -1;
#pragma unroll
        for (int k = 0; k < kDistGroup; ++k) {
            if (jj[k] >= 0) {
                const int slot = base + (e0 + k) * kTile;
                rl[k] = sa.rest[slot];
                al[k] = uniformAlpha ? B.distAlphaU : sa.alpha[slot];
                ll[k] = kFirst ? 0.0f : sa.lam[slot];
                pj[k] = Pin[jj[k]];
            }
        }
#pragma unroll
        for (int k = 0; k < kDistGroup; ++k) {
            if (jj[k] < 0) continue;
            const int slot = base + (e0 + k) * kTile;
            float l = ll[k];
            const float3 diff = sub3(pi, xyz(pj[k]));
            const float d = len3(diff);
            if (d < 1e-7f) {
                if (!kLast && kFirst) sa.lam[slot] = l;
                continue;
            }
            const float C = d - rl[k];
            const float3 grad = mul3(diff, 1.0f / d);
            const float wj = pj[k].w * B.invMassScale;
            const float at = al[k];
            const float denom = (wi + wj) + at;
            if (denom < 1e-8f) {
                if (!kLast && kFirst) sa.lam[slot] = l;
                continue;
            }
            const float dl = (-C - at * l) / denom;
            l += dl;
            if (!kLast) sa.lam[slot] = l;
            if (wi > 0.0f) {
                acc = add3(acc, mul3(grad, wi * dl));
                valid++;
            }
        }
    }
    float3 out = pi;
    if (wi > 0.0f && valid > 0) {
        const float3 avg = mul3(acc, 1.0f / float(valid));
        out = add3(pi, mul3(avg, B.underRelax));
    }
    if (epi && epilogue_node(epi, i, out, pi4.w, B, ea)) return;
    Pout[i] = make_float4(out.x, out.y, out.z, pi4.w);
}

// Epilogue as its own pass (in place) when no cross-node pass hosts it.
__global__ void __launch_bounds__(256) epilogue_kernel(int nPad, const int* __restrict__ tileBody,
                                                       const BodyDev* __restrict__ bodies, float4* __restrict__ P,
                                                       unsigned epi, EpiArgs ea)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const BodyDev& B = bodies[wave_body(tileBody, i)];
    if (i >= B.nodeEnd) return;
    const float4 p4 = P[i];
    float3 p = xyz(p4);
    if (epilogue_node(epi, i, p, p4.w, B, ea)) return;
    P[i] = make_float4(p.x, p.y, p.z, p4.w);
}

// alpha tilde per directed entry (the expression of forward.cu:6533-6545,
// evaluated once per settings change instead of every iteration), written to
// the sliced-ELL slot of the entry.
__global__ void distance_alpha_kernel(int nPad, const int* __restrict__ tileBody, const BodyDev* __restrict__ bodies,
                                      const int* __restrict__ rowOff, const float* __restrict__ nstiff, SellArgs sa)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const BodyDev& B = bodies[wave_body(tileBody, i)];
    if (i >= B.nodeEnd || B.distAlphaUniform) return;
    const float compliance = B.distCompliance, stiffScale = B.stiffnessScale;
    const float dt2 = fmaxf(B.dt * B.dt, 1e-8f);
    const int tile = i / kTile, lane = i % kTile;
    const int base = sa.sliceOff[tile] + lane;
    for (int e = rowOff[i]; e < rowOff[i + 1]; ++e) {
        const float st = fmaxf(nstiff[e] * stiffScale, 1e-6f);
        const float alpha = compliance / st;
        sa.alpha[base + (e - rowOff[i]) * kTile] = alpha / dt2;
    }
}

// attachApplyKernel (forward.cu:6714-6725), node ids already offset.
__global__ void attach_kernel(int count, const int* __restrict__ idx, const float* __restrict__ pos, int nodeBegin,
                              int nodeEnd, float3* X, float3* V, float* invMass, float* objW, float weight)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int li = idx[k];
    if (li < 0 || li >= nodeEnd - nodeBegin) return;
    const int i = nodeBegin + li;
    invMass[i] = 0.0f;
    X[i] = f3(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]);
    V[i] = f3(0.0f, 0.0f, 0.0f);
    objW[i] = weight;
}

// addRigidVelocityAllKernel (forward.cu:6701-6712)
__global__ void rigid_velocity_kernel(int nodeBegin, int nodeEnd, const float3* X, float3* V, const float* invMass,
                                      float3 dv, float3 dw, float3 c)
{
    const int i = nodeBegin + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nodeEnd || invMass[i] <= 0.0f) return;
    const float3 p = X[i];
    const float rx = p.x - c.x, ry = p.y - c.y, rz = p.z - c.z;
    const float3 v = V[i];
    V[i] = f3(v.x + dv.x + (dw.y * rz - dw.z * ry), v.y + dv.y + (dw.z * rx - dw.x * rz),
              v.z + dv.z + (dw.x * ry - dw.y * rx));
}

__global__ void scatter_caller_kernel(int nodeBegin, int nodeEnd, const int* __restrict__ callerPos,
                                      const float3* __restrict__ X, float3* __restrict__ xCaller)
{
    const int i = nodeBegin + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nodeEnd) xCaller[callerPos[i]] = X[i];
}

__global__ void fill_float_kernel(int n, float* a, float v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void fill_int_kernel(int n, int* a, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline int blocks_for(int n, int bs) { return (n + bs - 1) / bs; }

void launch_predict(hipStream_t s, const NodeArgs& na, const float3* X, float3* V, const float* invMass, float4* P)
{
    predict_kernel<<<blocks_for(na.nPad, 256), 256, 0, s>>>(na.nPad, na.tileBody, na.bodies, X, V, invMass, P);
}

void launch_distance_jacobi(hipStream_t s, const NodeArgs& na, bool first, bool last, const float4* Pin, float4* Pout,
                            const SellArgs& sa, unsigned epi, const EpiArgs& ea, const Tuning& tu)
{
    const int tiles = na.nPad / kTile;
    const int per = (tu.jacobiTiles == 1 || tu.jacobiTiles == 2) ? tu.jacobiTiles : 4;
    const dim3 g(xcd_grid(blocks_for(tiles, per))), b(64 * per);
    if (first && last)
        distance_jacobi_kernel<true, true><<<g, b, 0, s>>>(tiles, na.tileBody, na.bodies, Pin, Pout, sa, epi, ea);
    else if (first)
        distance_jacobi_kernel<true, false><<<g, b, 0, s>>>(tiles, na.tileBody, na.bodies, Pin, Pout, sa, epi, ea);
    else if (last)
        distance_jacobi_kernel<false, true><<<g, b, 0, s>>>(tiles, na.tileBody, na.bodies, Pin, Pout, sa, epi, ea);
    else
        distance_jacobi_kernel<false, false><<<g, b, 0, s>>>(tiles, na.tileBody, na.bodies, Pin, Pout, sa, epi, ea);
}

void launch_epilogue(hipStream_t s, const NodeArgs& na, float4* P, unsigned epi, const EpiArgs& ea)
{
    epilogue_kernel<<<blocks_for(na.nPad, 256), 256, 0, s>>>(na.nPad, na.tileBody, na.bodies, P, epi, ea);
}

void launch_distance_alpha(hipStream_t s, const NodeArgs& na, const int* rowOff, const float* nstiff, const SellArgs& sa)
{
    distance_alpha_kernel<<<blocks_for(na.nPad, 256), 256, 0, s>>>(na.nPad, na.tileBody, na.bodies, rowOff, nstiff, sa);
}

void launch_attach(hipStream_t s, int count, const int* idx, const float* pos, int nodeBegin, int nodeEnd, float3* X,
                   float3* V, float* invMass, float* objW, float weight)
{
    if (count <= 0) return;
    attach_kernel<<<blocks_for(count, 256), 256, 0, s>>>(count, idx, pos, nodeBegin, nodeEnd, X, V, invMass, objW, weight);
}

void launch_rigid_velocity(hipStream_t s, int nodeBegin, int nodeEnd, const float3* X, float3* V, const float* invMass,
                           float3 dv, float3 dw, float3 c)
{
    const int n = nodeEnd - nodeBegin;
    if (n <= 0) return;
    rigid_velocity_kernel<<<blocks_for(n, 256), 256, 0, s>>>(nodeBegin, nodeEnd, X, V, invMass, dv, dw, c);
}

void launch_scatter_caller(hipStream_t s, int nodeBegin, int nodeEnd, const int* callerPos, const float3* X,
                           float3* xCaller)
{
    const int n = nodeEnd - nodeBegin;
    if (n > 0) scatter_caller_kernel<<<blocks_for(n, 256), 256, 0, s>>>(nodeBegin, nodeEnd, callerPos, X, xCaller);
}

void launch_fill_float(hipStream_t s, int n, float* a, float v)
{
    if (n > 0) fill_float_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, a, v);
}

void launch_fill_int(hipStream_t s, int n, int* a, int v)
{
    if (n > 0) fill_int_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, a, v);
}

__global__ void fill_double_zero_kernel(int n, double* a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

void launch_fill_double(hipStream_t s, int n, double* a)
{
    if (n > 0) fill_double_zero_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, a);
}

}  // namespace xpbd
