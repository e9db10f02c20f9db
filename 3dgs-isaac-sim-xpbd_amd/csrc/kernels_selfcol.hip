// kernels_selfcol.hip — object (connected-component) shape matching and the
// spatial-hash self-collision of the MI355X XPBD solver.
//
// Object shape (forward.cu:6801-6937, 11671-11702): per object, fp64 sums of
// (p - s) and (p - s) q^T (q = x_rest - c0, s = last centroid), wave-reduced
// when a wave is one object, then fp64 polar decomposition and
// p += stiffness * (c + R q - p).  Self-collision (forward.cu:6986-7322,
// 11796-11849): max predicted displacement -> cell size, cell keys of the
// frame-start positions, stable radix sort, cell ranges, per-node candidate
// lists (27 cells, 32 slots, deterministic worst replacement), and per
// iteration a read pass (separation to d_c / averaged-normal push for fast
// pairs) and a write pass that also hosts the iteration epilogue.
#include "kernels.h"
#include "mat3.h"
#include "node_ops.h"
#include "polar.h"

namespace xpbd {


__device__ __forceinline__ const BodyDev& body_at(const SelfColArgs& a, int i)
{
    return a.bodies[__builtin_amdgcn_readfirstlane(a.tileBody[i / kTile])];
}

// ---------------------------------------------------------------------------
// Object shape matching
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) obj_accum_kernel(int nPad, SelfColArgs a, const float4* __restrict__ P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int id = -1;
    double v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = 0.0;
    int stride = 12;
    if (i < nPad) {
        const BodyDev& B = body_at(a, i);
        stride = B.objWeighted ? 16 : 12;
        if (B.objOn && i < B.nodeEnd) id = a.objId[i];
        if (id >= 0) {
            const float4 p = P[i];
            const double dx = (double)p.x - a.objShift[3 * id], dy = (double)p.y - a.objShift[3 * id + 1],
                         dz = (double)p.z - a.objShift[3 * id + 2];
            const double qx = a.objRestRel[3 * (size_t)i], qy = a.objRestRel[3 * (size_t)i + 1],
                         qz = a.objRestRel[3 * (size_t)i + 2];
            v[0] = dx; v[1] = dy; v[2] = dz;
            v[3] = dx * qx; v[4] = dx * qy; v[5] = dx * qz;
            v[6] = dy * qx; v[7] = dy * qy; v[8] = dy * qz;
            v[9] = dz * qx; v[10] = dz * qy; v[11] = dz * qz;
            if (B.objWeighted) {
                const double w = (double)a.objW[i];
#pragma unroll
                for (int k = 0; k < 12; ++k) v[k] *= w;
                v[12] = w; v[13] = w * qx; v[14] = w * qy; v[15] = w * qz;
            }
        }
    }
    const int id0 = __shfl(id, 0, 64);
    const bool uniform = __all(id == id0);
    if (uniform) {
        if (id0 < 0) return;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (k >= stride) break;
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) v[k] += __shfl_xor(v[k], m, 64);
        }
        if (lane < stride) {
            double mine = 0.0;
#pragma unroll
            for (int k = 0; k < 16; ++k)
                if (k == lane) mine = v[k];
            atomicAdd(&a.objSums[16 * (size_t)id0 + lane], mine);
        }
    } else if (id >= 0) {
        for (int k = 0; k < stride; ++k) atomicAdd(&a.objSums[16 * (size_t)id + k], v[k]);
    }
}

// One thread per object (forward.cu:6887-6937).
__global__ void obj_solve_kernel(int K, SelfColArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int b = a.objBody[k];
    const BodyDev& B = a.bodies[b];
    if (!B.objOn) return;
    float* Pm = &a.objParams[16 * (size_t)k];
    const int n = a.objCount[k];
    if (n < 4) { Pm[15] = 0.0f; a.objDisp[k] = 0.0f; return; }
    const double* S = &a.objSums[16 * (size_t)k];
    const bool wOn = B.objWeighted != 0;
    const double W = wOn ? S[12] : (double)n;
    if (!(W > 0.0)) { Pm[15] = 0.0f; a.objDisp[k] = 0.0f; return; }
    const double inv = 1.0 / W;
    const double dc[3] = {S[0] * inv, S[1] * inv, S[2] * inv};
    const double* sq = wOn ? &S[13] : &a.objSumQ[3 * (size_t)k];
    double m[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m[3 * r + c] = S[3 + 3 * r + c] - dc[r] * sq[c];
    double cNow[3];
    for (int t = 0; t < 3; ++t) {
        cNow[t] = a.objShift[3 * k + t] + dc[t];
        a.objShift[3 * k + t] = cNow[t];
    }
    a.objDisp[k] = (float)sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
    double fn = 0.0;
    for (int t = 0; t < 9; ++t) fn += m[t] * m[t];
    fn = sqrt(fn);
    if (!(fn > 1e-300) || !isfinite(fn)) { Pm[15] = 0.0f; return; }
    M3 A, R;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) A.m[c][r] = (float)(m[3 * r + c] / fn);
    if (!polar_rotation_f64(A, R)) { Pm[15] = 0.0f; return; }
    for (int t = 0; t < 3; ++t) {
        Pm[t] = wOn ? (float)((double)a.objC0[3 * k + t] + S[13 + t] * inv) : a.objC0[3 * k + t];
        Pm[3 + t] = (float)cNow[t];
        for (int cc = 0; cc < 3; ++cc) Pm[6 + 3 * t + cc] = R.m[cc][t];
    }
    Pm[15] = 1.0f;
}

__device__ __forceinline__ void atomic_min_f(float* p, float v)
{
    v += 0.0f;
    if (v >= 0.0f) atomicMin((int*)p, __float_as_int(v));
    else atomicMax((unsigned int*)p, __float_as_uint(v));
}

__device__ __forceinline__ void atomic_max_f(float* p, float v)
{
    v += 0.0f;
    if (v >= 0.0f) atomicMax((int*)p, __float_as_int(v));
    else atomicMin((unsigned int*)p, __float_as_uint(v));
}

__global__ void obj_boxes_init_kernel(int K, float* boxes)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 6 * K) boxes[t] = ((t % 6) < 3) ? 3.402823466e+38f : -3.402823466e+38f;
}

// Object boxes of the predicted positions (forward.cu:6959-6981), then the
// shape-matching move (6801-6823).  Boxes come first: they see the positions
// before the move, as in the reference.
__global__ void __launch_bounds__(256) obj_boxes_kernel(int nPad, SelfColArgs a, const float4* __restrict__ P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const BodyDev& B = body_at(a, i);
    if (!B.objOn || !B.objBoxes || i >= B.nodeEnd) return;
    const int id = a.objId[i];
    const float4 p = P[i];
    if (id < 0 || !isfinite(p.x) || !isfinite(p.y) || !isfinite(p.z)) return;
    float* b = &a.objBoxesArr[6 * (size_t)id];
    atomic_min_f(&b[0], p.x); atomic_min_f(&b[1], p.y); atomic_min_f(&b[2], p.z);
    atomic_max_f(&b[3], p.x); atomic_max_f(&b[4], p.y); atomic_max_f(&b[5], p.z);
}

__global__ void __launch_bounds__(256) obj_match_kernel(int nPad, SelfColArgs a, float4* __restrict__ P,
                                                       const float3* __restrict__ rest)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const BodyDev& B = body_at(a, i);
    if (!B.objOn || i >= B.nodeEnd) return;
    const int id = a.objId[i];
    const float4 p = P[i];
    if (id < 0 || p.w <= 0.0f) return;
    const float* Pm = &a.objParams[16 * (size_t)id];
    if (Pm[15] == 0.0f) return;
    const float3 r = rest[i];
    const float qx = r.x - Pm[0], qy = r.y - Pm[1], qz = r.z - Pm[2];
    const float gx = Pm[3] + Pm[6] * qx + Pm[7] * qy + Pm[8] * qz;
    const float gy = Pm[4] + Pm[9] * qx + Pm[10] * qy + Pm[11] * qz;
    const float gz = Pm[5] + Pm[12] * qx + Pm[13] * qy + Pm[14] * qz;
    const float s = B.objStiffness;
    P[i] = make_float4(p.x + s * (gx - p.x), p.y + s * (gy - p.y), p.z + s * (gz - p.z), p.w);
}

// ---------------------------------------------------------------------------
// Self-collision build
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned int sc_hash(int cx, int cy, int cz, unsigned int mask)
{
    return (((unsigned int)cx * 73856093u) ^ ((unsigned int)cy * 19349663u) ^ ((unsigned int)cz * 83492791u)) & mask;
}

__device__ __forceinline__ int sc_cell(float v, float invCell)
{
    return (int)floorf(fminf(fmaxf(v * invCell, -1.0e9f), 1.0e9f));
}

// Key of a cell of the predicted-position grid: a hash of the row (cy, cz)
// plus cx, so the three x-neighbours of a row have consecutive keys and sit
// next to each other in the sorted order.
__device__ __forceinline__ unsigned int sc_fine_key(int cx, int cy, int cz, unsigned int mask)
{
    return ((((unsigned int)cy * 19349663u) ^ ((unsigned int)cz * 83492791u)) + (unsigned int)cx) & mask;
}

// Per body: clear the build counters and the max displacement, and decide
// whether the active restriction applies (boxes valid, finite body motion).
__global__ void sc_prepare_kernel(int B, SelfColArgs a)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const BodyDev& D = a.bodies[b];
    a.scMax2[b] = 0u;
    a.scCounters[3 * b] = 0; a.scCounters[3 * b + 1] = 0; a.scCounters[3 * b + 2] = 0;
    if (b == 0 && a.accUsed) *a.accUsed = 0;   // the deferred-list pool starts empty
    float maxDisp = 0.0f;
    for (int k = 0; k < D.objK; ++k) maxDisp = fmaxf(maxDisp, a.objDisp[D.objBase + k]);
    a.scUseActive[b] = (D.scOn && D.objActiveAllowed && isfinite(maxDisp)) ? 1 : 0;
    a.scBodyDisp[b] = maxDisp;
}

// Active mask (forward.cu:6998-7030) and max |p - x|^2 over active nodes (7058-7068).
// Blocks of 1024 threads: each wave (one 64-node tile of one body) reduces its
// largest move, then the block combines its waves, one atomic per run of
// waves of the same body (atomics on one address serialise).
constexpr int kDispThreads = 1024;

__global__ void __launch_bounds__(kDispThreads) sc_active_disp_kernel(int nPad, SelfColArgs a,
                                                                     const float3* __restrict__ X,
                                                                     const float4* __restrict__ P)
{
    __shared__ unsigned int wmax[kDispThreads / 64];
    __shared__ int wbody[kDispThreads / 64];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int w = threadIdx.x >> 6;
    int b = -1;
    unsigned int m = 0u;
    if (i < nPad) {   // nPad is a multiple of 64: whole waves
        b = __builtin_amdgcn_readfirstlane(a.tileBody[i / kTile]);
        const BodyDev& B = a.bodies[b];
        if (B.scOn) {
            const bool node = i < B.nodeEnd;
            float3 x = make_float3(0.0f, 0.0f, 0.0f);
            float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (node) { x = X[i]; p = P[i]; }
            unsigned char act = node ? 1 : 0;
            if (node && a.scUseActive[b]) {
                const int id = a.objId[i];
                act = 0;
                if (id < 0) {
                    act = 1;
                } else {
                    const float inflate = B.scSearch + 2.0f * a.scBodyDisp[b] + 2.0f * B.scDc;
                    for (int k = 0; k < B.objK; ++k) {
                        const int gk = B.objBase + k;
                        if (gk == id) continue;
                        const float* bx = &a.objBoxesArr[6 * (size_t)gk];
                        const float lx = bx[0] - inflate, ly = bx[1] - inflate, lz = bx[2] - inflate;
                        const float hx = bx[3] + inflate, hy = bx[4] + inflate, hz = bx[5] + inflate;
                        const bool inP = p.x >= lx && p.x <= hx && p.y >= ly && p.y <= hy && p.z >= lz && p.z <= hz;
                        const bool inX = x.x >= lx && x.x <= hx && x.y >= ly && x.y <= hy && x.z >= lz && x.z <= hz;
                        if (inP || inX) { act = 1; break; }
                    }
                }
            }
            a.scActive[i] = act;
            const float dx = p.x - x.x, dy = p.y - x.y, dz = p.z - x.z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            // non-negative floats order like their bit patterns
            m = (act && isfinite(d2) && d2 > 0.0f) ? __float_as_uint(d2) : 0u;
        } else {
            b = -1;
        }
    }
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) m = max(m, (unsigned int)__shfl_xor((int)m, s, 64));
    if ((threadIdx.x & 63) == 0) {
        wmax[w] = m;
        wbody[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (int)(blockDim.x >> 6);
        unsigned int run = 0u;
        for (int k = 0; k < nw; ++k) {
            if (wbody[k] < 0) continue;
            run = max(run, wmax[k]);
            if (k + 1 == nw || wbody[k + 1] != wbody[k]) {
                if (run > 0u) atomicMax(&a.scMax2[wbody[k]], run);
                run = 0u;
            }
        }
    }
}

// Cell size per body: search + 2 max displacement, capped at search + 16 d_c
// (forward.cu:11820-11831), or the fixed override.
__global__ void sc_reach_kernel(int B, SelfColArgs a)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const BodyDev& D = a.bodies[b];
    const float maxDisp2 = __uint_as_float(a.scMax2[b]);
    const float maxDisp = (isfinite(maxDisp2) && maxDisp2 > 0.0f) ? sqrtf(maxDisp2) : 0.0f;
    float reach = D.scSearch + 2.0f * maxDisp;
    const float cap = D.scSearch + 16.0f * D.scDc;
    if (D.scCellOverride > 0.0f) reach = D.scCellOverride;
    else if (reach > cap) reach = cap;
    a.scInvCell[b] = 1.0f / reach;
}

// Keys of the frame-start positions (forward.cu:7034-7044) in the body's table,
// stored at the node's caller-order position with the device index as value,
// so that the stable sort breaks ties by the caller's node order.
__global__ void __launch_bounds__(256) sc_keys_kernel(int nPad, SelfColArgs a, const float3* __restrict__ X)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const int b = __builtin_amdgcn_readfirstlane(a.tileBody[i / kTile]);
    const BodyDev& B = a.bodies[b];
    unsigned int key = a.sentinel;
    if (B.scOn && i < B.nodeEnd && a.scActive[i]) {
        const float3 p = X[i];
        if (isfinite(p.x) && isfinite(p.y) && isfinite(p.z)) {
            const float inv = a.scInvCell[b];
            key = (unsigned int)B.scTableBase +
                  sc_hash(sc_cell(p.x, inv), sc_cell(p.y, inv), sc_cell(p.z, inv), (unsigned int)B.scTableMask);
        }
    }
    const int pos = a.callerPos[i];
    a.keys[pos] = key;
    a.idx[pos] = i;
}

__global__ void __launch_bounds__(256) sc_range_kernel(int nPad, SelfColArgs a, const unsigned int* __restrict__ keysSorted)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const unsigned int k = keysSorted[i];
    if (k >= a.sentinel) return;
    if (i == 0 || keysSorted[i - 1] != k) a.cellStart[k] = i;
    if (i == nPad - 1 || keysSorted[i + 1] != k) a.cellEnd[k] = i + 1;
}

// Sorted-order copies of what the coarse candidate scan reads, so that each
// cell of the stencil is one contiguous run: (p, object id), (x, node id),
// rest.  Bodies with the predicted-position grid get only the sorted rank:
// the coarse scan walks few of their nodes and reads those through the
// permutation.
__global__ void __launch_bounds__(256) sc_stage_kernel(int nPad, SelfColArgs a, const float3* __restrict__ X,
                                                      const float4* __restrict__ P, const float3* __restrict__ rest,
                                                      const int* __restrict__ sortedIdx)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nPad) return;
    const int j = sortedIdx[t];
    if (a.coarseRank) {
        a.coarseRank[j] = t;
        if (a.bodies[a.tileBody[j / kTile]].scFine) return;
    }
    const float4 p = P[j];
    const float3 x = X[j], r = rest[j];
    a.sortP[t] = make_float4(p.x, p.y, p.z, __int_as_float(a.objId[j]));
    a.sortX[t] = make_float4(x.x, x.y, x.z, 0.0f);
    a.sortR[t] = make_float4(r.x, r.y, r.z, __int_as_float(j));
}

// Whether two of the 27 coarse cells around x hash to one bucket.  The hash
// is X ^ Y ^ Z of per-axis products, so two cells collide exactly when
// (dX ^ dY ^ dZ) & mask == 0, where each d is 0 (same coordinate) or the xor
// of two of the three products of that axis: 63 cases instead of 351 pairs.
__device__ __forceinline__ bool sc_stencil_shares_bucket(const float3& x, float inv, unsigned int mask)
{
    const int cx = sc_cell(x.x, inv), cy = sc_cell(x.y, inv), cz = sc_cell(x.z, inv);
    const unsigned int x0 = (unsigned int)(cx - 1) * 73856093u, x1 = (unsigned int)cx * 73856093u,
                       x2 = (unsigned int)(cx + 1) * 73856093u;
    const unsigned int y0 = (unsigned int)(cy - 1) * 19349663u, y1 = (unsigned int)cy * 19349663u,
                       y2 = (unsigned int)(cy + 1) * 19349663u;
    const unsigned int z0 = (unsigned int)(cz - 1) * 83492791u, z1 = (unsigned int)cz * 83492791u,
                       z2 = (unsigned int)(cz + 1) * 83492791u;
    const unsigned int dx[4] = {0u, x0 ^ x1, x0 ^ x2, x1 ^ x2};
    const unsigned int dy[4] = {0u, y0 ^ y1, y0 ^ y2, y1 ^ y2};
    const unsigned int dz[4] = {0u, z0 ^ z1, z0 ^ z2, z1 ^ z2};
    bool dup = false;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int w = 0; w < 4; ++w)
                if (u | v | w) dup |= ((dx[u] ^ dy[v] ^ dz[w]) & mask) == 0u;
    return dup;
}

// Keys of the predicted positions in the fine grid (bodies with scFine): the
// nodes that have a coarse key (active, finite x) and a finite p.
__global__ void __launch_bounds__(256) sc_fine_keys_kernel(int nPad, SelfColArgs a, const float3* __restrict__ X,
                                                          const float4* __restrict__ P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const int b = __builtin_amdgcn_readfirstlane(a.tileBody[i / kTile]);
    const BodyDev& B = a.bodies[b];
    unsigned int key = (unsigned int)a.fineTotal;
    if (B.scOn && B.scFine && i < B.nodeEnd && a.scActive[i]) {
        const float3 x = X[i];
        const float4 p = P[i];
        if (isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(p.x) && isfinite(p.y) && isfinite(p.z)) {
            const float g = B.scFineInv;
            key = (unsigned int)B.scFineBase +
                  sc_fine_key(sc_cell(p.x, g), sc_cell(p.y, g), sc_cell(p.z, g), (unsigned int)B.scFineMask);
            // bit 1 of the active flag: two of the 27 coarse stencil cells
            // around x share a bucket (the fine scans then look up the first
            // offset of each candidate's bucket)
            int flags = 1;
            if (sc_stencil_shares_bucket(x, a.scInvCell[b], (unsigned int)B.scTableMask)) flags |= 2;
            // bit 2: the nine row stretches of the 27 fine cells around p
            // (keys k_r - 1 .. k_r + 1 around the middle-cell key k_r) cannot
            // be walked as stretches, because one wraps around the table or two
            // come within two keys of each other; the fine scans then walk each
            // distinct cell once
            const int gx = sc_cell(p.x, g), gy = sc_cell(p.y, g), gz = sc_cell(p.z, g);
            const unsigned int fmask = (unsigned int)B.scFineMask;
            unsigned int kr[9];
            bool perCell = false;
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                kr[r] = sc_fine_key(gx, gy + r % 3 - 1, gz + r / 3 - 1, fmask);
                perCell |= kr[r] == 0u || kr[r] == fmask;
            }
#pragma unroll
            for (int r = 1; r < 9; ++r)
#pragma unroll
                for (int r2 = 0; r2 < r; ++r2) perCell |= (kr[r] > kr[r2] ? kr[r] - kr[r2] : kr[r2] - kr[r]) <= 2u;
            if (perCell) flags |= 4;
            if (flags != 1) a.scActive[i] = flags;
        }
    }
    a.fkeys[i] = key;
    a.fidx[i] = i;
}

__global__ void __launch_bounds__(256) sc_fine_range_kernel(int nPad, SelfColArgs a,
                                                           const unsigned int* __restrict__ keysSorted)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const unsigned int k = keysSorted[i];
    if (k >= (unsigned int)a.fineTotal) return;
    if (i == 0 || keysSorted[i - 1] != k) a.fcellStart[k] = i;
    if (i == nPad - 1 || keysSorted[i + 1] != k) a.fcellEnd[k] = i + 1;
}

// Fine-order copies for the fine scans: (p, node id), (rest, coarse sorted
// position: the candidate's place in the coarse scan) and (coarse cell of x,
// coarse key).
__global__ void __launch_bounds__(256) sc_fine_stage_kernel(int nPad, SelfColArgs a, const float3* __restrict__ X,
                                                           const float4* __restrict__ P,
                                                           const float3* __restrict__ rest,
                                                           const unsigned int* __restrict__ keysSorted,
                                                           const int* __restrict__ sortedIdx,
                                                           const unsigned int* __restrict__ coarseSorted)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nPad || keysSorted[t] >= (unsigned int)a.fineTotal) return;
    const int j = sortedIdx[t];
    const float4 p = P[j];
    const float3 r = rest[j], x = X[j];
    const int rk = a.coarseRank[j];
    const float inv = a.scInvCell[a.tileBody[j / kTile]];
    a.fineP[t] = make_float4(p.x, p.y, p.z, __int_as_float(j));
    a.fineR[t] = make_float4(r.x, r.y, r.z, __int_as_float(rk));
    a.fineC[t] = make_int4(sc_cell(x.x, inv), sc_cell(x.y, inv), sc_cell(x.z, inv), (int)coarseSorted[rk]);   // j's coarse key
}

// Distance from x to the stencil cell at offset o along one axis, given the
// position f in [0, 1) of x inside its own cell of edge e.
__device__ __forceinline__ float sc_axis_gap(int o, float f, float e)
{
    return o < 0 ? f * e : (o > 0 ? (1.0f - f) * e : 0.0f);
}

__device__ __forceinline__ int lane_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

// Candidate lists are slot-major inside each 64-node tile: slot k of node i
// at tile * 64 * 32 + 64 k + (i mod 64), so a wave reads one slot of its 64
// nodes as one contiguous 256 B request.
__device__ __forceinline__ size_t sc_slot(int i, int k)
{
    return (size_t)(i / kTile) * (kTile * kSelfColCap) + (size_t)k * kTile + (i % kTile);
}
__device__ __forceinline__ float lane_fl(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Lane holding the worst of the 32 list slots (lanes 0-31): highest tier, then
// largest priority, then the lowest slot (the scan of forward.cu:7160-7172).
__device__ __forceinline__ int sc_worst_lane(int tier, float prio, int lane)
{
    int wl = lane < kSelfColCap ? lane : 0;
    int wt = lane < kSelfColCap ? tier : -1;
    float wp = prio;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const int ol = __shfl_xor(wl, m, 64), ot = __shfl_xor(wt, m, 64);
        const float op = __shfl_xor(wp, m, 64);
        const bool take = ot > wt || (ot == wt && (op > wp || (op == wp && ol < wl)));
        if (take) { wl = ol; wt = ot; wp = op; }
    }
    return __builtin_amdgcn_readfirstlane(wl);
}

// Candidate-list slots of one node: lanes 0-31 hold slot (code, tier,
// priority); the worst slot (highest tier, then largest priority, then lowest
// slot, the scan of forward.cu:7160-7172) is cached until a slot changes.
struct ScSlots {
    int cnt = 0, overflow = 0;
    int code = 0, tier = 0;
    float prio = 0.0f;
    int worst = -1, worstTier = 0;
    float worstPrio = 0.0f;

    // One accepted candidate, in scan order (wave-uniform arguments).
    __device__ __forceinline__ void insert(int cc, int ct, float cp, int lane)
    {
        if (cnt < kSelfColCap) {
            if (lane == cnt) { code = cc; tier = ct; prio = cp; }
            ++cnt;
            return;
        }
        if (worst < 0) {
            worst = sc_worst_lane(tier, prio, lane);
            worstTier = lane_i(tier, worst);
            worstPrio = lane_fl(prio, worst);
        }
        if (ct < worstTier || (ct == worstTier && cp < worstPrio)) {
            if (lane == worst) { code = cc; tier = ct; prio = cp; }
            worst = -1;
        }
        ++overflow;
    }
};

// Walk of a virtual run of up to 27 bucket ranges (lane c < 27 holds range c
// as start rs and length rn): returns the exclusive prefix of the lengths in
// `pre` and the run length.
__device__ __forceinline__ int sc_run_prefix(int rn, int lane, int& pre)
{
    pre = rn;
#pragma unroll
    for (int m = 1; m < 32; m <<= 1) {
        const int o = __shfl_up(pre, m, 64);
        if (lane >= m) pre += o;
    }
    const int total = lane_i(pre, 26);
    pre -= rn;
    return total;
}

// Sorted-array position of run position v: the last range c with pre_c <= v.
__device__ __forceinline__ int sc_run_pos(int v, int rs, int pre)
{
    int c0 = 0;
#pragma unroll
    for (int step = 16; step >= 1; step >>= 1) {
        const int cand = c0 + step;
        const int pc = __shfl(pre, cand < 27 ? cand : 26, 64);
        if (cand < 27 && pc <= v) c0 = cand;
    }
    return __shfl(rs, c0, 64) + (v - __shfl(pre, c0, 64));
}

constexpr int kFineList = 128;   // accepted candidates a wave sorts in LDS (a multiple of 64)
constexpr int kAccChunk = 1024;  // list-pool entries a wave reserves at once (>= kFineList)
constexpr int kScNeedCoarse = -2;  // list length: the node takes the coarse scan
constexpr int kScNeedWave = -3;    // list length: the node takes the wave fine scan
constexpr int kScDone = -4;        // list length: the group fine scan wrote the slots

// Block order for the candidate kernels.  Blocks go to the XCDs round-robin.
// xcd > 1: each XCD takes runs of xcd consecutive blocks (nodes in Morton
// order, so neighbouring cells stay in that XCD's L2 while crowded regions
// spread over every XCD); xcd == 1: one contiguous run per XCD.
__device__ __forceinline__ int sc_block(int xcd)
{
    int blk = (int)blockIdx.x;
    if (xcd == 1) {
        blk = xcd_block();
    } else if (xcd > 1) {
        const int x = blk & 7, r = blk >> 3;
        blk = ((r / xcd) * 8 + x) * xcd + r % xcd;
    }
    return blk;
}

// Fine scan by lane groups: G lanes per node, 64 / G consecutive nodes per
// wave, for bodies whose nodes all belong to one object with self-contact
// inside it on (the only test there is |p_i - p_j| < search).  A group walks
// the 27 cells (edge >= search) around p_i in the predicted-position grid as
// nine row stretches of the sorted order, G consecutive entries per step (so
// one load of a group touches one or two cache lines), and appends the
// scan-order keys of its accepted candidates (first stencil offset of the
// coarse bucket << 27 | coarse sorted position, as in the wave scan below) to
// the node's LDS list.  The keys are then ranked: with at most 32 the ranks
// are the reference's slot order and the group writes the slots itself; with
// up to kGroupList the ranked list (code, squared distance) goes to the list
// pool for the slot kernel, which replays the reference's replacement order.
// Longer lists, crowded cells and rows that wrap around the table or overlap
// go to the wave fine scan.
constexpr int kGroupList = 64;               // accepted candidates a group keeps
constexpr int kGroupStride = kGroupList + 4;  // LDS row of one node: rows of the 8 groups of a wave start on distinct banks
constexpr int kGroupScan = 2048;             // most grid entries a group walks

template <int G>
__global__ void __launch_bounds__(256) sc_fine_group_kernel(int nPad, SelfColArgs a, const float3* __restrict__ X,
                                                           const float4* __restrict__ P,
                                                           const float3* __restrict__ rest,
                                                           const int* __restrict__ sortedIdx)
{
    constexpr int NPW = 64 / G;
    __shared__ __align__(16) unsigned int gkey[4 * NPW][kGroupStride];
    const int lane = threadIdx.x & 63;
    const int sub = lane % G, grp = lane / G;
    const int wave = (xcd_block() * (int)blockDim.x + (int)threadIdx.x) >> 6;
    const int i0 = wave * NPW;
    if (i0 >= nPad) return;   // whole waves
    const int b = __builtin_amdgcn_readfirstlane(a.tileBody[i0 / kTile]);
    const BodyDev& B = a.bodies[b];
    if (!B.scOn || !B.scFine) return;
    const int i = i0 + grp;
    const bool node = i < B.nodeEnd;
    unsigned int* __restrict__ list = gkey[(threadIdx.x >> 6) * NPW + grp];
    int n = 0;
    bool over = false, scanned = false;
    float3 x = make_float3(0.0f, 0.0f, 0.0f);
    float4 p4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (node) {
        x = X[i];
        p4 = P[i];
        scanned = a.scActive[i] && isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(p4.x) &&
                  isfinite(p4.y) && isfinite(p4.z);
    }
    if (scanned) {
        const float3 q = rest[i];
        const float search = B.scSearch, search2 = search * search;
        const float excl2 = B.scExcl2;
        const float inv = a.scInvCell[b];
        const unsigned int mask = (unsigned int)B.scTableMask;
        const int cx = sc_cell(x.x, inv), cy = sc_cell(x.y, inv), cz = sc_cell(x.z, inv);
        // two stencil offsets whose coarse cells share a bucket: the bucket is
        // scanned at the first of them (flag set by the fine-keys pass)
        const bool anyDup = (a.scActive[i] & 2) != 0;
        const float ginv = B.scFineInv;
        const unsigned int fmask = (unsigned int)B.scFineMask;
        const int* __restrict__ fs = a.fcellStart + B.scFineBase;
        const int* __restrict__ fe = a.fcellEnd + B.scFineBase;
        const int gx = sc_cell(p4.x, ginv), gy = sc_cell(p4.y, ginv), gz = sc_cell(p4.z, ginv);
        // middle-cell key of row r = (oy + 1) + 3 (oz + 1); its cells are
        // keys kr - 1, kr, kr + 1, adjacent in the sorted order unless kr is
        // at either end of the table
        // (stretches that wrap or overlap: flagged by the fine-keys pass)
        unsigned int kr[9];
        const bool perCell = (a.scActive[i] & 4) != 0;
#pragma unroll
        for (int r = 0; r < 9; ++r) kr[r] = sc_fine_key(gx, gy + r % 3 - 1, gz + r / 3 - 1, fmask);
        // the nine stretches as one run: run entry v sits at v + rowD[r] for
        // the last row r with rowP[r] <= v (empty rows are overridden)
        int rowP[9], rowD[9];
        int total = 0;
        if (!perCell) {
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                const unsigned int k = kr[r];
                const int s0 = fs[k - 1], s1 = fs[k], s2 = fs[k + 1];
                const int e0 = fe[k - 1], e1 = fe[k], e2 = fe[k + 1];
                int lo = s0 >= 0 ? s0 : (s1 >= 0 ? s1 : s2);
                int hi = s2 >= 0 ? e2 : (s1 >= 0 ? e1 : (s0 >= 0 ? e0 : -1));
                if (lo < 0) { lo = 0; hi = 0; }
                rowP[r] = total;
                rowD[r] = lo - total;
                total += hi - lo;
            }
        }
        over = perCell || total > kGroupScan;
        // run entry v -> sorted position
        auto runPos = [&](int v) {
            int d = rowD[0];
#pragma unroll
            for (int r = 1; r < 9; ++r)
                if (v >= rowP[r]) d = rowD[r];
            return v + d;
        };
        // scan-order key of an entry that passed the distance and rest tests,
        // or false when its bucket is not one of the stencil's
        auto keyOf = [&](int t, const float4& qj, unsigned int& key) {
            const int4 cj = a.fineC[t];   // only for accepted entries
            const int dx = cj.x - cx, dy = cj.y - cy, dz = cj.z - cz;
            int off = -1;
            if (((unsigned int)(dx + 1) <= 2u) & ((unsigned int)(dy + 1) <= 2u) & ((unsigned int)(dz + 1) <= 2u)) {
                off = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
                if (anyDup) {
                    const unsigned int hc = sc_hash(cx + dx, cy + dy, cz + dz, mask);
                    for (int c2 = 0; c2 < off; ++c2)
                        if (sc_hash(cx + c2 % 3 - 1, cy + (c2 / 3) % 3 - 1, cz + c2 / 9 - 1, mask) == hc) {
                            off = c2;
                            break;
                        }
                }
            } else {
                // outside the 27 cells (capped or fixed cell): a candidate
                // only through a bucket shared with a stencil cell
                const unsigned int ck = (unsigned int)cj.w - (unsigned int)B.scTableBase;
                for (int c2 = 0; c2 < 27; ++c2)
                    if (sc_hash(cx + c2 % 3 - 1, cy + (c2 / 3) % 3 - 1, cz + c2 / 9 - 1, mask) == ck) {
                        off = c2;
                        break;
                    }
            }
            key = ((unsigned int)(off < 0 ? 0 : off) << 27) | (unsigned int)__float_as_int(qj.w);
            return off >= 0;
        };
        auto passes = [&](const float4& pj, const float4& qj) {
            const float ex = p4.x - pj.x, ey = p4.y - pj.y, ez = p4.z - pj.z;
            const float d2 = ex * ex + ey * ey + ez * ez;
            const float qx = q.x - qj.x, qy = q.y - qj.y, qz = q.z - qj.z;
            const float q2 = qx * qx + qy * qy + qz * qz;
            // bitwise: no branch for the compiler to sink the loads into
            return (d2 < search2) & (__float_as_int(pj.w) != i) & (q2 >= excl2);
        };
        // append in lane order (the order is restored by the ranking)
        const unsigned long long gmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
        auto append = [&](bool acc, unsigned int key) {
            const unsigned long long gm = (__ballot(acc) >> (grp * G)) & gmask;
            const int slot = n + __popcll(gm & ((1ull << sub) - 1ull));
            if (acc && slot < kGroupList) list[slot] = key;
            n += __popcll(gm);
        };
        // 2 G entries per step, both loads of both issued together
        for (int base = 0; base < total && !over; base += 2 * G) {
            const int v0 = base + sub, v1 = v0 + G;
            const bool in0 = v0 < total, in1 = v1 < total;
            const int t0 = runPos(in0 ? v0 : 0), t1 = runPos(in1 ? v1 : 0);
            const float4 pa = a.fineP[t0], ra = a.fineR[t0];
            const float4 pb = a.fineP[t1], rb = a.fineR[t1];
            unsigned int k0 = 0u, k1 = 0u;
            bool acc0 = in0 & passes(pa, ra);
            bool acc1 = in1 & passes(pb, rb);
            if (acc0) acc0 = keyOf(t0, ra, k0);
            if (acc1) acc1 = keyOf(t1, rb, k1);
            append(acc0, k0);
            append(acc1, k1);
            over = n > kGroupList;
        }
    }
    // pad the node's keys to whole uint4 reads with keys above all
    const int n4 = (n + 3) & ~3;
    if (scanned && !over && sub < n4 - n) list[n + sub] = 0xffffffffu;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    // the contact push of object shape matching needs to know whether the body
    // has any candidate (every lane of the wave is active here)
    const bool anyCand = __ballot(node && !over && n > 0) != 0ull;
    if (lane == 0 && anyCand && B.objOn) a.scCounters[3 * b] = 1;
    // pool space for the lists longer than 32: one reservation per wave, each
    // group's share from a prefix sum of the groups' lengths
    const bool pool = node && !over && n > kSelfColCap;
    const int need = (pool && sub == 0) ? n : 0;
    int incl = need;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const int o = __shfl_up(incl, m, 64);
        if (lane >= m) incl += o;
    }
    const int waveNeed = __shfl(incl, 63, 64);
    int off = 0;
    if (waveNeed > 0) {
        int base = 0;
        if (lane == 0) base = atomicAdd(a.accUsed, waveNeed);
        base = __shfl(base, 0, 64);
        off = __shfl(base + incl - need, grp * G, 64);   // the group leader's exclusive prefix
    }
    if (!node) return;
    if (over) {
        if (sub == 0) a.accCount[i] = kScNeedWave;
        return;
    }
    if (pool && (long long)off + n > a.accCap) {
        if (sub == 0) a.accCount[i] = kScNeedCoarse;   // the pool is full
        return;
    }
    // rank of each key (all distinct): its place in the scan order
    for (int e = sub; e < n; e += G) {
        const unsigned int ke = list[e];
        int rank = 0;
        for (int m = 0; m < n4; m += 4) {
            const uint4 km = *reinterpret_cast<const uint4*>(&list[m]);
            rank += (km.x < ke ? 1 : 0) + (km.y < ke ? 1 : 0) + (km.z < ke ? 1 : 0) + (km.w < ke ? 1 : 0);
        }
        const int j = sortedIdx[ke & 0x07ffffffu];
        if (!pool) {
            a.contacts[sc_slot(i, rank)] = j;
        } else {
            // the slot kernel's priority: the distance the test used
            const float4 pj = P[j];
            const float ex = p4.x - pj.x, ey = p4.y - pj.y, ez = p4.z - pj.z;
            const float d2 = ex * ex + ey * ey + ez * ez;
            a.accList[(size_t)off + rank] = make_int2(j, __float_as_int(d2));
        }
    }
    if (sub == 0) {
        if (pool) {
            a.accOff[i] = off;
            a.accCount[i] = n;
        } else {
            a.counts[i] = n;
            a.overflow[i] = 0;
            a.accCount[i] = kScDone;
        }
    }
}

// Fine scan, one wave64 per node at a time (NPW consecutive nodes per wave),
// for bodies whose nodes all belong to one object with self-contact inside it
// on, walking only the nodes the group fine scan handed over.  There the only
// test is |p_i - p_j| < search, so every candidate lies in
// the 27 cells of edge >= search around p_i in a second grid built on the
// predicted positions.  Those few cells (nine rows of three cells, each row
// one stretch of the sorted order) are tested instead of the coarse stencil;
// a passing j is a reference candidate when its coarse bucket is one of the
// stencil's buckets, and it takes the scan position (first stencil offset of
// that bucket, sorted position inside it), read off its coarse cell when that
// is one of the 27 around x_i.  The accepted ones are ranked by that position
// (a 32-bit key) in LDS and written in that order to the list pool; the slot
// kernel fills the 32 slots from there, so the lists and counters equal the
// coarse scan's.  A node with more than kFineList accepted candidates, or
// that finds the pool full, is handed to the coarse kernel.
template <int NPW>
__global__ void __launch_bounds__(256) sc_fine_kernel(int nPad, SelfColArgs a, const float3* __restrict__ X,
                                                     const float4* __restrict__ P, const float3* __restrict__ rest,
                                                     int xcd)
{
    __shared__ __align__(16) unsigned int lkey[4][kFineList];   // accepted: scan-order key
    __shared__ int2 lpay[4][kFineList];                           // accepted: node id, squared distance
    const int wave = (sc_block(xcd) * (int)blockDim.x + (int)threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int wl = threadIdx.x >> 6;
    const int i0 = wave * NPW;
    if (i0 >= nPad) return;   // whole waves
    const int b = __builtin_amdgcn_readfirstlane(a.tileBody[i0 / kTile]);
    const BodyDev& B = a.bodies[b];
    if (!B.scOn || !B.scFine) return;
    if (!a.fineWaveAll &&
        __ballot(lane < NPW && i0 + lane < B.nodeEnd && a.accCount[i0 + lane] == kScNeedWave) == 0ull)
        return;
    int chunkPos = 0, chunkEnd = 0;   // the wave's share of the list pool
    bool anyDirect = false;           // a list written to the slots here
    for (int ii = 0; ii < NPW; ++ii) {
        const int i = i0 + ii;
        if (i >= B.nodeEnd) break;
        if (!a.fineWaveAll && a.accCount[i] != kScNeedWave) continue;
        const float3 x = X[i];
        const float4 p4 = P[i];
        const int act = a.scActive[i];
        const float3 q = rest[i];
        int deferred = 0;   // list length in the pool, or kScNeedCoarse
        if (act && isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(p4.x) && isfinite(p4.y) &&
            isfinite(p4.z)) {
            const float search = B.scSearch, search2 = search * search;
            const float inv = a.scInvCell[b];
            const int cx = sc_cell(x.x, inv), cy = sc_cell(x.y, inv), cz = sc_cell(x.z, inv);
            // lane c < 27: stencil offset c and its (global) coarse bucket
            const int c = lane < 27 ? lane : 26;
            const int ox = c % 3 - 1, oy = (c / 3) % 3 - 1, oz = c / 9 - 1;
            const unsigned int hg =
                (unsigned int)B.scTableBase + sc_hash(cx + ox, cy + oy, cz + oz, (unsigned int)B.scTableMask);
            const float ginv = B.scFineInv;
            const unsigned int fmask = (unsigned int)B.scFineMask;
            const int fbase = B.scFineBase;
            const float excl2 = B.scExcl2;
            const int gx = sc_cell(p4.x, ginv), gy = sc_cell(p4.y, ginv), gz = sc_cell(p4.z, ginv);
            // lane c < 27: fine cell (gx + ox, gy + oy, gz + oz) and its range
            const unsigned int fk = sc_fine_key(gx + ox, gy + oy, gz + oz, fmask);
            // start and end fetched together (an end is only used with its
            // start, which marks the cells that are occupied this build)
            int cs = -1, ce = 0;
            if (lane < 27) {
                cs = a.fcellStart[fbase + fk];
                ce = a.fcellEnd[fbase + fk];
                if (cs < 0) ce = 0;
            }
            // first stencil offset holding each coarse bucket (two offsets
            // share one when their cells collide in the hash)
            // (flagged per node by the fine-keys pass; rare)
            const bool dup = (act & 2) != 0;
            int firstC = c;
            if (dup)
                for (int c2 = 26; c2 >= 0; --c2)
                    if ((unsigned int)lane_i((int)hg, c2) == hg) firstC = c2;
            // Row r = c / 3 (one (oy, oz)) covers keys k_r - 1 .. k_r + 1 (k_r
            // the key of its middle cell), one contiguous stretch of the
            // sorted order.  Rows whose stretches wrap around the table or
            // come within two keys of each other would visit a key twice:
            // then each distinct cell key is walked once instead (flagged by
            // the fine-keys pass).
            const bool perCell = (act & 4) != 0;
            // row r: first run position rowP (past every run position when
            // the row is empty) and sorted index minus run position rowD
            int rowP[9], rowD[9];
            int rs = 0, pre = 0, total = 0;
            if (!perCell) {
                // the non-empty cells of a row are adjacent in the sorted order
                const int lo = cs >= 0 ? cs : 0x7fffffff, hi = cs >= 0 ? ce : 0;
                const int lo1 = __shfl_down(lo, 1, 64), lo2 = __shfl_down(lo, 2, 64);
                const int hi1 = __shfl_down(hi, 1, 64), hi2 = __shfl_down(hi, 2, 64);
                const int rlo = min(lo, min(lo1, lo2)), rhi = max(hi, max(hi1, hi2));
#pragma unroll
                for (int r = 0; r < 9; ++r) {
                    const int l = lane_i(rlo, 3 * r), h2 = lane_i(rhi, 3 * r);
                    const int n = h2 > l ? h2 - l : 0;
                    rowP[r] = n > 0 ? total : 0x7fffffff;
                    rowD[r] = l - total;
                    total += n;
                }
            } else {
                bool ffirst = lane < 27;
                for (int c2 = 0; c2 < 27; ++c2)
                    if ((unsigned int)lane_i((int)fk, c2) == fk && c2 < c) ffirst = false;
                int rn = 0;
                if (ffirst && cs >= 0) {
                    rs = cs;
                    rn = ce - cs;
                }
                total = sc_run_prefix(rn, lane, pre);
            }
            int nacc = 0;
            auto runPos = [&](int v) {
                if (perCell) return sc_run_pos(v, rs, pre);
                int d = 0;
#pragma unroll
                for (int r = 0; r < 9; ++r)
                    if (v >= rowP[r]) d = rowD[r];
                return v + d;
            };
            // test one run entry per lane and append the accepted ones in
            // lane order
            auto scanEntry = [&](int v, const float4& pj, const float4& qj, const int4& cj) {
                bool acc = false, inside = true;
                unsigned int key = 0u, ck = 0u;
                int j = 0, rk = 0;
                float d2 = 0.0f;
                if (v < total) {
                    j = __float_as_int(pj.w);
                    ck = (unsigned int)cj.w;
                    rk = __float_as_int(qj.w);
                    const float ex = p4.x - pj.x, ey = p4.y - pj.y, ez = p4.z - pj.z;
                    d2 = ex * ex + ey * ey + ez * ez;
                    const float qx = q.x - qj.x, qy = q.y - qj.y, qz = q.z - qj.z;
                    const float q2 = qx * qx + qy * qy + qz * qz;
                    acc = (j != i) & (d2 < search2) & (q2 >= excl2);
                    // scan position: a candidate whose coarse cell is one of
                    // the 27 around x_i sits in that offset's bucket
                    const int dx = cj.x - cx, dy = cj.y - cy, dz = cj.z - cz;
                    inside = ((unsigned int)(dx + 1) <= 2u) & ((unsigned int)(dy + 1) <= 2u) &
                             ((unsigned int)(dz + 1) <= 2u);
                    key = ((unsigned int)((dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)) << 27) | (unsigned int)rk;
                }
                if (dup) {
                    const int f = __shfl(firstC, inside ? (int)(key >> 27) : 0, 64);
                    key = ((unsigned int)f << 27) | (unsigned int)rk;
                }
                // any other candidate is one only when its bucket collides
                // with a stencil bucket: lanes 0-26 hold those, so each is
                // matched against all of them at once
                unsigned long long om = __ballot(acc && !inside);
                while (om) {
                    const int l = __builtin_ctzll(om);
                    om &= om - 1ull;
                    const unsigned int ckl = (unsigned int)lane_i((int)ck, l);
                    const unsigned long long hit = __ballot(lane < 27 && hg == ckl);
                    if (lane == l) {
                        if (hit) key = ((unsigned int)__builtin_ctzll(hit) << 27) | (unsigned int)rk;
                        else acc = false;
                    }
                }
                const unsigned long long am = __ballot(acc);
                const int slot = nacc + (int)__builtin_amdgcn_mbcnt_hi((unsigned int)(am >> 32),
                                                                      __builtin_amdgcn_mbcnt_lo((unsigned int)am, 0u));
                if (acc && slot < kFineList) {
                    lkey[wl][slot] = key;
                    lpay[wl][slot] = make_int2(j, __float_as_int(d2));
                }
                nacc += __popcll(am);
            };
            for (int base = 0; base < total; base += 64) {
                const int v = base + lane;
                const int t = runPos(v);   // every lane: the per-cell walk shuffles across the wave
                float4 pj = make_float4(0.0f, 0.0f, 0.0f, 0.0f), qj = pj;
                int4 cj = make_int4(0, 0, 0, 0);
                if (v < total) {
                    pj = a.fineP[t];
                    qj = a.fineR[t];
                    cj = a.fineC[t];
                }
                scanEntry(v, pj, qj, cj);
            }
            deferred = kScNeedCoarse;   // unless the ranked list is written below
            if (nacc <= kFineList) {
                // pad the key list to whole uint4 reads with keys above all
                const int n4 = (nacc + 3) & ~3;
                if (lane < n4 - nacc) lkey[wl][nacc + lane] = 0xffffffffu;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                // rank of each accepted entry by its scan-order key (first
                // offset, coarse sorted position: all distinct)
                constexpr int R = kFineList / 64;
                unsigned int ek[R];
                int2 ep[R];
                int er[R];
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int s = k * 64 + lane;
                    ek[k] = s < nacc ? lkey[wl][s] : 0xffffffffu;
                    ep[k] = s < nacc ? lpay[wl][s] : make_int2(0, 0);
                    er[k] = 0;
                }
                for (int m = 0; m < n4; m += 4) {
                    const uint4 km = *reinterpret_cast<const uint4*>(&lkey[wl][m]);
#pragma unroll
                    for (int k = 0; k < R; ++k)
                        er[k] += (km.x < ek[k] ? 1 : 0) + (km.y < ek[k] ? 1 : 0) + (km.z < ek[k] ? 1 : 0) +
                                 (km.w < ek[k] ? 1 : 0);
                }
                if (nacc <= kSelfColCap) {
                    // the whole list fits: the ranks are the slots
#pragma unroll
                    for (int k = 0; k < R; ++k)
                        if (k * 64 + lane < nacc) a.contacts[sc_slot(i, er[k])] = ep[k].x;
                    if (lane == 0) {
                        a.counts[i] = nacc;
                        a.overflow[i] = 0;
                    }
                    anyDirect |= nacc > 0;
                    deferred = kScDone;
                } else {
                    // the sorted list goes to the slot kernel (one thread per
                    // node) when the shared list pool has room; the wave takes
                    // pool space kAccChunk entries at a time
                    if (chunkPos + nacc > chunkEnd) {
                        int off = 0;
                        if (lane == 0) off = atomicAdd(a.accUsed, kAccChunk);
                        chunkPos = __builtin_amdgcn_readfirstlane(off);
                        chunkEnd = chunkPos + kAccChunk;
                    }
                    const int off = chunkPos;
                    if ((long long)chunkEnd <= a.accCap) {
                        chunkPos += nacc;
#pragma unroll
                        for (int k = 0; k < R; ++k)
                            if (k * 64 + lane < nacc) a.accList[(size_t)off + er[k]] = ep[k];
                        if (lane == 0) a.accOff[i] = off;
                        deferred = nacc;
                    }
                }
                __builtin_amdgcn_wave_barrier();   // the next node reuses the wave's LDS rows
            }
        }
        if (lane == 0) a.accCount[i] = deferred;
    }
    // the contact push of object shape matching needs to know whether the body
    // has any candidate (the slot kernel marks the pooled lists)
    if (lane == 0 && anyDirect && B.objOn) a.scCounters[3 * b] = 1;
}

// Candidate lists (forward.cu:7079-7180), coarse scan (the reference's own
// walk), one wave64 per node at a time; each wave takes NPW consecutive nodes
// (a divisor of the 64-node tile, so a wave never spans two bodies).  The 27
// stencil cells of the frame-start position, each hash bucket once at its
// first stencil offset, in sorted order, are tested 64 candidates at a time;
// accepted candidates enter the slots in lane order, which is the scan order.
// Stencil cells that cannot hold a candidate are skipped: a node j passes
// either test only if |x_i - x_j| < search + |p_i - x_i| + max |p - x|, so a
// cell farther than that bound (with a relative margin for rounding) holds
// none; a bucket is still scanned at its first offset whenever any offset
// mapping to it is needed.  In fine-scan bodies only the nodes the fine
// kernel handed over are walked here.
template <int NPW>
__global__ void __launch_bounds__(256) sc_build_kernel(int nPad, SelfColArgs a, const float3* __restrict__ X,
                                                      const float4* __restrict__ P, const float3* __restrict__ rest,
                                                      int xcd)
{
    const int wave = (sc_block(xcd) * (int)blockDim.x + (int)threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int i0 = wave * NPW;
    if (i0 >= nPad) return;   // whole waves
    const int b = __builtin_amdgcn_readfirstlane(a.tileBody[i0 / kTile]);
    const BodyDev& B = a.bodies[b];
    if (!B.scOn) return;
    // fine-scan bodies: only the nodes the fine kernel handed over
    const bool fine = B.scFine != 0;
    const bool fineBody = fine && a.sortedIdx != nullptr && a.coarseRank != nullptr;
    if (fine && __ballot(lane < NPW && i0 + lane < B.nodeEnd && a.accCount[i0 + lane] == kScNeedCoarse) == 0ull)
        return;
    bool anyCand = false;
    for (int ii = 0; ii < NPW; ++ii) {
        const int i = i0 + ii;
        if (i >= B.nodeEnd) break;
        if (fine && a.accCount[i] != kScNeedCoarse) continue;
        const float3 x = X[i];
        const float4 p4 = P[i];
        ScSlots S;
        if (a.scActive[i] && isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(p4.x) && isfinite(p4.y) &&
            isfinite(p4.z)) {
            const float3 q = rest[i];
            const float dix = p4.x - x.x, diy = p4.y - x.y, diz = p4.z - x.z;
            const int oi = a.objId[i];
            const float search = B.scSearch, search2 = search * search;
            const float inv = a.scInvCell[b];
            const unsigned int mask = (unsigned int)B.scTableMask;
            const int cx = sc_cell(x.x, inv), cy = sc_cell(x.y, inv), cz = sc_cell(x.z, inv);
            // lane c < 27: stencil offset c and its (global) coarse bucket
            const int c = lane < 27 ? lane : 26;
            const int ox = c % 3 - 1, oy = (c / 3) % 3 - 1, oz = c / 9 - 1;
            const unsigned int h = sc_hash(cx + ox, cy + oy, cz + oz, mask);
            const unsigned int hg = (unsigned int)B.scTableBase + h;
            {
                const float edge = 1.0f / inv;
                const float tx = fminf(fmaxf(x.x * inv, -1.0e9f), 1.0e9f), ty = fminf(fmaxf(x.y * inv, -1.0e9f), 1.0e9f),
                            tz = fminf(fmaxf(x.z * inv, -1.0e9f), 1.0e9f);
                const float fx = tx - floorf(tx), fy = ty - floorf(ty), fz = tz - floorf(tz);
                const float di = sqrtf(dix * dix + diy * diy + diz * diz);
                const float maxMove = sqrtf(fmaxf(__uint_as_float(a.scMax2[b]), 0.0f));
                const float bound = (search + di + maxMove) * 1.0001f + edge * 1.0e-4f + 1.0e-6f;
                const float bound2 = bound * bound;
                const float gx = sc_axis_gap(ox, fx, edge), gy = sc_axis_gap(oy, fy, edge), gz = sc_axis_gap(oz, fz, edge);
                // nodes outside the 27 cells are at least one edge away; when
                // the edge is shorter than the bound (a capped or fixed cell)
                // such a node could pass through a colliding bucket, so no
                // cell is skipped
                const bool cull = edge >= bound;
                const bool need = lane < 27 && (!cull || gx * gx + gy * gy + gz * gz <= bound2);
                bool first = lane < 27, scan = false;
                for (int c2 = 0; c2 < 27; ++c2) {
                    const unsigned int h2 = (unsigned int)lane_i((int)h, c2);
                    const bool need2 = (__ballot(need) >> c2) & 1ull;
                    if (h2 == h && c2 < c) first = false;
                    if (h2 == h && c2 >= c && need2) scan = true;
                }
                int rs = 0, rn = 0;
                if (first && scan) {
                    rs = a.cellStart[hg];
                    rn = rs >= 0 ? a.cellEnd[hg] - rs : 0;
                }
                int pre;
                const int total = sc_run_prefix(rn, lane, pre);
                for (int base = 0; base < total; base += 64) {
                    const int v = base + lane;
                    const int t = sc_run_pos(v, rs, pre);
                    bool acc = false;
                    int code = 0, tr = 0;
                    float pr = 0.0f;
                    if (v < total) {
                        float4 pj, qj;
                        int jp = -1;   // fine-scan body: the candidate's node
                        if (fineBody) {
                            // not staged (see sc_stage_kernel): through the permutation
                            jp = a.sortedIdx[t];
                            const float4 pp = P[jp];
                            const float3 rr = rest[jp];
                            pj = make_float4(pp.x, pp.y, pp.z, __int_as_float(a.objId[jp]));
                            qj = make_float4(rr.x, rr.y, rr.z, __int_as_float(jp));
                        } else {
                            pj = a.sortP[t];
                            qj = a.sortR[t];   // issued with pj: no dependent second trip
                        }
                        const bool cross = __float_as_int(pj.w) != oi;
                        if (cross || B.scWithin) {
                            float d2 = 0.0f, distX = 0.0f, rel = 0.0f;
                            bool pass;
                            if (cross) {
                                float4 xj;
                                if (jp >= 0) {
                                    const float3 xx = X[jp];
                                    xj = make_float4(xx.x, xx.y, xx.z, 0.0f);
                                } else {
                                    xj = a.sortX[t];
                                }
                                const float ex = x.x - xj.x, ey = x.y - xj.y, ez = x.z - xj.z;
                                distX = sqrtf(ex * ex + ey * ey + ez * ez);
                                const float rx = dix - (pj.x - xj.x), ry = diy - (pj.y - xj.y),
                                            rz = diz - (pj.z - xj.z);
                                rel = sqrtf(rx * rx + ry * ry + rz * rz);
                                pass = distX < search + rel;
                            } else {
                                const float ex = p4.x - pj.x, ey = p4.y - pj.y, ez = p4.z - pj.z;
                                d2 = ex * ex + ey * ey + ez * ez;
                                pass = d2 < search2;
                            }
                            if (pass) {
                                const int j = __float_as_int(qj.w);
                                const float qx = q.x - qj.x, qy = q.y - qj.y, qz = q.z - qj.z;
                                const float q2 = qx * qx + qy * qy + qz * qz;
                                if (j != i) {
                                    if (cross) {
                                        acc = q2 >= B.scDc * B.scDc;
                                        pr = distX - rel;
                                        code = (rel > B.scFastRel) ? -(j + 1) : j;
                                    } else {
                                        acc = q2 >= B.scExcl2;
                                        pr = sqrtf(d2);
                                        code = j;
                                    }
                                    tr = cross ? 0 : 1;
                                }
                            }
                        }
                    }
                    unsigned long long am = __ballot(acc);
                    while (am) {
                        const int l = __builtin_ctzll(am);
                        am &= am - 1ull;
                        S.insert(lane_i(code, l), lane_i(tr, l), lane_fl(pr, l), lane);
                    }
                }
            }
        }
        if (lane < S.cnt) a.contacts[sc_slot(i, lane)] = S.code;
        if (lane == 0) {
            a.counts[i] = S.cnt;
            a.overflow[i] = S.overflow;
        }
        anyCand |= S.cnt > 0;
    }
    // the contact push of object shape matching needs to know whether the body
    // has any candidate; the totals are summed only when asked for
    if (lane == 0 && anyCand && B.objOn) a.scCounters[3 * b] = 1;
}

// Slots of the nodes whose sorted candidate list the fine scan deferred, one
// thread per node: the first 32 fill the slots in order, each later one
// replaces the worst slot (largest priority, lowest slot among equals) when
// strictly better, as the wave insertion does (all fine candidates are
// same-object ones, one tier).
__global__ void __launch_bounds__(256) sc_slots_kernel(int nPad, SelfColArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;   // nPad is a multiple of 64: whole waves
    const int b = __builtin_amdgcn_readfirstlane(a.tileBody[i / kTile]);
    const BodyDev& B = a.bodies[b];
    if (!B.scOn || !B.scFine) return;
    const int n = i < B.nodeEnd ? a.accCount[i] : -1;
    const int2* list = a.accList + (n > 0 ? a.accOff[i] : 0);
    constexpr int CAP = kSelfColCap;
    const int cnt = n < CAP ? n : CAP;
    if (n >= 0 && n <= CAP) {
        // the whole list fits: it is the slot order
        for (int r = 0; r < n; ++r) a.contacts[sc_slot(i, r)] = list[r].x;
    } else if (n > CAP) {
        int code[CAP];
        float prio[CAP];
#pragma unroll
        for (int s = 0; s < CAP; ++s) {
            const int2 e = list[s];
            code[s] = e.x;
            prio[s] = sqrtf(__int_as_float(e.y));   // the list holds squared distances
        }
        // worst slot, kept until a slot changes
        int w = 0;
        float wp = prio[0];
#pragma unroll
        for (int s = 1; s < CAP; ++s)
            if (prio[s] > wp) { wp = prio[s]; w = s; }
        for (int r = CAP; r < n; ++r) {
            const int2 e = list[r];
            const float p = sqrtf(__int_as_float(e.y));
            if (p < wp) {
#pragma unroll
                for (int s = 0; s < CAP; ++s)
                    if (s == w) { code[s] = e.x; prio[s] = p; }
                w = 0;
                wp = prio[0];
#pragma unroll
                for (int s = 1; s < CAP; ++s)
                    if (prio[s] > wp) { wp = prio[s]; w = s; }
            }
        }
#pragma unroll
        for (int s = 0; s < CAP; ++s) a.contacts[sc_slot(i, s)] = code[s];
    }
    if (n >= 0) {
        a.counts[i] = cnt;
        a.overflow[i] = n - cnt;
    }
    const bool any = __ballot(n > 0 && cnt > 0) != 0ull;
    if ((threadIdx.x & 63) == 0 && any && B.objOn) a.scCounters[3 * b] = 1;
}

// Statistics of one body's last build (forward.cu:11837-11849 reports them):
// candidates (sum of list lengths), overflow (accepted beyond the 32 slots)
// and active nodes, summed over [nodeBegin, nodeEnd) into out[3].
__global__ void __launch_bounds__(256) sc_stats_kernel(int nodeBegin, int nodeEnd, SelfColArgs a, int* out)
{
    const int i = nodeBegin + blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0, o = 0, act = 0;
    if (i < nodeEnd) {
        c = a.counts[i];
        o = a.overflow[i];
        act = a.scActive[i] ? 1 : 0;
    }
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
        c += __shfl_xor(c, s, 64);
        o += __shfl_xor(o, s, 64);
        act += __shfl_xor(act, s, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (c) atomicAdd(&out[0], c);
        if (o) atomicAdd(&out[1], o);
        if (act) atomicAdd(&out[2], act);
    }
}

void launch_selfcol_stats(hipStream_t s, int nodeBegin, int nodeEnd, const SelfColArgs& a, int* out)
{
    launch_fill_int(s, 3, out, 0);
    const int n = nodeEnd - nodeBegin;
    if (n > 0) sc_stats_kernel<<<(n + 255) / 256, 256, 0, s>>>(nodeBegin, nodeEnd, a, out);
}

__global__ void sc_fill_accum_kernel(int nPad, float3* accum)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nPad) accum[i] = make_float3(0.0f, 0.0f, 0.0f);
}

// ---------------------------------------------------------------------------
// Self-collision per iteration
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool sc_start_normal(float3 xi, float3 xj, float3 qi, float3 qj, float& nx, float& ny, float& nz)
{
    nx = xi.x - xj.x; ny = xi.y - xj.y; nz = xi.z - xj.z;
    float l2 = nx * nx + ny * ny + nz * nz;
    if (!(l2 > 1e-24f)) {
        nx = qi.x - qj.x; ny = qi.y - qj.y; nz = qi.z - qj.z;
        l2 = nx * nx + ny * ny + nz * nz;
        if (!(l2 > 1e-24f)) return false;
    }
    const float inv = 1.0f / sqrtf(l2);
    nx *= inv; ny *= inv; nz *= inv;
    return true;
}

// Contacts of one node fetched in groups (several gathers in flight per lane).
constexpr int kScSolveGroup = 8;

// Contact solve and apply of one iteration (forward.cu:7205-7307 read pass,
// 7311-7322 write pass), one thread per node: the node's push from its
// contact list is computed from the input positions (partners included) and
// added to its own output position, so the two passes of the reference fuse
// into one (input and output are different buffers); the iteration epilogue
// follows.
__device__ __forceinline__ void sc_node_push(int i, const SelfColArgs& a, const BodyDev& B,
                                             const float4* __restrict__ P, const float3* __restrict__ X,
                                             const float3* __restrict__ rest, float3& dpOut, float3& crossOut)
{
    const int cnt = a.counts[i];
    constexpr int CAP = kSelfColCap;
    const float4 pi4 = P[i];
    const float wi = pi4.w * B.invMassScale;
    const float dc = B.scDc;
    float ax = 0.0f, ay = 0.0f, az = 0.0f, cpx = 0.0f, cpy = 0.0f, cpz = 0.0f;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f, wFast = 0.0f;
    int act = 0, nFast = 0;
    const int* list = &a.contacts[sc_slot(i, 0)];   // slot k at list[k * kTile]
    const float3 xi = X[i], qi = rest[i];
    if (wi > 0.0f) {
        const float dc2 = dc * dc;
        // contacts in groups of kScSolveGroup: the group's codes and partner
        // positions are fetched back to back (several gathers in flight per
        // lane), then the group is applied in slot order
        for (int k0 = 0; k0 < cnt; k0 += kScSolveGroup) {
            int codes[kScSolveGroup];
            float4 pjs[kScSolveGroup];
#pragma unroll
            for (int u = 0; u < kScSolveGroup; ++u) codes[u] = (k0 + u < cnt) ? list[(k0 + u) * kTile] : -1;
#pragma unroll
            for (int u = 0; u < kScSolveGroup; ++u) {
                const int jj = (k0 + u < cnt) ? (codes[u] >= 0 ? codes[u] : -codes[u] - 1) : i;
                pjs[u] = P[jj];
            }
#pragma unroll
            for (int u = 0; u < kScSolveGroup; ++u) {
                if (k0 + u >= cnt) break;
                const int code = codes[u];
                const float4 pj = pjs[u];
                if (code < 0) {
                    const int j = -code - 1;
                    float nx, ny, nz;
                    if (!sc_start_normal(xi, X[j], qi, rest[j], nx, ny, nz)) continue;
                    const float proj = (pi4.x - pj.x) * nx + (pi4.y - pj.y) * ny + (pi4.z - pj.z) * nz;
                    if (!(proj < dc)) continue;
                    const float pen = dc - proj;
                    sx += nx * pen; sy += ny * pen; sz += nz * pen;
                    const float wj = pj.w * B.invMassScale;
                    wFast += wi / (wi + wj);
                    ++nFast;
                    continue;
                }
                const int j = code;
                const float dx = pi4.x - pj.x, dy = pi4.y - pj.y, dz = pi4.z - pj.z;
                const float d2 = dx * dx + dy * dy + dz * dz;
                if (!(d2 < dc2)) continue;
                const float wj = pj.w * B.invMassScale;
                const float wr = wi / (wi + wj);
                if (d2 < 1e-24f) {
                    const float3 qj = rest[j];
                    const float rx = qi.x - qj.x, ry = qi.y - qj.y, rz = qi.z - qj.z;
                    const float r2 = rx * rx + ry * ry + rz * rz;
                    if (!(r2 > 1e-24f)) continue;
                    const float s = dc * wr / sqrtf(r2);
                    ax += rx * s; ay += ry * s; az += rz * s;
                    ++act;
                    continue;
                }
                const float d = sqrtf(d2);
                const float s = (dc - d) / d * wr;
                ax += dx * s; ay += dy * s; az += dz * s;
                ++act;
            }
        }
    }
    float3 out = (act > 0) ? make_float3(ax / act, ay / act, az / act) : make_float3(0.0f, 0.0f, 0.0f);
    if (nFast > 0) {
        const float sl = sqrtf(sx * sx + sy * sy + sz * sz);
        if (sl > 1e-20f) {
            const float bx = sx / sl, by = sy / sl, bz = sz / sl;
            float delta = 0.0f;
            for (int k = 0; k < cnt; ++k) {
                const int code = list[k * kTile];
                if (code >= 0) continue;
                const int j = -code - 1;
                float nx, ny, nz;
                if (!sc_start_normal(xi, X[j], qi, rest[j], nx, ny, nz)) continue;
                const float4 pj = P[j];
                const float ddx = pi4.x - pj.x, ddy = pi4.y - pj.y, ddz = pi4.z - pj.z;
                if (!(ddx * nx + ddy * ny + ddz * nz < dc)) continue;
                const float pen = dc - (ddx * bx + ddy * by + ddz * bz);
                if (pen > delta) delta = pen;
            }
            if (delta > 0.0f) {
                const float sc = delta * (wFast / (float)nFast);
                cpx = bx * sc; cpy = by * sc; cpz = bz * sc;
                out = make_float3(out.x + cpx, out.y + cpy, out.z + cpz);
            }
        }
    }
    dpOut = out;
    crossOut = make_float3(cpx, cpy, cpz);
}

__global__ void __launch_bounds__(256) sc_contact_kernel(int nPad, SelfColArgs a, const float4* __restrict__ Pin,
                                                        float4* __restrict__ Pout, const float3* __restrict__ X,
                                                        const float3* __restrict__ rest, unsigned epi, EpiArgs ea)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const BodyDev& B = body_at(a, i);
    if (i >= B.nodeEnd) return;
    const float4 p4 = Pin[i];
    float3 p = make_float3(p4.x, p4.y, p4.z);
    if (B.scOn && a.counts[i] > 0) {
        float3 d, c;
        sc_node_push(i, a, B, Pin, X, rest, d, c);
        p = make_float3(p.x + d.x, p.y + d.y, p.z + d.z);
        const float3 acc = a.accum[i];
        a.accum[i] = make_float3(acc.x + c.x, acc.y + c.y, acc.z + c.z);
    }
    if (epi && epilogue_node(epi, i, p, p4.w, B, ea)) return;
    Pout[i] = make_float4(p.x, p.y, p.z, p4.w);
}

// ---------------------------------------------------------------------------
// Contact-average push after the iterations (forward.cu:11858-11890, 7326-7353)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) push_sum_kernel(int nPad, SelfColArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const int b = __builtin_amdgcn_readfirstlane(a.tileBody[i / kTile]);
    const BodyDev& B = a.bodies[b];
    if (!B.scOn || !B.objOn || i >= B.nodeEnd || a.scCounters[3 * b] <= 0) return;
    const int id = a.objId[i];
    if (id < 0) return;
    const float3 c = a.accum[i];
    if (!(c.x * c.x + c.y * c.y + c.z * c.z > 1e-30f)) return;
    atomicAdd(&a.pushSums[4 * (size_t)id], c.x);
    atomicAdd(&a.pushSums[4 * (size_t)id + 1], c.y);
    atomicAdd(&a.pushSums[4 * (size_t)id + 2], c.z);
    atomicAdd(&a.pushSums[4 * (size_t)id + 3], 1.0f);
}

__global__ void push_t_kernel(int K, SelfColArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const BodyDev& B = a.bodies[a.objBody[k]];
    float* T = &a.pushT[4 * (size_t)k];
    T[0] = T[1] = T[2] = T[3] = 0.0f;
    const float* S = &a.pushSums[4 * (size_t)k];
    const float n = S[3];
    if (!(n > 0.0f)) return;
    const float sc = B.objStiffness / n;
    const float tx = S[0] * sc, ty = S[1] * sc, tz = S[2] * sc;
    if (!isfinite(tx) || !isfinite(ty) || !isfinite(tz)) return;
    T[0] = tx; T[1] = ty; T[2] = tz; T[3] = 1.0f;
}

// Push untouched nodes of pushed objects, re-project on the ground (12588-12596),
// then the velocity commit.
__global__ void __launch_bounds__(256) push_apply_kernel(int nPad, SelfColArgs a, float4* __restrict__ P, EpiArgs ea)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const int b = __builtin_amdgcn_readfirstlane(a.tileBody[i / kTile]);
    const BodyDev& B = a.bodies[b];
    if (i >= B.nodeEnd) return;
    const float4 p4 = P[i];
    float3 p = make_float3(p4.x, p4.y, p4.z);
    const bool pushStep = B.scOn && B.objOn && a.scCounters[3 * b] > 0;
    if (pushStep) {
        const int id = a.objId[i];
        if (id >= 0 && p4.w > 0.0f) {
            const float* T = &a.pushT[4 * (size_t)id];
            const float3 c = a.accum[i];
            if (T[3] != 0.0f && !(c.x * c.x + c.y * c.y + c.z * c.z > 1e-30f))
                p = make_float3(p.x + T[0], p.y + T[1], p.z + T[2]);
        }
        if (B.groundOn && p4.w > 0.0f) {
            const float* rg = B.contactGround ? ea.rg : nullptr;
            const float c = B.gn[0] * p.x + B.gn[1] * p.y + B.gn[2] * p.z - B.planeD - (rg ? rg[i] : 0.0f);
            if (c < 0.0f) p = make_float3(p.x - c * B.gn[0], p.y - c * B.gn[1], p.z - c * B.gn[2]);
        }
    }
    commit_node(i, p, p4.w, B, ea, B.contactGround ? ea.rg : nullptr);
}

// ---------------------------------------------------------------------------
static inline int nbs(int n, int bs) { return (n + bs - 1) / bs; }

void launch_object_shape(hipStream_t s, const NodeArgs& na, const SelfColArgs& a, int K, float4* P, const float3* rest,
                         bool boxes)
{
    if (K <= 0) return;
    launch_fill_double(s, 16 * K, a.objSums);
    obj_accum_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, P);
    obj_solve_kernel<<<nbs(K, 64), 64, 0, s>>>(K, a);
    if (boxes) {
        obj_boxes_init_kernel<<<nbs(6 * K, 256), 256, 0, s>>>(K, a.objBoxesArr);
        obj_boxes_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, P);
    }
    obj_match_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, P, rest);
}

bool launch_selfcol_build(hipStream_t s, const NodeArgs& na, int numBodies, const SelfColArgs& a, const float3* X,
                          const float4* P, const float3* rest, int sortBits, unsigned int* keysAlt, int* idxAlt,
                          unsigned int* hist, const int** sortedIdxOut, const Tuning& tu,
                          void (*mark)(void*, const char*), void* markCtx)
{
    sc_prepare_kernel<<<nbs(numBodies, 64), 64, 0, s>>>(numBodies, a);
    sc_active_disp_kernel<<<nbs(na.nPad, kDispThreads), kDispThreads, 0, s>>>(na.nPad, a, X, P);
    sc_reach_kernel<<<nbs(numBodies, 64), 64, 0, s>>>(numBodies, a);
    sc_keys_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, X);
    const bool inAlt = launch_radix_sort(s, na.nPad, sortBits, a.keys, a.idx, keysAlt, idxAlt, hist, false);
    const unsigned int* ks = inAlt ? keysAlt : a.keys;
    const int* vs = inAlt ? idxAlt : a.idx;
    launch_fill_int(s, a.tableTotal, a.cellStart, -1);
    sc_range_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, ks);
    sc_stage_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, X, P, rest, vs);
    if (a.fineTotal > 0) {
        sc_fine_keys_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, X, P);
        const bool fAlt = launch_radix_sort(s, na.nPad, a.fineBits, a.fkeys, a.fidx, a.fkeysAlt, a.fidxAlt, hist, false);
        const unsigned int* fks = fAlt ? a.fkeysAlt : a.fkeys;
        const int* fvs = fAlt ? a.fidxAlt : a.fidx;
        launch_fill_int(s, a.fineTotal, a.fcellStart, -1);
        sc_fine_range_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, fks);
        sc_fine_stage_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, X, P, rest, fks, fvs, ks);
    }
    const int wpb = (tu.scBuildWaves == 1 || tu.scBuildWaves == 2) ? tu.scBuildWaves : 4;
    const int npw = (tu.scBuildNodes == 4 || tu.scBuildNodes == 8 || tu.scBuildNodes == 16 || tu.scBuildNodes == 64)
                        ? tu.scBuildNodes
                        : 32;
    const int blocks = nbs(nbs(na.nPad, npw), wpb);
    // the XCD remaps need whole rounds of 8 blocks (of 8 runs of xcd blocks);
    // the blocks past the node space return at once
    const int xcd = tu.scBuildXcd < 0 ? 0 : (tu.scBuildXcd > 256 ? 256 : tu.scBuildXcd);
    const int grid = xcd == 0 ? blocks : (xcd == 1 ? xcd_grid(blocks) : (blocks + 8 * xcd - 1) / (8 * xcd) * (8 * xcd));
#define XPBD_SC_LAUNCH_A(K, A)                                                                          \
    do {                                                                                                \
        if (npw == 4) K<4><<<grid, 64 * wpb, 0, s>>>(na.nPad, A, X, P, rest, xcd);                       \
        else if (npw == 8) K<8><<<grid, 64 * wpb, 0, s>>>(na.nPad, A, X, P, rest, xcd);                  \
        else if (npw == 16) K<16><<<grid, 64 * wpb, 0, s>>>(na.nPad, A, X, P, rest, xcd);                \
        else if (npw == 64) K<64><<<grid, 64 * wpb, 0, s>>>(na.nPad, A, X, P, rest, xcd);                \
        else K<32><<<grid, 64 * wpb, 0, s>>>(na.nPad, A, X, P, rest, xcd);                               \
    } while (0)
#define XPBD_SC_LAUNCH(K) XPBD_SC_LAUNCH_A(K, a)
    // fine scans first (thread, then wave for the nodes with more than 32
    // candidates): the coarse kernel walks the nodes they hand over
    if (a.fineTotal > 0) {
        if (mark) mark(markCtx, "selfcol_build_fine_group");
        // G lanes per node, 64 / G nodes per wave, 4 waves per block; with
        // G = 0 the wave scan takes every node
        const int g = tu.scGroupLanes;
        if (g == 8 || g == 16 || g == 32) {
            const int gridG = xcd_grid(nbs(na.nPad, 4 * (64 / g)));
            if (g == 16) sc_fine_group_kernel<16><<<gridG, 256, 0, s>>>(na.nPad, a, X, P, rest, vs);
            else if (g == 32) sc_fine_group_kernel<32><<<gridG, 256, 0, s>>>(na.nPad, a, X, P, rest, vs);
            else sc_fine_group_kernel<8><<<gridG, 256, 0, s>>>(na.nPad, a, X, P, rest, vs);
        }
        if (mark) mark(markCtx, "selfcol_build_fine_wave");
        SelfColArgs aw = a;
        aw.fineWaveAll = (g == 8 || g == 16 || g == 32) ? 0 : 1;
        XPBD_SC_LAUNCH_A(sc_fine_kernel, aw);
    }
    if (mark) mark(markCtx, "selfcol_build_coarse");
    SelfColArgs ac = a;
    ac.sortedIdx = vs;
    XPBD_SC_LAUNCH_A(sc_build_kernel, ac);
#undef XPBD_SC_LAUNCH
#undef XPBD_SC_LAUNCH_A
    if (mark) mark(markCtx, "selfcol_build_slots");
    if (a.fineTotal > 0) sc_slots_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a);
    sc_fill_accum_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a.accum);
    *sortedIdxOut = vs;
    return inAlt;
}

void launch_selfcol_iteration(hipStream_t s, const NodeArgs& na, const SelfColArgs& a, const float4* Pin,
                              float4* Pout, const float3* X, const float3* rest, unsigned epi, const EpiArgs& ea)
{
    sc_contact_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, Pin, Pout, X, rest, epi, ea);
}

void launch_contact_push(hipStream_t s, const NodeArgs& na, const SelfColArgs& a, int K, float4* P, const EpiArgs& ea)
{
    if (K > 0) {
        launch_fill_float(s, 4 * K, a.pushSums, 0.0f);
        push_sum_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a);
        push_t_kernel<<<nbs(K, 64), 64, 0, s>>>(K, a);
    }
    push_apply_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, P, ea);
}

}  // namespace xpbd
