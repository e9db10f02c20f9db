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
// accCount of a node of a fine-scan body: what its candidate search needs
// (0: nothing, the node is not in the grid)
constexpr int kScNeedCoarse = kScAccCoarse;  // the node takes the coarse walk
constexpr int kScLane = kScAccLane;          // the lane fine scan lists its candidates
constexpr int kScRule = kScAccRule;          // more than 32 candidates: the slot-rule kernel fills its slots
                                             // from the ranked list at list entry 0
constexpr int kScLong = kScAccLong;          // a list too long to rank in one lane: the group ranking kernel
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
    float maxDisp = 0.0f;
    for (int k = 0; k < D.objK; ++k) maxDisp = fmaxf(maxDisp, a.objDisp[D.objBase + k]);
    a.scUseActive[b] = (D.scOn && D.objActiveAllowed && isfinite(maxDisp)) ? 1 : 0;
    a.scBodyDisp[b] = maxDisp;
    if (b == 0) {
        // the coarse sort is needed from the start when a body takes the
        // coarse walk for all its nodes; the fine scans raise the gate for
        // the nodes they hand over
        int gate = a.fineTotal > 0 ? 0 : 1;
        for (int k = 0; k < B; ++k)
            if (a.bodies[k].scOn && !a.bodies[k].scFine) gate = 1;
        *a.scSortGate = gate;
        if (a.scLong) a.scLong[0] = 0;   // the long-list queue of this build
    }
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
    const float need = D.scSearch + 2.0f * maxDisp;
    float reach = need;
    const float cap = D.scSearch + 16.0f * D.scDc;
    if (D.scCellOverride > 0.0f) reach = D.scCellOverride;
    else if (reach > cap) reach = cap;
    a.scInvCell[b] = 1.0f / reach;
    a.scExact[b] = reach < need ? 1 : 0;
}

// Keys of the frame-start positions (forward.cu:7034-7044) in the body's table,
// in node order; the coarse sort reads them in caller order with the node
// index as value, so that the stable sort breaks ties by the caller's node
// order.
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
    a.keysDev[i] = key;
}

// The coarse sort's input in caller order, only when the sort runs.
__global__ void __launch_bounds__(256) sc_caller_order_kernel(int nPad, const unsigned int* __restrict__ keysDev,
                                                             const int* __restrict__ devOfCaller,
                                                             unsigned int* __restrict__ keys, int* __restrict__ idx,
                                                             const int* __restrict__ gate)
{
    if (gate && *gate == 0) return;
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= nPad) return;
    const int i = devOfCaller[pos];
    keys[pos] = keysDev[i];
    idx[pos] = i;
}

// Empty cells of the coarse tables (-1 start), when the coarse walk runs.
__global__ void __launch_bounds__(256) sc_clear_cells_kernel(SelfColArgs a)
{
    if (*a.scSortGate == 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.tableTotal) a.cellStart[i] = -1;
}

__global__ void __launch_bounds__(256) sc_range_kernel(int nPad, SelfColArgs a, const unsigned int* __restrict__ keysSorted)
{
    if (*a.scSortGate == 0) return;
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
    if (*a.scSortGate == 0) return;
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
// nodes that have a coarse key (active, finite x) and a finite p.  Each
// node's rank inside its cell comes from the cell's counter (the order
// inside a cell does not matter).  laneScan: the grid nodes take the lane
// fine scan, else the coarse walk.
__global__ void __launch_bounds__(256) sc_fine_keys_kernel(int nPad, SelfColArgs a, const float3* __restrict__ X,
                                                          const float4* __restrict__ P, int laneScan)
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
            const int gx = sc_cell(p.x, g), gy = sc_cell(p.y, g), gz = sc_cell(p.z, g);
            const unsigned int fmask = (unsigned int)B.scFineMask;
            key = (unsigned int)B.scFineBase + sc_fine_key(gx, gy, gz, fmask);
            // bit 1 of the active flag: two of the 27 coarse stencil cells
            // around x share a bucket (the scan then looks up the first
            // offset of each candidate's bucket)
            int flags = 1;
            if (sc_stencil_shares_bucket(x, a.scInvCell[b], (unsigned int)B.scTableMask)) flags |= 2;
            // bit 2: the nine row stretches of the 27 cells around p (keys k_r
            // - 1 .. k_r + 1 around the middle-cell key k_r) cannot be walked
            // as stretches when one wraps around the table or two come within
            // two keys of each other: the lane scan then walks cell by cell
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
            if (flags != 1) a.scActive[i] = (unsigned char)flags;
        }
    }
    a.fkeys[i] = key;
    if (key < (unsigned int)a.fineTotal) a.frank[i] = atomicAdd(&a.fcount[key], 1);
    if (B.scOn && B.scFine && i < B.nodeEnd) {
        int path = key < (unsigned int)a.fineTotal ? (laneScan ? kScLane : kScNeedCoarse) : 0;
        // the lane scan reads coarse-cell differences from 3 bits per axis,
        // exact while float rounding keeps neighbours' cells within two of
        // each other (cell indices below 2^22); farther out the coarse walk
        if (path == kScLane) {
            const float3 x = X[i];
            const float inv = a.scInvCell[b];
            const int cx = sc_cell(x.x, inv), cy = sc_cell(x.y, inv), cz = sc_cell(x.z, inv);
            if (max(abs(cx), max(abs(cy), abs(cz))) >= (1 << 22)) path = kScNeedCoarse;
        }
        a.accCount[i] = path;
        if (path == kScNeedCoarse) *a.scSortGate = 1;
    }
}

// Fine-order copies for the fine scans, written at each node's place in the
// counting sort (cell start + rank): (p, rest, caller-order position, node
// id), the low 3 bits of each coarse cell coordinate of x in the spare high
// bits of the last two words (both below 2^27), and (coarse cell of x, node
// id); finePos maps node to place.
__global__ void __launch_bounds__(256) sc_fine_scatter_kernel(int nPad, SelfColArgs a, const float3* __restrict__ X,
                                                             const float4* __restrict__ P,
                                                             const float3* __restrict__ rest)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const unsigned int key = a.fkeys[i];
    if (key >= (unsigned int)a.fineTotal) {
        a.finePos[i] = -1;
        return;
    }
    const int t = a.fstart[key] + a.frank[i];
    const int b = a.tileBody[i / kTile];
    const float4 p = P[i];
    const float3 x = X[i], r = rest[i];
    const float inv = a.scInvCell[b];
    const int cx = sc_cell(x.x, inv), cy = sc_cell(x.y, inv), cz = sc_cell(x.z, inv);
    a.fineE[2 * (size_t)t] = make_float4(p.x, p.y, p.z, r.x);
    const unsigned int code = ((unsigned int)cx & 7u) | (((unsigned int)cy & 7u) << 3) | (((unsigned int)cz & 7u) << 6);
    a.fineE[2 * (size_t)t + 1] = make_float4(r.y, r.z, __uint_as_float((unsigned int)a.callerPos[i] | (code << 27)),
                                             __uint_as_float((unsigned int)i | ((code >> 5) << 27)));
    a.fineC[t] = make_int4(cx, cy, cz, i);
    a.finePos[i] = t;
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

// Entry k of the lane fine scan's list of fine position t: each list is
// contiguous (a lane appends to its own lines, and the ranking kernels read a
// list as one stretch).
constexpr int kLaneStride = 1;
__device__ __forceinline__ size_t sc_lane_at(int t, int k)
{
    return (size_t)t * kLaneList + (size_t)k * kLaneStride;
}

// Fine scan, one lane per node, in the predicted-position grid's own order,
// for bodies whose nodes all belong to one object with self-contact inside
// it on.  There the only tests are |p_i - p_j| < search and
// |q_i - q_j|^2 >= excl^2, so every candidate lies in the 27 cells (edge >=
// search) around p_i: nine rows of three cells, each row one stretch
// [S[k - 1], S[k + 2]) of the counting sort (k the key of the row's middle
// cell).  Consecutive lanes hold nodes of one row, so their stretches
// overlap and their loads share cache lines; each lane keeps kLaneU entries
// (32 B: p, rest, caller-order position, packed coarse cell) in flight.  An
// accepted candidate is keyed at once: its coarse bucket's first stencil
// offset << 27 | its caller-order position.  The coarse walk meets exactly
// the candidates within search whose coarse bucket is one of x_i's 27
// stencil buckets.  A candidate within search lies within search + 2 max
// move of x_i in x, so with the reference's cell edge its coarse cell
// differs from x_i's by at most one per axis and the low bits of the packed
// cell give the offset (when two stencil cells share a bucket, the ranking
// kernel first moves each key to the first offset holding its bucket); with
// a capped or fixed edge the full cells give it.  A candidate more than one
// cell away (capped or fixed edge, or within rounding of the bound) counts
// only if its bucket is a stencil bucket, keyed at the first offset holding
// it.  Nodes whose row stretches wrap around the table or meet in the key
// space walk their 27 cells one by one instead, counting each entry only in
// its own cell.  The (key, node id) pairs go unsorted to the lane's list;
// lists past kLaneList go to the coarse walk.
constexpr int kLaneU = 4;   // entries in flight per lane

// Scan key of an accepted candidate of node i: the first stencil offset of
// its coarse bucket << 27 | its caller-order position cp.  The candidate's
// coarse cell differs from x_i's by at most one per axis (the node takes
// this scan only when the cell edge covers search + 2 max move), so the low
// 3 bits per axis of the cells (codes ci, cj) give the offset.  ~0 when the
// cells differ by more (possible only within rounding of that bound, by two
// at most below 2^22 cells): the node then takes the coarse walk, which
// decides such a candidate by its bucket.
__device__ __forceinline__ unsigned int sc_lane_key(unsigned int ci, unsigned int cj, unsigned int cp)
{
    const unsigned int ux = (cj - ci + 1u) & 7u, uy = ((cj >> 3) - (ci >> 3) + 1u) & 7u,
                       uz = ((cj >> 6) - (ci >> 6) + 1u) & 7u;
    return (ux <= 2u && uy <= 2u && uz <= 2u) ? (((ux + 3u * uy + 9u * uz) << 27) | cp) : 0xffffffffu;
}

// The fine record's second half: rest.y, rest.z, caller position | low code
// bits << 27, node id | high code bits << 27.
__device__ __forceinline__ unsigned int sc_rec_code(float4 r)
{
    return (__float_as_uint(r.z) >> 27) | ((__float_as_uint(r.w) >> 27) << 5);
}
__device__ __forceinline__ unsigned int sc_rec_caller(float4 r) { return __float_as_uint(r.z) & 0x07ffffffu; }
__device__ __forceinline__ int sc_rec_node(float4 r) { return (int)(__float_as_uint(r.w) & 0x07ffffffu); }

// |p_i - p_j|^2 as the scan tests it (products and sums unfused, so the
// ranking of long lists recomputes the same value)
__device__ __forceinline__ float sc_sq_dist(float4 pi, float4 pj)
{
    const float ex = pi.x - pj.x, ey = pi.y - pj.y, ez = pi.z - pj.z;
    return ex * ex + ey * ey + ez * ez;
}

// Scan key from the full coarse cells of x_i (ci) and the candidate (cj):
// the stencil offset of their difference when it is at most one cell per
// axis, else ~0.
__device__ __forceinline__ unsigned int sc_cell_key(int4 ci, int4 cj, unsigned int cp)
{
    const int dx = cj.x - ci.x + 1, dy = cj.y - ci.y + 1, dz = cj.z - ci.z + 1;
    return ((unsigned int)dx <= 2u && (unsigned int)dy <= 2u && (unsigned int)dz <= 2u)
               ? (((unsigned int)(dx + 3 * dy + 9 * dz) << 27) | cp)
               : 0xffffffffu;
}

// Scan key of a candidate whose coarse cell cj is more than one cell from
// x_i's cell ci on some axis (possible when the cell edge is capped or fixed
// below search + 2 max move): the coarse walk meets it only if its bucket is
// one of the 27 stencil buckets, at the first stencil offset holding that
// bucket; ~0 when it is none of them (not a candidate).
__device__ __forceinline__ unsigned int sc_far_key(int4 ci, int4 cj, unsigned int mask, unsigned int cp)
{
    const unsigned int hj = sc_hash(cj.x, cj.y, cj.z, mask);
    for (int o = 0; o < 27; ++o)
        if (sc_hash(ci.x + o % 3 - 1, ci.y + (o / 3) % 3 - 1, ci.z + o / 9 - 1, mask) == hj)
            return ((unsigned int)o << 27) | cp;
    return 0xffffffffu;
}

// EXACT: the nodes of bodies whose coarse cell edge is capped or fixed below
// search + 2 max move (full-cell keys, far candidates by bucket); otherwise
// the other nodes (packed-cell keys; a candidate two cells away, possible
// only within rounding, hands the node to the coarse walk).  Two
// instantiations keep the common one's register count low.
template <bool EXACT>
__global__ void __launch_bounds__(256) sc_fine_lane_kernel(SelfColArgs a, const float4* __restrict__ E,
                                                          const int4* __restrict__ FC, const int* __restrict__ FS,
                                                          int2* __restrict__ LK)
{
    const int t = xcd_block() * (int)blockDim.x + (int)threadIdx.x;
    if (t >= FS[a.fineTotal]) return;   // fine positions in use
    const int i = FC[t].w;
    // the capped-cell instantiation skips the (usual) other nodes before
    // loading their records
    if (EXACT && a.scExact[a.tileBody[i / kTile]] == 0) return;
    const float4 m0 = E[2 * (size_t)t], m1 = E[2 * (size_t)t + 1];
    if (a.accCount[i] != kScLane) return;   // flagged for the coarse walk
    const int b = a.tileBody[i / kTile];
    const BodyDev& B = a.bodies[b];
    const float search2 = B.scSearch * B.scSearch, excl2 = B.scExcl2;
    const float g = B.scFineInv;
    const unsigned int fmask = (unsigned int)B.scFineMask;
    const int* __restrict__ S = FS + B.scFineBase;
    const int gx = sc_cell(m0.x, g), gy = sc_cell(m0.y, g), gz = sc_cell(m0.z, g);
    // rows whose stretches wrap or meet in the key space: walk the 27 cells
    // one by one, each entry counted only in its own cell
    if ((a.scExact[b] != 0) != EXACT) return;   // the other instantiation's node
    const bool percell = (a.scActive[i] & 4) != 0;
    const unsigned int ci = sc_rec_code(m1);
    const size_t L0 = sc_lane_at(t, 0);   // entry k at L0 + k
    int n = 0;
    bool far = false;   // (not EXACT) a candidate two coarse cells away: the coarse walk decides
#pragma unroll 1
    for (int w = 0; w < (percell ? 27 : 9); ++w) {
        int lo, h, cx = 0, cy = 0, cz = 0;
        if (!percell) {
            const unsigned int k = sc_fine_key(gx, gy + w % 3 - 1, gz + w / 3 - 1, fmask);
            lo = S[k - 1];
            h = S[k + 2];
        } else {
            cx = gx + w % 3 - 1;
            cy = gy + (w / 3) % 3 - 1;
            cz = gz + w / 9 - 1;
            const unsigned int k = sc_fine_key(cx, cy, cz, fmask);
            lo = S[k];
            h = S[k + 1];
        }
        for (int e0 = lo; e0 < h; e0 += kLaneU) {
            float4 A[kLaneU], C[kLaneU];
#pragma unroll
            for (int u = 0; u < kLaneU; ++u)
                if (e0 + u < h) {
                    A[u] = E[2 * (size_t)(e0 + u)];
                    C[u] = E[2 * (size_t)(e0 + u) + 1];
                }
#pragma unroll
            for (int u = 0; u < kLaneU; ++u) {
                const int e = e0 + u;
                if (e >= h) break;
                const float d2 = sc_sq_dist(m0, A[u]);
                const float qx = m0.w - A[u].w, qy = m1.x - C[u].x, qz = m1.y - C[u].y;
                const float q2 = qx * qx + qy * qy + qz * qz;
                if (!((int)(d2 < search2) & (int)(q2 >= excl2) & (int)(e != t))) continue;
                if (percell && !((int)(sc_cell(A[u].x, g) == cx) & (int)(sc_cell(A[u].y, g) == cy) &
                                 (int)(sc_cell(A[u].z, g) == cz)))
                    continue;
                const unsigned int cp = sc_rec_caller(C[u]);
                unsigned int key;
                if (EXACT) {
                    // the coarse cells may lie far apart: the full cells
                    // decide (the packed low bits could wrap), and a
                    // candidate more than one cell away by its bucket
                    key = sc_cell_key(FC[t], FC[e], cp);
                    if (key == 0xffffffffu) {
                        key = sc_far_key(FC[t], FC[e], (unsigned int)B.scTableMask, cp);
                        if (key == 0xffffffffu) continue;
                    }
                } else {
                    key = sc_lane_key(ci, sc_rec_code(C[u]), cp);
                    far |= key == 0xffffffffu;
                }
                if (n < kLaneList) LK[L0 + (size_t)n * kLaneStride] = make_int2((int)key, sc_rec_node(C[u]));
                ++n;
            }
        }
    }
    if (n > kLaneList || far) {
        a.accCount[i] = kScNeedCoarse;
        *a.scSortGate = 1;
        n = -1;
    }
    a.laneN[t] = n;
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
    if (*a.scSortGate == 0) return;   // no node takes the coarse walk
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

// The reference's slot rule (forward.cu:7149-7174) over a list of n > 32
// entries already in scan order: the first 32 fill the slots in order, each
// later one replaces the worst slot (largest priority, lowest slot among
// equals) when strictly better (all fine candidates are same-object ones,
// one tier).  entry(r) gives the r-th candidate's code and priority bits;
// entries are fetched eight at a time.
template <typename Entry>
__device__ __forceinline__ void sc_fill_slots(int i, int n, SelfColArgs& a, Entry entry)
{
    constexpr int CAP = kSelfColCap;
    int code[CAP];
    float prio[CAP];
#pragma unroll
    for (int s = 0; s < CAP; ++s) {
        const int2 e = entry(s);
        code[s] = e.x;
        prio[s] = __int_as_float(e.y);
    }
    // worst slot, kept until a slot changes
    int w = 0;
    float wp = prio[0];
#pragma unroll
    for (int s = 1; s < CAP; ++s)
        if (prio[s] > wp) { wp = prio[s]; w = s; }
    constexpr int G = 8;
    for (int r0 = CAP; r0 < n; r0 += G) {
        int2 ev[G];
#pragma unroll
        for (int g = 0; g < G; ++g) ev[g] = r0 + g < n ? entry(r0 + g) : make_int2(0, 0);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (r0 + g >= n) break;
            const float p = __int_as_float(ev[g].y);
            if (p < wp) {
#pragma unroll
                for (int s = 0; s < CAP; ++s)
                    if (s == w) { code[s] = ev[g].x; prio[s] = p; }
                w = 0;
                wp = prio[0];
#pragma unroll
                for (int s = 1; s < CAP; ++s)
                    if (prio[s] > wp) { wp = prio[s]; w = s; }
            }
        }
    }
#pragma unroll
    for (int s = 0; s < CAP; ++s) a.contacts[sc_slot(i, s)] = code[s];
}

// Two stencil cells of node i share a bucket: a candidate's bucket is scanned
// at the first offset holding it (its cell is x_i's cell + its offset), so
// the keys of its list move to that offset.
__device__ __forceinline__ void sc_lane_first_offsets(int4 ci, unsigned int tmask, int m, size_t L0,
                                                      int2* __restrict__ LK)
{
    // first[o] for the 27 offsets, 5 bits each, six per word
    unsigned int hs[27];
#pragma unroll
    for (int c = 0; c < 27; ++c) hs[c] = sc_hash(ci.x + c % 3 - 1, ci.y + (c / 3) % 3 - 1, ci.z + c / 9 - 1, tmask);
    unsigned int first[5] = {0u, 0u, 0u, 0u, 0u};
#pragma unroll
    for (int o = 0; o < 27; ++o) {
        unsigned int f = (unsigned int)o;
#pragma unroll
        for (int c = o - 1; c >= 0; --c)
            if (hs[c] == hs[o]) f = (unsigned int)c;
        first[o / 6] |= f << (5 * (o % 6));
    }
    for (int k = 0; k < m; ++k) {
        const int2 en = LK[L0 + (size_t)k * kLaneStride];
        const unsigned int o = (unsigned int)en.x >> 27;
        const unsigned int wsel = o / 6u;
        const unsigned int word = wsel == 0u ? first[0] : wsel == 1u ? first[1] : wsel == 2u ? first[2]
                                : wsel == 3u ? first[3] : first[4];
        const unsigned int f = (word >> (5u * (o % 6u))) & 31u;
        if (f != o) LK[L0 + (size_t)k * kLaneStride] = make_int2((int)((f << 27) | ((unsigned int)en.x & 0x07ffffffu)), en.y);
    }
}

// Ranking of the lane fine scan's lists of at most 32 candidates, one lane
// per list in node order (a wave's 64 nodes are one slot tile, so each slot
// is written as one contiguous 256 B row; the lists are read from the nodes'
// fine positions).  Each lane sorts its (key, node id) pairs as
// 64-bit values key << 32 | id (keys all distinct; padding ~0 sorts last)
// with a bitonic network in registers: N pairs, N the wave's power of two at
// or above its longest such list, compare-exchange steps only; the ids in
// sorted order are the slots.  Longer lists are flagged for the group
// ranking kernel.
constexpr int kRankNet = kSelfColCap;

template <int N>
__device__ __forceinline__ void sc_bitonic(unsigned long long* k)
{
#pragma unroll
    for (int size = 2; size <= N; size <<= 1)
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
            for (int x = 0; x < N; ++x) {
                const int y = x ^ stride;
                if (y > x) {
                    const bool sw = k[y] < k[x];
                    const unsigned long long lo = sw ? k[y] : k[x], hi = sw ? k[x] : k[y];
                    const bool up = (x & size) == 0;
                    k[x] = up ? lo : hi;
                    k[y] = up ? hi : lo;
                }
            }
}

// Load, sort and hand out one lane's list (m <= 32) with an N-key network.
template <int N>
__device__ __forceinline__ void sc_rank_list(SelfColArgs& a, int i, int m, size_t L0, const int2* __restrict__ LK)
{
    // rows past the list (N <= 32 < kLaneList, in the buffer) are read and
    // replaced by padding
    unsigned long long k[N];
#pragma unroll
    for (int x = 0; x < N; ++x) {
        const int2 e = LK[L0 + (size_t)x * kLaneStride];
        k[x] = x < m ? ((unsigned long long)(unsigned int)e.x << 32) | (unsigned int)e.y : ~0ull;
    }
    sc_bitonic<N>(k);
#pragma unroll
    for (int x = 0; x < N; ++x)
        if (x < m) a.contacts[sc_slot(i, x)] = (int)(unsigned int)k[x];
}

__global__ void __launch_bounds__(256) sc_lane_rank_kernel(int nPad, SelfColArgs a, const int4* __restrict__ FC,
                                                          int2* __restrict__ LK)
{
    const int i = blockIdx.x * blockDim.x + (int)threadIdx.x;
    bool listed = false;
    int m = 0, t = 0, b = 0;
    if (i < nPad) {
        b = a.tileBody[i / kTile];
        const BodyDev& B = a.bodies[b];
        if (B.scOn && B.scFine && i < B.nodeEnd && a.accCount[i] == kScLane) {
            listed = true;
            t = a.finePos[i];   // in the grid: the fine keys gave the node the lane scan
            m = a.laneN[t];
        }
    }
    const size_t L0 = sc_lane_at(t, 0);
    if (listed && (a.scActive[i] & 2))
        sc_lane_first_offsets(FC[t], (unsigned int)a.bodies[b].scTableMask, m, L0, LK);
    const bool here = listed && m <= kRankNet;
    // longer lists go to the queue of the group ranking and the slot rule
    // (one atomic per wave)
    const bool lng = listed && !here;
    if (lng) a.accCount[i] = kScLong;
    const unsigned long long lb = __ballot(lng);
    if (lb != 0ull) {
        const int lane = (int)(threadIdx.x & 63);
        const int lead = __builtin_ctzll(lb);
        int base = 0;
        if (lane == lead) base = atomicAdd(&a.scLong[0], __popcll(lb));
        base = __shfl(base, lead, 64);
        if (lng) a.scLong[1 + base + __popcll(lb & ((1ull << lane) - 1ull))] = i;
    }
    // the wave's network size
    int need = here ? m : 0;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) need = max(need, __shfl_xor(need, s, 64));
    need = __builtin_amdgcn_readfirstlane(need);
    if (here && m > 0) {
        if (need <= 8) sc_rank_list<8>(a, i, m, L0, LK);
        else if (need <= 16) sc_rank_list<16>(a, i, m, L0, LK);
        else sc_rank_list<32>(a, i, m, L0, LK);
    }
    if (here) {
        a.counts[i] = m;
        a.overflow[i] = 0;
    }
    // the contact push of object shape matching needs to know whether the
    // body has any candidate
    if (listed && m > 0 && a.bodies[b].objOn) a.scCounters[3 * b] = 1;
}

// Ranking of the lane-scan lists of more than 32 candidates, 32 lanes per
// list, the lists taken from the build's long-list queue by a fixed grid
// (each list is one contiguous stretch).  The node's (key, node id) list is
// spread over the group's lanes; its keys (all distinct) are ranked by
// counting in LDS; the list goes back to its entries in rank order as (id,
// priority = |p_i - p_j|, from the squared distance recomputed as the scan
// tested it) and the node is flagged for the slot-rule kernel.
constexpr int kLaneGroup = 32;
constexpr int kLanePer = (kLaneList + kLaneGroup - 1) / kLaneGroup;
constexpr int kLaneGroups = 256 / kLaneGroup;

__global__ void __launch_bounds__(256) sc_slots_kernel(SelfColArgs a, const float4* __restrict__ P)
{
    __shared__ unsigned int skey[kLaneGroups][kLaneList];
    const int sub = threadIdx.x % kLaneGroup, grp = threadIdx.x / kLaneGroup;
    const int count = a.scLong[0];
    // the loop bound is the same for the whole block
    for (int q0 = (int)blockIdx.x * kLaneGroups; q0 < count; q0 += (int)gridDim.x * kLaneGroups) {
        const int q = q0 + grp;
        const bool lane = q < count;
        int i = 0, t = 0, m = 0;
        if (lane) {
            i = a.scLong[1 + q];
            t = a.finePos[i];
            m = a.laneN[t];
        }
        const size_t L0 = sc_lane_at(t, 0);
        // keys per lane in use: the wave's largest
        int kn = (m + kLaneGroup - 1) / kLaneGroup;
#pragma unroll
        for (int s = kLaneGroup; s < 64; s <<= 1) kn = max(kn, __shfl_xor(kn, s, 64));
        kn = __builtin_amdgcn_readfirstlane(kn);
        if (kn == 0) continue;   // wave-uniform: this wave's groups are past the queue
        unsigned int key[kLanePer];
        int id[kLanePer], rank[kLanePer];
#pragma unroll
        for (int k = 0; k < kLanePer; ++k) {
            key[k] = 0xffffffffu;
            id[k] = 0;
            rank[k] = 0;
            const int idx = sub + kLaneGroup * k;
            if (k < kn && idx < m) {
                const int2 e = a.laneKey[L0 + (size_t)idx * kLaneStride];
                key[k] = (unsigned int)e.x;
                id[k] = e.y;
                skey[grp][idx] = key[k];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (int f = 0; f < m; ++f) {
            const unsigned int kf = skey[grp][f];
#pragma unroll
            for (int k = 0; k < kLanePer; ++k)
                if (k < kn) rank[k] += kf < key[k] ? 1 : 0;
        }
        if (lane) {
            // the list in rank order with its priorities, for the slot rule (the
            // candidates' positions first, all in flight together)
            const float4 pi = a.fineE[2 * (size_t)t];
            float4 pj[kLanePer];
#pragma unroll
            for (int k = 0; k < kLanePer; ++k)
                if (k < kn && sub + kLaneGroup * k < m) pj[k] = P[id[k]];
#pragma unroll
            for (int k = 0; k < kLanePer; ++k)
                if (k < kn && sub + kLaneGroup * k < m)
                    a.laneKey[L0 + (size_t)rank[k] * kLaneStride] = make_int2(id[k], __float_as_int(sqrtf(sc_sq_dist(pi, pj[k]))));
            if (sub == 0) {
                a.counts[i] = kSelfColCap;
                a.overflow[i] = m - kSelfColCap;
                a.accCount[i] = kScRule;
            }
        }
        // the group's keys in LDS are read before the next list overwrites them
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
}

// The slot rule for the nodes of the long-list queue (more than 32
// candidates), one thread per node, over the (id, priority) list the group
// ranking left in rank order.
__global__ void __launch_bounds__(256) sc_slot_rule_kernel(SelfColArgs a)
{
    const int count = a.scLong[0];
    for (int q = (int)(blockIdx.x * blockDim.x + threadIdx.x); q < count; q += (int)(gridDim.x * blockDim.x)) {
        const int i = a.scLong[1 + q];
        const size_t L0 = sc_lane_at(a.finePos[i], 0);
        sc_fill_slots(i, kSelfColCap + a.overflow[i], a,
                      [&](int r) { return a.laneKey[L0 + (size_t)r * kLaneStride]; });
    }
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

// Clears the contact-push accumulators; nodes of predicted-position-grid
// bodies that have no grid key (inactive, non-finite) get empty lists.
__global__ void __launch_bounds__(256) sc_finish_kernel(int nPad, SelfColArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    a.accum[i] = make_float3(0.0f, 0.0f, 0.0f);
    const BodyDev& B = a.bodies[a.tileBody[i / kTile]];
    if (B.scOn && B.scFine && i < B.nodeEnd && a.accCount[i] == 0) {
        a.counts[i] = 0;
        a.overflow[i] = 0;
    }
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
    const float4 pi4 = P[i];
    const float wi = pi4.w * B.invMassScale;
    const float dc = B.scDc;
    float ax = 0.0f, ay = 0.0f, az = 0.0f, cpx = 0.0f, cpy = 0.0f, cpz = 0.0f;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f, wFast = 0.0f;
    int act = 0, nFast = 0;
    const int* list = &a.contacts[sc_slot(i, 0)];   // slot k at list[k * kTile]
    // x_i and the rest position only on the rare paths that need them
    // (cross-object pairs, coincident positions)
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
                    if (!sc_start_normal(X[i], X[j], rest[i], rest[j], nx, ny, nz)) continue;
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
                    const float3 qi = rest[i], qj = rest[j];
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
                if (!sc_start_normal(X[i], X[j], rest[i], rest[j], nx, ny, nz)) continue;
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
        // the cross-object part accumulates for the contact push; adding a
        // zero would leave the sum as it is
        if (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f) {
            const float3 acc = a.accum[i];
            a.accum[i] = make_float3(acc.x + c.x, acc.y + c.y, acc.z + c.z);
        }
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

void launch_sc_caller_order(hipStream_t s, int nPad, const unsigned int* keysDev, const int* devOfCaller,
                            unsigned int* keys, int* idx, const int* gate)
{
    sc_caller_order_kernel<<<nbs(nPad, 256), 256, 0, s>>>(nPad, keysDev, devOfCaller, keys, idx, gate);
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
    const bool laneScan = tu.scGroupLanes != 0;
    if (a.fineTotal > 0) {
        // counting sort of the predicted-position grid: cell counts (and each
        // node's rank in its cell), their prefix sum, then the fine-order copies
        if (mark) mark(markCtx, "selfcol_build_fine_grid");
        launch_fill_int(s, a.fineTotal, a.fcount, 0);
        sc_fine_keys_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, X, P, laneScan ? 1 : 0);
        launch_exclusive_scan(s, a.fineTotal, a.fcount, a.fstart, a.fscan);
        sc_fine_scatter_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, X, P, rest);
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
    // the lane fine scan first: the coarse kernel walks the nodes it hands over
    if (a.fineTotal > 0 && laneScan) {
        if (mark) mark(markCtx, "selfcol_build_fine_lane");
        sc_fine_lane_kernel<false><<<xcd_grid(nbs(na.nPad, 256)), 256, 0, s>>>(a, a.fineE, a.fineC, a.fstart, a.laneKey);
        sc_fine_lane_kernel<true><<<xcd_grid(nbs(na.nPad, 256)), 256, 0, s>>>(a, a.fineE, a.fineC, a.fstart, a.laneKey);
    }
    // the coarse walk and what it reads (stable sort of the frame-start keys,
    // cell ranges, sorted copies), only when a node needs it: the kernels
    // return at once when the gate reads 0 (the keys stay in place for an
    // export, which then sorts them)
    if (mark) mark(markCtx, "selfcol_build_coarse");
    sc_caller_order_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a.keysDev, a.devOfCaller, a.keys, a.idx,
                                                             a.scSortGate);
    const bool inAlt = launch_radix_sort(s, na.nPad, sortBits, a.keys, a.idx, keysAlt, idxAlt, hist, false,
                                         a.scSortGate);
    const unsigned int* ks = inAlt ? keysAlt : a.keys;
    const int* vs = inAlt ? idxAlt : a.idx;
    if (a.tableTotal > 0) sc_clear_cells_kernel<<<nbs(a.tableTotal, 256), 256, 0, s>>>(a);
    sc_range_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, ks);
    sc_stage_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, X, P, rest, vs);
    SelfColArgs ac = a;
    ac.sortedIdx = vs;
    XPBD_SC_LAUNCH_A(sc_build_kernel, ac);
#undef XPBD_SC_LAUNCH
#undef XPBD_SC_LAUNCH_A
    if (mark) mark(markCtx, "selfcol_build_slots");
    if (a.fineTotal > 0) {
        if (laneScan)
            sc_lane_rank_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a, a.fineC, a.laneKey);
        if (mark) mark(markCtx, "selfcol_build_rank_long");
        // fixed grids over the long-list queue (at most one entry per node)
        sc_slots_kernel<<<nbs(na.nPad, kLaneGroups) < 16384 ? nbs(na.nPad, kLaneGroups) : 16384, 256, 0, s>>>(a, P);
        if (mark) mark(markCtx, "selfcol_build_slot_rule");
        sc_slot_rule_kernel<<<nbs(na.nPad, 256) < 1024 ? nbs(na.nPad, 256) : 1024, 256, 0, s>>>(a);
    }
    sc_finish_kernel<<<nbs(na.nPad, 256), 256, 0, s>>>(na.nPad, a);
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
