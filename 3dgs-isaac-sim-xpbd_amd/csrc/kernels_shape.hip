// kernels_shape.hip — per-node 1-ring shape matching, the per-Gaussian
// scale / orientation update from the local deformation gradient, and the
// contact-ellipsoid packing.
//
// Shape matching (forward.cu:6088-6383): centroids of {node} + neighbours at
// rest and now, a boundary-bias reject (mean unit direction > 0.4), the
// Frobenius-normalised A = sum p q^T, its rotation by the fp64 polar
// decomposition (forward.cu:5991-6086; relative Jacobi stop 1e-12, flat fix
// S0 < 0.08 Smax, reflection fix) or the float path, goal = c + R q_i and an
// XPBD pull toward the goal (4972-4988).  One thread per node; Jacobi.
//
// Deformed shapes (forward.cu:13063-13225): G = (Q P^T + eI)(P P^T + eI)^-1
// from the 1-ring, singular values clamped to [1e-3, 1.5], M = S0 R0 S_f R^T,
// Sigma' = M^T M, then a cyclic fp64 Jacobi eigen-decomposition (13022-13061)
// of Sigma' / (tr / 3) gives the scales and the quaternion.
#include <mutex>
#include <string>

#include "kernels.h"
#include "mat3.h"
#include "node_ops.h"
#include "jacobi3.h"
#include "polar.h"

namespace xpbd {

// ---------------------------------------------------------------------------
__device__ bool polar_rotation_f32(const M3& A, M3& R)
{
    const M3 AtA = m3_mul(m3_transpose(A), A);
    float S2[3];
    M3 V;
    m3_eig_sym(AtA, S2, V);
    const float s0 = sqrtf(fmaxf(0.0f, S2[0])), s1 = sqrtf(fmaxf(0.0f, S2[1])), s2 = sqrtf(fmaxf(0.0f, S2[2]));
    if (!isfinite(s0) || !isfinite(s1) || !isfinite(s2)) return false;
    const float e = 1e-7f;
    const M3 Si = m3_diag(s0 > e ? 1.0f / s0 : 0.0f, s1 > e ? 1.0f / s1 : 0.0f, s2 > e ? 1.0f / s2 : 0.0f);
    M3 U = m3_mul(m3_mul(A, V), Si);
    if (!m3_finite(U)) return false;
    if (s0 < 0.08f * fmaxf(s2, 1e-7f)) {
        const float c0 = U.m[1][1] * U.m[2][2] - U.m[2][1] * U.m[1][2];
        const float c1 = U.m[1][2] * U.m[2][0] - U.m[2][2] * U.m[1][0];
        const float c2 = U.m[1][0] * U.m[2][1] - U.m[2][0] * U.m[1][1];
        const float l = sqrtf(c0 * c0 + c1 * c1 + c2 * c2);
        if (!(l > 1e-7f)) return false;
        U.m[0][0] = c0 / l; U.m[0][1] = c1 / l; U.m[0][2] = c2 / l;
    }
    R = m3_mul(U, m3_transpose(V));
    if (!m3_finite(R)) return false;
    if (m3_det(R) < 0.0f) {
        U.m[0][0] *= -1.0f; U.m[0][1] *= -1.0f; U.m[0][2] *= -1.0f;
        R = m3_mul(U, m3_transpose(V));
        if (!m3_finite(R)) return false;
    }
    return true;
}

// Butterfly sum over the 64 lanes of a wave: every lane ends with the same
// total, added in a fixed tree (lane partials l and l ^ m, m = 32 .. 1).
__device__ __forceinline__ float wave_tree_sum(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = v + __shfl_xor(v, m, 64);
    return v;
}

// Cluster shape matching.  A block takes 4 clusters per wave.  A wave first
// loads the member indices of its clusters' first 64 members (one per lane),
// then their current and rest positions, all clusters at once so the gathers
// overlap.  Lane l sums members l, l + 64, ... (member order, caller order
// inside the cluster) and the 64 partials are combined by wave_tree_sum,
// first for the centroid, then for the nine entries of
// A = sum (p - c)(q - c0)^T; lane 0 leaves centroid and A in LDS.  After the
// barrier the first wave takes one cluster per lane: A is normalised by its
// Frobenius norm and replaced by its rotation (fp64 polar when robust), and
// centroid, rotation and the valid flag go to the cluster's parameter row for
// the per-node projection.  The polar chain thus runs once per cluster, on
// as many lanes as the block has clusters.
constexpr int kClustersPerWave = 4;
constexpr int kClusterRounds = 2;   // members per lane loaded up front

__global__ void __launch_bounds__(256) shape_cluster_kernel(int K, ShapeClusterArgs a, const float4* __restrict__ P,
                                                           const float3* __restrict__ rest,
                                                           const int* __restrict__ tileBody,
                                                           const BodyDev* __restrict__ bodies)
{
    constexpr int C = kClustersPerWave, NR = kClusterRounds;
    __shared__ float sums[4 * C][12];   // centroid, then A entry (c, r) at 3 + 3 c + r
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int perBlock = (int)(blockDim.x >> 6) * C;
    const int k0 = xcd_block() * perBlock;   // neighbouring clusters share an XCD
    int b0[C], cnt[C], j[C][NR];
    float4 p[C][NR];
    float3 q[C][NR];
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const int k = k0 + wave * C + i;
        b0[i] = k < K ? a.off[k] : 0;
        cnt[i] = k < K ? a.off[k + 1] - b0[i] : 0;
        if (cnt[i] < 4) cnt[i] = 0;   // too small: no rotation
    }
#pragma unroll
    for (int i = 0; i < C; ++i)
#pragma unroll
        for (int u = 0; u < NR; ++u) j[i][u] = lane + 64 * u < cnt[i] ? a.mem[b0[i] + lane + 64 * u] : -1;
#pragma unroll
    for (int i = 0; i < C; ++i)
#pragma unroll
        for (int u = 0; u < NR; ++u) {
            p[i][u] = j[i][u] >= 0 ? P[j[i][u]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            q[i][u] = j[i][u] >= 0 ? rest[j[i][u]] : make_float3(0.0f, 0.0f, 0.0f);
        }
#pragma unroll
    for (int i = 0; i < C; ++i) {
        if (cnt[i] == 0) continue;   // wave-uniform
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
        for (int u = 0; u < NR; ++u)
            if (j[i][u] >= 0) { sx = sx + p[i][u].x; sy = sy + p[i][u].y; sz = sz + p[i][u].z; }
        for (int t = lane + 64 * NR; t < cnt[i]; t += 64) {
            const float4 pt = P[a.mem[b0[i] + t]];
            sx = sx + pt.x; sy = sy + pt.y; sz = sz + pt.z;
        }
        const float invC = 1.0f / (float)cnt[i];
        const float cx = wave_tree_sum(sx) * invC, cy = wave_tree_sum(sy) * invC, cz = wave_tree_sum(sz) * invC;
        const float3 cr = a.comRest[k0 + wave * C + i];
        float acc[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) acc[e] = 0.0f;
        for (int t = lane, u = 0; t < cnt[i]; t += 64, ++u) {
            float4 pt;
            float3 qt;
            if (u < NR) {
                pt = p[i][0]; qt = q[i][0];
#pragma unroll
                for (int v = 1; v < NR; ++v)
                    if (u == v) { pt = p[i][v]; qt = q[i][v]; }
            } else {
                const int jt = a.mem[b0[i] + t];
                pt = P[jt];
                qt = rest[jt];
            }
            const float pv[3] = {pt.x - cx, pt.y - cy, pt.z - cz}, qv[3] = {qt.x - cr.x, qt.y - cr.y, qt.z - cr.z};
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int r = 0; r < 3; ++r) acc[3 * c + r] = acc[3 * c + r] + pv[r] * qv[c];
        }
#pragma unroll
        for (int e = 0; e < 9; ++e) acc[e] = wave_tree_sum(acc[e]);
        if (lane == 0) {
            const int kl = wave * C + i;
            sums[kl][0] = cx; sums[kl][1] = cy; sums[kl][2] = cz;
#pragma unroll
            for (int e = 0; e < 9; ++e) sums[kl][3 + e] = acc[e];
        }
    }
    __syncthreads();
    const int k = k0 + lane;
    if (wave != 0 || lane >= perBlock || k >= K) return;
    const int first = a.off[k];
    float* prm = &a.params[16 * (size_t)k];
    if (a.off[k + 1] - first < 4) {
        prm[15] = 0.0f;
        return;
    }
    M3 A;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) A.m[c][r] = sums[lane][3 + 3 * c + r];
    float nsq = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) nsq += A.m[c][r] * A.m[c][r];
    const float an = sqrtf(nsq);
    bool ok = isfinite(an) && !(an < 1e-8f);
    M3 R = m3_diag(1.0f, 1.0f, 1.0f);
    if (ok) {
        const BodyDev& B = bodies[tileBody[a.mem[first] / kTile]];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int r = 0; r < 3; ++r) A.m[c][r] /= an;
        ok = B.shapeRobust ? polar_rotation_f64(A, R) : polar_rotation_f32(A, R);
    }
    prm[0] = sums[lane][0]; prm[1] = sums[lane][1]; prm[2] = sums[lane][2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) prm[3 + 3 * c + r] = R.m[c][r];
    prm[15] = ok ? 1.0f : 0.0f;
}

// Per-node shape matching over the graph 1-ring (or the node's cluster),
// plus the iteration epilogue.
__global__ void __launch_bounds__(256) shape_match_kernel(
    int nPad, const int* __restrict__ tileBody, const BodyDev* __restrict__ bodies, const float4* __restrict__ Pin,
    float4* __restrict__ Pout, const float3* __restrict__ rest, const int* __restrict__ rowOff,
    const int* __restrict__ nbr, ShapeClusterArgs sca, unsigned epi, EpiArgs ea)
{
    const int idx = xcd_block() * blockDim.x + threadIdx.x;
    if (idx >= nPad) return;
    const BodyDev& B = bodies[__builtin_amdgcn_readfirstlane(tileBody[idx / kTile])];
    if (idx >= B.nodeEnd) return;
    const float4 p4 = Pin[idx];
    float3 out = make_float3(p4.x, p4.y, p4.z);
    const float wi = p4.w * B.invMassScale;
    const int off = rowOff[idx], cnt = rowOff[idx + 1] - off;
    // cluster goal (the cluster pass above stored centroid and rotation)
    const int kc = (B.shapeOn && B.shapeClusters && wi > 0.0f) ? sca.clusterOf[idx] : -1;
    while (kc >= 0) {
        const float* prm = &sca.params[16 * (size_t)kc];
        if (prm[15] == 0.0f) break;
        M3 R;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int r = 0; r < 3; ++r) R.m[c][r] = prm[3 + 3 * c + r];
        const float3 r0 = rest[idx];
        const float3 cr = sca.comRest[kc];
        const float3 rq = m3_mul_vec(R, r0.x - cr.x, r0.y - cr.y, r0.z - cr.z);
        const float gx = prm[0] + rq.x, gy = prm[1] + rq.y, gz = prm[2] + rq.z;
        const float ddx = gx - p4.x, ddy = gy - p4.y, ddz = gz - p4.z;
        const float gd2 = ddx * ddx + ddy * ddy + ddz * ddz;
        if (!isfinite(gd2) || gd2 < 1e-18f) break;
        const float dt2 = fmaxf(B.dt * B.dt, 1e-8f);
        const float at = B.shapeCompliance / dt2;
        const float denom = wi + at;
        if (denom < 1e-8f) break;
        const float sc = B.shapeBlend * (wi / denom);
        out = make_float3(p4.x + ddx * sc, p4.y + ddy * sc, p4.z + ddz * sc);
        break;
    }
    while (B.shapeOn && !B.shapeClusters && wi > 0.0f && cnt >= 4) {
        const float3 r0 = rest[idx];
        float crx = r0.x, cry = r0.y, crz = r0.z;
        float ccx = p4.x, ccy = p4.y, ccz = p4.z;
        float cc = 1.0f;
        for (int k = 0; k < cnt; ++k) {
            const int j = nbr[off + k];
            const float3 rj = rest[j];
            const float4 pj = Pin[j];
            crx = crx + rj.x; cry = cry + rj.y; crz = crz + rj.z;
            ccx = ccx + pj.x; ccy = ccy + pj.y; ccz = ccz + pj.z;
            cc += 1.0f;
        }
        const float invC = 1.0f / cc;
        crx = crx * invC; cry = cry * invC; crz = crz * invC;
        ccx = ccx * invC; ccy = ccy * invC; ccz = ccz * invC;
        float dsx = 0.0f, dsy = 0.0f, dsz = 0.0f;
        for (int k = 0; k < cnt; ++k) {
            const float3 rj = rest[nbr[off + k]];
            const float qx = rj.x - crx, qy = rj.y - cry, qz = rj.z - crz;
            const float l = sqrtf(qx * qx + qy * qy + qz * qz);
            if (l > 1e-7f) {
                const float il = 1.0f / l;
                dsx = dsx + qx * il; dsy = dsy + qy * il; dsz = dsz + qz * il;
            }
        }
        const float bias = sqrtf(dsx * dsx + dsy * dsy + dsz * dsz) / fmaxf((float)cnt, 1.0f);
        if (bias > 0.4f) break;
        M3 A = m3_zero();
        auto acc = [&](float px, float py, float pz, float qx, float qy, float qz) {
            const float pv[3] = {px, py, pz}, qv[3] = {qx, qy, qz};
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int r = 0; r < 3; ++r) A.m[c][r] += pv[r] * qv[c];
        };
        acc(p4.x - ccx, p4.y - ccy, p4.z - ccz, r0.x - crx, r0.y - cry, r0.z - crz);
        for (int k = 0; k < cnt; ++k) {
            const int j = nbr[off + k];
            const float3 rj = rest[j];
            const float4 pj = Pin[j];
            acc(pj.x - ccx, pj.y - ccy, pj.z - ccz, rj.x - crx, rj.y - cry, rj.z - crz);
        }
        float nsq = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int r = 0; r < 3; ++r) nsq += A.m[c][r] * A.m[c][r];
        const float an = sqrtf(nsq);
        if (!isfinite(an) || an < 1e-8f) break;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int r = 0; r < 3; ++r) A.m[c][r] /= an;
        M3 R;
        const bool ok = B.shapeRobust ? polar_rotation_f64(A, R) : polar_rotation_f32(A, R);
        if (!ok) break;
        const float3 rq = m3_mul_vec(R, r0.x - crx, r0.y - cry, r0.z - crz);
        const float gx = ccx + rq.x, gy = ccy + rq.y, gz = ccz + rq.z;
        const float ddx = gx - p4.x, ddy = gy - p4.y, ddz = gz - p4.z;
        const float gd2 = ddx * ddx + ddy * ddy + ddz * ddz;
        if (!isfinite(gd2) || gd2 < 1e-18f) break;
        const float dt2 = fmaxf(B.dt * B.dt, 1e-8f);
        const float at = B.shapeCompliance / dt2;
        const float denom = wi + at;
        if (denom < 1e-8f) break;
        const float sc = B.shapeBlend * (wi / denom);
        out = make_float3(p4.x + ddx * sc, p4.y + ddy * sc, p4.z + ddz * sc);
        break;
    }
    if (epi && epilogue_node(epi, idx, out, p4.w, B, ea)) return;
    Pout[idx] = make_float4(out.x, out.y, out.z, p4.w);
}

// ---------------------------------------------------------------------------
// Deformed shapes, in four steps per Gaussian (forward.cu:13063-13225).

// 1. Local deformation gradient from the 1-ring: with P the rest offsets and
// Q the current offsets of the neighbours, G = (Q P^T + e I)(P P^T + e I)^-1,
// e = 1e-4 tr(P P^T) (1e-6 when the trace is 0).  The inverse is taken of
// (P P^T + e I) / m, m its mean diagonal, and rescaled by 1/m.  False when
// that normalised matrix is singular (|det| < 1e-9).
__device__ __forceinline__ bool ring_gradient(float3 x0, float3 x, int first, int count,
                                              const float3* __restrict__ rest, const float3* __restrict__ cur,
                                              const int* __restrict__ nbr, M3& G)
{
    M3 rr = m3_zero(), cr = m3_zero();   // sum p p^T, sum q p^T
    for (int k = 0; k < count; ++k) {
        const int j = nbr[first + k];
        const float3 a = rest[j], b = cur[j];
        const float pr[3] = {a.x - x0.x, a.y - x0.y, a.z - x0.z};
        const float qc[3] = {b.x - x.x, b.y - x.y, b.z - x.z};
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                rr.m[c][r] += pr[r] * pr[c];
                cr.m[c][r] += qc[r] * pr[c];
            }
    }
    const float trace = rr.m[0][0] + rr.m[1][1] + rr.m[2][2];
    const float ridge = trace > 0.0f ? trace * 1e-4f : 1e-6f;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {   // every entry takes the sum (so -0 off the diagonal becomes +0)
            const float d = c == r ? ridge : 0.0f;
            rr.m[c][r] += d;
            cr.m[c][r] += d;
        }
    const float mean = fmaxf((rr.m[0][0] + rr.m[1][1] + rr.m[2][2]) * (1.0f / 3.0f), 1e-30f);
    const float invMean = 1.0f / mean;
    M3 unit = rr, inv;
    m3_scale(unit, invMean);
    if (!m3_inverse_safe(unit, inv, 1e-9f)) return false;
    m3_scale(inv, invMean);
    G = m3_mul(cr, inv);
    return true;
}

// 2. Stretch and rotation of G with its singular values clamped to
// [1e-3, 1.5]: G^T G = V diag(s^2) V^T (float Jacobi), U = G V diag(1/s)
// (1/s = 0 below 1e-7), R = U V^T with U's first column negated when R is a
// reflection, stretch = V diag(s) V^T.
__device__ __forceinline__ void clamped_stretch(const M3& G, M3& stretch, M3& R)
{
    float s2[3];
    M3 V;
    m3_eig_sym(m3_mul(m3_transpose(G), G), s2, V);
    float sv[3], inv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        sv[k] = fminf(fmaxf(sqrtf(fmaxf(0.0f, s2[k])), 1e-3f), 1.5f);
        inv[k] = sv[k] > 1.0e-7f ? 1.0f / sv[k] : 0.0f;
    }
    M3 U = m3_mul(m3_mul(G, V), m3_diag(inv[0], inv[1], inv[2]));
    const M3 Vt = m3_transpose(V);
    R = m3_mul(U, Vt);
    if (m3_det(R) < 0.0f) {
#pragma unroll
        for (int r = 0; r < 3; ++r) U.m[0][r] *= -1.0f;
        R = m3_mul(U, Vt);
    }
    stretch = m3_mul(m3_mul(V, m3_diag(sv[0], sv[1], sv[2])), Vt);
}

// 3. Deformed covariance Sigma' = M^T M, M = diag(s0) R0 stretch R^T, R0 the
// rotation of the unit quaternion (w, x, y, z).
__device__ __forceinline__ M3 deformed_covariance(float sx, float sy, float sz, float w, float x, float y, float z,
                                                  const M3& stretch, const M3& R)
{
    M3 R0;
    R0.m[0][0] = 1.f - 2.f * (y * y + z * z); R0.m[0][1] = 2.f * (x * y - w * z); R0.m[0][2] = 2.f * (x * z + w * y);
    R0.m[1][0] = 2.f * (x * y + w * z); R0.m[1][1] = 1.f - 2.f * (x * x + z * z); R0.m[1][2] = 2.f * (y * z - w * x);
    R0.m[2][0] = 2.f * (x * z - w * y); R0.m[2][1] = 2.f * (y * z + w * x); R0.m[2][2] = 1.f - 2.f * (x * x + y * y);
    const M3 M = m3_mul(m3_mul(m3_mul(m3_diag(sx, sy, sz), R0), stretch), m3_transpose(R));
    return m3_mul(m3_transpose(M), M);
}

// Unit quaternion (w, x, y, z) of a proper rotation matrix F (rows r, columns
// c), by the largest of w^2, x^2, y^2, z^2 (Shepperd, J. Guidance and Control
// 1(3), 1978), normalised at the end.
__device__ __forceinline__ void rotation_to_quat(const double (&F)[3][3], double (&q)[4])
{
    const double tr = F[0][0] + F[1][1] + F[2][2];
    if (tr > 0.0) {
        const double d = 2.0 * sqrt(tr + 1.0);
        q[0] = 0.25 * d; q[1] = (F[2][1] - F[1][2]) / d; q[2] = (F[0][2] - F[2][0]) / d; q[3] = (F[1][0] - F[0][1]) / d;
    } else if (F[0][0] > F[1][1] && F[0][0] > F[2][2]) {
        const double d = 2.0 * sqrt(1.0 + F[0][0] - F[1][1] - F[2][2]);
        q[0] = (F[2][1] - F[1][2]) / d; q[1] = 0.25 * d; q[2] = (F[0][1] + F[1][0]) / d; q[3] = (F[0][2] + F[2][0]) / d;
    } else if (F[1][1] > F[2][2]) {
        const double d = 2.0 * sqrt(1.0 + F[1][1] - F[0][0] - F[2][2]);
        q[0] = (F[0][2] - F[2][0]) / d; q[1] = (F[0][1] + F[1][0]) / d; q[2] = 0.25 * d; q[3] = (F[1][2] + F[2][1]) / d;
    } else {
        const double d = 2.0 * sqrt(1.0 + F[2][2] - F[0][0] - F[1][1]);
        q[0] = (F[1][0] - F[0][1]) / d; q[1] = (F[0][2] + F[2][0]) / d; q[2] = (F[1][2] + F[2][1]) / d; q[3] = 0.25 * d;
    }
    const double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = q[k] / n;
}

// 4. Scales and orientation of a covariance with trace tr > 0: the
// symmetrised Sigma / (tr / 3) is diagonalised in fp64 by cyclic Jacobi
// (jacobi3.h; at most 32 sweeps, relative stop 1e-15); the eigenvector frame
// is made proper by negating its third axis; scale k = sqrt(lambda_k tr / 3),
// at least floorScale.
__device__ __forceinline__ void covariance_shape(const M3& Sigma, double tr, double floorScale, float* scale,
                                                 float* quat)
{
    const double norm = 3.0 / tr;
    double a[3][3], F[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) a[r][c] = 0.5 * ((double)Sigma.m[c][r] + (double)Sigma.m[r][c]) * norm;
    set_identity(F);
    jacobi_cyclic(a, F, 32, 1e-15, 1e-300);
    if (det_rows(F) < 0.0)
#pragma unroll
        for (int r = 0; r < 3; ++r) F[r][2] = -F[r][2];
#pragma unroll
    for (int k = 0; k < 3; ++k) scale[k] = (float)fmax(sqrt(fmax(a[k][k], 0.0) / norm), floorScale);
    double q[4];
    rotation_to_quat(F, q);
#pragma unroll
    for (int k = 0; k < 4; ++k) quat[k] = (float)q[k];
}

// One thread per Gaussian of the body: Gaussians that moved more than
// deformEps from rest and have at least 3 neighbours get the deformed scale
// and orientation; the others keep their own.  ctr[block] receives the
// block's count of deformed Gaussians.
__global__ void __launch_bounds__(256) deformed_shape_kernel(
    int nodeBegin, int n, const float3* __restrict__ rest, const float3* __restrict__ cur,
    const int* __restrict__ rowOff, const int* __restrict__ nbr, const float* __restrict__ scales,
    const float* __restrict__ rots, float deformEps, float* __restrict__ outScale, float* __restrict__ outQuat,
    unsigned int* __restrict__ ctr)
{
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    bool counted = false;
    if (li < n) {
        const int i = nodeBegin + li;
        const float* s0 = scales + 3 * li;
        const float* q0 = rots + 4 * li;
        const float3 x = cur[i], x0 = rest[i];
        const float mx = x.x - x0.x, my = x.y - x0.y, mz = x.z - x0.z;
        const float moved = sqrtf(mx * mx + my * my + mz * mz);
        const int first = rowOff[i], count = rowOff[i + 1] - first;
        M3 Sigma = m3_zero(), G, stretch, R;
        bool ok = moved > deformEps && count >= 3 && ring_gradient(x0, x, first, count, rest, cur, nbr, G);
        if (ok) {
            clamped_stretch(G, stretch, R);
            Sigma = deformed_covariance(s0[0], s0[1], s0[2], q0[0], q0[1], q0[2], q0[3], stretch, R);
            ok = m3_finite(Sigma);
        }
        const double tr = ok ? ((double)Sigma.m[0][0] + Sigma.m[1][1] + Sigma.m[2][2]) : 0.0;
        if (ok && tr > 0.0) {
            counted = true;
            covariance_shape(Sigma, tr, 1e-3 * (double)fmaxf(s0[0], fmaxf(s0[1], s0[2])), outScale + 3 * li,
                             outQuat + 4 * li);
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) outScale[3 * li + k] = s0[k];
#pragma unroll
            for (int k = 0; k < 4; ++k) outQuat[4 * li + k] = q0[k];
        }
    }
    // deformed count of the block: wave ballots summed in LDS, one partial per
    // block (every partial is written, so nothing needs clearing first)
    __shared__ unsigned int waveCount[4];
    const unsigned long long ballot = __ballot(counted);
    if ((threadIdx.x & 63) == 0) waveCount[threadIdx.x >> 6] = (unsigned int)__popcll(ballot);
    __syncthreads();
    if (threadIdx.x == 0) ctr[blockIdx.x] = waveCount[0] + waveCount[1] + waveCount[2] + waveCount[3];
}

// Sum of the per-block deformed counts into partials[blocks] (one block,
// fixed order: the total does not depend on timing).
__global__ void __launch_bounds__(256) shape_count_total_kernel(int blocks, unsigned int* __restrict__ partials)
{
    __shared__ unsigned int part[4];
    unsigned int s = 0;
    for (int k = threadIdx.x; k < blocks; k += 256) s += partials[k];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += (unsigned int)__shfl_xor((int)s, m, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partials[blocks] = part[0] + part[1] + part[2] + part[3];
}

// contactShapePackKernel (forward.cu:7358-7384)
__global__ void contact_pack_kernel(int nodeBegin, int n, const float* __restrict__ sc, const float* __restrict__ qt,
                                    const float* __restrict__ op, float tau, float rcap, float* __restrict__ M,
                                    float* __restrict__ rmax)
{
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= n) return;
    const int i = nodeBegin + li;
    const float a = op[li];
    const float k = (a > tau && tau > 0.0f) ? sqrtf(2.0f * logf(a / tau)) : 0.0f;
    float e0 = k * sc[3 * li], e1 = k * sc[3 * li + 1], e2 = k * sc[3 * li + 2];
    if (rcap > 0.0f) { e0 = fminf(e0, rcap); e1 = fminf(e1, rcap); e2 = fminf(e2, rcap); }
    if (!(isfinite(e0) && isfinite(e1) && isfinite(e2))) { e0 = e1 = e2 = 0.0f; }
    float w = qt[4 * li], x = qt[4 * li + 1], y = qt[4 * li + 2], z = qt[4 * li + 3];
    const float qn = sqrtf(w * w + x * x + y * y + z * z);
    if (qn > 1e-20f) { w /= qn; x /= qn; y /= qn; z /= qn; } else { w = 1.0f; x = y = z = 0.0f; }
    const float r00 = 1 - 2 * (y * y + z * z), r01 = 2 * (x * y - w * z), r02 = 2 * (x * z + w * y);
    const float r10 = 2 * (x * y + w * z), r11 = 1 - 2 * (x * x + z * z), r12 = 2 * (y * z - w * x);
    const float r20 = 2 * (x * z - w * y), r21 = 2 * (y * z + w * x), r22 = 1 - 2 * (x * x + y * y);
    const float l0 = e0 * e0, l1 = e1 * e1, l2 = e2 * e2;
    float* m = &M[6 * (size_t)i];
    m[0] = r00 * r00 * l0 + r01 * r01 * l1 + r02 * r02 * l2;
    m[1] = r00 * r10 * l0 + r01 * r11 * l1 + r02 * r12 * l2;
    m[2] = r00 * r20 * l0 + r01 * r21 * l1 + r02 * r22 * l2;
    m[3] = r10 * r10 * l0 + r11 * r11 * l1 + r12 * r12 * l2;
    m[4] = r10 * r20 * l0 + r11 * r21 * l1 + r12 * r22 * l2;
    m[5] = r20 * r20 * l0 + r21 * r21 * l1 + r22 * r22 * l2;
    rmax[i] = fmaxf(e0, fmaxf(e1, e2));
}

// contactGroundRadiusKernel (forward.cu:7394-7399) for bodies with ground contact
__global__ void contact_ground_kernel(int nPad, const int* __restrict__ tileBody, const BodyDev* __restrict__ bodies,
                                      const float* __restrict__ M, float* __restrict__ rg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad) return;
    const BodyDev& B = bodies[__builtin_amdgcn_readfirstlane(tileBody[i / kTile])];
    if (i >= B.nodeEnd || !B.contactGround) return;
    rg[i] = contact_support(&M[6 * (size_t)i], B.gn[0], B.gn[1], B.gn[2]);
}

// ---------------------------------------------------------------------------
static inline int nbk(int n, int bs) { return (n + bs - 1) / bs; }

void launch_shape_match(hipStream_t s, const NodeArgs& na, const float4* Pin, float4* Pout, const float3* rest,
                        const int* rowOff, const int* nbr, const ShapeClusterArgs& sca, unsigned epi,
                        const EpiArgs& ea, const Tuning& tu)
{
    const int per = (tu.shapeClusterWaves == 1 || tu.shapeClusterWaves == 2) ? tu.shapeClusterWaves : 4;
    if (sca.K > 0)
        shape_cluster_kernel<<<xcd_grid(nbk(sca.K, per * kClustersPerWave)), 64 * per, 0, s>>>(sca.K, sca, Pin, rest, na.tileBody,
                                                                         na.bodies);
    shape_match_kernel<<<xcd_grid(nbk(na.nPad, 256)), 256, 0, s>>>(na.nPad, na.tileBody, na.bodies, Pin, Pout, rest, rowOff,
                                                        nbr, sca, epi, ea);
}

int deformed_shape_blocks(int n) { return nbk(n, 256); }

void launch_deformed_shapes(hipStream_t s, int nodeBegin, int n, const float3* rest, const float3* cur,
                            const int* rowOff, const int* nbr, const float* scales, const float* rots, float deformEps,
                            float* outScale, float* outQuat, unsigned int* counts)
{
    const int blocks = nbk(n, 256);
    if (n > 0)
        deformed_shape_kernel<<<blocks, 256, 0, s>>>(nodeBegin, n, rest, cur, rowOff, nbr, scales, rots, deformEps,
                                                    outScale, outQuat, counts);
    shape_count_total_kernel<<<1, 256, 0, s>>>(blocks, counts);
}

// ---------------------------------------------------------------------------
// Deformed shapes without allocation: the per-block counts go to a
// file-scope device array instead of an engine buffer.  Each device holds its
// own copy of the array; its address on the current device is looked up once
// per device (hipGetSymbolAddress) and kept.  Calls on one device share the
// array, so calls that overlap on one device race.
constexpr int kShapeScratchBlocks = 1 << 16;   // blocks of 256 Gaussians: up to 16,777,216 per call
static __device__ __align__(256) unsigned int scratch[kShapeScratchBlocks + 1];

int deformed_shape_scratch_max() { return kShapeScratchBlocks * 256; }

bool deformed_shape_scratch(unsigned int** out, std::string& err)
{
    static std::mutex mu;
    static unsigned int* perDevice[256] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= 256) {
        err = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return false;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!perDevice[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return false;
        }
        perDevice[dev] = static_cast<unsigned int*>(p);
    }
    *out = perDevice[dev];
    return true;
}

void launch_contact_pack(hipStream_t s, int nodeBegin, int n, const float* sc, const float* qt, const float* op,
                         float tau, float rcap, float* M, float* rmax)
{
    if (n > 0) contact_pack_kernel<<<nbk(n, 256), 256, 0, s>>>(nodeBegin, n, sc, qt, op, tau, rcap, M, rmax);
}

void launch_contact_ground(hipStream_t s, const NodeArgs& na, const float* M, float* rg)
{
    contact_ground_kernel<<<nbk(na.nPad, 256), 256, 0, s>>>(na.nPad, na.tileBody, na.bodies, M, rg);
}

}  // namespace xpbd
