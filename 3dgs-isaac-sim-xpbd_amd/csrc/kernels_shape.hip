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
// Deformed shapes (forward.cu:13063-13225): A = (Q P^T + aI)(P P^T + aI)^-1
// from the 1-ring, singular values clamped to [1e-3, 1.5], M = S0 R0 S_f R^T,
// Sigma' = M^T M, then a cyclic fp64 Jacobi eigen-decomposition (13022-13061)
// of Sigma' / (tr / 3) gives the scales and the quaternion.
#include <mutex>
#include <string>

#include "kernels.h"
#include "mat3.h"
#include "node_ops.h"
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
__device__ void sym_eig3_f64(double a[3][3], double ev[3], double V[3][3])
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) V[r][c] = (r == c) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 32; ++sweep) {
        const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        const double diag = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
        if (off <= 1e-15 * diag + 1e-300) break;
        for (int k = 0; k < 3; ++k) {
            const int p = (k == 2) ? 1 : 0, q = (k == 0) ? 1 : 2;
            const double apq = a[p][q];
            if (fabs(apq) <= 1e-300) continue;
            const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
            const double t = ((theta >= 0.0) ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            const double c = 1.0 / sqrt(t * t + 1.0);
            const double s = t * c;
            for (int m = 0; m < 3; ++m) {
                const double amp = a[m][p], amq = a[m][q];
                a[m][p] = c * amp - s * amq;
                a[m][q] = s * amp + c * amq;
            }
            for (int m = 0; m < 3; ++m) {
                const double apm = a[p][m], aqm = a[q][m];
                a[p][m] = c * apm - s * aqm;
                a[q][m] = s * apm + c * aqm;
            }
            for (int m = 0; m < 3; ++m) {
                const double vmp = V[m][p], vmq = V[m][q];
                V[m][p] = c * vmp - s * vmq;
                V[m][q] = s * vmp + c * vmq;
            }
        }
    }
    ev[0] = a[0][0];
    ev[1] = a[1][1];
    ev[2] = a[2][2];
}

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
        const float s0x = scales[3 * li], s0y = scales[3 * li + 1], s0z = scales[3 * li + 2];
        const float q0w = rots[4 * li], q0x = rots[4 * li + 1], q0y = rots[4 * li + 2], q0z = rots[4 * li + 3];
        const float3 x = cur[i], x0 = rest[i];
        const float dxv = x.x - x0.x, dyv = x.y - x0.y, dzv = x.z - x0.z;
        const float diff = sqrtf(dxv * dxv + dyv * dyv + dzv * dzv);
        const int off = rowOff[i], kn = rowOff[i + 1] - off;
        bool deformed = false;
        M3 Sigma = m3_zero();
        if (diff > deformEps && kn >= 3) {
            M3 PPt = m3_zero(), QPt = m3_zero();
            for (int k = 0; k < kn; ++k) {
                const int j = nbr[off + k];
                const float3 rj = rest[j], cj = cur[j];
                const float p[3] = {rj.x - x0.x, rj.y - x0.y, rj.z - x0.z};
                const float q[3] = {cj.x - x.x, cj.y - x.y, cj.z - x.z};
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        PPt.m[c][r] += p[r] * p[c];
                        QPt.m[c][r] += q[r] * p[c];
                    }
            }
            const float tr = PPt.m[0][0] + PPt.m[1][1] + PPt.m[2][2];
            const float alpha = (tr > 0.0f) ? (tr * 1e-4f) : 1e-6f;
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const float d = (c == r) ? alpha : 0.0f;
                    PPt.m[c][r] += d;
                    QPt.m[c][r] += d;
                }
            const float sP = fmaxf((PPt.m[0][0] + PPt.m[1][1] + PPt.m[2][2]) * (1.0f / 3.0f), 1e-30f);
            const float invSP = 1.0f / sP;
            M3 Ph = PPt;
            m3_scale(Ph, invSP);
            M3 invHat;
            if (m3_inverse_safe(Ph, invHat, 1e-9f)) {
                M3 invPPt = invHat;
                m3_scale(invPPt, invSP);
                const M3 A = m3_mul(QPt, invPPt);
                const M3 AtA = m3_mul(m3_transpose(A), A);
                float S2[3];
                M3 V;
                m3_eig_sym(AtA, S2, V);
                float sv[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) sv[k] = fminf(fmaxf(sqrtf(fmaxf(0.0f, S2[k])), 1e-3f), 1.5f);
                const float e = 1.0e-7f;
                const M3 Sinv = m3_diag(sv[0] > e ? 1.0f / sv[0] : 0.0f, sv[1] > e ? 1.0f / sv[1] : 0.0f,
                                        sv[2] > e ? 1.0f / sv[2] : 0.0f);
                M3 U = m3_mul(m3_mul(A, V), Sinv);
                M3 R = m3_mul(U, m3_transpose(V));
                if (m3_det(R) < 0.0f) {
                    U.m[0][0] *= -1.0f; U.m[0][1] *= -1.0f; U.m[0][2] *= -1.0f;
                    R = m3_mul(U, m3_transpose(V));
                }
                const M3 Sf = m3_mul(m3_mul(V, m3_diag(sv[0], sv[1], sv[2])), m3_transpose(V));
                const M3 S = m3_diag(s0x, s0y, s0z);
                const float r = q0w, qx = q0x, qy = q0y, qz = q0z;
                M3 R0;
                R0.m[0][0] = 1.f - 2.f * (qy * qy + qz * qz); R0.m[0][1] = 2.f * (qx * qy - r * qz); R0.m[0][2] = 2.f * (qx * qz + r * qy);
                R0.m[1][0] = 2.f * (qx * qy + r * qz); R0.m[1][1] = 1.f - 2.f * (qx * qx + qz * qz); R0.m[1][2] = 2.f * (qy * qz - r * qx);
                R0.m[2][0] = 2.f * (qx * qz - r * qy); R0.m[2][1] = 2.f * (qy * qz + r * qx); R0.m[2][2] = 1.f - 2.f * (qx * qx + qy * qy);
                const M3 M = m3_mul(m3_mul(m3_mul(S, R0), Sf), m3_transpose(R));
                Sigma = m3_mul(m3_transpose(M), M);
                deformed = m3_finite(Sigma);
            }
        }
        const double tr = deformed ? ((double)Sigma.m[0][0] + Sigma.m[1][1] + Sigma.m[2][2]) : 0.0;
        if (!deformed || !(tr > 0.0)) {
            outScale[3 * li] = s0x; outScale[3 * li + 1] = s0y; outScale[3 * li + 2] = s0z;
            outQuat[4 * li] = q0w; outQuat[4 * li + 1] = q0x; outQuat[4 * li + 2] = q0y; outQuat[4 * li + 3] = q0z;
        } else {
            counted = true;
            const double inv = 3.0 / tr;
            double a[3][3];
            for (int rr = 0; rr < 3; ++rr)
                for (int c = 0; c < 3; ++c) a[rr][c] = 0.5 * ((double)Sigma.m[c][rr] + (double)Sigma.m[rr][c]) * inv;
            double ev[3], E[3][3];
            sym_eig3_f64(a, ev, E);
            const double det = E[0][0] * (E[1][1] * E[2][2] - E[1][2] * E[2][1]) -
                               E[0][1] * (E[1][0] * E[2][2] - E[1][2] * E[2][0]) +
                               E[0][2] * (E[1][0] * E[2][1] - E[1][1] * E[2][0]);
            if (det < 0.0)
                for (int rr = 0; rr < 3; ++rr) E[rr][2] = -E[rr][2];
            const double sFloor = 1e-3 * (double)fmaxf(s0x, fmaxf(s0y, s0z));
            for (int k = 0; k < 3; ++k) outScale[3 * li + k] = (float)fmax(sqrt(fmax(ev[k], 0.0) / inv), sFloor);
            double w, qx, qy, qz;
            const double t = E[0][0] + E[1][1] + E[2][2];
            if (t > 0.0) {
                const double sq = 2.0 * sqrt(t + 1.0);
                w = 0.25 * sq; qx = (E[2][1] - E[1][2]) / sq; qy = (E[0][2] - E[2][0]) / sq; qz = (E[1][0] - E[0][1]) / sq;
            } else if (E[0][0] > E[1][1] && E[0][0] > E[2][2]) {
                const double sq = 2.0 * sqrt(1.0 + E[0][0] - E[1][1] - E[2][2]);
                w = (E[2][1] - E[1][2]) / sq; qx = 0.25 * sq; qy = (E[0][1] + E[1][0]) / sq; qz = (E[0][2] + E[2][0]) / sq;
            } else if (E[1][1] > E[2][2]) {
                const double sq = 2.0 * sqrt(1.0 + E[1][1] - E[0][0] - E[2][2]);
                w = (E[0][2] - E[2][0]) / sq; qx = (E[0][1] + E[1][0]) / sq; qy = 0.25 * sq; qz = (E[1][2] + E[2][1]) / sq;
            } else {
                const double sq = 2.0 * sqrt(1.0 + E[2][2] - E[0][0] - E[1][1]);
                w = (E[1][0] - E[0][1]) / sq; qx = (E[0][2] + E[2][0]) / sq; qy = (E[1][2] + E[2][1]) / sq; qz = 0.25 * sq;
            }
            const double qn = sqrt(w * w + qx * qx + qy * qy + qz * qz);
            outQuat[4 * li] = (float)(w / qn);
            outQuat[4 * li + 1] = (float)(qx / qn);
            outQuat[4 * li + 2] = (float)(qy / qn);
            outQuat[4 * li + 3] = (float)(qz / qn);
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
