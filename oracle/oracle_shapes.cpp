/*
 * oracle_shapes.cpp — TEST INFRASTRUCTURE ONLY (see xpbd_oracle.cpp).
 * Serial restatement of:
 *   per-node 1-ring shape matching  xpbdShapeMatchingKernel (forward.cu:6088-6383)
 *     with shapeMatchRotationDouble (5991-6086) or the float path
 *   deformed Gaussian shapes        deformedShapeKernel (13063-13225) + symEig3Double (13022-13061)
 *   contact ellipsoids              contactShapePackKernel (7358-7384)
 *
 * Pinned by tests/test_oracle_shapes.py: rigid rotations are tracked exactly
 * (shape-matching goal == current position), the rotation equals the polar
 * factor of numpy's SVD, rest positions return the input shape unchanged, a
 * rigid rotation R maps each Gaussian covariance to R S R^T, and a uniform
 * scale s (inside the clamp) multiplies every scale by s.
 */
#include <algorithm>
#include <cmath>
#include <vector>

#include "oracle_shapes.h"
#include "oracle_volume.h"

namespace xo {

static M3 mzero() { M3 a; for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) a.m[c][r] = 0.0f; return a; }
static M3 mdiag(float a, float b, float c) { M3 m = mzero(); m.m[0][0] = a; m.m[1][1] = b; m.m[2][2] = c; return m; }

// glm mat3 * mat3 (type_mat3x3.inl:486-518)
M3 mmul(const M3& A, const M3& B)
{
    M3 R;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            R.m[c][r] = A.m[0][r] * B.m[c][0] + A.m[1][r] * B.m[c][1] + A.m[2][r] * B.m[c][2];
    return R;
}

M3 mtrans(const M3& A)
{
    M3 R;
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) R.m[c][r] = A.m[r][c];
    return R;
}

static V3 mvec(const M3& M, V3 v)
{
    return V3{M.m[0][0] * v.x + M.m[1][0] * v.y + M.m[2][0] * v.z,
              M.m[0][1] * v.x + M.m[1][1] * v.y + M.m[2][1] * v.z,
              M.m[0][2] * v.x + M.m[1][2] * v.y + M.m[2][2] * v.z};
}

// eigenDecomposition_glm (forward.cu:782-897): ascending values, vectors as columns
void eig_sym3_vec(const M3& S, float ev[3], M3& Vout)
{
    float A[3][3];
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) A[r][c] = S.m[c][r];
    float V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int iter = 0; iter < 50; ++iter) {
        int p = 0, q = 1;
        float maxOff = std::fabs(A[0][1]);
        if (std::fabs(A[0][2]) > maxOff) { p = 0; q = 2; maxOff = std::fabs(A[0][2]); }
        if (std::fabs(A[1][2]) > maxOff) { p = 1; q = 2; maxOff = std::fabs(A[1][2]); }
        if (maxOff < 1e-6f) break;
        const float app = A[p][p], aqq = A[q][q], apq = A[p][q];
        const float phi = 0.5f * (aqq - app) / (apq + 1e-20f);
        const float t = (phi >= 0.0f) ? (1.0f / (phi + std::sqrt(1.0f + phi * phi)))
                                      : (1.0f / (phi - std::sqrt(1.0f + phi * phi)));
        const float c = 1.0f / std::sqrt(1.0f + t * t);
        const float s = t * c;
        A[p][p] = app - t * apq;
        A[q][q] = aqq + t * apq;
        A[p][q] = A[q][p] = 0.0f;
        for (int r = 0; r < 3; ++r) {
            if (r == p || r == q) continue;
            const float arp = A[r][p], arq = A[r][q];
            A[r][p] = A[p][r] = c * arp - s * arq;
            A[r][q] = A[q][r] = c * arq + s * arp;
        }
        for (int r = 0; r < 3; ++r) {
            const float vrp = V[r][p], vrq = V[r][q];
            V[r][p] = c * vrp - s * vrq;
            V[r][q] = s * vrp + c * vrq;
        }
    }
    ev[0] = A[0][0]; ev[1] = A[1][1]; ev[2] = A[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (ev[i] > ev[j]) {
                std::swap(ev[i], ev[j]);
                for (int r = 0; r < 3; ++r) std::swap(V[r][i], V[r][j]);
            }
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) Vout.m[c][r] = V[r][c];
}

// shapeMatchRotationDouble (forward.cu:5991-6086)
bool polar_double(const M3& A, M3& Rout)
{
    double a[3][3];
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) a[r][c] = (double)A.m[c][r];
    double m[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m[r][c] = a[0][r] * a[0][c] + a[1][r] * a[1][c] + a[2][r] * a[2][c];
    double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int iter = 0; iter < 50; ++iter) {
        bool done = true;
        int p = 0, q = 1;
        double maxOff = -1.0;
        for (int k = 0; k < 3; ++k) {
            const double off = std::fabs(m[P[k]][Q[k]]);
            if (off > 1e-12 * std::sqrt(std::fabs(m[P[k]][P[k]] * m[Q[k]][Q[k]])) + 1e-300) done = false;
            if (off > maxOff) { maxOff = off; p = P[k]; q = Q[k]; }
        }
        if (done) break;
        const double app = m[p][p], aqq = m[q][q], apq = m[p][q];
        const double phi = 0.5 * (aqq - app) / (apq + 1e-300);
        const double t = (phi >= 0.0) ? 1.0 / (phi + std::sqrt(1.0 + phi * phi)) : 1.0 / (phi - std::sqrt(1.0 + phi * phi));
        const double c = 1.0 / std::sqrt(1.0 + t * t);
        const double s = t * c;
        m[p][p] = app - t * apq;
        m[q][q] = aqq + t * apq;
        m[p][q] = m[q][p] = 0.0;
        for (int r = 0; r < 3; ++r) {
            if (r == p || r == q) continue;
            const double arp = m[r][p], arq = m[r][q];
            m[r][p] = m[p][r] = c * arp - s * arq;
            m[r][q] = m[q][r] = c * arq + s * arp;
        }
        for (int r = 0; r < 3; ++r) {
            const double vrp = v[r][p], vrq = v[r][q];
            v[r][p] = c * vrp - s * vrq;
            v[r][q] = s * vrp + c * vrq;
        }
    }
    double ev[3] = {m[0][0], m[1][1], m[2][2]};
    for (int i = 0; i < 2; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (ev[i] > ev[j]) {
                std::swap(ev[i], ev[j]);
                for (int r = 0; r < 3; ++r) std::swap(v[r][i], v[r][j]);
            }
    double S[3], Sinv[3];
    for (int i = 0; i < 3; ++i) S[i] = std::sqrt(ev[i] > 0.0 ? ev[i] : 0.0);
    const double Smax = (S[2] > 1e-300) ? S[2] : 1e-300;
    for (int i = 0; i < 3; ++i) Sinv[i] = (S[i] > 1e-12 * Smax) ? 1.0 / S[i] : 0.0;
    double u[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) u[r][c] = (a[r][0] * v[0][c] + a[r][1] * v[1][c] + a[r][2] * v[2][c]) * Sinv[c];
    if (S[0] < 0.08 * Smax) {
        const double cx = u[1][1] * u[2][2] - u[2][1] * u[1][2];
        const double cy = u[2][1] * u[0][2] - u[0][1] * u[2][2];
        const double cz = u[0][1] * u[1][2] - u[1][1] * u[0][2];
        const double l = std::sqrt(cx * cx + cy * cy + cz * cz);
        if (!(l > 1e-7)) return false;
        u[0][0] = cx / l; u[1][0] = cy / l; u[2][0] = cz / l;
    }
    double R[3][3];
    for (int pass = 0; pass < 2; ++pass) {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) R[r][c] = u[r][0] * v[c][0] + u[r][1] * v[c][1] + u[r][2] * v[c][2];
        const double det = R[0][0] * (R[1][1] * R[2][2] - R[1][2] * R[2][1]) -
                           R[0][1] * (R[1][0] * R[2][2] - R[1][2] * R[2][0]) +
                           R[0][2] * (R[1][0] * R[2][1] - R[1][1] * R[2][0]);
        if (det >= 0.0 || pass == 1) break;
        for (int r = 0; r < 3; ++r) u[r][0] = -u[r][0];
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            if (!std::isfinite(R[r][c])) return false;
            Rout.m[c][r] = (float)R[r][c];
        }
    return true;
}

static bool finite9(const M3& M)
{
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) if (!std::isfinite(M.m[c][r])) return false;
    return true;
}

// float polar path of xpbdShapeMatchingKernel (6228-6300)
static bool polar_float(const M3& A, M3& R)
{
    const M3 AtA = mmul(mtrans(A), A);
    float S2[3];
    M3 V;
    eig_sym3_vec(AtA, S2, V);
    const float Sv[3] = {std::sqrt(std::fmax(0.0f, S2[0])), std::sqrt(std::fmax(0.0f, S2[1])), std::sqrt(std::fmax(0.0f, S2[2]))};
    if (!std::isfinite(Sv[0]) || !std::isfinite(Sv[1]) || !std::isfinite(Sv[2])) return false;
    const float e = 1e-7f;
    const M3 Si = mdiag(Sv[0] > e ? 1.0f / Sv[0] : 0.0f, Sv[1] > e ? 1.0f / Sv[1] : 0.0f, Sv[2] > e ? 1.0f / Sv[2] : 0.0f);
    M3 U = mmul(mmul(A, V), Si);
    if (!finite9(U)) return false;
    if (Sv[0] < 0.08f * std::fmax(Sv[2], 1e-7f)) {
        const float* u1 = U.m[1];
        const float* u2 = U.m[2];
        const float c0 = u1[1] * u2[2] - u2[1] * u1[2];
        const float c1 = u1[2] * u2[0] - u2[2] * u1[0];
        const float c2 = u1[0] * u2[1] - u2[0] * u1[1];
        const float l = std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        if (!(l > 1e-7f)) return false;
        U.m[0][0] = c0 / l; U.m[0][1] = c1 / l; U.m[0][2] = c2 / l;
    }
    R = mmul(U, mtrans(V));
    if (!finite9(R)) return false;
    if (det3(R) < 0.0f) {
        for (int r = 0; r < 3; ++r) U.m[0][r] *= -1.0f;
        R = mmul(U, mtrans(V));
        if (!finite9(R)) return false;
    }
    return true;
}

// xpbdShapeMatchingKernel (6088-6383) as one Jacobi pass pin -> pout
void shape_matching(const std::vector<V3>& pin, std::vector<V3>& pout)
{
    const int n = g.n;
    XO_PARALLEL_FOR
    for (int idx = 0; idx < n; ++idx) {
        const V3 pi = pin[idx];
        const float wi = g.invMass[idx] * s.invMassScale;
        pout[idx] = pi;
        if (wi <= 0.0f) continue;
        const int off = g.off[idx], cnt = g.off[idx + 1] - off;
        if (cnt < 4) continue;
        V3 comRest = g.rest[idx], comCurr = pi;
        float cc = 1.0f;
        for (int k = 0; k < cnt; ++k) {
            const int j = g.nbr[off + k];
            comRest = add(comRest, g.rest[j]);
            comCurr = add(comCurr, pin[j]);
            cc += 1.0f;
        }
        if (cc < 2.0f) continue;
        const float invC = 1.0f / cc;
        comRest = mul(comRest, invC);
        comCurr = mul(comCurr, invC);
        {
            V3 dirSum{0, 0, 0};
            for (int k = 0; k < cnt; ++k) {
                const V3 q = sub(g.rest[g.nbr[off + k]], comRest);
                const float l = len(q);
                if (l > 1e-7f) dirSum = add(dirSum, mul(q, 1.0f / l));
            }
            const float bias = len(dirSum) / std::fmax((float)cnt, 1.0f);
            if (bias > 0.4f) continue;
        }
        M3 A = mzero();
        auto acc = [&](V3 p, V3 q) {
            const float pv[3] = {p.x, p.y, p.z}, qv[3] = {q.x, q.y, q.z};
            for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) A.m[c][r] += pv[r] * qv[c];
        };
        acc(sub(pi, comCurr), sub(g.rest[idx], comRest));
        for (int k = 0; k < cnt; ++k) {
            const int j = g.nbr[off + k];
            acc(sub(pin[j], comCurr), sub(g.rest[j], comRest));
        }
        float nsq = 0.0f;
        for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) nsq += A.m[c][r] * A.m[c][r];
        const float an = std::sqrt(nsq);
        if (!std::isfinite(an) || an < 1e-8f) continue;
        for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) A.m[c][r] /= an;
        M3 R = mdiag(1, 1, 1);
        if (s.shapeRobust) { if (!polar_double(A, R)) continue; }
        else if (!polar_float(A, R)) continue;
        const V3 rq = mvec(R, sub(g.rest[idx], comRest));
        const V3 goal = add(comCurr, rq);
        const V3 gd = sub(goal, pi);
        const float gd2 = len2(gd);
        if (!std::isfinite(gd2) || gd2 < 1e-18f) continue;
        // xpbdProjectTowardGoal (4972-4988)
        const float dt2 = std::fmax(s.dt * s.dt, 1e-8f);
        const float at = s.shapeCompliance / dt2;
        const float denom = wi + at;
        if (denom < 1e-8f) continue;
        const float scale = s.shapeBlend * (wi / denom);
        pout[idx] = add(pi, mul(sub(goal, pi), scale));
    }
}

// ---------------------------------------------------------------------------
// Cluster shape matching (extension; the reference's shape constraint with
// node clusters, e.g. k-means, in place of the graph 1-ring).  Per cluster of
// >= 4 members (members in increasing node order): rest centroid (once),
// current centroid, A = sum (p - c)(q - c0)^T normalised by its Frobenius norm,
// polar rotation (fp64 when robust), then every member is projected toward
// c + R (q - c0) with xpbdProjectTowardGoal (forward.cu:4972-4988), like
// the per-node pass (6088-6383).  The sums follow the GPU's wave layout: 64
// lane partials in member order, combined by a fixed butterfly tree.
// ---------------------------------------------------------------------------
ShapeClusters shc;

void shape_clusters_set(const int* clusterOf, int n)
{
    shc = ShapeClusters();
    if (!clusterOf) return;
    int K = 0;
    for (int i = 0; i < n; ++i) K = std::max(K, clusterOf[i] + 1);
    std::vector<int> cnt(K, 0);
    for (int i = 0; i < n; ++i)
        if (clusterOf[i] >= 0) ++cnt[clusterOf[i]];
    shc.off.assign(K + 1, 0);
    for (int k = 0; k < K; ++k) shc.off[k + 1] = shc.off[k] + cnt[k];
    shc.mem.resize(shc.off[K]);
    std::vector<int> fill(shc.off.begin(), shc.off.end() - 1);
    for (int i = 0; i < n; ++i)
        if (clusterOf[i] >= 0) shc.mem[fill[clusterOf[i]]++] = i;
    shc.clusterOf.assign(clusterOf, clusterOf + n);
    shc.comRest.assign(K, V3{0, 0, 0});
    for (int k = 0; k < K; ++k) {
        const int c = shc.off[k + 1] - shc.off[k];
        if (c <= 0) continue;
        V3 sr{0, 0, 0};
        for (int t = shc.off[k]; t < shc.off[k + 1]; ++t) sr = add(sr, g.rest[shc.mem[t]]);
        shc.comRest[k] = mul(sr, 1.0f / (float)c);
    }
    shc.on = true;
}

// Butterfly combination of 64 partials (entries [e0, e1) of each): round m
// (32, 16, ..., 1) replaces partial l by partial l + partial (l xor m).
static void tree_sum64(float (&part)[64][12], int e0, int e1)
{
    for (int m = 32; m >= 1; m >>= 1) {
        float next[64][12];
        for (int l = 0; l < 64; ++l)
            for (int e = e0; e < e1; ++e) next[l][e] = part[l][e] + part[l ^ m][e];
        for (int l = 0; l < 64; ++l)
            for (int e = e0; e < e1; ++e) part[l][e] = next[l][e];
    }
}

void shape_cluster_matching(const std::vector<V3>& pin, std::vector<V3>& pout)
{
    const int n = g.n, K = (int)shc.comRest.size();
    std::vector<V3> comCurr(K);
    std::vector<M3> rot(K);
    std::vector<char> ok(K, 0);
    XO_PARALLEL_FOR
    for (int k = 0; k < K; ++k) {
        const int b0 = shc.off[k], cnt = shc.off[k + 1] - b0;
        if (cnt < 4) continue;
        // sums in the GPU's fixed order: 64 lane partials (members l, l + 64,
        // ... in member order), combined by a butterfly tree
        float part[64][12];
        for (int l = 0; l < 64; ++l)
            for (int e = 0; e < 12; ++e) part[l][e] = 0.0f;
        for (int l = 0; l < 64; ++l)
            for (int t = l; t < cnt; t += 64) {
                const V3 p = pin[shc.mem[b0 + t]];
                part[l][0] = part[l][0] + p.x; part[l][1] = part[l][1] + p.y; part[l][2] = part[l][2] + p.z;
            }
        tree_sum64(part, 0, 3);
        const float invC = 1.0f / (float)cnt;
        const V3 cc{part[0][0] * invC, part[0][1] * invC, part[0][2] * invC};
        const V3 cr = shc.comRest[k];
        for (int l = 0; l < 64; ++l)
            for (int t = l; t < cnt; t += 64) {
                const int j = shc.mem[b0 + t];
                const V3 p = pin[j], q = g.rest[j];
                const float pv[3] = {p.x - cc.x, p.y - cc.y, p.z - cc.z}, qv[3] = {q.x - cr.x, q.y - cr.y, q.z - cr.z};
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r) part[l][3 + 3 * c + r] = part[l][3 + 3 * c + r] + pv[r] * qv[c];
            }
        tree_sum64(part, 3, 12);
        M3 A;
        for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) A.m[c][r] = part[0][3 + 3 * c + r];
        float nsq = 0.0f;
        for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) nsq += A.m[c][r] * A.m[c][r];
        const float an = std::sqrt(nsq);
        if (!std::isfinite(an) || an < 1e-8f) continue;
        for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) A.m[c][r] /= an;
        M3 R = mdiag(1, 1, 1);
        if (s.shapeRobust) { if (!polar_double(A, R)) continue; }
        else if (!polar_float(A, R)) continue;
        comCurr[k] = cc;
        rot[k] = R;
        ok[k] = 1;
    }
    XO_PARALLEL_FOR
    for (int idx = 0; idx < n; ++idx) {
        const V3 pi = pin[idx];
        pout[idx] = pi;
        const float wi = g.invMass[idx] * s.invMassScale;
        const int k = shc.clusterOf[idx];
        if (wi <= 0.0f || k < 0 || !ok[k]) continue;
        const V3 rq = mvec(rot[k], sub(g.rest[idx], shc.comRest[k]));
        const V3 goal = add(comCurr[k], rq);
        const V3 gd = sub(goal, pi);
        const float gd2 = len2(gd);
        if (!std::isfinite(gd2) || gd2 < 1e-18f) continue;
        const float dt2 = std::fmax(s.dt * s.dt, 1e-8f);
        const float at = s.shapeCompliance / dt2;
        const float denom = wi + at;
        if (denom < 1e-8f) continue;
        const float scale = s.shapeBlend * (wi / denom);
        pout[idx] = add(pi, mul(sub(goal, pi), scale));
    }
}

// symEig3Double (13022-13061)
static void sym_eig3_double(double a[3][3], double ev[3], double V[3][3])
{
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) V[r][c] = (r == c) ? 1.0 : 0.0;
    const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int sweep = 0; sweep < 32; ++sweep) {
        const double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
        const double diag = std::fabs(a[0][0]) + std::fabs(a[1][1]) + std::fabs(a[2][2]);
        if (off <= 1e-15 * diag + 1e-300) break;
        for (int k = 0; k < 3; ++k) {
            const int p = P[k], q = Q[k];
            const double apq = a[p][q];
            if (std::fabs(apq) <= 1e-300) continue;
            const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
            const double t = ((theta >= 0.0) ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
            const double c = 1.0 / std::sqrt(t * t + 1.0);
            const double sn = t * c;
            for (int m = 0; m < 3; ++m) {
                const double amp = a[m][p], amq = a[m][q];
                a[m][p] = c * amp - sn * amq;
                a[m][q] = sn * amp + c * amq;
            }
            for (int m = 0; m < 3; ++m) {
                const double apm = a[p][m], aqm = a[q][m];
                a[p][m] = c * apm - sn * aqm;
                a[q][m] = sn * apm + c * aqm;
            }
            for (int m = 0; m < 3; ++m) {
                const double vmp = V[m][p], vmq = V[m][q];
                V[m][p] = c * vmp - sn * vmq;
                V[m][q] = sn * vmp + c * vmq;
            }
        }
    }
    ev[0] = a[0][0]; ev[1] = a[1][1]; ev[2] = a[2][2];
}

// deformedShapeKernel (13063-13225), robust = 1 (g_renderFRobustHost)
int deformed_shapes(float deformEps, std::vector<float>& outScale, std::vector<float>& outQuat)
{
    const int n = g.n;
    outScale.assign((size_t)3 * n, 0.0f);
    outQuat.assign((size_t)4 * n, 0.0f);
    int deformedCount = 0;
#pragma omp parallel for num_threads(xo::g_threads) schedule(static) reduction(+ : deformedCount)
    for (int i = 0; i < n; ++i) {
        const float s0[3] = {g.scales[3 * i], g.scales[3 * i + 1], g.scales[3 * i + 2]};
        const float q0[4] = {g.rots[4 * i], g.rots[4 * i + 1], g.rots[4 * i + 2], g.rots[4 * i + 3]};
        const V3 x = g.x[i], x0 = g.rest[i];
        const float dx = x.x - x0.x, dy = x.y - x0.y, dz = x.z - x0.z;
        const float diff = std::sqrt(dx * dx + dy * dy + dz * dz);
        const int kn = g.off[i + 1] - g.off[i];
        bool deformed = false;
        M3 Sigma = mzero();
        if (diff > deformEps && kn >= 3) {
            M3 PPt = mzero(), QPt = mzero();
            for (int k = 0; k < kn; ++k) {
                const int j = g.nbr[g.off[i] + k];
                const float p[3] = {g.rest[j].x - x0.x, g.rest[j].y - x0.y, g.rest[j].z - x0.z};
                const float q[3] = {g.x[j].x - x.x, g.x[j].y - x.y, g.x[j].z - x.z};
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r) { PPt.m[c][r] += p[r] * p[c]; QPt.m[c][r] += q[r] * p[c]; }
            }
            const float tr = PPt.m[0][0] + PPt.m[1][1] + PPt.m[2][2];
            const float alpha = (tr > 0.0f) ? (tr * 1e-4f) : 1e-6f;
            for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) {
                const float d = (c == r) ? alpha : 0.0f;
                PPt.m[c][r] += d;
                QPt.m[c][r] += d;
            }
            const float sP = std::fmax((PPt.m[0][0] + PPt.m[1][1] + PPt.m[2][2]) * (1.0f / 3.0f), 1e-30f);
            const float invSP = 1.0f / sP;
            M3 Ph = PPt;
            for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) Ph.m[c][r] *= invSP;
            M3 invHat, invPPt;
            const bool ok = inverse_safe(Ph, invHat, 1e-9f);
            if (ok) {
                invPPt = invHat;
                for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) invPPt.m[c][r] *= invSP;
                const M3 A = mmul(QPt, invPPt);
                const M3 AtA = mmul(mtrans(A), A);
                float S2[3];
                M3 V;
                eig_sym3_vec(AtA, S2, V);
                float Sv[3];
                for (int k = 0; k < 3; ++k) {
                    Sv[k] = std::sqrt(std::fmax(0.0f, S2[k]));
                    Sv[k] = std::fmin(std::fmax(Sv[k], 1e-3f), 1.5f);
                }
                const float e = 1.0e-7f;
                const M3 Sinv = mdiag(Sv[0] > e ? 1.0f / Sv[0] : 0.0f, Sv[1] > e ? 1.0f / Sv[1] : 0.0f,
                                      Sv[2] > e ? 1.0f / Sv[2] : 0.0f);
                M3 U = mmul(mmul(A, V), Sinv);
                M3 R = mmul(U, mtrans(V));
                if (det3(R) < 0.0f) {
                    for (int r = 0; r < 3; ++r) U.m[0][r] *= -1.0f;
                    R = mmul(U, mtrans(V));
                }
                const M3 Sf = mmul(mmul(V, mdiag(Sv[0], Sv[1], Sv[2])), mtrans(V));
                const M3 S = mdiag(s0[0], s0[1], s0[2]);
                const float r = q0[0], qx = q0[1], qy = q0[2], qz = q0[3];
                M3 R0;
                // glm::mat3(9 scalars) fills columns
                R0.m[0][0] = 1.f - 2.f * (qy * qy + qz * qz); R0.m[0][1] = 2.f * (qx * qy - r * qz); R0.m[0][2] = 2.f * (qx * qz + r * qy);
                R0.m[1][0] = 2.f * (qx * qy + r * qz); R0.m[1][1] = 1.f - 2.f * (qx * qx + qz * qz); R0.m[1][2] = 2.f * (qy * qz - r * qx);
                R0.m[2][0] = 2.f * (qx * qz - r * qy); R0.m[2][1] = 2.f * (qy * qz + r * qx); R0.m[2][2] = 1.f - 2.f * (qx * qx + qy * qy);
                const M3 M = mmul(mmul(mmul(S, R0), Sf), mtrans(R));
                Sigma = mmul(mtrans(M), M);
                deformed = finite9(Sigma);
            }
        }
        const double tr = deformed ? ((double)Sigma.m[0][0] + Sigma.m[1][1] + Sigma.m[2][2]) : 0.0;
        if (!deformed || !(tr > 0.0)) {
            for (int k = 0; k < 3; ++k) outScale[3 * i + k] = s0[k];
            for (int k = 0; k < 4; ++k) outQuat[4 * i + k] = q0[k];
            continue;
        }
        const double inv = 3.0 / tr;
        double a[3][3];
        for (int rr = 0; rr < 3; ++rr)
            for (int c = 0; c < 3; ++c) a[rr][c] = 0.5 * ((double)Sigma.m[c][rr] + (double)Sigma.m[rr][c]) * inv;
        double ev[3], E[3][3];
        sym_eig3_double(a, ev, E);
        const double det = E[0][0] * (E[1][1] * E[2][2] - E[1][2] * E[2][1]) -
                           E[0][1] * (E[1][0] * E[2][2] - E[1][2] * E[2][0]) +
                           E[0][2] * (E[1][0] * E[2][1] - E[1][1] * E[2][0]);
        if (det < 0.0) for (int rr = 0; rr < 3; ++rr) E[rr][2] = -E[rr][2];
        const double sFloor = 1e-3 * (double)std::fmax(s0[0], std::fmax(s0[1], s0[2]));
        for (int k = 0; k < 3; ++k) outScale[3 * i + k] = (float)std::fmax(std::sqrt(std::fmax(ev[k], 0.0) / inv), sFloor);
        double w, qx, qy, qz;
        const double t = E[0][0] + E[1][1] + E[2][2];
        if (t > 0.0) {
            const double sq = 2.0 * std::sqrt(t + 1.0);
            w = 0.25 * sq; qx = (E[2][1] - E[1][2]) / sq; qy = (E[0][2] - E[2][0]) / sq; qz = (E[1][0] - E[0][1]) / sq;
        } else if (E[0][0] > E[1][1] && E[0][0] > E[2][2]) {
            const double sq = 2.0 * std::sqrt(1.0 + E[0][0] - E[1][1] - E[2][2]);
            w = (E[2][1] - E[1][2]) / sq; qx = 0.25 * sq; qy = (E[0][1] + E[1][0]) / sq; qz = (E[0][2] + E[2][0]) / sq;
        } else if (E[1][1] > E[2][2]) {
            const double sq = 2.0 * std::sqrt(1.0 + E[1][1] - E[0][0] - E[2][2]);
            w = (E[0][2] - E[2][0]) / sq; qx = (E[0][1] + E[1][0]) / sq; qy = 0.25 * sq; qz = (E[1][2] + E[2][1]) / sq;
        } else {
            const double sq = 2.0 * std::sqrt(1.0 + E[2][2] - E[0][0] - E[1][1]);
            w = (E[1][0] - E[0][1]) / sq; qx = (E[0][2] + E[2][0]) / sq; qy = (E[1][2] + E[2][1]) / sq; qz = 0.25 * sq;
        }
        const double qn = std::sqrt(w * w + qx * qx + qy * qy + qz * qz);
        outQuat[4 * i + 0] = (float)(w / qn);
        outQuat[4 * i + 1] = (float)(qx / qn);
        outQuat[4 * i + 2] = (float)(qy / qn);
        outQuat[4 * i + 3] = (float)(qz / qn);
        ++deformedCount;
    }
    return deformedCount;
}

// contactShapePackKernel (7358-7384)
void contact_pack(const float* sc, const float* qt, const float* op, float tau, float rcap, std::vector<float>& M,
                  std::vector<float>& rmax)
{
    const int n = g.n;
    M.assign((size_t)6 * n, 0.0f);
    rmax.assign(n, 0.0f);
    for (int i = 0; i < n; ++i) {
        const float a = op[i];
        const float k = (a > tau && tau > 0.0f) ? std::sqrt(2.0f * std::log(a / tau)) : 0.0f;
        float e0 = k * sc[3 * i], e1 = k * sc[3 * i + 1], e2 = k * sc[3 * i + 2];
        if (rcap > 0.0f) { e0 = std::fmin(e0, rcap); e1 = std::fmin(e1, rcap); e2 = std::fmin(e2, rcap); }
        if (!(std::isfinite(e0) && std::isfinite(e1) && std::isfinite(e2))) e0 = e1 = e2 = 0.0f;
        float w = qt[4 * i], x = qt[4 * i + 1], y = qt[4 * i + 2], z = qt[4 * i + 3];
        const float qn = std::sqrt(w * w + x * x + y * y + z * z);
        if (qn > 1e-20f) { w /= qn; x /= qn; y /= qn; z /= qn; } else { w = 1.0f; x = y = z = 0.0f; }
        const float r00 = 1 - 2 * (y * y + z * z), r01 = 2 * (x * y - w * z), r02 = 2 * (x * z + w * y);
        const float r10 = 2 * (x * y + w * z), r11 = 1 - 2 * (x * x + z * z), r12 = 2 * (y * z - w * x);
        const float r20 = 2 * (x * z - w * y), r21 = 2 * (y * z + w * x), r22 = 1 - 2 * (x * x + y * y);
        const float l0 = e0 * e0, l1 = e1 * e1, l2 = e2 * e2;
        float* m = &M[(size_t)6 * i];
        m[0] = r00 * r00 * l0 + r01 * r01 * l1 + r02 * r02 * l2;
        m[1] = r00 * r10 * l0 + r01 * r11 * l1 + r02 * r12 * l2;
        m[2] = r00 * r20 * l0 + r01 * r21 * l1 + r02 * r22 * l2;
        m[3] = r10 * r10 * l0 + r11 * r11 * l1 + r12 * r12 * l2;
        m[4] = r10 * r20 * l0 + r11 * r21 * l1 + r12 * r22 * l2;
        m[5] = r20 * r20 * l0 + r21 * r21 * l1 + r22 * r22 * l2;
        rmax[i] = std::fmax(e0, std::fmax(e1, e2));
    }
}

}  // namespace xo

// Test hooks on the oracle's 3x3 routines (column-major m[col][row] arrays).
extern "C" {

int oracle_polar_double(const float* A9, float* R9)
{
    xo::M3 A, R;
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) A.m[c][r] = A9[3 * c + r];
    if (!xo::polar_double(A, R)) return 0;
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) R9[3 * c + r] = R.m[c][r];
    return 1;
}

void oracle_eig_sym(const float* S9, float* ev3, float* V9)
{
    xo::M3 S, V;
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) S.m[c][r] = S9[3 * c + r];
    xo::eig_sym3_vec(S, ev3, V);
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) V9[3 * c + r] = V.m[c][r];
}

}  // extern "C"
