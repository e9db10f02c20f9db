/*
 * oracle_volume.cpp — TEST INFRASTRUCTURE ONLY (see xpbd_oracle.cpp).
 * Serial restatement of the reference volume constraint:
 *   cluster build  rebuildVolumeClusters (forward.cu:9814-10403)
 *   rest precompute precomputeVolumeGaussianKernel (7691-7844)
 *   solve          xpbdVolumeClusterSolveKernel (8531-8664), thread-per-cluster
 *   apply          xpbdVolumeGatherApplyKernel (9187-9243)
 *   J statistics   computeVolumeJKernel (9428-9481) + host stats (12658-12736)
 * glm semantics are restated with a column-major 3x3 (m[col][row]).
 *
 * Pinned by tests/test_oracle_volume.py: rest J == 1 for every active cluster
 * (the reference's own self-check, 10370-10397), J of a uniformly scaled
 * cluster == s^3, 4-point clusters against |V/V0| of the tetrahedron, k-ring
 * membership against scipy breadth-first order, the reverse CSR against a
 * transpose, and eigenvalues against numpy.linalg.eigh.
 */
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "oracle_volume.h"

namespace xo {

VolumeState vol;

static inline M3 zero3() { M3 m; for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) m.m[c][r] = 0.0f; return m; }

// glm::outerProduct(v, v) accumulated: m[c] += v * v[c]
static inline void add_outer(M3& S, float x, float y, float z)
{
    const float v[3] = {x, y, z};
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) S.m[c][r] += v[r] * v[c];
}

// glm determinant of mat3 (func_matrix.inl:253-262)
float det3(const M3& M)
{
    const auto& m = M.m;
    return +m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2])
           - m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2])
           + m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
}

// sigmaScaleAndShapeDet (forward.cu:7666-7676)
bool scale_shape_det(const M3& S, float& s, float& detHat)
{
    const float sc = (S.m[0][0] + S.m[1][1] + S.m[2][2]) * (1.0f / 3.0f);
    if (!std::isfinite(sc) || sc <= 1e-24f) return false;
    M3 h;
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) h.m[c][r] = S.m[c][r] / sc;
    const float d = det3(h);
    if (!std::isfinite(d) || d <= 1e-6f) return false;
    s = sc;
    detHat = d;
    return true;
}

// inverse3x3_safe (forward.cu:1696-1720)
bool inverse_safe(const M3& M, M3& inv, float epsDet)
{
    const auto& m = M.m;
    const float a00 = m[0][0], a01 = m[0][1], a02 = m[0][2];
    const float a10 = m[1][0], a11 = m[1][1], a12 = m[1][2];
    const float a20 = m[2][0], a21 = m[2][1], a22 = m[2][2];
    const float det = a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) + a02 * (a10 * a21 - a11 * a20);
    if (std::fabs(det) < epsDet) return false;
    const float id = 1.0f / det;
    inv.m[0][0] = (a11 * a22 - a12 * a21) * id;
    inv.m[0][1] = -(a01 * a22 - a02 * a21) * id;
    inv.m[0][2] = (a01 * a12 - a02 * a11) * id;
    inv.m[1][0] = -(a10 * a22 - a12 * a20) * id;
    inv.m[1][1] = (a00 * a22 - a02 * a20) * id;
    inv.m[1][2] = -(a00 * a12 - a02 * a10) * id;
    inv.m[2][0] = (a10 * a21 - a11 * a20) * id;
    inv.m[2][1] = -(a00 * a21 - a01 * a20) * id;
    inv.m[2][2] = (a00 * a11 - a01 * a10) * id;
    return true;
}

// glm mat3 * vec3 (type_mat3x3.inl:468-474)
static inline V3 mat_vec(const M3& M, V3 v)
{
    const auto& m = M.m;
    return V3{m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z,
              m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z,
              m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z};
}

// jacobiEigenDecomposition3x3 + eigenDecomposition_glm (forward.cu:782-897):
// eigenvalues ascending.
void eig_sym3(const M3& S, float ev[3])
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
            if (ev[i] > ev[j]) std::swap(ev[i], ev[j]);
}

// ---------------------------------------------------------------------------
// Cluster build (forward.cu:9814-10240), memberSelect = 1 (farthest point),
// no cover maximum (the C ABI path).
// ---------------------------------------------------------------------------
static std::vector<int> compute_members(int i, int k, int cap, std::vector<int>& stamp)
{
    const int n = g.n;
    const int HARD_CAP = 8192;
    std::vector<int> collected;
    if (k <= 1) {
        for (int e = g.off[i]; e < g.off[i + 1]; ++e) {
            const int j = g.nbr[e];
            if (j >= 0 && j < n && j != i) collected.push_back(j);
        }
    } else {
        std::vector<int> frontier{i}, next;
        stamp[i] = i;
        for (int d = 0; d < k && !frontier.empty() && (int)collected.size() < HARD_CAP; ++d) {
            next.clear();
            bool capped = false;
            for (int f : frontier) {
                for (int e = g.off[f]; e < g.off[f + 1]; ++e) {
                    const int j = g.nbr[e];
                    if (j < 0 || j >= n || stamp[j] == i) continue;
                    stamp[j] = i;
                    collected.push_back(j);
                    next.push_back(j);
                    if ((int)collected.size() >= HARD_CAP) { capped = true; break; }
                }
                if (capped) break;
            }
            frontier.swap(next);
            if (capped) break;
        }
    }
    if ((int)collected.size() <= cap) return collected;
    // farthest-point subsample (9999-10031)
    auto dist2 = [](V3 a, V3 b) {
        const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
        return dx * dx + dy * dy + dz * dz;
    };
    const V3 cpos = g.rest[i];
    const int Mc = (int)collected.size();
    std::vector<float> md(Mc);
    std::vector<char> taken(Mc, 0);
    int cur = 0;
    float best = -1.0f;
    for (int m = 0; m < Mc; ++m) {
        const float dd = dist2(g.rest[collected[m]], cpos);
        md[m] = dd;
        if (dd > best) { best = dd; cur = m; }
    }
    std::vector<int> pick;
    for (int s2 = 0; s2 < cap; ++s2) {
        taken[cur] = 1;
        pick.push_back(collected[cur]);
        const V3 cp = g.rest[collected[cur]];
        int nxt = -1;
        float nbest = -1.0f;
        for (int m = 0; m < Mc; ++m) {
            if (taken[m]) continue;
            const float dd = dist2(g.rest[collected[m]], cp);
            if (dd < md[m]) md[m] = dd;
            if (md[m] > nbest) { nbest = md[m]; nxt = m; }
        }
        if (nxt < 0) break;
        cur = nxt;
    }
    return pick;
}

void volume_build_reference(int k, int cap, int leaderR)
{
    const int n = g.n;
    std::vector<char> isLeader(n, 0);
    std::vector<std::vector<int>> members(n);
    std::vector<int> coverCnt(n, 0), stamp(n, -1);
    if (leaderR <= 1) {
        for (int i = 0; i < n; ++i) { isLeader[i] = 1; members[i] = compute_members(i, k, cap, stamp); }
    } else {
        const int target = std::max(1, std::min(6, leaderR));
        std::vector<int> stampR(n, -1), frontier, next;
        for (int seed = 0; seed < n; ++seed) {
            if (coverCnt[seed] >= target) continue;
            if (stampR[seed] == -2) continue;
            isLeader[seed] = 1;
            members[seed] = compute_members(seed, k, cap, stamp);
            ++coverCnt[seed];
            for (int j : members[seed]) ++coverCnt[j];
            frontier.assign(1, seed);
            stampR[seed] = -2;
            for (int d = 0; d < leaderR - 1 && !frontier.empty(); ++d) {
                next.clear();
                for (int f : frontier)
                    for (int e = g.off[f]; e < g.off[f + 1]; ++e) {
                        const int j = g.nbr[e];
                        if (j < 0 || j >= n || stampR[j] == -2) continue;
                        stampR[j] = -2;
                        next.push_back(j);
                    }
                frontier.swap(next);
            }
        }
        bool progressed = true;
        while (progressed) {
            progressed = false;
            for (int i = 0; i < n; ++i) {
                if (coverCnt[i] >= target || isLeader[i]) continue;
                isLeader[i] = 1;
                members[i] = compute_members(i, k, cap, stamp);
                ++coverCnt[i];
                for (int j : members[i]) ++coverCnt[j];
                progressed = true;
            }
        }
    }
    volume_set_topology(isLeader, members, 4);
}

void volume_set_topology(const std::vector<char>& isLeader, const std::vector<std::vector<int>>& members,
                         int minMembers)
{
    static long long versions = 0;
    const int n = g.n;
    vol.version = ++versions;
    vol.isLeader = isLeader;
    vol.minMembers = minMembers;
    vol.memOff.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) vol.memOff[i + 1] = vol.memOff[i] + (int)members[i].size();
    vol.memIdx.resize(vol.memOff[n]);
    for (int i = 0; i < n; ++i) std::copy(members[i].begin(), members[i].end(), vol.memIdx.begin() + vol.memOff[i]);
    // reverse CSR: self entry for leaders first, then clusters in index order (10199-10240)
    std::vector<int> revCnt(n, 0);
    for (int i = 0; i < n; ++i) if (isLeader[i]) ++revCnt[i];
    for (int i = 0; i < n; ++i) for (int j : members[i]) ++revCnt[j];
    vol.revOff.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) vol.revOff[i + 1] = vol.revOff[i] + revCnt[i];
    vol.revIdx.resize(vol.revOff[n]);
    std::vector<int> cursor(vol.revOff.begin(), vol.revOff.end() - 1);
    for (int i = 0; i < n; ++i) if (isLeader[i]) vol.revIdx[cursor[i]++] = i;
    for (int i = 0; i < n; ++i) for (int j : members[i]) vol.revIdx[cursor[j]++] = i;
    volume_precompute();
    vol.lam.assign(n, 0.0f);
    vol.cC.assign(n, V3{0, 0, 0});
    vol.cS.assign((size_t)n * 6, 0.0f);
    vol.coef.assign(n, 0.0f);
    vol.built = true;
}

// precomputeVolumeGaussianKernel (7691-7844) without the mixture rest (off by default)
void volume_precompute()
{
    const int n = g.n;
    vol.detRest.assign(n, 0.0f);
    vol.Vrest.assign(n, 0.0f);
    vol.restLen.assign(n, 0.0f);
    vol.matType.assign(n, 3);
    for (int idx = 0; idx < n; ++idx) {
        const V3 p0g = g.rest[idx];
        float dmin = FLT_MAX;
        for (int e = g.off[idx]; e < g.off[idx + 1]; ++e) {
            const int j = g.nbr[e];
            if (j < 0 || j >= n || j == idx) continue;
            const float dl = len(sub(g.rest[j], p0g));
            if (dl > 1e-12f) dmin = std::fmin(dmin, dl);
        }
        vol.restLen[idx] = (dmin < FLT_MAX) ? dmin : 0.0f;
        const int off = vol.memOff[idx], cnt = vol.memOff[idx + 1] - off;
        if (cnt < vol.minMembers) continue;
        V3 c = g.rest[idx];
        int used = 0;
        for (int k = 0; k < cnt; ++k) { c = add(c, g.rest[vol.memIdx[off + k]]); ++used; }
        if (used < 3) continue;
        const float invN = 1.0f / (float)(used + 1);
        c = mul(c, invN);
        M3 S = zero3();
        { const V3 d = sub(g.rest[idx], c); add_outer(S, d.x, d.y, d.z); }
        for (int k = 0; k < cnt; ++k) { const V3 d = sub(g.rest[vol.memIdx[off + k]], c); add_outer(S, d.x, d.y, d.z); }
        for (int cc = 0; cc < 3; ++cc) for (int r = 0; r < 3; ++r) S.m[cc][r] *= invN;
        float sR, dh;
        if (!scale_shape_det(S, sR, dh)) { vol.matType[idx] = 4; continue; }
        float ev[3];
        eig_sym3(S, ev);
        const float lmax = std::fmax(ev[2], 1e-24f);
        const float a1 = ev[0] / lmax, a2 = ev[1] / lmax;
        const float thr = s.volAniso;
        vol.matType[idx] = (a1 > thr) ? 0 : (a2 > thr) ? 1 : 2;
        const float d = sR * sR * sR * dh;
        vol.detRest[idx] = d;
        vol.Vrest[idx] = (4.0f / 3.0f) * 3.14159265f * std::sqrt(std::fmax(d, 0.0f));
    }
}

// xpbdVolumeClusterSolveKernel (8531-8664) with the isotropic regulariser (mode 1)
static void volume_solve(const std::vector<V3>& pin, float dt)
{
    const int n = g.n;
    XO_PARALLEL_FOR
    for (int idx = 0; idx < n; ++idx) {
        vol.coef[idx] = 0.0f;
        if (vol.matType[idx] != 0) continue;
        const V3 pi = pin[idx];
        const int off = vol.memOff[idx], cnt = vol.memOff[idx + 1] - off;
        V3 c = pi;
        int used = 0;
        for (int k = 0; k < cnt; ++k) { c = add(c, pin[vol.memIdx[off + k]]); ++used; }
        if (used < 3) continue;
        const float nf = (float)(used + 1);
        const float invN = 1.0f / nf;
        c = mul(c, invN);
        M3 S = zero3();
        { const V3 d = sub(pi, c); add_outer(S, d.x, d.y, d.z); }
        for (int k = 0; k < cnt; ++k) { const V3 d = sub(pin[vol.memIdx[off + k]], c); add_outer(S, d.x, d.y, d.z); }
        for (int cc = 0; cc < 3; ++cc) for (int r = 0; r < 3; ++r) S.m[cc][r] *= invN;
        float sC, dhC;
        if (!scale_shape_det(S, sC, dhC)) continue;
        const float detCur = sC * sC * sC * dhC;
        const float dRest = vol.detRest[idx];
        if (dRest < 1e-30f) continue;
        float J = std::sqrt(detCur / dRest);
        if (!std::isfinite(J) || J <= 0.0f) continue;
        if (J > 3.0f || J < 0.333f) J = std::fmin(3.0f, std::fmax(0.333f, J));
        const float C = J - 1.0f;
        if (std::fabs(C) < 1e-4f) continue;
        M3 Sreg = S;
        if (dhC < 0.2f) {
            const float d = 0.06f * sC;
            Sreg.m[0][0] += d; Sreg.m[1][1] += d; Sreg.m[2][2] += d;
        }
        M3 Si;
        if (!inverse_safe(Sreg, Si, 1e-20f)) continue;
        const float gs = J * invN;
        float denom = 0.0f;
        {
            const V3 gv = mul(mat_vec(Si, sub(pi, c)), gs);
            denom += (g.invMass[idx] * s.invMassScale) * (gv.x * gv.x + gv.y * gv.y + gv.z * gv.z);
        }
        for (int k = 0; k < cnt; ++k) {
            const int j = vol.memIdx[off + k];
            const V3 gv = mul(mat_vec(Si, sub(pin[j], c)), gs);
            denom += (g.invMass[j] * s.invMassScale) * (gv.x * gv.x + gv.y * gv.y + gv.z * gv.z);
        }
        const float dt2 = std::fmax(dt * dt, 1e-8f);
        const float alphaT = s.volCompliance / dt2;
        const float den = denom + alphaT;
        if (!(den > 1e-12f)) continue;
        const float lam = vol.lam[idx];
        const float dLam = (-C - alphaT * lam) / den;
        if (!std::isfinite(dLam)) continue;
        vol.lam[idx] = lam + dLam;
        vol.cC[idx] = c;
        float* o = &vol.cS[(size_t)6 * idx];
        o[0] = Si.m[0][0]; o[1] = Si.m[1][1]; o[2] = Si.m[2][2];
        o[3] = Si.m[1][0]; o[4] = Si.m[2][0]; o[5] = Si.m[2][1];
        vol.coef[idx] = dLam * gs;
    }
}

// xpbdVolumeGatherApplyKernel (9187-9243)
static void volume_apply(const std::vector<V3>& pin, std::vector<V3>& pout, float underRelax)
{
    const int n = g.n;
    XO_PARALLEL_FOR
    for (int idx = 0; idx < n; ++idx) {
        const V3 pi = pin[idx];
        const float wi = g.invMass[idx] * s.invMassScale;
        if (wi <= 0.0f) { pout[idx] = pi; continue; }
        V3 sum{0, 0, 0};
        int received = 0;
        for (int k = vol.revOff[idx]; k < vol.revOff[idx + 1]; ++k) {
            const int i = vol.revIdx[k];
            const float coef = vol.coef[i];
            if (coef == 0.0f) continue;
            const float* o = &vol.cS[(size_t)6 * i];
            M3 Si;
            Si.m[0][0] = o[0]; Si.m[1][1] = o[1]; Si.m[2][2] = o[2];
            Si.m[0][1] = Si.m[1][0] = o[3];
            Si.m[0][2] = Si.m[2][0] = o[4];
            Si.m[1][2] = Si.m[2][1] = o[5];
            const V3 gv = mat_vec(Si, sub(pi, vol.cC[i]));
            const V3 dp{gv.x * wi * coef, gv.y * wi * coef, gv.z * wi * coef};
            if (!std::isfinite(dp.x) || !std::isfinite(dp.y) || !std::isfinite(dp.z)) continue;
            sum = add(sum, dp);
            ++received;
        }
        if (received <= 0) { pout[idx] = pi; continue; }
        V3 avg = mul(sum, 1.0f / (float)received);
        if (!std::isfinite(avg.x) || !std::isfinite(avg.y) || !std::isfinite(avg.z)) { pout[idx] = pi; continue; }
        const float rl = vol.restLen[idx];
        if (rl > 1e-12f) {
            const float maxStep = 0.25f * rl;
            const float l = len(avg);
            if (l > maxStep && l > 1e-20f) avg = mul(avg, maxStep / l);
        }
        pout[idx] = add(pi, mul(avg, underRelax));
    }
}

void volume_iteration(std::vector<V3>*& pin, std::vector<V3>*& pout)
{
    volume_solve(*pin, s.dt);
    volume_apply(*pin, *pout, s.underRelax);
    std::swap(pin, pout);
}

// computeVolumeJKernel (9428-9481)
void volume_J(const std::vector<V3>& pos, std::vector<float>& J)
{
    const int n = g.n;
    J.assign(n, 1.0f);
    for (int idx = 0; idx < n; ++idx) {
        if (vol.matType[idx] != 0) continue;
        const int off = vol.memOff[idx], cnt = vol.memOff[idx + 1] - off;
        V3 c = pos[idx];
        int used = 0;
        for (int k = 0; k < cnt; ++k) { c = add(c, pos[vol.memIdx[off + k]]); ++used; }
        if (used < 3) continue;
        const float invN = 1.0f / (float)(used + 1);
        c = mul(c, invN);
        M3 S = zero3();
        { const V3 d = sub(pos[idx], c); add_outer(S, d.x, d.y, d.z); }
        for (int k = 0; k < cnt; ++k) { const V3 d = sub(pos[vol.memIdx[off + k]], c); add_outer(S, d.x, d.y, d.z); }
        for (int cc = 0; cc < 3; ++cc) for (int r = 0; r < 3; ++r) S.m[cc][r] *= invN;
        float sC, dh;
        if (!scale_shape_det(S, sC, dh)) continue;
        const float detCur = sC * sC * sC * dh;
        const float dRest = vol.detRest[idx];
        if (dRest < 1e-30f) continue;
        const float j = std::sqrt(detCur / dRest);
        if (std::isfinite(j) && j > 0.0f) J[idx] = j;
    }
}

// Volume-weighted J over active clusters (12671-12700)
void volume_stats(const std::vector<V3>& pos, float& mean, float& stdv)
{
    std::vector<float> J;
    volume_J(pos, J);
    double wSum = 0, wJ = 0, wJ2 = 0;
    for (int i = 0; i < g.n; ++i) {
        if (vol.matType[i] != 0) continue;
        const float j = J[i];
        if (!std::isfinite(j)) continue;
        const double v = (double)vol.Vrest[i];
        if (v > 0.0 && std::isfinite(v)) { wSum += v; wJ += v * j; wJ2 += v * (double)j * j; }
    }
    if (wSum > 0.0) {
        const double m = wJ / wSum;
        mean = (float)m;
        stdv = (float)std::sqrt(std::fmax(0.0, wJ2 / wSum - m * m));
    }
}

}  // namespace xo
