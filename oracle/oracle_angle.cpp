/*
 * oracle_angle.cpp — TEST INFRASTRUCTURE ONLY (see xpbd_oracle.cpp).
 *
 * Umbrella angle constraint restated serially:
 *   - angle_build: the neighbour pairing of ensureChainmailGPU
 *     (forward.cu:10708-10819): neighbours of each centre sorted by their angle
 *     in a tangent basis (u = the longest neighbour direction, v = the most
 *     perpendicular one), each paired with the next one around; a gap wider
 *     than 120 degrees leaves the entry unpaired;
 *   - angle_pass: xpbdAngleAccumulateKernel (4990-5096) with the
 *     contributions added in serial order (centre, entry, role) instead of by
 *     atomics, then xpbdAngleApplyKernel (5098-5130).
 */
#include "oracle_angle.h"

#include <algorithm>
#include <cmath>

namespace xo {

AnglePairs ang;

void angle_build(const std::vector<V3>& pos)
{
    const int n = g.n;
    const int m = g.off[n];
    ang.pairNext.assign(m, -1);
    ang.restCos.assign(m, 2.0f);
    struct Sample {
        int edge, nidx;
        V3 vec;
        float len, angle;
    };
    std::vector<Sample> samples;
    for (int i = 0; i < n; ++i) {
        const int off = g.off[i], cnt = g.off[i + 1] - off;
        if (cnt < 2) continue;
        const V3 p0 = pos[i];
        samples.clear();
        for (int k = 0; k < cnt; ++k) {
            const int e = off + k, j = g.nbr[e];
            if (j < 0 || j >= n) continue;
            const V3 pn = pos[j];
            const V3 v{pn.x - p0.x, pn.y - p0.y, pn.z - p0.z};
            const float l = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
            if (l <= 1e-8f) continue;
            samples.push_back({e, j, v, l, 0.0f});
        }
        if (samples.size() < 2) continue;
        size_t iU = 0;
        for (size_t q = 1; q < samples.size(); ++q)
            if (samples[q].len > samples[iU].len) iU = q;
        const V3 u{samples[iU].vec.x / samples[iU].len, samples[iU].vec.y / samples[iU].len,
                   samples[iU].vec.z / samples[iU].len};
        float bestL2 = -1.0f;
        V3 best{0, 0, 0};
        for (const Sample& sm : samples) {
            const float du = sm.vec.x * u.x + sm.vec.y * u.y + sm.vec.z * u.z;
            const V3 perp{sm.vec.x - du * u.x, sm.vec.y - du * u.y, sm.vec.z - du * u.z};
            const float l2 = perp.x * perp.x + perp.y * perp.y + perp.z * perp.z;
            if (l2 > bestL2) { bestL2 = l2; best = perp; }
        }
        V3 va;
        if (bestL2 > 1e-12f) {
            const float invL = 1.0f / std::sqrt(bestL2);
            va = V3{best.x * invL, best.y * invL, best.z * invL};
        } else {
            const V3 a = (std::fabs(u.x) < 0.9f) ? V3{1.0f, 0.0f, 0.0f} : V3{0.0f, 1.0f, 0.0f};
            const V3 c{u.y * a.z - u.z * a.y, u.z * a.x - u.x * a.z, u.x * a.y - u.y * a.x};
            const float cl2 = c.x * c.x + c.y * c.y + c.z * c.z;
            if (cl2 <= 1e-12f) continue;
            const float invL = 1.0f / std::sqrt(cl2);
            va = V3{c.x * invL, c.y * invL, c.z * invL};
        }
        for (Sample& sm : samples) {
            const float x = sm.vec.x * u.x + sm.vec.y * u.y + sm.vec.z * u.z;
            const float y = sm.vec.x * va.x + sm.vec.y * va.y + sm.vec.z * va.z;
            sm.angle = std::atan2(y, x);
        }
        std::sort(samples.begin(), samples.end(), [](const Sample& a, const Sample& b) { return a.angle < b.angle; });
        const float twoPi = 2.0f * 3.14159265358979323846f;
        const float maxGap = 2.0f * 3.14159265358979323846f / 3.0f;
        for (size_t q = 0; q < samples.size(); ++q) {
            const Sample& cur = samples[q];
            const Sample& nxt = samples[(q + 1) % samples.size()];
            float gap = nxt.angle - cur.angle;
            if (gap < 0.0f) gap += twoPi;
            if (gap > maxGap) {
                ang.pairNext[cur.edge] = -1;
                ang.restCos[cur.edge] = 2.0f;
                continue;
            }
            ang.pairNext[cur.edge] = nxt.nidx;
            const V3 n1{cur.vec.x / cur.len, cur.vec.y / cur.len, cur.vec.z / cur.len};
            const V3 n2{nxt.vec.x / nxt.len, nxt.vec.y / nxt.len, nxt.vec.z / nxt.len};
            float c = n1.x * n2.x + n1.y * n2.y + n1.z * n2.z;
            c = std::fmin(1.0f, std::fmax(-1.0f, c));
            ang.restCos[cur.edge] = c;
        }
    }
    ang.built = true;
}

static bool finite3(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

void angle_pass(const std::vector<V3>& pin, std::vector<V3>& pout)
{
    const int n = g.n;
    std::vector<V3> sum(n, V3{0, 0, 0});
    std::vector<int> count(n, 0);
    const float dt2 = std::fmax(s.dt * s.dt, 1e-8f);
    const float alpha = std::fmax(s.angleCompliance, 0.0f) / dt2;
    for (int i = 0; i < n; ++i) {
        const int off = g.off[i], cnt = g.off[i + 1] - off;
        if (cnt < 2) continue;
        const V3 p0 = pin[i];
        const float w0 = g.invMass[i] * s.invMassScale;
        for (int k = 0; k < cnt; ++k) {
            const int e0 = off + k;
            const int i1 = g.nbr[e0], i2 = ang.pairNext[e0];
            if (i1 < 0 || i1 >= n || i2 < 0 || i2 >= n || i1 == i2) continue;
            const float rc = ang.restCos[e0];
            if (!std::isfinite(rc) || rc < -1.0f || rc > 1.0f) continue;
            const V3 v1 = sub(pin[i1], p0), v2 = sub(pin[i2], p0);
            const float l1 = len(v1), l2 = len(v2);
            if (l1 <= 1e-8f || l2 <= 1e-8f) continue;
            const V3 n1 = mul(v1, 1.0f / l1), n2 = mul(v2, 1.0f / l2);
            float cosCurr = dot(n1, n2);
            cosCurr = std::fmin(1.0f, std::fmax(-1.0f, cosCurr));
            const float C = cosCurr - rc;
            if (!std::isfinite(C)) continue;
            const V3 g1 = mul(sub(n2, mul(n1, dot(n2, n1))), 1.0f / l1);
            const V3 g2 = mul(sub(n1, mul(n2, dot(n1, n2))), 1.0f / l2);
            const V3 g0 = mul(add(g1, g2), -1.0f);
            const float w1 = g.invMass[i1] * s.invMassScale, w2 = g.invMass[i2] * s.invMassScale;
            const float denom = w0 * len2(g0) + w1 * len2(g1) + w2 * len2(g2);
            if (denom < 1e-8f) continue;
            float dl = -C / (denom + alpha);
            if (!std::isfinite(dl)) continue;
            dl = std::fmax(std::fmin(dl, 0.1f), -0.1f);
            V3 dp0 = mul(g0, dl * w0), dp1 = mul(g1, dl * w1), dp2 = mul(g2, dl * w2);
            if (!finite3(dp0) || !finite3(dp1) || !finite3(dp2)) continue;
            const float maxStep = std::fmax(1e-6f, 0.25f * std::fmin(l1, l2));
            const float a0 = len(dp0);
            if (a0 > maxStep && a0 > 1e-8f) dp0 = mul(dp0, maxStep / a0);
            const float a1 = len(dp1);
            if (a1 > maxStep && a1 > 1e-8f) dp1 = mul(dp1, maxStep / a1);
            const float a2 = len(dp2);
            if (a2 > maxStep && a2 > 1e-8f) dp2 = mul(dp2, maxStep / a2);
            if (w0 > 0.0f) { sum[i] = add(sum[i], dp0); ++count[i]; }
            if (w1 > 0.0f) { sum[i1] = add(sum[i1], dp1); ++count[i1]; }
            if (w2 > 0.0f) { sum[i2] = add(sum[i2], dp2); ++count[i2]; }
        }
    }
    const float blend = std::fmin(std::fmax(s.angleBlend, 0.0f), 1.0f);
    for (int i = 0; i < n; ++i) {
        const V3 p = pin[i];
        if (g.invMass[i] <= 0.0f || count[i] <= 0) { pout[i] = p; continue; }
        const V3 avg = mul(sum[i], 1.0f / float(count[i]));
        if (!finite3(avg)) { pout[i] = p; continue; }
        pout[i] = add(p, mul(avg, blend));
    }
}

}  // namespace xo
