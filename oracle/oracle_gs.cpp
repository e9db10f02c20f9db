/*
 * oracle_gs.cpp — TEST INFRASTRUCTURE ONLY (see xpbd_oracle.cpp).
 *
 * Graph-coloured Gauss-Seidel distance and volume passes, restated serially.
 *
 * Pinning (tests/test_oracle_gs.py):
 *   - colouring validity: no two edges of one colour share a node, no two
 *     clusters of one colour share a node; colour counts within the greedy
 *     bounds (maxDeg <= C <= 2 maxDeg - 1 for edges; C >= max clusters per
 *     node for volumes);
 *   - the edge set equals numpy's dedupe of the directed CSR, with the rest
 *     length the mean of the directed duplicates;
 *   - first-fit order: every edge / cluster of colour c > 0 conflicts with
 *     some earlier-processed element of each colour below c (greedy
 *     minimality, which a wrong processing order or colour choice breaks);
 *   - a single edge: one GS iteration reaches the two-particle closed form;
 *   - with one colour holding everything (a graph whose edges share no node)
 *     GS equals one Jacobi iteration without under-relaxation.
 */
#include "oracle_gs.h"

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "oracle_volume.h"

namespace xo {

DistanceGS dgs;
VolumeGS vgs;

// buildDistanceGSEdges (forward.cu:10407-10517).  The reference sorts the
// directed entries with std::sort by (a, b); duplicates of one pair are then
// summed in sorted order.  A stable sort gives the same result whenever a pair
// has at most two directed entries (the sum of two floats commutes), which is
// what a symmetric CSR holds.
void distance_gs_build()
{
    dgs = DistanceGS();
    const int n = g.n;
    struct E { int a, b; float rest, stiff; int dup; };
    std::vector<E> dir;
    for (int i = 0; i < n; ++i)
        for (int e = g.off[i]; e < g.off[i + 1]; ++e) {
            const int j = g.nbr[e];
            if (j < 0 || j >= n || j == i) continue;
            dir.push_back({std::min(i, j), std::max(i, j), g.ndist[e], g.nstiff[e], 1});
        }
    std::stable_sort(dir.begin(), dir.end(), [](const E& x, const E& y) { return x.a != y.a ? x.a < y.a : x.b < y.b; });
    std::vector<E> und;
    for (const E& d : dir) {
        if (!und.empty() && und.back().a == d.a && und.back().b == d.b) {
            und.back().rest += d.rest;
            und.back().stiff += d.stiff;
            und.back().dup++;
        } else {
            und.push_back(d);
        }
    }
    for (E& u : und) {
        if (u.dup == 1) ++dgs.oneWay;
        u.rest /= float(u.dup);
        u.stiff /= float(u.dup);
    }
    const int m = (int)und.size();
    dgs.built = true;
    if (m == 0) { dgs.colorOff.assign(1, 0); return; }
    std::vector<int> deg(n, 0);
    for (const E& u : und) { deg[u.a]++; deg[u.b]++; }
    dgs.maxDeg = *std::max_element(deg.begin(), deg.end());
    // first-fit: the lowest colour used by neither end node
    const int words = (2 * dgs.maxDeg + 63) / 64;
    std::vector<unsigned long long> used((size_t)n * words, 0ull);
    std::vector<int> color(m);
    int nc = 0;
    for (int e = 0; e < m; ++e) {
        unsigned long long* ua = &used[(size_t)und[e].a * words];
        unsigned long long* ub = &used[(size_t)und[e].b * words];
        int c = -1;
        for (int w = 0; w < words && c < 0; ++w) {
            const unsigned long long fr = ~(ua[w] | ub[w]);
            if (!fr) continue;
            int bit = 0;
            while (!((fr >> bit) & 1ull)) ++bit;
            c = w * 64 + bit;
        }
        ua[c >> 6] |= 1ull << (c & 63);
        ub[c >> 6] |= 1ull << (c & 63);
        color[e] = c;
        nc = std::max(nc, c + 1);
    }
    // colour-major counting sort, (a, b) order inside a colour
    dgs.colorOff.assign(nc + 1, 0);
    for (int e = 0; e < m; ++e) dgs.colorOff[color[e] + 1]++;
    for (int c = 0; c < nc; ++c) dgs.colorOff[c + 1] += dgs.colorOff[c];
    std::vector<int> fill(dgs.colorOff.begin(), dgs.colorOff.end() - 1);
    dgs.a.resize(m); dgs.b.resize(m); dgs.rest.resize(m); dgs.stiff.resize(m);
    for (int e = 0; e < m; ++e) {
        const int p = fill[color[e]]++;
        dgs.a[p] = und[e].a; dgs.b[p] = und[e].b;
        dgs.rest[p] = und[e].rest; dgs.stiff[p] = und[e].stiff;
    }
    dgs.lam.assign(m, 0.0f);
}

// xpbdDistanceGSColorKernel (forward.cu:6576-6621), colour by colour in place;
// no under-relaxation (12252-12262).
void distance_gs_iteration(std::vector<V3>& pos, float dt)
{
    const int nc = (int)dgs.colorOff.size() - 1;
    for (int c = 0; c < nc; ++c)
        for (int e = dgs.colorOff[c]; e < dgs.colorOff[c + 1]; ++e) {
            const int i = dgs.a[e], j = dgs.b[e];
            const float wi = g.invMass[i] * s.invMassScale;
            const float wj = g.invMass[j] * s.invMassScale;
            if (wi + wj <= 0.0f) continue;
            const V3 pi = pos[i], pj = pos[j];
            const V3 diff = sub(pi, pj);
            const float d = len(diff);
            if (d < 1e-7f) continue;
            const float dt2 = std::fmax(dt * dt, 1e-8f);
            const float st = std::fmax(dgs.stiff[e] * s.stiffnessScale, 1e-6f);
            const float at = (s.distCompliance / st) / dt2;
            const float denom = (wi + wj) + at;
            if (denom < 1e-8f) continue;
            const float C = d - dgs.rest[e];
            const float lam = dgs.lam[e];
            const float dl = (-C - at * lam) / denom;
            dgs.lam[e] = lam + dl;
            const V3 nrm = mul(diff, 1.0f / d);
            if (wi > 0.0f) pos[i] = add(pi, mul(nrm, wi * dl));
            if (wj > 0.0f) pos[j] = sub(pj, mul(nrm, wj * dl));
        }
}

// buildVolumeGSColors (forward.cu:9745-9812): leaders in index order take the
// lowest colour not used by any cluster containing the leader or a member.
void volume_gs_build()
{
    vgs = VolumeGS();
    vgs.built = true;
    vgs.version = vol.version;
    const int n = g.n;
    double work = 0.0;
    for (int j = 0; j < n; ++j) {
        const int h = vol.revOff[j + 1] - vol.revOff[j];
        vgs.lowerBound = std::max(vgs.lowerBound, h);
        work += (double)h * h;
    }
    std::vector<int> leaders;
    for (int i = 0; i < n; ++i)
        if (vol.isLeader[i]) leaders.push_back(i);
    if (leaders.empty()) return;
    if (work > 4.0e9) { vgs.skipped = true; return; }
    std::vector<int> color(n, -1), stamp;
    int nc = 0;
    for (int l : leaders) {
        if ((int)stamp.size() < nc + 1) stamp.resize(nc + 1, -1);
        auto forbid = [&](int j) {
            for (int k = vol.revOff[j]; k < vol.revOff[j + 1]; ++k) {
                const int m = vol.revIdx[k];
                if (m >= 0 && m < n && color[m] >= 0) stamp[color[m]] = l;
            }
        };
        forbid(l);
        for (int k = vol.memOff[l]; k < vol.memOff[l + 1]; ++k) forbid(vol.memIdx[k]);
        int c = 0;
        while (c < nc && stamp[c] == l) ++c;
        color[l] = c;
        nc = std::max(nc, c + 1);
    }
    vgs.colorOff.assign(nc + 1, 0);
    for (int l : leaders) vgs.colorOff[color[l] + 1]++;
    for (int c = 0; c < nc; ++c) vgs.colorOff[c + 1] += vgs.colorOff[c];
    std::vector<int> fill(vgs.colorOff.begin(), vgs.colorOff.end() - 1);
    vgs.order.resize(leaders.size());
    for (int l : leaders) vgs.order[fill[color[l]]++] = l;
}

namespace {
M3 zero_m3()
{
    M3 m;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) m.m[c][r] = 0.0f;
    return m;
}
void outer_add(M3& S, float x, float y, float z)
{
    const float v[3] = {x, y, z};
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) S.m[c][r] += v[r] * v[c];
}
V3 mv(const M3& M, V3 v)
{
    return V3{M.m[0][0] * v.x + M.m[1][0] * v.y + M.m[2][0] * v.z,
              M.m[0][1] * v.x + M.m[1][1] * v.y + M.m[2][1] * v.z,
              M.m[0][2] * v.x + M.m[1][2] * v.y + M.m[2][2] * v.z};
}
}  // namespace

// xpbdVolumeGSColorKernel (forward.cu:9041-9183) with the isotropic
// regulariser (the default, 413-415), colour by colour, moving every member in
// place with the 0.25 restLen step clamp.
void volume_gs_iteration(std::vector<V3>& pos, float dt)
{
    const int nc = (int)vgs.colorOff.size() - 1;
    for (int c = 0; c < nc; ++c)
        for (int t = vgs.colorOff[c]; t < vgs.colorOff[c + 1]; ++t) {
            const int idx = vgs.order[t];
            if (vol.matType[idx] != 0) continue;
            const V3 pi = pos[idx];
            const int off = vol.memOff[idx], cnt = vol.memOff[idx + 1] - off;
            V3 cc = pi;
            int used = 0;
            for (int k = 0; k < cnt; ++k) { cc = add(cc, pos[vol.memIdx[off + k]]); ++used; }
            if (used < 3) continue;
            const float invN = 1.0f / (float)(used + 1);
            cc = mul(cc, invN);
            M3 S = zero_m3();
            { const V3 d = sub(pi, cc); outer_add(S, d.x, d.y, d.z); }
            for (int k = 0; k < cnt; ++k) { const V3 d = sub(pos[vol.memIdx[off + k]], cc); outer_add(S, d.x, d.y, d.z); }
            for (int a = 0; a < 3; ++a)
                for (int r = 0; r < 3; ++r) S.m[a][r] *= invN;
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
                const V3 gv = mv(Si, sub(pi, cc));
                const float gx = gs * gv.x, gy = gs * gv.y, gz = gs * gv.z;
                denom += (g.invMass[idx] * s.invMassScale) * (gx * gx + gy * gy + gz * gz);
            }
            for (int k = 0; k < cnt; ++k) {
                const int j = vol.memIdx[off + k];
                const V3 gv = mv(Si, sub(pos[j], cc));
                const float gx = gs * gv.x, gy = gs * gv.y, gz = gs * gv.z;
                denom += (g.invMass[j] * s.invMassScale) * (gx * gx + gy * gy + gz * gz);
            }
            const float dt2 = std::fmax(dt * dt, 1e-8f);
            const float alphaT = s.volCompliance / dt2;
            const float den = denom + alphaT;
            if (!(den > 1e-12f)) continue;
            const float lam = vol.lam[idx];
            const float dLam = (-C - alphaT * lam) / den;
            if (!std::isfinite(dLam)) continue;
            vol.lam[idx] = lam + dLam;
            const float coef = dLam * gs;
            for (int k = -1; k < cnt; ++k) {
                const int j = (k < 0) ? idx : vol.memIdx[off + k];
                const float wj = g.invMass[j] * s.invMassScale;
                if (wj <= 0.0f) continue;
                const V3 xj = pos[j];
                const V3 gv = mv(Si, sub(xj, cc));
                V3 dp{gv.x * wj * coef, gv.y * wj * coef, gv.z * wj * coef};
                if (!std::isfinite(dp.x) || !std::isfinite(dp.y) || !std::isfinite(dp.z)) continue;
                const float rl = vol.restLen[j];
                if (rl > 1e-12f) {
                    const float maxStep = 0.25f * rl;
                    const float l = len(dp);
                    if (l > maxStep && l > 1e-20f) dp = mul(dp, maxStep / l);
                }
                pos[j] = add(xj, dp);
            }
        }
}

}  // namespace xo
