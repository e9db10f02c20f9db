// angle_build.cpp — neighbour pairing of the umbrella angle constraint and the
// per-node contribution lists of the MI355X angle pass.
//
// Pairing (forward.cu:10708-10819, ensureChainmailGPU): around each centre the
// neighbours are ordered by their angle in a tangent basis (u = direction of
// the longest neighbour, v = the neighbour direction most perpendicular to u,
// or a fixed axis crossed with u when all are collinear) and each is paired
// with the next one around; a gap wider than 120 degrees leaves the entry
// unpaired (pairNext -1, rest cosine 2).  The positions are those the graph is
// loaded with.
#include <algorithm>
#include <cmath>

#include "engine.h"

namespace xpbd {

void build_angle_pairs(const BodyInput& in, AngleTopo& t)
{
    const int n = in.n, m = in.rowOff[n];
    t.pairNext.assign(m, -1);
    t.restCos.assign(m, 2.0f);
    struct Sample {
        int edge, nidx;
        float vx, vy, vz, len, angle;
    };
    std::vector<Sample> smp;
    for (int i = 0; i < n; ++i) {
        const int off = in.rowOff[i], cnt = in.rowOff[i + 1] - off;
        if (cnt < 2) continue;
        const float px = in.pos[3 * i], py = in.pos[3 * i + 1], pz = in.pos[3 * i + 2];
        smp.clear();
        for (int k = 0; k < cnt; ++k) {
            const int e = off + k, j = in.nbr[e];
            if (j < 0 || j >= n) continue;
            const float vx = in.pos[3 * j] - px, vy = in.pos[3 * j + 1] - py, vz = in.pos[3 * j + 2] - pz;
            const float l = std::sqrt(vx * vx + vy * vy + vz * vz);
            if (l <= 1e-8f) continue;
            smp.push_back({e, j, vx, vy, vz, l, 0.0f});
        }
        if (smp.size() < 2) continue;
        size_t iU = 0;
        for (size_t q = 1; q < smp.size(); ++q)
            if (smp[q].len > smp[iU].len) iU = q;
        const float ux = smp[iU].vx / smp[iU].len, uy = smp[iU].vy / smp[iU].len, uz = smp[iU].vz / smp[iU].len;
        float bestL2 = -1.0f, bx = 0.0f, by = 0.0f, bz = 0.0f;
        for (const Sample& s : smp) {
            const float du = s.vx * ux + s.vy * uy + s.vz * uz;
            const float qx = s.vx - du * ux, qy = s.vy - du * uy, qz = s.vz - du * uz;
            const float l2 = qx * qx + qy * qy + qz * qz;
            if (l2 > bestL2) { bestL2 = l2; bx = qx; by = qy; bz = qz; }
        }
        float vax, vay, vaz;
        if (bestL2 > 1e-12f) {
            const float invL = 1.0f / std::sqrt(bestL2);
            vax = bx * invL; vay = by * invL; vaz = bz * invL;
        } else {
            const bool xAxis = std::fabs(ux) < 0.9f;
            const float ax = xAxis ? 1.0f : 0.0f, ay = xAxis ? 0.0f : 1.0f, az = 0.0f;
            const float cx = uy * az - uz * ay, cy = uz * ax - ux * az, cz = ux * ay - uy * ax;
            const float cl2 = cx * cx + cy * cy + cz * cz;
            if (cl2 <= 1e-12f) continue;
            const float invL = 1.0f / std::sqrt(cl2);
            vax = cx * invL; vay = cy * invL; vaz = cz * invL;
        }
        for (Sample& s : smp) {
            const float x = s.vx * ux + s.vy * uy + s.vz * uz;
            const float y = s.vx * vax + s.vy * vay + s.vz * vaz;
            s.angle = std::atan2(y, x);
        }
        std::sort(smp.begin(), smp.end(), [](const Sample& a, const Sample& b) { return a.angle < b.angle; });
        const float twoPi = 2.0f * 3.14159265358979323846f;
        const float maxGap = 2.0f * 3.14159265358979323846f / 3.0f;
        for (size_t q = 0; q < smp.size(); ++q) {
            const Sample& cur = smp[q];
            const Sample& nxt = smp[(q + 1) % smp.size()];
            float gap = nxt.angle - cur.angle;
            if (gap < 0.0f) gap += twoPi;
            if (gap > maxGap) {
                t.pairNext[cur.edge] = -1;
                t.restCos[cur.edge] = 2.0f;
                continue;
            }
            t.pairNext[cur.edge] = nxt.nidx;
            const float n1x = cur.vx / cur.len, n1y = cur.vy / cur.len, n1z = cur.vz / cur.len;
            const float n2x = nxt.vx / nxt.len, n2y = nxt.vy / nxt.len, n2z = nxt.vz / nxt.len;
            float c = n1x * n2x + n1y * n2y + n1z * n2z;
            c = std::fmin(1.0f, std::fmax(-1.0f, c));
            t.restCos[cur.edge] = c;
        }
    }
    t.built = true;
}

}  // namespace xpbd
