/*
 * oracle_selfcol.cpp — TEST INFRASTRUCTURE ONLY (see xpbd_oracle.cpp).
 * Serial restatement of:
 *   object components   objectEnsureComponents / objectEnsureRestCache (forward.cu:11303-11439)
 *   object shape (GPU)  objectShapeMatchGPU (11671-11702): accumulate (6829-6880, double sums in
 *                       node order), solve (6887-6937), boxes (6953-6981), match (6801-6823)
 *   self-collision      selfColBuild (11796-11849): spacing (11765-11792), max displacement, keys
 *                       (7034-7044), stable sort on the key (the cub radix sort, 11836), cell ranges
 *                       (7047-7055), candidate lists (7079-7180); solve/apply (7205-7322);
 *                       contact-average push (11858-11890, 7326-7353)
 *
 * Pinned by tests/test_oracle_selfcol.py: hash keys against a numpy
 * restatement of the closed-form hash, the sorted order against
 * numpy.argsort(kind="stable"), every candidate against an O(N^2) brute force
 * of the pair rules, zero candidates at rest, and separation to d_c for two
 * overlapping unconnected bodies.
 */
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "oracle_selfcol.h"
#include "oracle_shapes.h"

namespace xo {

ObjectState obj;
SelfColState sc;

static int find_root(std::vector<int>& parent, int x)
{
    while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; }
    return x;
}

bool object_ensure_components()
{
    const int n = g.n;
    if (obj.valid && obj.minComp == s.objMinComponent) return obj.K > 0;
    std::vector<int> parent(n);
    for (int i = 0; i < n; ++i) parent[i] = i;
    for (int i = 0; i < n; ++i)
        for (int e = g.off[i]; e < g.off[i + 1]; ++e) {
            const int j = g.nbr[e];
            const int ri = find_root(parent, i), rj = find_root(parent, j);
            if (ri != rj) parent[ri] = rj;
        }
    std::vector<int> size(n, 0);
    for (int i = 0; i < n; ++i) ++size[find_root(parent, i)];
    std::vector<int> rootId(n, -1);
    int K = 0;
    obj.id.assign(n, -1);
    for (int i = 0; i < n; ++i) {
        const int r = find_root(parent, i);
        if (rootId[r] == -1) rootId[r] = (size[r] >= s.objMinComponent) ? K++ : -2;
        if (rootId[r] >= 0) obj.id[i] = rootId[r];
    }
    obj.K = K;
    obj.count.assign(K, 0);
    for (int i = 0; i < n; ++i) if (obj.id[i] >= 0) ++obj.count[obj.id[i]];
    // rest cache (11376-11439)
    std::vector<double> sum(3 * (size_t)K, 0.0);
    for (int i = 0; i < n; ++i) {
        const int id = obj.id[i];
        if (id < 0) continue;
        sum[3 * id] += g.rest[i].x; sum[3 * id + 1] += g.rest[i].y; sum[3 * id + 2] += g.rest[i].z;
    }
    obj.c0.assign(3 * (size_t)K, 0.0f);
    for (int k = 0; k < K; ++k)
        for (int a = 0; a < 3; ++a) obj.c0[3 * k + a] = (float)(sum[3 * k + a] / std::max(obj.count[k], 1));
    obj.restRel.assign(3 * (size_t)n, 0.0);
    for (int i = 0; i < n; ++i) {
        const int id = obj.id[i];
        if (id < 0) continue;
        obj.restRel[3 * i] = (double)(g.rest[i].x - obj.c0[3 * id]);
        obj.restRel[3 * i + 1] = (double)(g.rest[i].y - obj.c0[3 * id + 1]);
        obj.restRel[3 * i + 2] = (double)(g.rest[i].z - obj.c0[3 * id + 2]);
    }
    obj.sumQ.assign(3 * (size_t)K, 0.0);
    for (int i = 0; i < n; ++i) {
        const int id = obj.id[i];
        if (id < 0) continue;
        for (int a = 0; a < 3; ++a) obj.sumQ[3 * id + a] += obj.restRel[3 * (size_t)i + a];
    }
    obj.shift.assign(3 * (size_t)K, 0.0);
    for (int t = 0; t < 3 * K; ++t) obj.shift[t] = (double)obj.c0[t];
    obj.params.assign(16 * (size_t)K, 0.0f);
    obj.disp.assign(K, 0.0f);
    obj.shiftValid = false;
    obj.boxesValid = false;
    obj.valid = true;
    obj.minComp = s.objMinComponent;
    return K > 0;
}

// objectShapeMatchGPU (11671-11702) on the predicted positions p
void object_shape_match(std::vector<V3>& p)
{
    const int n = g.n, K = obj.K;
    // weights only on the device path (12164: objShapeGPU and K <= 64); the
    // host path (11573-11668) fits without them
    const bool wOn = g.attActive && !g.objW.empty() && s.objShapeGPU && K <= 64;
    const int stride = wOn ? 16 : 12;
    std::vector<double> sums((size_t)stride * K, 0.0);
    for (int i = 0; i < n; ++i) {
        const int id = obj.id[i];
        if (id < 0) continue;
        const double dx = (double)p[i].x - obj.shift[3 * id], dy = (double)p[i].y - obj.shift[3 * id + 1],
                     dz = (double)p[i].z - obj.shift[3 * id + 2];
        const double qx = obj.restRel[3 * (size_t)i], qy = obj.restRel[3 * (size_t)i + 1], qz = obj.restRel[3 * (size_t)i + 2];
        double v[16] = {dx, dy, dz, dx * qx, dx * qy, dx * qz, dy * qx, dy * qy, dy * qz, dz * qx, dz * qy, dz * qz, 0, 0, 0, 0};
        if (wOn) {
            const double w = (double)g.objW[i];
            for (int k = 0; k < 12; ++k) v[k] *= w;
            v[12] = w; v[13] = w * qx; v[14] = w * qy; v[15] = w * qz;
        }
        for (int k = 0; k < stride; ++k) sums[(size_t)stride * id + k] += v[k];
    }
    for (int k = 0; k < K; ++k) {
        float* P = &obj.params[16 * (size_t)k];
        const int cnt = obj.count[k];
        if (cnt < 4) { P[15] = 0.0f; obj.disp[k] = 0.0f; continue; }
        const double* S = &sums[(size_t)stride * k];
        const double W = wOn ? S[12] : (double)cnt;
        if (!(W > 0.0)) { P[15] = 0.0f; obj.disp[k] = 0.0f; continue; }
        const double inv = 1.0 / W;
        const double dc[3] = {S[0] * inv, S[1] * inv, S[2] * inv};
        const double* sq = wOn ? &S[13] : &obj.sumQ[3 * (size_t)k];
        double a[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) a[3 * r + c] = S[3 + 3 * r + c] - dc[r] * sq[c];
        double cNow[3];
        for (int t = 0; t < 3; ++t) { cNow[t] = obj.shift[3 * k + t] + dc[t]; obj.shift[3 * k + t] = cNow[t]; }
        obj.disp[k] = (float)std::sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
        double fn = 0.0;
        for (int t = 0; t < 9; ++t) fn += a[t] * a[t];
        fn = std::sqrt(fn);
        if (!(fn > 1e-300) || !std::isfinite(fn)) { P[15] = 0.0f; continue; }
        M3 A, R;
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) A.m[c][r] = (float)(a[3 * r + c] / fn);
        if (!polar_double(A, R)) { P[15] = 0.0f; continue; }
        for (int t = 0; t < 3; ++t) {
            P[t] = wOn ? (float)((double)obj.c0[3 * k + t] + S[13 + t] * inv) : obj.c0[3 * k + t];
            P[3 + t] = (float)cNow[t];
            for (int cc = 0; cc < 3; ++cc) P[6 + 3 * t + cc] = R.m[cc][t];
        }
        P[15] = 1.0f;
    }
    if (s.scEnabled && !s.sqActive) {
        obj.boxes.assign(6 * (size_t)K, 0.0f);
        for (int k = 0; k < K; ++k)
            for (int t = 0; t < 6; ++t) obj.boxes[6 * k + t] = (t < 3) ? FLT_MAX : -FLT_MAX;
        for (int i = 0; i < n; ++i) {
            const int id = obj.id[i];
            const V3 q = p[i];
            if (id < 0 || !std::isfinite(q.x) || !std::isfinite(q.y) || !std::isfinite(q.z)) continue;
            float* b = &obj.boxes[6 * (size_t)id];
            b[0] = std::fmin(b[0], q.x + 0.0f); b[1] = std::fmin(b[1], q.y + 0.0f); b[2] = std::fmin(b[2], q.z + 0.0f);
            b[3] = std::fmax(b[3], q.x + 0.0f); b[4] = std::fmax(b[4], q.y + 0.0f); b[5] = std::fmax(b[5], q.z + 0.0f);
        }
        float maxDisp = 0.0f;
        for (int k = 0; k < K; ++k) maxDisp = std::fmax(maxDisp, obj.disp[k]);
        obj.maxBodyDisp = maxDisp;
        obj.boxesValid = obj.shiftValid && std::isfinite(maxDisp);
    }
    obj.shiftValid = true;
    const float st = s.objShapeStiffness;
    for (int i = 0; i < n; ++i) {
        const int id = obj.id[i];
        if (id < 0 || g.invMass[i] <= 0.0f) continue;
        const float* P = &obj.params[16 * (size_t)id];
        if (P[15] == 0.0f) continue;
        const V3 r = g.rest[i];
        const float qx = r.x - P[0], qy = r.y - P[1], qz = r.z - P[2];
        const V3 q = p[i];
        const float gx = P[3] + P[6] * qx + P[7] * qy + P[8] * qz;
        const float gy = P[4] + P[9] * qx + P[10] * qy + P[11] * qz;
        const float gz = P[5] + P[12] * qx + P[13] * qy + P[14] * qz;
        p[i] = V3{q.x + st * (gx - q.x), q.y + st * (gy - q.y), q.z + st * (gz - q.z)};
    }
}

// ---------------------------------------------------------------------------
// Self-collision
// ---------------------------------------------------------------------------
unsigned int hash_cell(int cx, int cy, int cz, unsigned int mask)
{
    return (((unsigned int)cx * 73856093u) ^ ((unsigned int)cy * 19349663u) ^ ((unsigned int)cz * 83492791u)) & mask;
}

int cell_of(float v, float invCell)
{
    return (int)std::floor(std::fmin(std::fmax(v * invCell, -1.0e9f), 1.0e9f));
}

static void ensure_spacing()
{
    if (sc.spacing > 0.0f) return;
    std::vector<float> mins;
    for (int i = 0; i < g.n; ++i) {
        float m = FLT_MAX;
        for (int e = g.off[i]; e < g.off[i + 1]; ++e) {
            const float d = g.ndist[e];
            if (d > 0.0f && std::isfinite(d) && d < m) m = d;
        }
        if (m < FLT_MAX) mins.push_back(m);
    }
    if (mins.empty()) return;
    std::nth_element(mins.begin(), mins.begin() + mins.size() / 2, mins.end());
    sc.spacing = mins[mins.size() / 2];
}

void selfcol_build(const std::vector<V3>& p)
{
    const int n = g.n;
    ensure_spacing();
    const float dc = s.scRadiusScale * sc.spacing;
    sc.contactsTotal = 0;
    if (!(dc > 0.0f)) return;
    if (sc.table == 0) {
        int t = 1024;
        while (t < 2 * n) t <<= 1;
        if (s.scTableOverride > 0) t = s.scTableOverride;
        int bits = 0;
        while ((1 << bits) < t) ++bits;
        sc.table = t;
        sc.bits = bits;
    }
    const float search = s.scSearchScale * dc;
    const float excl = s.scExcludeScale * dc;
    object_ensure_components();
    const int K = obj.K;
    const bool useActive = !s.scWithinBody && obj.boxesValid && K >= 1 && K <= 64;
    sc.active.assign(n, 1);
    int activeCount = 0;
    if (useActive) {
        const float inflate = search + 2.0f * obj.maxBodyDisp + 2.0f * dc;
        for (int i = 0; i < n; ++i) {
            const int id = obj.id[i];
            unsigned char a = 0;
            if (id < 0) a = 1;
            else {
                for (int k = 0; k < K; ++k) {
                    if (k == id) continue;
                    const float* b = &obj.boxes[6 * (size_t)k];
                    const float lx = b[0] - inflate, ly = b[1] - inflate, lz = b[2] - inflate;
                    const float hx = b[3] + inflate, hy = b[4] + inflate, hz = b[5] + inflate;
                    const V3 P = p[i], X = g.x[i];
                    const bool inP = P.x >= lx && P.x <= hx && P.y >= ly && P.y <= hy && P.z >= lz && P.z <= hz;
                    const bool inX = X.x >= lx && X.x <= hx && X.y >= ly && X.y <= hy && X.z >= lz && X.z <= hz;
                    if (inP || inX) { a = 1; break; }
                }
            }
            sc.active[i] = a;
            activeCount += a;
        }
    }
    obj.boxesValid = false;
    float maxDisp2 = 0.0f;
    for (int i = 0; i < n; ++i) {
        if (useActive && !sc.active[i]) continue;
        const float dx = p[i].x - g.x[i].x, dy = p[i].y - g.x[i].y, dz = p[i].z - g.x[i].z;
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (std::isfinite(d2)) maxDisp2 = std::fmax(maxDisp2, d2);
    }
    const float maxDisp = (std::isfinite(maxDisp2) && maxDisp2 > 0.0f) ? std::sqrt(maxDisp2) : 0.0f;
    float reach = search + 2.0f * maxDisp;
    const float cap = search + 16.0f * dc;
    if (s.scCellOverride > 0.0f) reach = s.scCellOverride;
    else if (reach > cap) reach = cap;
    const float invCell = 1.0f / reach;
    const unsigned int mask = (unsigned int)sc.table - 1u, sentinel = (unsigned int)sc.table;
    sc.keys.assign(n, 0u);
    for (int i = 0; i < n; ++i) {
        const V3 X = g.x[i];
        if (useActive && !sc.active[i]) { sc.keys[i] = sentinel; continue; }
        if (!std::isfinite(X.x) || !std::isfinite(X.y) || !std::isfinite(X.z)) { sc.keys[i] = sentinel; continue; }
        sc.keys[i] = hash_cell(cell_of(X.x, invCell), cell_of(X.y, invCell), cell_of(X.z, invCell), mask);
    }
    sc.order.resize(n);
    for (int i = 0; i < n; ++i) sc.order[i] = i;
    std::stable_sort(sc.order.begin(), sc.order.end(), [&](int a, int b) { return sc.keys[a] < sc.keys[b]; });
    sc.cellStart.assign(sc.table, -1);
    sc.cellEnd.assign(sc.table, 0);
    for (int t = 0; t < n; ++t) {
        const unsigned int k = sc.keys[sc.order[t]];
        if (k > mask) continue;
        if (t == 0 || sc.keys[sc.order[t - 1]] != k) sc.cellStart[k] = t;
        if (t == n - 1 || sc.keys[sc.order[t + 1]] != k) sc.cellEnd[k] = t + 1;
    }
    const int CAP = 32;
    sc.contacts.assign((size_t)n * CAP, 0);
    sc.counts.assign(n, 0);
    int overflow = 0;
    const float excludeR2 = excl * excl, crossR2 = dc * dc, fastRel = s.scFastRel * dc;
    long long contactsTotal = 0;
#pragma omp parallel for num_threads(xo::g_threads) schedule(static) reduction(+ : overflow, contactsTotal)
    for (int i = 0; i < n; ++i) {
        const V3 x = g.x[i], pp = p[i];
        int cnt = 0;
        if ((!useActive || sc.active[i]) && std::isfinite(x.x) && std::isfinite(x.y) && std::isfinite(x.z) &&
            std::isfinite(pp.x) && std::isfinite(pp.y) && std::isfinite(pp.z)) {
            const V3 q = g.rest[i];
            const float dix = pp.x - x.x, diy = pp.y - x.y, diz = pp.z - x.z;
            const int oi = obj.id.empty() ? 0 : obj.id[i];
            const float search2 = search * search;
            const int cx = cell_of(x.x, invCell), cy = cell_of(x.y, invCell), cz = cell_of(x.z, invCell);
            unsigned int hs[27];
            float prio[CAP];
            int tier[CAP];
            int nh = 0;
            for (int oz = -1; oz <= 1; ++oz)
                for (int oy = -1; oy <= 1; ++oy)
                    for (int ox = -1; ox <= 1; ++ox) {
                        const unsigned int h = hash_cell(cx + ox, cy + oy, cz + oz, mask);
                        bool dup = false;
                        for (int m = 0; m < nh; ++m) if (hs[m] == h) { dup = true; break; }
                        if (dup) continue;
                        hs[nh++] = h;
                        const int st = sc.cellStart[h];
                        if (st < 0) continue;
                        for (int t = st; t < sc.cellEnd[h]; ++t) {
                            const int j = sc.order[t];
                            if (j == i) continue;
                            const bool cross = !obj.id.empty() && obj.id[j] != oi;
                            if (!cross && !s.scWithinBody) continue;
                            const V3 pj = p[j], qj = g.rest[j];
                            const float qx = q.x - qj.x, qy = q.y - qj.y, qz = q.z - qj.z;
                            const float q2 = qx * qx + qy * qy + qz * qz;
                            float pr;
                            int code = j;
                            if (cross) {
                                const V3 xj = g.x[j];
                                const float ex = x.x - xj.x, ey = x.y - xj.y, ez = x.z - xj.z;
                                const float distX = std::sqrt(ex * ex + ey * ey + ez * ez);
                                const float rx = dix - (pj.x - xj.x), ry = diy - (pj.y - xj.y), rz = diz - (pj.z - xj.z);
                                const float rel = std::sqrt(rx * rx + ry * ry + rz * rz);
                                if (!(distX < search + rel)) continue;
                                if (!(q2 >= crossR2)) continue;
                                pr = distX - rel;
                                if (rel > fastRel) code = -(j + 1);
                            } else {
                                const float ex = pp.x - pj.x, ey = pp.y - pj.y, ez = pp.z - pj.z;
                                const float d2 = ex * ex + ey * ey + ez * ez;
                                if (!(d2 < search2)) continue;
                                if (!(q2 >= excludeR2)) continue;
                                pr = std::sqrt(d2);
                            }
                            const int tr = cross ? 0 : 1;
                            if (cnt < CAP) {
                                sc.contacts[(size_t)i * CAP + cnt] = code;
                                prio[cnt] = pr;
                                tier[cnt] = tr;
                                ++cnt;
                            } else {
                                int worst = 0;
                                for (int m = 1; m < CAP; ++m)
                                    if (tier[m] > tier[worst] || (tier[m] == tier[worst] && prio[m] > prio[worst])) worst = m;
                                if (tr < tier[worst] || (tr == tier[worst] && pr < prio[worst])) {
                                    sc.contacts[(size_t)i * CAP + worst] = code;
                                    prio[worst] = pr;
                                    tier[worst] = tr;
                                }
                                ++overflow;
                            }
                        }
                    }
        }
        sc.counts[i] = cnt;
        contactsTotal += cnt;
    }
    sc.contactsTotal += contactsTotal;
    sc.overflow = overflow;
    sc.activeCount = useActive ? activeCount : n;
    sc.dc = dc;
    sc.invCell = invCell;
    sc.accum.assign(n, V3{0, 0, 0});
}

static bool start_normal(V3 xi, V3 xj, V3 qi, V3 qj, float& nx, float& ny, float& nz)
{
    nx = xi.x - xj.x; ny = xi.y - xj.y; nz = xi.z - xj.z;
    float l2 = nx * nx + ny * ny + nz * nz;
    if (!(l2 > 1e-24f)) {
        nx = qi.x - qj.x; ny = qi.y - qj.y; nz = qi.z - qj.z;
        l2 = nx * nx + ny * ny + nz * nz;
        if (!(l2 > 1e-24f)) return false;
    }
    const float inv = 1.0f / std::sqrt(l2);
    nx *= inv; ny *= inv; nz *= inv;
    return true;
}

// selfColSolveKernel + selfColApplyKernel (7205-7322) on pos in place
void selfcol_iteration(std::vector<V3>& pos)
{
    const int n = g.n, CAP = 32;
    const float dc = sc.dc;
    std::vector<V3> dp(n, V3{0, 0, 0}), dpc(n, V3{0, 0, 0});
    XO_PARALLEL_FOR
    for (int i = 0; i < n; ++i) {
        const int cnt = sc.counts[i];
        if (cnt <= 0) continue;
        const float wi = g.invMass[i] * s.invMassScale;
        float ax = 0, ay = 0, az = 0, cpx = 0, cpy = 0, cpz = 0, sx = 0, sy = 0, sz = 0, wFast = 0;
        int act = 0, nFast = 0;
        if (wi > 0.0f) {
            const V3 pi = pos[i], xi = g.x[i], qi = g.rest[i];
            const float dc2 = dc * dc;
            for (int k = 0; k < cnt; ++k) {
                const int code = sc.contacts[(size_t)i * CAP + k];
                if (code < 0) {
                    const int j = -code - 1;
                    float nx, ny, nz;
                    if (!start_normal(xi, g.x[j], qi, g.rest[j], nx, ny, nz)) continue;
                    const V3 pj = pos[j];
                    const float proj = (pi.x - pj.x) * nx + (pi.y - pj.y) * ny + (pi.z - pj.z) * nz;
                    if (!(proj < dc)) continue;
                    const float pen = dc - proj;
                    sx += nx * pen; sy += ny * pen; sz += nz * pen;
                    const float wj = g.invMass[j] * s.invMassScale;
                    wFast += wi / (wi + wj);
                    ++nFast;
                    continue;
                }
                const int j = code;
                const V3 pj = pos[j];
                const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                const float d2 = dx * dx + dy * dy + dz * dz;
                if (!(d2 < dc2)) continue;
                const float wj = g.invMass[j] * s.invMassScale;
                const float wr = wi / (wi + wj);
                if (d2 < 1e-24f) {
                    const V3 qj = g.rest[j];
                    const float rx = qi.x - qj.x, ry = qi.y - qj.y, rz = qi.z - qj.z;
                    const float r2 = rx * rx + ry * ry + rz * rz;
                    if (!(r2 > 1e-24f)) continue;
                    const float sf = dc * wr / std::sqrt(r2);
                    ax += rx * sf; ay += ry * sf; az += rz * sf;
                    ++act;
                    continue;
                }
                const float d = std::sqrt(d2);
                const float sf = (dc - d) / d * wr;
                ax += dx * sf; ay += dy * sf; az += dz * sf;
                ++act;
            }
        }
        V3 out = (act > 0) ? V3{ax / act, ay / act, az / act} : V3{0, 0, 0};
        if (nFast > 0) {
            const float sl = std::sqrt(sx * sx + sy * sy + sz * sz);
            if (sl > 1e-20f) {
                const float bx = sx / sl, by = sy / sl, bz = sz / sl;
                const V3 pi = pos[i], xi = g.x[i], qi = g.rest[i];
                float delta = 0.0f;
                for (int k = 0; k < cnt; ++k) {
                    const int code = sc.contacts[(size_t)i * CAP + k];
                    if (code >= 0) continue;
                    const int j = -code - 1;
                    float nx, ny, nz;
                    if (!start_normal(xi, g.x[j], qi, g.rest[j], nx, ny, nz)) continue;
                    const V3 pj = pos[j];
                    const float ddx = pi.x - pj.x, ddy = pi.y - pj.y, ddz = pi.z - pj.z;
                    if (!(ddx * nx + ddy * ny + ddz * nz < dc)) continue;
                    const float pen = dc - (ddx * bx + ddy * by + ddz * bz);
                    if (pen > delta) delta = pen;
                }
                if (delta > 0.0f) {
                    const float scl = delta * (wFast / (float)nFast);
                    cpx = bx * scl; cpy = by * scl; cpz = bz * scl;
                    out = V3{out.x + cpx, out.y + cpy, out.z + cpz};
                }
            }
        }
        dp[i] = out;
        dpc[i] = V3{cpx, cpy, cpz};
    }
    for (int i = 0; i < n; ++i) {
        if (sc.counts[i] <= 0) continue;
        pos[i] = V3{pos[i].x + dp[i].x, pos[i].y + dp[i].y, pos[i].z + dp[i].z};
        sc.accum[i] = V3{sc.accum[i].x + dpc[i].x, sc.accum[i].y + dpc[i].y, sc.accum[i].z + dpc[i].z};
    }
}

// objectContactPush (11858-11890): returns true when a push was applied
bool object_contact_push(std::vector<V3>& pos)
{
    const int n = g.n, K = obj.K;
    if (K <= 0 || s.objShapeStiffness <= 0.0f) return false;
    std::vector<float> sums(4 * (size_t)K, 0.0f), T(4 * (size_t)K, 0.0f);
    for (int i = 0; i < n; ++i) {
        const int id = obj.id[i];
        if (id < 0) continue;
        const V3 a = sc.accum[i];
        if (!(a.x * a.x + a.y * a.y + a.z * a.z > 1e-30f)) continue;
        sums[4 * id] += a.x; sums[4 * id + 1] += a.y; sums[4 * id + 2] += a.z; sums[4 * id + 3] += 1.0f;
    }
    bool any = false;
    for (int k = 0; k < K; ++k) {
        const float c = sums[4 * k + 3];
        if (!(c > 0.0f)) continue;
        const float scl = s.objShapeStiffness / c;
        T[4 * k] = sums[4 * k] * scl; T[4 * k + 1] = sums[4 * k + 1] * scl; T[4 * k + 2] = sums[4 * k + 2] * scl;
        if (!std::isfinite(T[4 * k]) || !std::isfinite(T[4 * k + 1]) || !std::isfinite(T[4 * k + 2])) continue;
        T[4 * k + 3] = 1.0f;
        any = true;
    }
    if (!any) return false;
    for (int i = 0; i < n; ++i) {
        const int id = obj.id[i];
        if (id < 0 || g.invMass[i] <= 0.0f) continue;
        const float* t = &T[4 * (size_t)id];
        if (t[3] == 0.0f) continue;
        const V3 a = sc.accum[i];
        if (a.x * a.x + a.y * a.y + a.z * a.z > 1e-30f) continue;
        pos[i] = V3{pos[i].x + t[0], pos[i].y + t[1], pos[i].z + t[2]};
    }
    return true;
}

}  // namespace xo
