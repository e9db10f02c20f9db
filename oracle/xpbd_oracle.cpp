/*
 * xpbd_oracle.cpp — TEST INFRASTRUCTURE ONLY.
 *
 * A serial CPU restatement of the reference XPBD step path
 * (native/xpbd_capi.cu + native/solver_core/cuda_rasterizer/forward.cu), used
 * by the test suite as the parity oracle and by bench.py as the timed CPU
 * baseline.  It is never linked into, loaded by, or called from the product
 * library.  It exports the same C ABI as libxpbd_isaac.so (include/xpbd_capi.h)
 * so the same ctypes bridge drives both.
 *
 * Each function cites the reference lines it restates.  Floating-point
 * expressions keep the reference's operand order; build with
 * -ffp-contract=off so no multiply-add is fused (the GPU library is built the
 * same way), which makes every order-independent (Jacobi / per-node) pass
 * reproducible to the bit.
 *
 * Parity pinning (what checks each function beyond restating it) is listed in
 * the header comment of each section and exercised in tests/test_oracle_*.py.
 */
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

#include "xpbd_capi.h"
#include "xpbd_oracle_internal.h"

namespace xo {

State g;   // the one global instance (native/xpbd_capi.cu:6, 17-29)
int g_threads = 1;
Settings s;

int fail(const std::string& msg)
{
    g.err = msg;
    std::fprintf(stderr, "[xpbd_oracle] %s\n", msg.c_str());
    return -1;
}

// ---------------------------------------------------------------------------
// Graph load: ChainMail::loadGraph (forward.cu:2702-2779).  Symmetric CSR in
// input order, both directions per edge; invalid or self edges dropped.
// Pinned by tests/test_oracle_graph.py (CSR against a numpy construction).
// ---------------------------------------------------------------------------
static void load_graph(int n, const float* pos, int e, const int* edges, const float* rest, const float* stiff)
{
    std::vector<std::vector<int>> nb(n);
    std::vector<std::vector<float>> nd(n), ns(n);
    for (int k = 0; k < e; ++k) {
        const int a = edges[2 * k], b = edges[2 * k + 1];
        if (a < 0 || b < 0 || a >= n || b >= n || a == b) continue;
        nb[a].push_back(b); nd[a].push_back(rest[k]); ns[a].push_back(stiff[k]);
        nb[b].push_back(a); nd[b].push_back(rest[k]); ns[b].push_back(stiff[k]);
    }
    g.off.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) g.off[i + 1] = g.off[i] + (int)nb[i].size();
    const int m = g.off[n];
    g.nbr.resize(m); g.ndist.resize(m); g.nstiff.resize(m);
    for (int i = 0; i < n; ++i)
        for (size_t k = 0; k < nb[i].size(); ++k) {
            g.nbr[g.off[i] + k] = nb[i][k];
            g.ndist[g.off[i] + k] = nd[i][k];
            g.nstiff[g.off[i] + k] = ns[i][k];
        }
    g.rest.resize(n);
    for (int i = 0; i < n; ++i) g.rest[i] = V3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
}

// ensureChainmailGPU (forward.cu:10519-10845): first step after create uploads
// x = rest, v = 0, invMass = 1, lambdas = 0, drops any attachment state.
static void ensure_init()
{
    if (g.initialised) return;
    const int n = g.n;
    g.x = g.rest;
    g.v.assign(n, V3{0, 0, 0});
    g.invMass.assign(n, 1.0f);
    g.invMassSaved.clear();
    g.objW.clear();
    g.attActive = false;
    g.p.assign(n, V3{0, 0, 0});
    g.q.assign(n, V3{0, 0, 0});
    g.lamA.assign(g.off[n], 0.0f);
    g.lamB.assign(g.off[n], 0.0f);
    extensions_on_init();
    g.initialised = true;
}

// ---------------------------------------------------------------------------
// Per-node passes.  Pinned by tests/test_oracle_known_answers.py: free fall
// closed form, two-particle distance closed form, ground / sphere / box /
// capsule SDF pushes against analytic surfaces, restitution and friction
// ratios of a single particle.
// ---------------------------------------------------------------------------

// xpbdPredictKernel (forward.cu:5954-5979)
static void predict(float dt, V3 grav)
{
    XO_PARALLEL_FOR
    for (int i = 0; i < g.n; ++i) {
        V3 x = g.x[i], v = g.v[i];
        if (g.invMass[i] > 0.0f) {
            v = add(v, mul(grav, dt));
            g.p[i] = add(x, mul(v, dt));
            g.v[i] = v;
        } else {
            g.p[i] = x;
            g.v[i] = V3{0, 0, 0};
        }
    }
}

// xpbdSolveJacobiKernel (forward.cu:6484-6571): reads pin/lamIn, writes
// pout/lamOut for every node from the same inputs.
static void distance_jacobi(const std::vector<V3>& pin, std::vector<V3>& pout,
                            const std::vector<float>& lamIn, std::vector<float>& lamOut, float dt, float underRelax)
{
    const int n = g.n;
    const float dt2 = std::fmax(dt * dt, 1e-8f);
    XO_PARALLEL_FOR
    for (int i = 0; i < n; ++i) {
        const V3 pi = pin[i];
        const float wi = g.invMass[i] * s.invMassScale;
        V3 acc{0, 0, 0};
        int valid = 0;
        for (int e = g.off[i]; e < g.off[i + 1]; ++e) {
            const int j = g.nbr[e];
            float lam = lamIn[e];
            const V3 d3 = sub(pi, pin[j]);
            const float d = len(d3);
            if (d < 1e-7f) { lamOut[e] = lam; continue; }
            const float C = d - g.ndist[e];
            const V3 grad = mul(d3, 1.0f / d);
            const float wj = g.invMass[j] * s.invMassScale;
            const float st = std::fmax(g.nstiff[e] * s.stiffnessScale, 1e-6f);
            const float alpha = s.distCompliance / st;
            const float at = alpha / dt2;
            const float denom = (wi + wj) + at;
            if (denom < 1e-8f) { lamOut[e] = lam; continue; }
            const float dl = (-C - at * lam) / denom;
            lam += dl;
            lamOut[e] = lam;
            if (wi > 0.0f) {
                acc = add(acc, mul(grad, wi * dl));
                valid++;
            }
        }
        V3 out = pi;
        if (wi > 0.0f && valid > 0) {
            const V3 avg = mul(acc, 1.0f / float(valid));
            out = add(pi, mul(avg, underRelax));
        }
        pout[i] = out;
    }
}

// kinColliderProjectKernel (forward.cu:7425-7518).  Statistics are summed in
// node order (the reference uses float atomics, so only their value, not the
// last bits, is comparable).
static void kinematic_colliders(std::vector<V3>& pos, bool countHits)
{
    const int nc = s.colCount;
    const float margin = s.colMargin, mu = s.colFriction;
    for (int i = 0; i < g.n; ++i) {
        if (g.invMass[i] <= 0.0f) continue;
        V3 p = pos[i];
        bool moved = false;
        const float rmax = contact_rmax(i);
        const float reach = margin + rmax;
        for (int c = 0; c < nc; ++c) {
            const Collider& C = s.col[c];
            const float dx = p.x - C.t[0], dy = p.y - C.t[1], dz = p.z - C.t[2];
            const float q0 = C.R[0] * dx + C.R[3] * dy + C.R[6] * dz;
            const float q1 = C.R[1] * dx + C.R[4] * dy + C.R[7] * dz;
            const float q2 = C.R[2] * dx + C.R[5] * dy + C.R[8] * dz;
            float n0 = 0.0f, n1 = 0.0f, n2 = 1.0f, dist;
            if (C.type == 1) {
                const float a0 = std::fabs(q0) - C.h[0], a1 = std::fabs(q1) - C.h[1], a2 = std::fabs(q2) - C.h[2];
                if (a0 > 0.0f || a1 > 0.0f || a2 > 0.0f) {
                    const float o0 = std::fmax(a0, 0.0f), o1 = std::fmax(a1, 0.0f), o2 = std::fmax(a2, 0.0f);
                    dist = std::sqrt(o0 * o0 + o1 * o1 + o2 * o2);
                    if (dist >= reach) continue;
                    const float inv = 1.0f / std::fmax(dist, 1e-20f);
                    n0 = std::copysign(o0 * inv, q0); n1 = std::copysign(o1 * inv, q1); n2 = std::copysign(o2 * inv, q2);
                } else {
                    if (a0 >= a1 && a0 >= a2) { dist = a0; n0 = (q0 >= 0.0f) ? 1.0f : -1.0f; n2 = 0.0f; }
                    else if (a1 >= a2)        { dist = a1; n1 = (q1 >= 0.0f) ? 1.0f : -1.0f; n2 = 0.0f; }
                    else                      { dist = a2; n2 = (q2 >= 0.0f) ? 1.0f : -1.0f; }
                }
            } else {
                float e2 = q2;
                if (C.type == 2) e2 = q2 - std::fmin(std::fmax(q2, -C.h[1]), C.h[1]);
                const float r = std::sqrt(q0 * q0 + q1 * q1 + e2 * e2);
                dist = r - C.h[0];
                if (dist >= reach) continue;
                if (r > 1e-20f) { n0 = q0 / r; n1 = q1 / r; n2 = e2 / r; }
            }
            const float wx = C.R[0] * n0 + C.R[1] * n1 + C.R[2] * n2;
            const float wy = C.R[3] * n0 + C.R[4] * n1 + C.R[5] * n2;
            const float wz = C.R[6] * n0 + C.R[7] * n1 + C.R[8] * n2;
            if (contact_colliders_on() && rmax > 0.0f) {
                dist -= contact_support(i, wx, wy, wz);
                if (dist >= margin) continue;
            }
            const float corr = margin - dist;
            const V3 pIn = p;
            p.x += wx * corr; p.y += wy * corr; p.z += wz * corr;
            V3 dp{wx * corr, wy * corr, wz * corr};
            if (mu > 0.0f) {
                const float ox = C.Rp[0] * q0 + C.Rp[1] * q1 + C.Rp[2] * q2 + C.tp[0];
                const float oy = C.Rp[3] * q0 + C.Rp[4] * q1 + C.Rp[5] * q2 + C.tp[1];
                const float oz = C.Rp[6] * q0 + C.Rp[7] * q1 + C.Rp[8] * q2 + C.tp[2];
                const V3 x0 = g.x[i];
                const float rx = (p.x - x0.x) - (pIn.x - ox), ry = (p.y - x0.y) - (pIn.y - oy), rz = (p.z - x0.z) - (pIn.z - oz);
                const float rn = rx * wx + ry * wy + rz * wz;
                const float tx = rx - rn * wx, ty = ry - rn * wy, tz = rz - rn * wz;
                const float lt = std::sqrt(tx * tx + ty * ty + tz * tz);
                if (lt > 1e-20f) {
                    const float k = std::fmin(1.0f, mu * corr / lt);
                    p.x -= k * tx; p.y -= k * ty; p.z -= k * tz;
                    dp.x -= k * tx; dp.y -= k * ty; dp.z -= k * tz;
                }
            }
            moved = true;
            float* acc = &g.colPush[7 * c];
            const float w = std::sqrt(dp.x * dp.x + dp.y * dp.y + dp.z * dp.z);
            acc[0] += dp.x; acc[1] += dp.y; acc[2] += dp.z;
            acc[3] += w * p.x; acc[4] += w * p.y; acc[5] += w * p.z;
            acc[6] += w;
            if (countHits) g.colHits[c] += 1;
        }
        if (moved) pos[i] = p;
    }
}

// groundProjectKernel (forward.cu:7405-7421)
static void ground_project(std::vector<V3>& pos)
{
    const float* n = s.groundN;
    const float planeD = s.groundHeight + s.groundRadius;
    XO_PARALLEL_FOR
    for (int i = 0; i < g.n; ++i) {
        if (g.invMass[i] <= 0.0f) continue;
        const V3 p = pos[i];
        const float c = n[0] * p.x + n[1] * p.y + n[2] * p.z - planeD - ground_contact_radius(i);
        if (c < 0.0f) pos[i] = V3{p.x - c * n[0], p.y - c * n[1], p.z - c * n[2]};
    }
}

// xpbdUpdateVelocityKernel (forward.cu:6623-6661; lines 6662-6671 are dead)
static void velocity_plain(const std::vector<V3>& pfin, float dt, float velDamping)
{
    XO_PARALLEL_FOR
    for (int i = 0; i < g.n; ++i) {
        const V3 x = g.x[i], p = pfin[i];
        if (g.invMass[i] > 0.0f) {
            const V3 dx = sub(p, x);
            const float dx2 = len2(dx);
            if (!std::isfinite(dx2) || dx2 < 1e-18f) { g.v[i] = V3{0, 0, 0}; g.x[i] = p; continue; }
            V3 vs = mul(dx, 1.0f / std::fmax(dt, 1e-8f));
            const float damping = std::fmin(std::fmax(velDamping, 0.0f), 1.0f);
            vs = mul(vs, 1.0f - damping);
            if (!std::isfinite(vs.x) || !std::isfinite(vs.y) || !std::isfinite(vs.z) || len2(vs) < 1e-16f) vs = V3{0, 0, 0};
            g.v[i] = vs;
            g.x[i] = p;
            continue;
        }
        g.v[i] = V3{0, 0, 0};
        g.x[i] = p;
    }
}

// xpbdUpdateVelocityGroundKernel (forward.cu:7528-7599)
static void velocity_ground(const std::vector<V3>& pfin, float dt, float velDamping)
{
    const float* n = s.groundN;
    const float planeD = s.groundHeight + s.groundRadius;
    const float restingSpeed = 2.0f * s.groundGravity * dt;
    const float restitution = s.groundRestitution;   // object body contact is never on through the C ABI
    XO_PARALLEL_FOR
    for (int i = 0; i < g.n; ++i) {
        const V3 x = g.x[i], p = pfin[i];
        if (g.invMass[i] <= 0.0f) { g.v[i] = V3{0, 0, 0}; g.x[i] = p; continue; }
        const V3 vPre = g.v[i];
        const V3 dx = sub(p, x);
        const float dx2 = len2(dx);
        if (!std::isfinite(dx2)) { g.v[i] = V3{0, 0, 0}; g.x[i] = p; continue; }
        V3 v = (dx2 < 1e-18f) ? V3{0, 0, 0} : mul(dx, 1.0f / std::fmax(dt, 1e-8f));
        const float gap = n[0] * p.x + n[1] * p.y + n[2] * p.z - planeD - ground_contact_radius(i);
        if (gap <= s.groundSlop) {
            const float vn = n[0] * v.x + n[1] * v.y + n[2] * v.z;
            const float vnPre = n[0] * vPre.x + n[1] * vPre.y + n[2] * vPre.z;
            float vnNew = vn, jn = 0.0f;
            if (vnPre < 0.0f) {
                const float e = (-vnPre <= restingSpeed) ? 0.0f : restitution;
                vnNew = std::fmax(vn, -e * vnPre);
                jn = -(1.0f + e) * vnPre;
            }
            V3 vt{v.x - vn * n[0], v.y - vn * n[1], v.z - vn * n[2]};
            const float vtLen = len(vt);
            const float dvt = s.groundFriction * jn;
            if (vtLen <= dvt || vtLen < 1e-12f) vt = V3{0, 0, 0};
            else vt = mul(vt, (vtLen - dvt) / vtLen);
            v = V3{vt.x + vnNew * n[0], vt.y + vnNew * n[1], vt.z + vnNew * n[2]};
        }
        const float damping = std::fmin(std::fmax(velDamping, 0.0f), 1.0f);
        v = mul(v, 1.0f - damping);
        if (!std::isfinite(v.x) || !std::isfinite(v.y) || !std::isfinite(v.z) || len2(v) < 1e-16f) v = V3{0, 0, 0};
        g.v[i] = v;
        g.x[i] = p;
    }
}

// Press plates (forward.cu:9411-9424) and slab pin (9351-9407).
static void press_project(std::vector<V3>& pos)
{
    const float lo = s.sqLo + (s.sqPressFromLo ? s.sqCurDisp : 0.0f);
    const float hi = s.sqHi - (s.sqPressFromLo ? 0.0f : s.sqCurDisp);
    for (int i = 0; i < g.n; ++i) {
        if (g.invMass[i] <= 0.0f) continue;
        V3 p = pos[i];
        const float v = (s.sqAxis == 0) ? p.x : (s.sqAxis == 1) ? p.y : p.z;
        if (!(v < lo) && !(v > hi)) continue;
        const float t = (v < lo) ? lo : hi;
        if (s.sqAxis == 0) p.x = t; else if (s.sqAxis == 1) p.y = t; else p.z = t;
        pos[i] = p;
    }
}

static V3 slab_target(bool top, int j)
{
    V3 r = top ? g.sqTopRest[j] : g.sqBotRest[j];
    if (top) {
        if (s.sqAxis == 0) r.x -= s.sqCurDisp; else if (s.sqAxis == 1) r.y -= s.sqCurDisp; else r.z -= s.sqCurDisp;
    }
    return r;
}

static void slab_apply()
{
    const int nt = (int)g.sqTop.size(), nb = (int)g.sqBot.size();
    for (int k = 0; k < nt + nb; ++k) {
        const bool top = k < nt;
        const int j = top ? k : k - nt;
        const int id = top ? g.sqTop[j] : g.sqBot[j];
        const V3 r = slab_target(top, j);
        if (s.sqSlip) {
            V3 p = g.x[id], v = g.v[id];
            if (s.sqAxis == 0) { p.x = r.x; v.x = 0; } else if (s.sqAxis == 1) { p.y = r.y; v.y = 0; } else { p.z = r.z; v.z = 0; }
            g.x[id] = p; g.v[id] = v;
        } else {
            g.x[id] = r; g.v[id] = V3{0, 0, 0}; g.invMass[id] = 0.0f;
        }
    }
}

static void slab_pin_axis(std::vector<V3>& pos)
{
    const int nt = (int)g.sqTop.size(), nb = (int)g.sqBot.size();
    for (int k = 0; k < nt + nb; ++k) {
        const bool top = k < nt;
        const int j = top ? k : k - nt;
        const int id = top ? g.sqTop[j] : g.sqBot[j];
        const V3 r = slab_target(top, j);
        V3 p = pos[id];
        if (s.sqAxis == 0) p.x = r.x; else if (s.sqAxis == 1) p.y = r.y; else p.z = r.z;
        pos[id] = p;
    }
}

// ---------------------------------------------------------------------------
// runXPBDSimulation (forward.cu:12038-12751) via stepPhysicsOnly (13253-13264)
// ---------------------------------------------------------------------------
static void run_step()
{
    ensure_init();
    const int n = g.n;
    const float dt = s.dt;
    const float grav = (s.groundGravity > 0.0f && !s.sqActive) ? s.groundGravity : 0.0f;
    const V3 gravity{-grav * s.groundN[0], -grav * s.groundN[1], -grav * s.groundN[2]};
    if (s.sqActive && grav == 0.0f) {
        // stepPhysicsOnly passes zero gravity while squash is active (13258-13262)
    }

    // 0: lambda reset per substep (12053-12074)
    std::fill(g.lamA.begin(), g.lamA.end(), 0.0f);
    std::fill(g.lamB.begin(), g.lamB.end(), 0.0f);
    extensions_begin_step();

    // 1: squash ramp / dwell and slab pin (12095-12119)
    if (s.sqActive && (s.sqPress || (!g.sqTop.empty() && !g.sqBot.empty()))) {
        if (s.sqAutoLog && s.sqDwell > 0) {
            if (--s.sqDwell == 0) ++s.sqCkptIdx;
        } else {
            s.sqCurDisp = std::fmin(s.sqCurDisp + s.sqRamp * dt, s.sqMaxDisp);
            if (s.sqAutoLog && s.sqCkptIdx < 4) {
                static const float ck[4] = {0.25f, 0.50f, 0.75f, 1.00f};
                const float target = ck[s.sqCkptIdx] * s.sqMaxDisp;
                if (s.sqCurDisp >= target) { s.sqCurDisp = target; s.sqDwell = s.sqDwellFrames; }
            }
        }
        if (!s.sqPress && !g.sqTop.empty()) slab_apply();
    }

    const bool groundOn = s.groundEnabled && !s.sqActive;

    // 3: predict (12155-12160), 4: object shape (12163-12166), 5: self-collision build (12179-12200)
    predict(dt, gravity);
    extensions_after_predict();

    // 6: collider statistics reset (12228-12244)
    const bool colOn = s.colCount > 0;
    if (colOn) {
        std::fill(g.colHits.begin(), g.colHits.end(), 0);
        std::fill(g.colPush.begin(), g.colPush.end(), 0.0f);
    }
    extensions_before_iterations();

    // 7: iterations (12250-12560)
    std::vector<V3>* pin = &g.p;
    std::vector<V3>* pout = &g.q;
    std::vector<float>* lin = &g.lamA;
    std::vector<float>* lout = &g.lamB;
    for (int it = 0; it < s.iters; ++it) {
        if (extensions_distance(*pin, *pout, it)) {
            // a coloured Gauss-Seidel distance pass updated pin in place
        } else if (s.useDistance) {
            distance_jacobi(*pin, *pout, *lin, *lout, dt, s.underRelax);
            std::swap(pin, pout);
            std::swap(lin, lout);
        }
        extensions_iteration(pin, pout, it);
        if (s.sqActive && !s.sqPress && s.sqSlip && !g.sqTop.empty() && !g.sqBot.empty()) slab_pin_axis(*pin);
        if (s.sqActive && s.sqPress) press_project(*pin);
        extensions_self_collision(*pin);
        if (colOn) kinematic_colliders(*pin, it == s.iters - 1);
        if (groundOn) ground_project(*pin);
    }
    // 8: contact-average push, then one more ground projection (12588-12596)
    if (extensions_after_iterations(*pin, groundOn) && s.groundEnabled) ground_project(*pin);

    // 9: velocity update and commit (12600-12628)
    if (groundOn) velocity_ground(*pin, dt, s.velDamping);
    else velocity_plain(*pin, dt, s.velDamping);
    (void)n;
}

}  // namespace xo

using namespace xo;

// ===========================================================================
// C ABI (native/xpbd_capi.cu:66-310)
// ===========================================================================
extern "C" {

const char* xpbd_last_error(void) { return g.err.c_str(); }

// Oracle only (no counterpart in the product or the reference): threads of
// the order-independent loops, for the timed multi-core CPU baseline.
int xpbd_oracle_set_threads(int threads)
{
    xo::g_threads = threads > 0 ? threads : 1;
    return xo::g_threads;
}
const char* xpbd_ext_backend(void) { return "oracle-cpu"; }

static void free_all()
{
    g.n = 0;
    g.stepped = false;
    g.initialised = false;
    g.contactOn = false;
    s.contactColliders = false;
    s.contactGround = false;
    g.x.clear(); g.v.clear(); g.p.clear(); g.q.clear(); g.rest.clear();
    g.scales.clear(); g.rots.clear(); g.opacity.clear();
    extensions_free();
}

int xpbd_create(int n, const float* pos, const float* scales, const float* rots, const float* opacity,
                int e, const int* edges, const float* rest, const float* stiff)
{
    free_all();
    if (n <= 0 || !pos || !scales || !rots || !opacity || e < 0 || (e > 0 && (!edges || !rest || !stiff)))
        return fail("xpbd_create: invalid arguments");
    load_graph(n, pos, e, edges, rest, stiff);
    g.scales.assign(scales, scales + 3 * (size_t)n);
    g.rots.assign(rots, rots + 4 * (size_t)n);
    g.opacity.assign(opacity, opacity + n);
    g.n = n;
    g.initialised = false;   // invalidatePhysicsGraph (native/xpbd_capi.cu:99)
    return 0;
}

void xpbd_destroy(void)
{
    xpbd_set_colliders(0, nullptr, nullptr, nullptr, 0.0f, 0.0f);
    if (g.initialised) {
        if (g.attActive) g.invMass = g.invMassSaved;
        g.attActive = false;
    }
    free_all();
}

int xpbd_count(void) { return g.n; }

// forward.cu:4896-4921
void xpbd_set_solver(int iters, float dt, float underRelax, float velDamping, float stiffnessScale,
                     float invMassScale, float distCompliance, float shapeCompliance, float shapeBlend,
                     float angleCompliance, float angleBlend)
{
    s.iters = iters < 1 ? 1 : iters;
    s.dt = std::fmax(1e-5f, dt);
    s.underRelax = std::fmax(0.0f, underRelax);
    s.velDamping = std::fmax(0.0f, velDamping);
    s.stiffnessScale = std::fmax(0.0f, stiffnessScale);
    s.invMassScale = std::fmax(0.0f, invMassScale);
    s.distCompliance = std::fmax(0.0f, distCompliance);
    s.shapeCompliance = std::fmax(0.0f, shapeCompliance);
    s.shapeBlend = std::fmin(std::fmax(0.0f, shapeBlend), 1.0f);
    s.angleCompliance = std::fmax(0.0f, angleCompliance);
    s.angleBlend = std::fmin(std::fmax(0.0f, angleBlend), 1.0f);
}

void xpbd_set_constraints(int distance, int shape, int angle, int volume)
{
    s.useDistance = distance != 0;
    s.useShape = shape != 0;
    s.useAngle = angle != 0;
    s.useVolume = volume != 0;
}

// forward.cu:4430-4446
void xpbd_set_ground(int enabled, float ux, float uy, float uz, float height, float friction,
                     float restitution, float contactRadius, float contactSlop, float gravity)
{
    float nx = ux, ny = uy, nz = uz;
    const float l = std::sqrt(nx * nx + ny * ny + nz * nz);
    if (!(l > 1e-8f) || !std::isfinite(l)) { nx = 0; ny = 0; nz = 1; }
    else { nx /= l; ny /= l; nz /= l; }
    s.groundN[0] = nx; s.groundN[1] = ny; s.groundN[2] = nz;
    s.groundEnabled = enabled != 0;
    s.groundHeight = std::isfinite(height) ? height : 0.0f;
    s.groundFriction = std::fmax(0.0f, friction);
    s.groundRestitution = std::fmin(std::fmax(restitution, 0.0f), 1.0f);
    s.groundRadius = std::fmax(0.0f, contactRadius);
    s.groundSlop = std::fmax(0.0f, contactSlop);
    s.groundGravity = std::fmax(0.0f, gravity);
}

// forward.cu:4478-4500
void xpbd_set_colliders(int count, const int* types, const float* poses, const float* dims, float margin,
                        float friction)
{
    if (count < 0) count = 0;
    if (count > 32) count = 32;
    const bool samePrev = (count == s.colCount);
    for (int c = 0; c < count; ++c) {
        Collider& K = s.col[c];
        const bool keep = samePrev && K.type == types[c];
        for (int a = 0; a < 9; ++a) K.Rp[a] = keep ? K.R[a] : poses[12 * c + a];
        for (int a = 0; a < 3; ++a) K.tp[a] = keep ? K.t[a] : poses[12 * c + 9 + a];
        K.type = types[c];
        for (int a = 0; a < 9; ++a) K.R[a] = poses[12 * c + a];
        for (int a = 0; a < 3; ++a) { K.t[a] = poses[12 * c + 9 + a]; K.h[a] = std::fmax(dims[3 * c + a], 0.0f); }
    }
    s.colCount = count;
    s.colMargin = (std::isfinite(margin) && margin > 0.0f) ? margin : 0.0f;
    s.colFriction = (std::isfinite(friction) && friction > 0.0f) ? friction : 0.0f;
}

// forward.cu:4502-4519
int xpbd_collider_stats(int* hits, float* push, float* centroid, int maxCount)
{
    const int n = (maxCount < s.colCount) ? maxCount : s.colCount;
    if (n <= 0 || !g.colStatsValid) return 0;
    for (int c = 0; c < n; ++c) {
        if (hits) hits[c] = g.colHits[c];
        const float* a = &g.colPush[7 * c];
        for (int k = 0; k < 3; ++k) {
            if (push) push[3 * c + k] = a[k];
            if (centroid) centroid[3 * c + k] = (a[6] > 0.0f) ? a[3 + k] / a[6] : s.col[c].t[k];
        }
    }
    return n;
}

// forward.cu:6701-6712, 6790-6796
int xpbd_add_rigid_velocity(const float* dv, const float* dw, const float* c)
{
    if (!g.stepped) return fail("xpbd_add_rigid_velocity: no step yet");
    for (int i = 0; i < g.n; ++i) {
        if (g.invMass[i] <= 0.0f) continue;
        const V3 p = g.x[i];
        const float rx = p.x - c[0], ry = p.y - c[1], rz = p.z - c[2];
        const V3 v = g.v[i];
        g.v[i] = V3{v.x + dv[0] + (dw[1] * rz - dw[2] * ry),
                    v.y + dv[1] + (dw[2] * rx - dw[0] * rz),
                    v.z + dv[2] + (dw[0] * ry - dw[1] * rx)};
    }
    return 0;
}

// forward.cu:6759-6788 (+ attachApplyKernel 6714-6725)
int xpbd_set_attached(int count, const int* idx, const float* pos, float weight)
{
    if (!g.stepped) return fail("xpbd_set_attached: no step yet");
    const int n = g.n;
    if (g.attActive) g.invMass = g.invMassSaved;
    if (count <= 0 || !idx || !pos) { g.attActive = false; return 0; }
    if (!g.attActive) g.invMassSaved = g.invMass;
    g.objW.assign(n, 1.0f);
    const float w = std::fmax(weight, 1.0f);
    for (int k = 0; k < count; ++k) {
        const int i = idx[k];
        if (i < 0 || i >= n) continue;
        g.invMass[i] = 0.0f;
        g.x[i] = V3{pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]};
        g.v[i] = V3{0, 0, 0};
        g.objW[i] = w;
    }
    g.attActive = true;
    return 0;
}

int xpbd_step(void)
{
    if (g.n <= 0) return fail("xpbd_step: call xpbd_create first");
    run_step();
    g.stepped = true;
    g.colStatsValid = true;
    return 0;
}

unsigned long long xpbd_positions_device(void) { return 0ull; }

int xpbd_positions_host(float* out, int n)
{
    if (!g.stepped) return fail("xpbd_positions_host: no step yet");
    if (!out || n != g.n) return fail("xpbd_positions_host: size mismatch");
    for (int i = 0; i < n; ++i) { out[3 * i] = g.x[i].x; out[3 * i + 1] = g.x[i].y; out[3 * i + 2] = g.x[i].z; }
    return 0;
}

// squashReset (forward.cu:4411-4427) after releasing attachments
int xpbd_reset(void)
{
    if (!g.stepped) return fail("xpbd_reset: no step yet");
    xpbd_set_attached(0, nullptr, nullptr, 1.0f);
    g.x = g.rest;
    std::fill(g.v.begin(), g.v.end(), V3{0, 0, 0});
    std::fill(g.invMass.begin(), g.invMass.end(), 1.0f);
    extensions_on_reset();
    s.sqActive = false; s.sqCurDisp = 0.0f; s.sqCkptIdx = 0; s.sqDwell = 0;
    return 0;
}

// groundLaunch (forward.cu:4551-4610)
int xpbd_launch(const float* lin, const float* ang, const float* tilt)
{
    if (!g.stepped) return fail("xpbd_launch: no step yet");
    const int n = g.n;
    double cx = 0, cy = 0, cz = 0;
    for (int i = 0; i < n; ++i) { cx += g.rest[i].x; cy += g.rest[i].y; cz += g.rest[i].z; }
    const float comX = float(cx / n), comY = float(cy / n), comZ = float(cz / n);
    const float ax = tilt ? tilt[0] : 0.0f, ay = tilt ? tilt[1] : 0.0f, az = tilt ? tilt[2] : 0.0f;
    const float th = std::sqrt(ax * ax + ay * ay + az * az);
    float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (th > 1e-8f) {
        const float kx = ax / th, ky = ay / th, kz = az / th;
        const float sn = std::sin(th), c = std::cos(th), t = 1.0f - c;
        R[0] = c + t * kx * kx;      R[1] = t * kx * ky - sn * kz; R[2] = t * kx * kz + sn * ky;
        R[3] = t * ky * kx + sn * kz; R[4] = c + t * ky * ky;      R[5] = t * ky * kz - sn * kx;
        R[6] = t * kz * kx - sn * ky; R[7] = t * kz * ky + sn * kx; R[8] = c + t * kz * kz;
    }
    const float vx = lin ? lin[0] : 0.0f, vy = lin ? lin[1] : 0.0f, vz = lin ? lin[2] : 0.0f;
    const float wx = ang ? ang[0] : 0.0f, wy = ang ? ang[1] : 0.0f, wz = ang ? ang[2] : 0.0f;
    for (int i = 0; i < n; ++i) {
        const float px = g.rest[i].x - comX, py = g.rest[i].y - comY, pz = g.rest[i].z - comZ;
        const float rx = R[0] * px + R[1] * py + R[2] * pz;
        const float ry = R[3] * px + R[4] * py + R[5] * pz;
        const float rz = R[6] * px + R[7] * py + R[8] * pz;
        g.x[i] = V3{comX + rx, comY + ry, comZ + rz};
        g.v[i] = V3{vx + (wy * rz - wz * ry), vy + (wz * rx - wx * rz), vz + (wx * ry - wy * rx)};
    }
    std::fill(g.invMass.begin(), g.invMass.end(), 1.0f);
    extensions_on_reset();
    s.sqActive = false; s.sqCurDisp = 0.0f; s.sqCkptIdx = 0; s.sqDwell = 0;
    return 0;
}

// squashStart (forward.cu:4310-4387) in press mode (native/xpbd_capi.cu:281-288)
int xpbd_press_start(int axis, float rampPerSec, float maxDispPct, int fromLo)
{
    if (!g.stepped) return fail("xpbd_press_start: step once first");
    s.sqPressNext = true;
    s.sqPressFromLo = fromLo != 0;
    if (axis < 0 || axis > 2) axis = 1;
    const float slabPct = 0.08f;
    maxDispPct = std::fmax(0.01f, std::fmin(1.0f, maxDispPct));
    s.sqAxis = axis;
    s.sqRamp = rampPerSec;
    const int n = g.n;
    float lo = 3.402823466e+38f, hi = -3.402823466e+38f;
    for (int i = 0; i < n; ++i) {
        const float v = (axis == 0) ? g.rest[i].x : (axis == 1) ? g.rest[i].y : g.rest[i].z;
        lo = std::fmin(lo, v); hi = std::fmax(hi, v);
    }
    const float extent = hi - lo;
    s.sqPress = s.sqPressNext;
    s.sqLo = lo; s.sqHi = hi;
    s.sqMaxDisp = maxDispPct * extent;
    s.sqCurDisp = 0.0f;
    const float topThr = hi - slabPct * extent, botThr = lo + slabPct * extent;
    g.sqTop.clear(); g.sqBot.clear(); g.sqTopRest.clear(); g.sqBotRest.clear();
    for (int i = 0; i < n; ++i) {
        const float v = (axis == 0) ? g.rest[i].x : (axis == 1) ? g.rest[i].y : g.rest[i].z;
        if (v >= topThr) { g.sqTop.push_back(i); g.sqTopRest.push_back(g.rest[i]); }
        else if (v <= botThr) { g.sqBot.push_back(i); g.sqBotRest.push_back(g.rest[i]); }
    }
    if (g.sqTop.empty() || g.sqBot.empty()) {
        g.sqTop.clear(); g.sqBot.clear();
        return fail("xpbd_press_start: squashStart failed");
    }
    s.sqActive = true;
    s.sqCkptIdx = 0; s.sqDwell = 0;
    return 0;
}

void xpbd_press_autolog(int enabled) { s.sqAutoLog = enabled != 0; }

// squashStop (forward.cu:4389-4400)
void xpbd_press_stop(void)
{
    if (!s.sqActive) return;
    s.sqActive = false;
    if (g.initialised) std::fill(g.invMass.begin(), g.invMass.end(), 1.0f);
}

// native/xpbd_capi.cu:297-303; forward.cu:4675-4689
int xpbd_press_state(float* lo, float* hi, float* curDisp, float* maxDisp)
{
    const bool moving = s.sqActive && s.sqPress;
    if (curDisp) *curDisp = s.sqCurDisp;
    if (maxDisp) *maxDisp = s.sqMaxDisp;
    if (lo) *lo = s.sqLo + ((moving && s.sqPressFromLo) ? s.sqCurDisp : 0.0f);
    if (hi) *hi = s.sqHi - ((moving && !s.sqPressFromLo) ? s.sqCurDisp : 0.0f);
    return s.sqActive ? 1 : 0;
}

int xpbd_ext_get_state(float* pos, float* vel, int n)
{
    if (!g.stepped) return fail("xpbd_ext_get_state: no step yet");
    if (n != g.n) return fail("xpbd_ext_get_state: size mismatch");
    for (int i = 0; i < n; ++i) {
        if (pos) { pos[3 * i] = g.x[i].x; pos[3 * i + 1] = g.x[i].y; pos[3 * i + 2] = g.x[i].z; }
        if (vel) { vel[3 * i] = g.v[i].x; vel[3 * i + 1] = g.v[i].y; vel[3 * i + 2] = g.v[i].z; }
    }
    return 0;
}

int xpbd_ext_set_state(const float* pos, const float* vel, int n)
{
    if (!g.stepped) return fail("xpbd_ext_set_state: no step yet");
    if (n != g.n) return fail("xpbd_ext_set_state: size mismatch");
    for (int i = 0; i < n; ++i) {
        if (pos) g.x[i] = V3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        if (vel) g.v[i] = V3{vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
    }
    return 0;
}

int xpbd_ext_profile(int substeps, char* json, int len)
{
    if (!g.stepped) return fail("xpbd_ext_profile: no step yet");
    for (int k = 0; k < substeps; ++k) run_step();
    if (json && len > 2) std::snprintf(json, len, "{}");
    return 0;
}

int xpbd_ext_get_settings(float* out, int len)
{
    const float v[42] = {
        (float)s.iters, s.dt, s.underRelax, s.velDamping, s.stiffnessScale, s.invMassScale, s.distCompliance,
        s.shapeCompliance, s.shapeBlend, s.angleCompliance, s.angleBlend, (float)s.useDistance, (float)s.useShape,
        (float)s.useAngle, (float)s.useVolume, s.volCompliance, (float)s.volRingK, (float)s.volMaxMembers,
        (float)s.volLeaderMinHop, (float)s.groundEnabled, s.groundN[0], s.groundN[1], s.groundN[2], s.groundHeight,
        s.groundFriction, s.groundRestitution, s.groundRadius, s.groundSlop, s.groundGravity, s.objShapeStiffness,
        (float)s.objMinComponent, (float)s.objShapeGPU, (float)s.scEnabled, s.scRadiusScale, s.scExcludeScale,
        (float)s.scWithinBody, (float)s.shapeRobust, (float)s.colCount, s.colMargin, s.colFriction,
        (float)s.sqAutoLog, (float)s.volJStats};
    if (!out || len < 42) return fail("xpbd_ext_get_settings: need 42 floats");
    std::memcpy(out, v, sizeof(v));
    return 42;
}

int xpbd_ext_get_inv_mass(float* invMass, int n)
{
    if (!g.stepped) return fail("xpbd_ext_get_inv_mass: no step yet");
    if (n != g.n || !invMass) return fail("xpbd_ext_get_inv_mass: size mismatch");
    for (int i = 0; i < n; ++i) invMass[i] = g.invMass[i];
    return 0;
}

}  // extern "C"
