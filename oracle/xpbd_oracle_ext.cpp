/*
 * xpbd_oracle_ext.cpp — TEST INFRASTRUCTURE ONLY (see xpbd_oracle.cpp).
 * Setters and hooks for the constraint families beyond distance / colliders /
 * ground: volume, shape matching, self-collision, object shape, contact shapes
 * and deformed shapes.
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "xpbd_capi.h"
#include "xpbd_oracle_internal.h"
#include "oracle_volume.h"
#include "oracle_shapes.h"
#include "oracle_selfcol.h"
#include "oracle_gs.h"
#include "oracle_angle.h"

namespace xo {

static std::vector<float> g_cM, g_cRmax, g_cRg;   // contact ellipsoids k^2 Sigma, max semi-axis, ground radius
void contact_ground_radius_update();

// Build (or rebuild after a parameter change) the volume clusters the way
// ensureChainmailGPU / g_volTopoDirty do (forward.cu:10524-10530, 10842).
static void volume_ensure()
{
    if (vol.explicitTopo) { if (!vol.built) volume_precompute(); return; }
    if (vol.built && vol.k == s.volRingK && vol.cap == s.volMaxMembers && vol.r == s.volLeaderMinHop) return;
    volume_build_reference(s.volRingK, s.volMaxMembers, s.volLeaderMinHop);
    vol.k = s.volRingK; vol.cap = s.volMaxMembers; vol.r = s.volLeaderMinHop;
}

void extensions_on_init() { if (!vol.explicitTopo) vol.built = false; }
void extensions_free()
{
    vol = VolumeState();
    g_cM.clear(); g_cRmax.clear(); g_cRg.clear();
    obj = ObjectState();
    sc = SelfColState();
    dgs = DistanceGS();
    vgs = VolumeGS();
    shc = ShapeClusters();
    ang = AnglePairs();
}
void extensions_on_reset() {}
// The colourings are built on first use after being enabled, and the volume
// colouring again after every topology rebuild (9745-9812 runs inside
// rebuildVolumeClusters).
static void gs_ensure()
{
    if (s.distGS && !dgs.built) distance_gs_build();
    if (s.volGS && vol.built && (!vgs.built || vgs.version != vol.version)) volume_gs_build();
}
void extensions_begin_step()
{
    if (s.useVolume) {
        volume_ensure();
        std::fill(vol.lam.begin(), vol.lam.end(), 0.0f);
    }
    gs_ensure();
    std::fill(dgs.lam.begin(), dgs.lam.end(), 0.0f);   // 12067-12069
}
static bool volume_gs_ready() { return s.volGS && vgs.built && !vgs.skipped && vgs.colorOff.size() > 1; }
void extensions_after_predict_impl();
void extensions_after_predict() { extensions_after_predict_impl(); }
void extensions_before_iterations() { contact_ground_radius_update(); }
// Distance GS per colour in place (12250-12262) when enabled and edges exist.
bool extensions_distance(std::vector<V3>& pin, std::vector<V3>&, int)
{
    if (!(s.useDistance && s.distGS && dgs.built && !dgs.a.empty())) return false;
    distance_gs_iteration(pin, s.dt);
    return true;
}
void extensions_iteration(std::vector<V3>*& pin, std::vector<V3>*& pout, int)
{
    if (s.useVolume && vol.built) {
        if (volume_gs_ready()) volume_gs_iteration(*pin, s.dt);   // 12310-12324
        else volume_iteration(pin, pout);
    }
    if (s.useShape) {
        if (shc.on) shape_cluster_matching(*pin, *pout);
        else shape_matching(*pin, *pout);
        std::swap(pin, pout);
    }
    // angle constraint after shape matching (forward.cu:12486-12517); the
    // pairing uses the positions the graph was uploaded with
    if (s.useAngle) {
        if (!ang.built) angle_build(g.rest);
        angle_pass(*pin, *pout);
        std::swap(pin, pout);
    }
}
void extensions_self_collision(std::vector<V3>& pin)
{
    if (s.scEnabled && sc.contactsTotal > 0) selfcol_iteration(pin);
}
// Object shape after predict (12124-12166) and the self-collision build (12179-12200)
static bool g_objShapeStep = false;

void extensions_after_predict_impl()
{
    g_objShapeStep = (s.objShapeStiffness > 0.0f) && !s.sqActive;
    if (g_objShapeStep) {
        if (object_ensure_components()) object_shape_match(g.p);
        else g_objShapeStep = false;
    }
    if (s.scEnabled && !s.sqActive) selfcol_build(g.p);
    else sc.contactsTotal = 0;
}

bool extensions_after_iterations(std::vector<V3>& pin, bool)
{
    if (s.scEnabled && sc.contactsTotal > 0 && g_objShapeStep) {
        object_contact_push(pin);
        return true;
    }
    return false;
}
// Contact ellipsoids (forward.cu:6733-6757, 7358-7399, 12219-12226)
static bool contact_ready() { return (s.contactColliders || s.contactGround) && (int)g_cRmax.size() == g.n && g.n > 0; }
bool contact_colliders_on() { return contact_ready() && s.contactColliders; }
float contact_rmax(int i) { return contact_colliders_on() ? g_cRmax[i] : 0.0f; }
float contact_support(int i, float nx, float ny, float nz)
{
    const float* m = &g_cM[(size_t)6 * i];
    const float q = m[0] * nx * nx + m[3] * ny * ny + m[5] * nz * nz + 2.0f * (m[1] * nx * ny + m[2] * nx * nz + m[4] * ny * nz);
    return std::sqrt(std::fmax(q, 0.0f));
}
float ground_contact_radius(int i)
{
    return (contact_ready() && s.contactGround && s.groundEnabled && (int)g_cRg.size() == g.n) ? g_cRg[i] : 0.0f;
}
void contact_ground_radius_update()
{
    if (!(contact_ready() && s.contactGround && s.groundEnabled)) return;
    g_cRg.assign(g.n, 0.0f);
    for (int i = 0; i < g.n; ++i) g_cRg[i] = contact_support(i, s.groundN[0], s.groundN[1], s.groundN[2]);
}

}  // namespace xo

using namespace xo;

extern "C" {

// native/xpbd_capi.cu:130-137 -> forward.cu:3849-3853, 3934-3944, 3953-3965
void xpbd_set_volume(float compliance, int ringK, int maxMembers, int leaderMinHop)
{
    s.volCompliance = compliance;   // setVolumeParams stores it unclamped (3849-3853)
    const int k = std::max(1, std::min(8, ringK));
    const int cap = std::max(8, std::min(8192, maxMembers));
    s.volRingK = k;
    s.volMaxMembers = cap;
    s.volLeaderMinHop = std::max(1, std::min(8, leaderMinHop));
    s.volMemberSelect = 1;
}

// forward.cu:4464-4476
void xpbd_set_object_shape(float stiffness)
{
    float v = std::fmin(std::fmax(stiffness, 0.0f), 1.0f);
    if (!std::isfinite(v)) v = 0.0f;
    s.objShapeStiffness = v;
    s.objMinComponent = 50;
}

void xpbd_set_object_shape_gpu(int enabled) { s.objShapeGPU = enabled != 0; }

void xpbd_set_shape_robust(int enabled) { s.shapeRobust = enabled != 0; }

// forward.cu:4532-4540
void xpbd_set_self_collision(int enabled, float radiusScale, float excludeScale, int withinBody)
{
    s.scEnabled = enabled != 0;
    s.scWithinBody = withinBody != 0;
    s.scRadiusScale = std::fmin(std::fmax(std::isfinite(radiusScale) ? radiusScale : 1.5f, 0.1f), 10.0f);
    s.scExcludeScale = std::fmin(std::fmax(std::isfinite(excludeScale) ? excludeScale : 2.0f, 1.05f), 20.0f);
    if (!s.scEnabled) { sc.contactsTotal = 0; sc.overflow = 0; }
}

// native/xpbd_capi.cu:209-212: candidates, active particles, overflow replacements
void xpbd_self_collision_stats(int* out)
{
    if (!out) return;
    out[0] = (int)sc.contactsTotal;
    out[1] = sc.activeCount;
    out[2] = sc.overflow;
}

// Extension: shape matching over node clusters (clusterOf[n], -1 = none)
// instead of the graph 1-ring; null returns to the 1-ring.  After create.
int xpbd_ext_set_shape_clusters(const int* clusterOf, int n)
{
    if (g.n <= 0) return fail("xpbd_ext_set_shape_clusters: call xpbd_create first");
    if (clusterOf && n != g.n) return fail("xpbd_ext_set_shape_clusters: size mismatch");
    if (clusterOf)
        for (int i = 0; i < n; ++i)
            if (clusterOf[i] < -1 || clusterOf[i] >= n) return fail("xpbd_ext_set_shape_clusters: cluster id out of range");
    shape_clusters_set(clusterOf, n);
    return 0;
}

// Extension: grid of the last self-collision build: cell edge, d_c, search radius.
int xpbd_ext_self_collision_grid_info(float* info)
{
    if (!g.stepped) return fail("xpbd_ext_self_collision_grid_info: no step yet");
    if (!info) return fail("xpbd_ext_self_collision_grid_info: null output");
    info[0] = sc.invCell > 0.0f ? 1.0f / sc.invCell : 0.0f;
    info[1] = s.scRadiusScale * sc.spacing;
    info[2] = s.scSearchScale * info[1];
    return 0;
}

// Extension: the oracle keeps the caller's order; accepted for ABI parity.
int xpbd_ext_set_node_order(int mode)
{
    if (mode != 0 && mode != 1) return fail("xpbd_ext_set_node_order: mode must be 0 (caller order) or 1 (Morton)");
    return 0;
}

// Extension: graph-coloured Gauss-Seidel passes (forward.cu:3780-3810; the
// reference reaches them through FORWARD::setDistanceGSEnabled /
// setVolumeGSEnabled, not through its C ABI).
int xpbd_ext_set_gauss_seidel(int distance, int volume)
{
    s.distGS = distance != 0;
    s.volGS = volume != 0;
    return 0;
}

// Extension: sizes of the colourings: distance edges, distance colours,
// one-way edges, volume clusters, volume colours, volume colouring skipped.
int xpbd_ext_gs_info(int* info)
{
    if (!g.stepped) return fail("xpbd_ext_gs_info: no step yet");
    if (!info) return fail("xpbd_ext_gs_info: null output");
    // the reference builds both colourings whether or not GS is enabled
    // (10834 at graph upload, 10245 with every cluster rebuild)
    if (!dgs.built) distance_gs_build();
    const bool volOn = s.useVolume && vol.built;
    if (volOn && (!vgs.built || vgs.version != vol.version)) volume_gs_build();
    info[0] = (int)dgs.a.size();
    info[1] = dgs.colorOff.empty() ? 0 : (int)dgs.colorOff.size() - 1;
    info[2] = dgs.oneWay;
    info[3] = volOn ? (int)vgs.order.size() : 0;
    info[4] = (volOn && !vgs.colorOff.empty()) ? (int)vgs.colorOff.size() - 1 : 0;
    info[5] = (volOn && vgs.skipped) ? 1 : 0;
    return 0;
}

// Extension: the distance colouring (edge ends ab[2E], rest[E], stiff[E],
// colour offsets colorOff[C+1]); any pointer may be null.
int xpbd_ext_distance_gs_edges(int* ab, float* rest, float* stiff, int* colorOff)
{
    if (!g.stepped) return fail("xpbd_ext_distance_gs_edges: no step yet");
    if (!dgs.built) distance_gs_build();
    for (size_t e = 0; e < dgs.a.size(); ++e) {
        if (ab) { ab[2 * e] = dgs.a[e]; ab[2 * e + 1] = dgs.b[e]; }
        if (rest) rest[e] = dgs.rest[e];
        if (stiff) stiff[e] = dgs.stiff[e];
    }
    if (colorOff) std::copy(dgs.colorOff.begin(), dgs.colorOff.end(), colorOff);
    return 0;
}

// Extension: the volume colouring (leaders in colour-major order, colour offsets).
int xpbd_ext_volume_gs_order(int* order, int* colorOff)
{
    if (!g.stepped) return fail("xpbd_ext_volume_gs_order: no step yet");
    if (!(s.useVolume && vol.built)) return 0;
    if (!vgs.built || vgs.version != vol.version) volume_gs_build();
    if (order) std::copy(vgs.order.begin(), vgs.order.end(), order);
    if (colorOff) std::copy(vgs.colorOff.begin(), vgs.colorOff.end(), colorOff);
    return 0;
}

// Extension: fixed hash grid (cell size > 0 replaces the displacement rule;
// table size, a power of two >= 1024, replaces nextpow2(2N)); 0 keeps the rule.
int xpbd_ext_set_self_collision_grid(float cellSize, int tableSize)
{
    if (tableSize != 0 && (tableSize < 1024 || (tableSize & (tableSize - 1))))
        return fail("xpbd_ext_set_self_collision_grid: table size must be a power of two >= 1024");
    s.scCellOverride = (std::isfinite(cellSize) && cellSize > 0.0f) ? cellSize : 0.0f;
    s.scTableOverride = tableSize;
    sc.table = 0;
    return 0;
}

// Extension: how the GPU library finds candidates; the serial walk here is the
// reference's in either case.
int xpbd_ext_set_self_collision_fine(int) { return 0; }

// Extension: the umbrella angle pairing (forward.cu:10708-10819), per entry.
int xpbd_ext_angle_pairs(int* pairNext, float* restCos, int m)
{
    if (g.n <= 0) return fail("xpbd_ext_angle_pairs: call xpbd_create first");
    if (m != g.off[g.n]) return fail("xpbd_ext_angle_pairs: size mismatch");
    if (!ang.built) angle_build(g.rest);
    if (pairNext && m) std::memcpy(pairNext, ang.pairNext.data(), sizeof(int) * m);
    if (restCos && m) std::memcpy(restCos, ang.restCos.data(), sizeof(float) * m);
    return 0;
}

// Extension: the last build's artefacts: cell keys (by node), node order
// after the stable key sort, candidate counts and candidate lists [n x 32].
int xpbd_ext_self_collision_lists(unsigned int* keys, int* sorted, int* counts, int* contacts, int n)
{
    if (!g.stepped) return fail("xpbd_ext_self_collision_lists: no step yet");
    if (n != g.n) return fail("xpbd_ext_self_collision_lists: size mismatch");
    if ((int)sc.keys.size() != n) return fail("xpbd_ext_self_collision_lists: no self-collision build yet");
    for (int i = 0; i < n; ++i) {
        if (keys) keys[i] = sc.keys[i];
        if (sorted) sorted[i] = sc.order[i];
        if (counts) counts[i] = sc.counts[i];
    }
    if (contacts) std::copy(sc.contacts.begin(), sc.contacts.end(), contacts);
    return 0;
}

static float cap_or_zero(float c) { return (std::isfinite(c) && c > 0.0f) ? c : 0.0f; }

// native/xpbd_capi.cu:179-191
int xpbd_set_contact_shape(int enabled, float tau, float radiusCap)
{
    if (g.n <= 0) return fail("xpbd_set_contact_shape: call xpbd_create first");
    g.contactOn = enabled != 0;
    s.contactTau = tau;
    s.contactCap = radiusCap;
    if (g.contactOn) contact_pack(g.scales.data(), g.rots.data(), g.opacity.data(), tau, cap_or_zero(radiusCap), g_cM, g_cRmax);
    s.contactColliders = (enabled & 1) != 0;
    s.contactGround = (enabled & 2) != 0;
    return 0;
}

// native/xpbd_capi.cu:248-260
int xpbd_compute_shapes(float deformEps, float* hostScales, float* hostQuats, int n)
{
    if (!g.stepped) return fail("xpbd_compute_shapes: no step yet");
    if (n != g.n) return fail("xpbd_compute_shapes: size mismatch");
    std::vector<float> sc, q;
    const int k = deformed_shapes(deformEps, sc, q);
    if (g.contactOn) contact_pack(sc.data(), q.data(), g.opacity.data(), s.contactTau, cap_or_zero(s.contactCap), g_cM, g_cRmax);
    if (hostScales) std::copy(sc.begin(), sc.end(), hostScales);
    if (hostQuats) std::copy(q.begin(), q.end(), hostQuats);
    return k;
}

static float g_jvw = 1.0f, g_jvwStd = 0.0f;

// xpbd_volume_stats (native/xpbd_capi.cu:306-310): J statistics of the last
// committed positions, evaluated on request.
void xpbd_volume_stats(int enable, float* jvw, float* jvwStd)
{
    s.volJStats = enable != 0;
    if (s.volJStats && g.stepped) {
        volume_ensure();
        volume_stats(g.x, g_jvw, g_jvwStd);
    }
    if (jvw) *jvw = g_jvw;
    if (jvwStd) *jvwStd = g_jvwStd;
}

// Explicit clusters: node i leads {i} + members[off[i]:off[i+1]] (empty = no cluster).
int xpbd_ext_set_volume_clusters(const int* off, const int* members, int minMembers)
{
    if (g.n <= 0) return fail("xpbd_ext_set_volume_clusters: call xpbd_create first");
    if (!off) {
        vol = VolumeState();
        return 0;
    }
    std::vector<char> isLeader(g.n, 0);
    std::vector<std::vector<int>> mem(g.n);
    for (int i = 0; i < g.n; ++i) {
        for (int k = off[i]; k < off[i + 1]; ++k) {
            const int j = members[k];
            if (j < 0 || j >= g.n || j == i) return fail("xpbd_ext_set_volume_clusters: member index out of range");
            mem[i].push_back(j);
        }
        isLeader[i] = off[i + 1] > off[i];
    }
    vol = VolumeState();
    vol.explicitTopo = true;
    volume_set_topology(isLeader, mem, std::max(3, minMembers));
    return 0;
}

int xpbd_ext_volume_members(int* members, int len)
{
    if (!g.stepped) return fail("xpbd_ext_volume_members: no step yet");
    volume_ensure();
    const int total = vol.memOff.empty() ? 0 : vol.memOff[g.n];
    if (len < total || (total > 0 && !members)) return fail("xpbd_ext_volume_members: buffer too small");
    for (int k = 0; k < total; ++k) members[k] = vol.memIdx[k];
    return total;
}

int xpbd_ext_volume_topology(int* leaders, int* memberCount, int* revCount, int* matType, float* detRest, int n)
{
    if (!g.stepped) return fail("xpbd_ext_volume_topology: no step yet");
    if (n != g.n) return fail("xpbd_ext_volume_topology: size mismatch");
    volume_ensure();
    for (int i = 0; i < n; ++i) {
        if (leaders) leaders[i] = vol.isLeader[i];
        if (memberCount) memberCount[i] = vol.memOff[i + 1] - vol.memOff[i];
        if (revCount) revCount[i] = vol.revOff[i + 1] - vol.revOff[i];
        if (matType) matType[i] = vol.matType[i];
        if (detRest) detRest[i] = vol.detRest[i];
    }
    return 0;
}

}  // extern "C"
