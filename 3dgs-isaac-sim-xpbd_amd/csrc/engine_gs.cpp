// engine_gs.cpp — graph-coloured Gauss-Seidel colourings of the Engine.
//
// The colourings are built on the host in the caller's node numbering with the
// reference's greedy rules, because the colour of every edge and cluster must
// match the reference's bit for bit (a device colouring would differ):
//   distance (forward.cu:10407-10517): directed entries -> undirected a < b,
//     sorted by (a, b), rest / stiffness averaged over duplicates, first-fit
//     colour not used at either end, colour-major counting sort;
//   volume (forward.cu:9745-9812): leaders in index order take the lowest
//     colour not used by any cluster containing the leader or a member;
//     skipped when sum over nodes of (clusters containing it)^2 > 4e9.
// Colour c of every body is then launched together (bodies share no nodes).
#include <algorithm>
#include <cstring>

#include "engine.h"

namespace xpbd {

#define XPBD_HIP(call)                                                                    \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            err = std::string(#call) + ": " + hipGetErrorString(e_);                      \
            return false;                                                                 \
        }                                                                                 \
    } while (0)

void build_distance_gs(const BodyInput& in, DistanceGSTopo& t)
{
    t = DistanceGSTopo();
    t.built = true;
    const int n = in.n;
    struct Dir { int a, b; float rest, stiff; int dup; };
    std::vector<Dir> dir;
    dir.reserve(in.nbr.size());
    for (int i = 0; i < n; ++i)
        for (int e = in.rowOff[i]; e < in.rowOff[i + 1]; ++e) {
            const int j = in.nbr[e];
            if (j < 0 || j >= n || j == i) continue;
            dir.push_back({std::min(i, j), std::max(i, j), in.nrest[e], in.nstiff[e], 1});
        }
    // a stable sort: duplicates of one pair keep the CSR order in the sums
    std::stable_sort(dir.begin(), dir.end(), [](const Dir& x, const Dir& y) { return x.a != y.a ? x.a < y.a : x.b < y.b; });
    std::vector<Dir> und;
    und.reserve(dir.size() / 2 + 1);
    for (const Dir& d : dir) {
        if (!und.empty() && und.back().a == d.a && und.back().b == d.b) {
            und.back().rest += d.rest;
            und.back().stiff += d.stiff;
            und.back().dup++;
        } else {
            und.push_back(d);
        }
    }
    for (Dir& u : und) {
        if (u.dup == 1) ++t.oneWay;
        u.rest /= float(u.dup);
        u.stiff /= float(u.dup);
    }
    const int m = (int)und.size();
    if (m == 0) { t.colorOff.assign(1, 0); return; }
    std::vector<int> deg(n, 0);
    for (const Dir& u : und) { deg[u.a]++; deg[u.b]++; }
    const int maxDeg = *std::max_element(deg.begin(), deg.end());
    const int words = (2 * maxDeg + 63) / 64;
    std::vector<unsigned long long> used((size_t)n * words, 0ull);
    std::vector<int> color(m);
    int nc = 0;
    for (int e = 0; e < m; ++e) {
        unsigned long long* ua = &used[(size_t)und[e].a * words];
        unsigned long long* ub = &used[(size_t)und[e].b * words];
        int c = -1;
        for (int w = 0; w < words && c < 0; ++w) {
            const unsigned long long fr = ~(ua[w] | ub[w]);
            if (fr) c = w * 64 + __builtin_ctzll(fr);
        }
        ua[c >> 6] |= 1ull << (c & 63);
        ub[c >> 6] |= 1ull << (c & 63);
        color[e] = c;
        nc = std::max(nc, c + 1);
    }
    t.colorOff.assign(nc + 1, 0);
    for (int e = 0; e < m; ++e) t.colorOff[color[e] + 1]++;
    for (int c = 0; c < nc; ++c) t.colorOff[c + 1] += t.colorOff[c];
    std::vector<int> fill(t.colorOff.begin(), t.colorOff.end() - 1);
    t.a.resize(m); t.b.resize(m); t.rest.resize(m); t.stiff.resize(m);
    for (int e = 0; e < m; ++e) {
        const int p = fill[color[e]]++;
        t.a[p] = und[e].a; t.b[p] = und[e].b;
        t.rest[p] = und[e].rest; t.stiff[p] = und[e].stiff;
    }
}

void build_volume_gs(const VolumeTopo& v, int n, VolumeGSTopo& t)
{
    t = VolumeGSTopo();
    t.built = true;
    t.version = v.version;
    if (v.memOff.empty()) return;
    double work = 0.0;
    for (int j = 0; j < n; ++j) {
        const int h = v.revOff[j + 1] - v.revOff[j];
        t.lowerBound = std::max(t.lowerBound, h);
        work += (double)h * h;
    }
    std::vector<int> leaders;
    for (int i = 0; i < n; ++i)
        if (v.isLeader[i]) leaders.push_back(i);
    if (leaders.empty()) return;
    if (work > 4.0e9) { t.skipped = true; return; }
    std::vector<int> color(n, -1), stamp;
    int nc = 0;
    for (int l : leaders) {
        if ((int)stamp.size() < nc + 1) stamp.resize(nc + 1, -1);
        auto forbid = [&](int j) {
            for (int k = v.revOff[j]; k < v.revOff[j + 1]; ++k) {
                const int c = color[v.revIdx[k]];
                if (c >= 0) stamp[c] = l;
            }
        };
        forbid(l);
        for (int k = v.memOff[l]; k < v.memOff[l + 1]; ++k) forbid(v.memIdx[k]);
        int c = 0;
        while (c < nc && stamp[c] == l) ++c;
        color[l] = c;
        nc = std::max(nc, c + 1);
    }
    t.colorOff.assign(nc + 1, 0);
    for (int l : leaders) t.colorOff[color[l] + 1]++;
    for (int c = 0; c < nc; ++c) t.colorOff[c + 1] += t.colorOff[c];
    std::vector<int> fill(t.colorOff.begin(), t.colorOff.end() - 1);
    t.order.resize(leaders.size());
    for (int l : leaders) t.order[fill[color[l]]++] = l;
}

// Device distance colouring: colour c of every participating body, in body
// order, edge ends renumbered to device nodes.
bool Engine::ensureDistanceGS(std::string& err)
{
    const int B = numBodies();
    std::vector<char> in(B, 0);
    for (int b = 0; b < B; ++b) in[b] = (set_[b].useDistance && set_[b].distGS) ? 1 : 0;
    if (dgsDevValid_ && in == dgsIn_) return true;
    XPBD_HIP(hipStreamSynchronize(stream_));   // queued substeps may still use the old buffers
    int nc = 0;
    for (int b = 0; b < B; ++b) {
        if (!in[b]) continue;
        if (!dgs_[b].built) build_distance_gs(*inputs_[b], dgs_[b]);
        nc = std::max(nc, (int)dgs_[b].colorOff.size() - 1);
    }
    std::vector<int2> ab;
    std::vector<float> rest, stiff;
    dgsColorOff_.assign(nc + 1, 0);
    for (int c = 0; c < nc; ++c) {
        for (int b = 0; b < B; ++b) {
            if (!in[b]) continue;
            const DistanceGSTopo& t = dgs_[b];
            if (c + 1 >= (int)t.colorOff.size()) continue;
            const int nb = bodyBegin_[b];
            for (int e = t.colorOff[c]; e < t.colorOff[c + 1]; ++e) {
                ab.push_back(make_int2(nb + inv_[b][t.a[e]], nb + inv_[b][t.b[e]]));
                rest.push_back(t.rest[e]);
                stiff.push_back(t.stiff[e]);
            }
        }
        dgsColorOff_[c + 1] = (int)ab.size();
    }
    const size_t m = std::max<size_t>(ab.size(), 1);
    dgsAB_.alloc(m); dgsRest_.alloc(m); dgsStiff_.alloc(m); dgsAlpha_.alloc(m); dgsLam_.alloc(m);
    if (!dgsAB_.p || !dgsRest_.p || !dgsStiff_.p || !dgsAlpha_.p || !dgsLam_.p) {
        err = "distance colouring: hipMalloc failed";
        return false;
    }
    if (!ab.empty()) {
        XPBD_HIP(hipMemcpyAsync(dgsAB_.p, ab.data(), sizeof(int2) * ab.size(), hipMemcpyHostToDevice, stream_));
        XPBD_HIP(hipMemcpyAsync(dgsRest_.p, rest.data(), sizeof(float) * rest.size(), hipMemcpyHostToDevice, stream_));
        XPBD_HIP(hipMemcpyAsync(dgsStiff_.p, stiff.data(), sizeof(float) * stiff.size(), hipMemcpyHostToDevice, stream_));
    }
    XPBD_HIP(hipStreamSynchronize(stream_));
    dgsIn_ = in;
    dgsAlphaKey_.clear();
    dgsDevValid_ = true;
    return true;
}

// Device volume colouring (after the volume topology is current).
bool Engine::ensureVolumeGS(std::string& err)
{
    const int B = numBodies();
    std::vector<char> in(B, 0);
    bool stale = !vgsDevValid_;
    for (int b = 0; b < B; ++b) {
        in[b] = (set_[b].useVolume && set_[b].volGS && !volTopo_[b].memOff.empty()) ? 1 : 0;
        if (!in[b]) continue;
        VolumeGSTopo& t = vgs_[b];
        if (!t.built || t.version != volTopo_[b].version) {
            if (!volumeHostLists(b, err)) return false;
            build_volume_gs(volTopo_[b], bodyCount_[b], t);
            stale = true;
        }
        if (t.skipped) in[b] = 0;   // the reference falls back to the gather pass (12066)
    }
    if (!stale && in == vgsIn_) return true;
    XPBD_HIP(hipStreamSynchronize(stream_));
    int nc = 0;
    for (int b = 0; b < B; ++b)
        if (in[b]) nc = std::max(nc, (int)vgs_[b].colorOff.size() - 1);
    std::vector<int> order;
    vgsColorOff_.assign(nc + 1, 0);
    for (int c = 0; c < nc; ++c) {
        for (int b = 0; b < B; ++b) {
            if (!in[b]) continue;
            const VolumeGSTopo& t = vgs_[b];
            if (c + 1 >= (int)t.colorOff.size()) continue;
            for (int k = t.colorOff[c]; k < t.colorOff[c + 1]; ++k) order.push_back(bodyBegin_[b] + inv_[b][t.order[k]]);
        }
        vgsColorOff_[c + 1] = (int)order.size();
    }
    vgsOrder_.alloc(std::max<size_t>(order.size(), 1));
    if (!vgsOrder_.p) { err = "volume colouring: hipMalloc failed"; return false; }
    if (!order.empty())
        XPBD_HIP(hipMemcpyAsync(vgsOrder_.p, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    vgsIn_ = in;
    vgsDevValid_ = true;
    return true;
}

bool Engine::gsInfo(int b, int info[6], std::string& err)
{
    XPBD_ALLOC_SCOPE;
    // both colourings exist whether or not GS is enabled, like the reference's
    // (forward.cu:10834, 10245); the volume one while the constraint is on
    if (!dgs_[b].built) build_distance_gs(*inputs_[b], dgs_[b]);
    const bool volOn = set_[b].useVolume;
    if (volOn) {
        if (!ensureVolume(false, err)) return false;
        if (!vgs_[b].built || vgs_[b].version != volTopo_[b].version) {
            if (!volumeHostLists(b, err)) return false;
            build_volume_gs(volTopo_[b], bodyCount_[b], vgs_[b]);
        }
    }
    const DistanceGSTopo& d = dgs_[b];
    const VolumeGSTopo& v = vgs_[b];
    info[0] = (int)d.a.size();
    info[1] = d.colorOff.empty() ? 0 : (int)d.colorOff.size() - 1;
    info[2] = d.oneWay;
    info[3] = volOn ? (int)v.order.size() : 0;
    info[4] = (volOn && !v.colorOff.empty()) ? (int)v.colorOff.size() - 1 : 0;
    info[5] = (volOn && v.skipped) ? 1 : 0;
    return true;
}

bool Engine::distanceGSEdges(int b, int* ab, float* rest, float* stiff, int* colorOff, std::string&)
{
    XPBD_ALLOC_SCOPE;
    if (!dgs_[b].built) build_distance_gs(*inputs_[b], dgs_[b]);
    const DistanceGSTopo& t = dgs_[b];
    for (size_t e = 0; e < t.a.size(); ++e) {
        if (ab) { ab[2 * e] = t.a[e]; ab[2 * e + 1] = t.b[e]; }
        if (rest) rest[e] = t.rest[e];
        if (stiff) stiff[e] = t.stiff[e];
    }
    if (colorOff) std::copy(t.colorOff.begin(), t.colorOff.end(), colorOff);
    return true;
}

bool Engine::volumeGSOrder(int b, int* order, int* colorOff, std::string& err)
{
    XPBD_ALLOC_SCOPE;
    if (!set_[b].useVolume) return true;
    if (!ensureVolume(false, err)) return false;
    if (!vgs_[b].built || vgs_[b].version != volTopo_[b].version) {
        if (!volumeHostLists(b, err)) return false;
        build_volume_gs(volTopo_[b], bodyCount_[b], vgs_[b]);
    }
    const VolumeGSTopo& t = vgs_[b];
    if (order) std::copy(t.order.begin(), t.order.end(), order);
    if (colorOff) std::copy(t.colorOff.begin(), t.colorOff.end(), colorOff);
    return true;
}

}  // namespace xpbd
