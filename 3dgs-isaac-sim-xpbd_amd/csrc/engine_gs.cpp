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
#include <cstdlib>
#include <cstdio>
#include <chrono>
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
    const auto tb0 = std::chrono::steady_clock::now();
    struct Tm { std::chrono::steady_clock::time_point t0; ~Tm() { if (getenv("XPBD_BUILD_TIMING")) fprintf(stderr, "[build] distance GS edges and colours %.1f ms\n", std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count()); } } tm{tb0};
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
    struct Tm { std::chrono::steady_clock::time_point t0; ~Tm() { if (getenv("XPBD_BUILD_TIMING")) fprintf(stderr, "[build] volume GS colours %.1f ms\n", std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count()); } } tm{std::chrono::steady_clock::now()};
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

namespace {

template <typename T>
void swap_dev(DevBuf<T>& a, DevBuf<T>& b)
{
    std::swap(a.p, b.p);
    std::swap(a.n, b.n);
}

int bits_for(long long v)   // bits to hold the values 0 .. v
{
    int b = 0;
    while ((1LL << b) <= v) ++b;
    return b;
}

// Colour-major order of items 0 .. m-1 with colours color[] (< 256): a stable
// one-pass sort by colour; off[c] = first position of colour c.
bool colour_major(hipStream_t s, int m, const int* color, DevBuf<int>& order, std::vector<int>& colorOff,
                  std::string& err)
{
    DevBuf<unsigned int> k0, k1, hist;
    DevBuf<int> v1;
    k0.alloc(m); k1.alloc(m); v1.alloc(m); hist.alloc(radix_sort_hist_size(m)); order.alloc(m);
    if (!k0.p || !k1.p || !v1.p || !hist.p || !order.p) { err = "colouring: hipMalloc failed"; return false; }
    XPBD_HIP(hipMemcpyAsync(k0.p, color, sizeof(int) * m, hipMemcpyDeviceToDevice, s));
    const bool alt = launch_radix_sort(s, m, 8, k0.p, order.p, k1.p, v1.p, hist.p, true);
    if (alt) swap_dev(order, v1);
    std::vector<int> hc(m);
    XPBD_HIP(hipMemcpyAsync(hc.data(), color, sizeof(int) * m, hipMemcpyDeviceToHost, s));
    XPBD_HIP(hipStreamSynchronize(s));
    int nc = 0;
    for (int c : hc) nc = std::max(nc, c + 1);
    colorOff.assign(nc + 1, 0);
    for (int c : hc) colorOff[c + 1]++;
    for (int c = 0; c < nc; ++c) colorOff[c + 1] += colorOff[c];
    return true;
}

// Rounds of a device colouring until every item holds a colour; false in
// `ok` when an item needed more colours than the rounds track.
template <typename Round>
bool colour_rounds(hipStream_t s, DevBuf<int>& flags, Round round, bool& ok, int& rounds, std::string& err)
{
    int hf[2] = {1, 0};
    ok = true;
    for (rounds = 0; hf[0]; ++rounds) {
        if (rounds >= 100000) { ok = false; return true; }   // a dependency chain this long: the host colours
        launch_fill_int(s, 2, flags.p, 0);
        round();
        XPBD_HIP(hipGetLastError());
        XPBD_HIP(hipMemcpyAsync(hf, flags.p, sizeof(hf), hipMemcpyDeviceToHost, s));
        XPBD_HIP(hipStreamSynchronize(s));
        if (hf[1]) { ok = false; return true; }
    }
    return true;
}

}  // namespace

// The undirected edges and their greedy colouring of build_distance_gs, on
// the GPU: the directed entries sorted stably by (min, max) (two stable
// radix sorts: by max, then by min), runs of one pair summed in CSR order,
// the colouring in rounds.  ok = false: the colouring needs more colours than
// the rounds track (the caller colours on the host).
bool device_distance_gs(const BodyInput& in, hipStream_t s, DistanceGSTopo& t, bool& ok, std::string& err)
{
    const auto t0 = std::chrono::steady_clock::now();
    t = DistanceGSTopo();
    t.built = true;
    ok = true;
    const int n = in.n, m = (int)in.nbr.size();
    if (n <= 0 || m <= 0) { t.colorOff.assign(1, 0); return true; }
    DevBuf<int> rowOff, nbr, val, val2, head, rank, scanTmp, ea, eb, deg, incOff, cursor, inc, color, flags, oneWay, order;
    DevBuf<unsigned int> ka, kb, key, key2, hist;
    DevBuf<float> nrest, nstiff, erest, estiff;
    rowOff.alloc(n + 1); nbr.alloc(m); nrest.alloc(m); nstiff.alloc(m);
    ka.alloc(m); kb.alloc(m); val.alloc(m); key.alloc(m); key2.alloc(m); val2.alloc(m); hist.alloc(radix_sort_hist_size(m));
    head.alloc(m); rank.alloc(m + 1); scanTmp.alloc(exclusive_scan_tmp_size(std::max(m, n)));
    flags.alloc(2); oneWay.alloc(1);
    if (!rowOff.p || !nbr.p || !nrest.p || !nstiff.p || !ka.p || !kb.p || !val.p || !key.p || !key2.p || !val2.p ||
        !hist.p || !head.p || !rank.p || !scanTmp.p || !flags.p || !oneWay.p) {
        err = "distance colouring: hipMalloc failed";
        return false;
    }
    XPBD_HIP(hipMemcpyAsync(rowOff.p, in.rowOff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    XPBD_HIP(hipMemcpyAsync(nbr.p, in.nbr.data(), sizeof(int) * m, hipMemcpyHostToDevice, s));
    XPBD_HIP(hipMemcpyAsync(nrest.p, in.nrest.data(), sizeof(float) * m, hipMemcpyHostToDevice, s));
    XPBD_HIP(hipMemcpyAsync(nstiff.p, in.nstiff.data(), sizeof(float) * m, hipMemcpyHostToDevice, s));
    launch_gs_entry_keys(s, n, rowOff.p, nbr.p, ka.p, kb.p, val.p);
    const int bits = bits_for(n);
    // stable by max, then stable by min: (min, max, CSR position) order
    XPBD_HIP(hipMemcpyAsync(key.p, kb.p, sizeof(unsigned int) * m, hipMemcpyDeviceToDevice, s));
    if (launch_radix_sort(s, m, bits, key.p, val.p, key2.p, val2.p, hist.p, false)) swap_dev(val, val2);
    launch_gs_gather_key(s, m, ka.p, val.p, key.p);
    if (launch_radix_sort(s, m, bits, key.p, val.p, key2.p, val2.p, hist.p, false)) {
        swap_dev(val, val2);
        swap_dev(key, key2);
    }
    launch_gs_heads(s, n, m, key.p, kb.p, val.p, head.p);
    launch_exclusive_scan(s, m, head.p, rank.p, scanTmp.p);
    int me = 0;
    XPBD_HIP(hipMemcpyAsync(&me, rank.p + m, sizeof(int), hipMemcpyDeviceToHost, s));
    XPBD_HIP(hipStreamSynchronize(s));
    if (me == 0) { t.colorOff.assign(1, 0); return true; }
    ea.alloc(me); eb.alloc(me); erest.alloc(me); estiff.alloc(me); color.alloc(me);
    deg.alloc(n); incOff.alloc(n + 1); cursor.alloc(n); inc.alloc(2 * (size_t)me);
    if (!ea.p || !eb.p || !erest.p || !estiff.p || !color.p || !deg.p || !incOff.p || !cursor.p || !inc.p) {
        err = "distance colouring: hipMalloc failed";
        return false;
    }
    launch_fill_int(s, 1, oneWay.p, 0);
    launch_gs_edges(s, n, m, key.p, kb.p, val.p, head.p, rank.p, nrest.p, nstiff.p, ea.p, eb.p, erest.p, estiff.p,
                    oneWay.p);
    launch_fill_int(s, n, deg.p, 0);
    launch_gs_incidence(s, me, ea.p, eb.p, deg.p, nullptr, nullptr, true);
    launch_exclusive_scan(s, n, deg.p, incOff.p, scanTmp.p);
    XPBD_HIP(hipMemcpyAsync(cursor.p, incOff.p, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    launch_gs_incidence(s, me, ea.p, eb.p, nullptr, cursor.p, inc.p, false);
    launch_fill_int(s, me, color.p, -1);
    int rounds = 0;
    if (!colour_rounds(s, flags, [&]() { launch_gs_edge_round(s, me, ea.p, eb.p, incOff.p, inc.p, color.p, flags.p); },
                       ok, rounds, err))
        return false;
    if (!ok) return true;
    if (!colour_major(s, me, color.p, order, t.colorOff, err)) return false;
    std::vector<int> hord(me), ha(me), hb(me);
    std::vector<float> hr(me), hs(me);
    XPBD_HIP(hipMemcpyAsync(hord.data(), order.p, sizeof(int) * me, hipMemcpyDeviceToHost, s));
    XPBD_HIP(hipMemcpyAsync(ha.data(), ea.p, sizeof(int) * me, hipMemcpyDeviceToHost, s));
    XPBD_HIP(hipMemcpyAsync(hb.data(), eb.p, sizeof(int) * me, hipMemcpyDeviceToHost, s));
    XPBD_HIP(hipMemcpyAsync(hr.data(), erest.p, sizeof(float) * me, hipMemcpyDeviceToHost, s));
    XPBD_HIP(hipMemcpyAsync(hs.data(), estiff.p, sizeof(float) * me, hipMemcpyDeviceToHost, s));
    XPBD_HIP(hipMemcpyAsync(&t.oneWay, oneWay.p, sizeof(int), hipMemcpyDeviceToHost, s));
    XPBD_HIP(hipStreamSynchronize(s));
    t.a.resize(me); t.b.resize(me); t.rest.resize(me); t.stiff.resize(me);
    for (int p = 0; p < me; ++p) {
        const int u = hord[p];
        t.a[p] = ha[u]; t.b[p] = hb[u]; t.rest[p] = hr[u]; t.stiff[p] = hs[u];
    }
    if (getenv("XPBD_BUILD_TIMING"))
        fprintf(stderr, "[build] device distance GS edges and colours %.1f ms (%d edges, %d colours, %d rounds)\n",
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(), me,
                (int)t.colorOff.size() - 1, rounds);
    return true;
}

// The leaders' greedy colouring of build_volume_gs, on the GPU (the lists of
// the topology on the host are uploaded).  ok = false: more colours than the
// rounds track.
bool device_volume_gs(const VolumeTopo& v, int n, hipStream_t s, VolumeGSTopo& t, bool& ok, std::string& err)
{
    const auto t0 = std::chrono::steady_clock::now();
    t = VolumeGSTopo();
    t.built = true;
    t.version = v.version;
    ok = true;
    if (v.memOff.empty()) return true;
    double work = 0.0;
    for (int j = 0; j < n; ++j) {
        const int h = v.revOff[j + 1] - v.revOff[j];
        t.lowerBound = std::max(t.lowerBound, h);
        work += (double)h * h;
    }
    std::vector<int> leaders;
    for (int i = 0; i < n; ++i)
        if (v.isLeader[i]) leaders.push_back(i);
    if (leaders.empty()) return true;
    if (work > 4.0e9) { t.skipped = true; return true; }
    const int L = (int)leaders.size();
    const size_t M = v.memIdx.size(), R = v.revIdx.size();
    DevBuf<int> dl, memOff, memIdx, revOff, revIdx, color, lcolor, flags, order;
    dl.alloc(L); memOff.alloc(n + 1); memIdx.alloc(std::max<size_t>(M, 1)); revOff.alloc(n + 1);
    revIdx.alloc(std::max<size_t>(R, 1)); color.alloc(n); lcolor.alloc(L); flags.alloc(2);
    if (!dl.p || !memOff.p || !memIdx.p || !revOff.p || !revIdx.p || !color.p || !lcolor.p || !flags.p) {
        err = "volume colouring: hipMalloc failed";
        return false;
    }
    XPBD_HIP(hipMemcpyAsync(dl.p, leaders.data(), sizeof(int) * L, hipMemcpyHostToDevice, s));
    XPBD_HIP(hipMemcpyAsync(memOff.p, v.memOff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    if (M) XPBD_HIP(hipMemcpyAsync(memIdx.p, v.memIdx.data(), sizeof(int) * M, hipMemcpyHostToDevice, s));
    XPBD_HIP(hipMemcpyAsync(revOff.p, v.revOff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    if (R) XPBD_HIP(hipMemcpyAsync(revIdx.p, v.revIdx.data(), sizeof(int) * R, hipMemcpyHostToDevice, s));
    launch_fill_int(s, n, color.p, -1);
    int rounds = 0;
    if (!colour_rounds(s, flags,
                       [&]() {
                           launch_gs_cluster_round(s, L, dl.p, memOff.p, memIdx.p, revOff.p, revIdx.p, color.p, flags.p);
                       },
                       ok, rounds, err))
        return false;
    if (!ok) return true;
    // the leaders' colours in leader order, then colour-major
    std::vector<int> hc(n);
    XPBD_HIP(hipMemcpyAsync(hc.data(), color.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    XPBD_HIP(hipStreamSynchronize(s));
    std::vector<int> lc(L);
    for (int q = 0; q < L; ++q) lc[q] = hc[leaders[q]];
    XPBD_HIP(hipMemcpyAsync(lcolor.p, lc.data(), sizeof(int) * L, hipMemcpyHostToDevice, s));
    if (!colour_major(s, L, lcolor.p, order, t.colorOff, err)) return false;
    std::vector<int> hord(L);
    XPBD_HIP(hipMemcpyAsync(hord.data(), order.p, sizeof(int) * L, hipMemcpyDeviceToHost, s));
    XPBD_HIP(hipStreamSynchronize(s));
    t.order.resize(L);
    for (int p = 0; p < L; ++p) t.order[p] = leaders[hord[p]];
    if (getenv("XPBD_BUILD_TIMING"))
        fprintf(stderr, "[build] device volume GS colours %.1f ms (%d leaders, %d colours, %d rounds)\n",
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(), L,
                (int)t.colorOff.size() - 1, rounds);
    return true;
}

bool Engine::buildDistanceGS(int b, std::string& err)
{
    if (dgs_[b].built) return true;
    if (deviceBuilders_ == 1 || (deviceBuilders_ < 0 && bodyCount_[b] >= kDeviceBuildMin)) {
        bool ok = true;
        if (!device_distance_gs(*inputs_[b], stream_, dgs_[b], ok, err)) return false;
        if (ok) return true;
    }
    build_distance_gs(*inputs_[b], dgs_[b]);
    return true;
}

bool Engine::buildVolumeGS(int b, std::string& err)
{
    if (vgs_[b].built && vgs_[b].version == volTopo_[b].version) return true;
    if (!volumeHostLists(b, err)) return false;
    if (deviceBuilders_ == 1 || (deviceBuilders_ < 0 && bodyCount_[b] >= kDeviceBuildMin)) {
        bool ok = true;
        if (!device_volume_gs(volTopo_[b], bodyCount_[b], stream_, vgs_[b], ok, err)) return false;
        if (ok) return true;
    }
    build_volume_gs(volTopo_[b], bodyCount_[b], vgs_[b]);
    return true;
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
        if (!buildDistanceGS(b, err)) return false;
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
            if (!buildVolumeGS(b, err)) return false;
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
    if (!buildDistanceGS(b, err)) return false;
    const bool volOn = set_[b].useVolume;
    if (volOn) {
        if (!ensureVolume(false, err)) return false;
        if (!buildVolumeGS(b, err)) return false;
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

bool Engine::distanceGSEdges(int b, int* ab, float* rest, float* stiff, int* colorOff, std::string& err)
{
    XPBD_ALLOC_SCOPE;
    if (!buildDistanceGS(b, err)) return false;
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
    if (!buildVolumeGS(b, err)) return false;
    const VolumeGSTopo& t = vgs_[b];
    if (order) std::copy(t.order.begin(), t.order.end(), order);
    if (colorOff) std::copy(t.colorOff.begin(), t.colorOff.end(), colorOff);
    return true;
}

}  // namespace xpbd
