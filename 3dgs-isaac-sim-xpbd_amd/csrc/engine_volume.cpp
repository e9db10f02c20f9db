// engine_volume.cpp — volume-constraint topology management of the Engine:
// per-body cluster construction (k-ring rule or explicit), upload into the
// padded node space, rest precompute on the device, the active-leader list,
// and J statistics on request.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
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

VolumeArgs Engine::volArgs() const
{
    VolumeArgs a;
    a.nActive = volActive_;
    a.active = volActiveList_.p;
    a.volOff = volOff_.p;
    a.volIdx = volIdx_.p;
    a.revOff = revOff_.p;
    a.revIdx = revIdx_.p;
    a.detRest = detRest_.p;
    a.Vrest = Vrest_.p;
    a.restLen = restLen_.p;
    a.matType = matType_.p;
    a.lam = lamVol_.p;
    a.bbC = bbC_.p;
    a.bbS = bbS_.p;
    a.bbS2 = bbS2_.p;
    return a;
}

bool Engine::setVolumeClusters(int b, const int* off, const int* members, int minMembers, std::string& err)
{
    const int n = bodyCount_[b];
    VolumeTopo t;
    if (off) {
        t.isLeader.assign(n, 0);
        t.memOff.assign(n + 1, 0);
        for (int i = 0; i < n; ++i) {
            if (off[i + 1] < off[i]) { err = "volume clusters: offsets must be non-decreasing"; return false; }
            t.memOff[i + 1] = t.memOff[i] + (off[i + 1] - off[i]);
            t.isLeader[i] = off[i + 1] > off[i];
        }
        t.memIdx.resize(t.memOff[n]);
        for (int i = 0; i < n; ++i)
            for (int k = off[i]; k < off[i + 1]; ++k) {
                const int j = members[k];
                if (j < 0 || j >= n || j == i) { err = "volume clusters: member index out of range"; return false; }
                t.memIdx[t.memOff[i] + (k - off[i])] = j;
            }
        t.minMembers = std::max(3, minMembers);
        t.explicitClusters = true;
        t.finish(n);
    }
    volTopo_[b] = std::move(t);
    volDeviceValid_ = false;
    return true;
}

void Engine::setDeviceBuilders(int mode)
{
    if (mode == deviceBuilders_) return;
    deviceBuilders_ = mode;
    for (VolumeTopo& t : volTopo_)
        if (!t.explicitClusters) t.memOff.clear();
    volDeviceValid_ = false;
}

// Make the device topology current.  needAll: build clusters for every body
// (statistics), otherwise only for bodies with the volume constraint on.
bool Engine::ensureVolume(bool needAll, std::string& err)
{
    const int B = numBodies();
    bool changed = !volDeviceValid_;
    for (int b = 0; b < B; ++b) {
        const Settings& s = set_[b];
        if (!needAll && !s.useVolume) continue;
        VolumeTopo& t = volTopo_[b];
        if (t.explicitClusters) continue;
        if (!t.memOff.empty() && t.k == s.volRingK && t.cap == s.volMaxMembers && t.r == s.volLeaderMinHop) continue;
        const BodyInput& in = *inputs_[b];
        const auto tb0 = std::chrono::steady_clock::now();
        VolDevLists& dl = *volDev_[b];
        if (deviceBuilders_ == 1 || (deviceBuilders_ < 0 && in.n >= kDeviceBuildMin)) {
            if (!device_volume_build(in, s.volRingK, s.volMaxMembers, s.volLeaderMinHop, stream_, t, dl.memIdx,
                                     dl.revIdx, err))
                return false;
        } else {
            build_volume_reference(in, s.volRingK, s.volMaxMembers, s.volLeaderMinHop, t);
            dl.memIdx.free();
            dl.revIdx.free();
        }
        t.k = s.volRingK; t.cap = s.volMaxMembers; t.r = s.volLeaderMinHop;
        changed = true;
        if (getenv("XPBD_BUILD_TIMING"))
            fprintf(stderr, "[build] volume topology body %d: %.1f ms\n", b,
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tb0).count());
    }
    if (!changed) return true;
    const auto tc0 = std::chrono::steady_clock::now();

    // concatenate into the padded node space (bodies without clusters get
    // empty rows).  Device node d is the caller's node ord[d]; cluster member
    // lists and the clusters containing a node keep the caller-order
    // topology's order.  Offsets first, then the rows: on the host for
    // host-built topologies, by a device permutation for device-built ones.
    std::vector<int> hOff(nPad_ + 1, 0), hRevOff(nPad_ + 1, 0);
    long long tot = 0, rtot = 0;
    for (int b = 0; b < B; ++b) {
        const VolumeTopo& t = volTopo_[b];
        const int nb = bodyBegin_[b], n = bodyCount_[b];
        const int padEnd = (b + 1 < B) ? bodyBegin_[b + 1] : nPad_;
        const std::vector<int>& ord = ord_[b];
        for (int d = 0; d < n; ++d) {
            const int i = ord[d];
            hOff[nb + d] = (int)tot;
            hRevOff[nb + d] = (int)rtot;
            if (t.memOff.empty()) continue;
            tot += t.memOff[i + 1] - t.memOff[i];
            rtot += t.revOff[i + 1] - t.revOff[i];
            if (tot >= 0x7fffffffLL || rtot >= 0x7fffffffLL) { err = "volume: cluster lists too large"; return false; }
        }
        for (int i = nb + n; i <= padEnd && i <= nPad_; ++i) {
            hOff[i] = (int)tot;
            hRevOff[i] = (int)rtot;
        }
    }
    hOff[nPad_] = (int)tot;
    hRevOff[nPad_] = (int)rtot;
    volOff_.alloc(nPad_ + 1);
    revOff_.alloc(nPad_ + 1);
    volIdx_.alloc(std::max<size_t>((size_t)tot, 1));
    revIdx_.alloc(std::max<size_t>((size_t)rtot, 1));
    if (!detRest_.p) {
        detRest_.alloc(nPad_); Vrest_.alloc(nPad_); restLen_.alloc(nPad_); matType_.alloc(nPad_);
        lamVol_.alloc(nPad_); bbC_.alloc(nPad_); bbS_.alloc(nPad_); bbS2_.alloc(nPad_); volJ_.alloc(nPad_);
    }
    if (!volOff_.p || !revOff_.p || !volIdx_.p || !revIdx_.p || !detRest_.p || !Vrest_.p || !restLen_.p ||
        !matType_.p || !lamVol_.p || !bbC_.p || !bbS_.p || !bbS2_.p || !volJ_.p) {
        err = "volume: hipMalloc failed";
        return false;
    }
    XPBD_HIP(hipMemcpyAsync(volOff_.p, hOff.data(), sizeof(int) * (nPad_ + 1), hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipMemcpyAsync(revOff_.p, hRevOff.data(), sizeof(int) * (nPad_ + 1), hipMemcpyHostToDevice, stream_));
    std::vector<int> seg, rseg;
    for (int b = 0; b < B; ++b) {
        const VolumeTopo& t = volTopo_[b];
        const int nb = bodyBegin_[b], n = bodyCount_[b];
        if (t.memOff.empty() || n == 0) continue;
        const std::vector<int>& ord = ord_[b];
        const std::vector<int>& inv = inv_[b];
        const int m0 = hOff[nb], m1 = hOff[nb + n], r0 = hRevOff[nb], r1 = hRevOff[nb + n];
        if (t.listsOnDevice) {
            const VolDevLists& dl = *volDev_[b];
            DevBuf<int> dOrd, dInv, dMemOff, dRevOff;
            dOrd.alloc(n); dInv.alloc(n); dMemOff.alloc(n + 1); dRevOff.alloc(n + 1);
            if (!dOrd.p || !dInv.p || !dMemOff.p || !dRevOff.p) { err = "volume: hipMalloc failed"; return false; }
            XPBD_HIP(hipMemcpyAsync(dOrd.p, ord.data(), sizeof(int) * n, hipMemcpyHostToDevice, stream_));
            XPBD_HIP(hipMemcpyAsync(dInv.p, inv.data(), sizeof(int) * n, hipMemcpyHostToDevice, stream_));
            XPBD_HIP(hipMemcpyAsync(dMemOff.p, t.memOff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, stream_));
            XPBD_HIP(hipMemcpyAsync(dRevOff.p, t.revOff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, stream_));
            if (m1 > m0) launch_permute_rows(stream_, n, nb, dOrd.p, dInv.p, dMemOff.p, dl.memIdx.p, volOff_.p, volIdx_.p);
            if (r1 > r0) launch_permute_rows(stream_, n, nb, dOrd.p, dInv.p, dRevOff.p, dl.revIdx.p, revOff_.p, revIdx_.p);
            XPBD_HIP(hipGetLastError());
            XPBD_HIP(hipStreamSynchronize(stream_));   // before the temporaries go
            continue;
        }
        seg.resize(m1 - m0);
        rseg.resize(r1 - r0);
        size_t w = 0, rw = 0;
        for (int d = 0; d < n; ++d) {
            const int i = ord[d];
            for (int k = t.memOff[i]; k < t.memOff[i + 1]; ++k) seg[w++] = nb + inv[t.memIdx[k]];
            for (int k = t.revOff[i]; k < t.revOff[i + 1]; ++k) rseg[rw++] = nb + inv[t.revIdx[k]];
        }
        if (!seg.empty())
            XPBD_HIP(hipMemcpyAsync(volIdx_.p + m0, seg.data(), sizeof(int) * seg.size(), hipMemcpyHostToDevice, stream_));
        if (!rseg.empty())
            XPBD_HIP(hipMemcpyAsync(revIdx_.p + r0, rseg.data(), sizeof(int) * rseg.size(), hipMemcpyHostToDevice, stream_));
        XPBD_HIP(hipStreamSynchronize(stream_));   // seg and rseg are reused
    }
    if (getenv("XPBD_BUILD_TIMING"))
        fprintf(stderr, "[build] volume concatenation: %.1f ms (%lld members, %lld reverse)\n",
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tc0).count(), tot, rtot);
    XPBD_HIP(hipMemsetAsync(lamVol_.p, 0, sizeof(float) * nPad_, stream_));
    XPBD_HIP(hipMemsetAsync(bbC_.p, 0, sizeof(float4) * nPad_, stream_));
    // parameters of every body must be on the device before the precompute
    std::string perr;
    if (!uploadParams(perr)) { err = perr; return false; }
    const NodeArgs na{nPad_, tileBody_.p, bodies_.p};
    VolumeArgs va = volArgs();
    launch_vol_precompute(stream_, na, rest_.p, rowOff_.p, nbr_.p, va, volAniso_);
    hMatType_.resize(nPad_);
    hVrest_.resize(nPad_);
    XPBD_HIP(hipMemcpyAsync(hMatType_.data(), matType_.p, sizeof(int) * nPad_, hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipMemcpyAsync(hVrest_.data(), Vrest_.p, sizeof(float) * nPad_, hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    XPBD_HIP(hipGetLastError());
    std::vector<int> act;
    long long mem = 0;
    for (int i = 0; i < nPad_; ++i)
        if (hMatType_[i] == 0) {
            act.push_back(i);
            mem += hOff[i + 1] - hOff[i];
        }
    volActive_ = (int)act.size();
    volAvgMembers_ = act.empty() ? 0.0 : (double)mem / (double)act.size();
    volActiveList_.alloc(std::max<size_t>(act.size(), 1));
    if (!act.empty())
        XPBD_HIP(hipMemcpyAsync(volActiveList_.p, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    volDeviceValid_ = true;
    return true;
}

// Volume-weighted J of the committed positions (forward.cu:12658-12700).
bool Engine::volumeStats(int b, float& jvw, float& jvwStd, std::string& err)
{
    XPBD_ALLOC_SCOPE;
    if (!ensureVolume(true, err)) return false;
    const NodeArgs na{nPad_, tileBody_.p, bodies_.p};
    VolumeArgs va = volArgs();
    launch_vol_J(stream_, na, x_.p, va, volJ_.p);
    const int nb = bodyBegin_[b], n = bodyCount_[b];
    std::vector<float> J(n);
    XPBD_HIP(hipMemcpyAsync(J.data(), volJ_.p + nb, sizeof(float) * n, hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    double wSum = 0.0, wJ = 0.0, wJ2 = 0.0;
    for (int c = 0; c < n; ++c) {   // caller order: the sums are order-sensitive
        const int i = inv_[b][c];
        if (hMatType_[nb + i] != 0) continue;
        const float j = J[i];
        if (!std::isfinite(j)) continue;
        const double v = (double)hVrest_[nb + i];
        if (v > 0.0 && std::isfinite(v)) { wSum += v; wJ += v * j; wJ2 += v * (double)j * j; }
    }
    if (wSum > 0.0) {
        const double m = wJ / wSum;
        jvw = (float)m;
        jvwStd = (float)std::sqrt(std::fmax(0.0, wJ2 / wSum - m * m));
    }
    return true;
}

bool Engine::volumeTopology(int b, int* leaders, int* memberCount, int* revCount, int* matType, float* detRest,
                            std::string& err)
{
    XPBD_ALLOC_SCOPE;
    if (!ensureVolume(true, err)) return false;
    const VolumeTopo& t = volTopo_[b];
    const int nb = bodyBegin_[b], n = bodyCount_[b];
    std::vector<float> dr(n);
    if (detRest) {
        XPBD_HIP(hipMemcpyAsync(dr.data(), detRest_.p + nb, sizeof(float) * n, hipMemcpyDeviceToHost, stream_));
        XPBD_HIP(hipStreamSynchronize(stream_));
    }
    for (int i = 0; i < n; ++i) {
        const bool has = !t.memOff.empty();
        if (leaders) leaders[i] = has ? t.isLeader[i] : 0;
        if (memberCount) memberCount[i] = has ? t.memOff[i + 1] - t.memOff[i] : 0;
        if (revCount) revCount[i] = has ? t.revOff[i + 1] - t.revOff[i] : 0;
        if (matType) matType[i] = hMatType_[nb + inv_[b][i]];
        if (detRest) detRest[i] = dr[inv_[b][i]];
    }
    return true;
}

bool Engine::volumeHostLists(int b, std::string& err)
{
    VolumeTopo& t = volTopo_[b];
    if (!t.listsOnDevice) return true;
    const int n = bodyCount_[b];
    VolDevLists& dl = *volDev_[b];
    const size_t m = (size_t)t.memOff[n], r = (size_t)t.revOff[n];
    if (t.memIdx.size() != m) {
        t.memIdx.resize(m);
        if (m) XPBD_HIP(hipMemcpyAsync(t.memIdx.data(), dl.memIdx.p, sizeof(int) * m, hipMemcpyDeviceToHost, stream_));
    }
    t.revIdx.resize(r);
    if (r) XPBD_HIP(hipMemcpyAsync(t.revIdx.data(), dl.revIdx.p, sizeof(int) * r, hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    t.listsOnDevice = false;
    dl.memIdx.free();
    dl.revIdx.free();
    return true;
}

bool Engine::volumeMembers(int b, int* members, int len, int& total, std::string& err)
{
    XPBD_ALLOC_SCOPE;
    if (!ensureVolume(true, err)) return false;
    if (!volumeHostLists(b, err)) return false;
    const VolumeTopo& t = volTopo_[b];
    const int n = bodyCount_[b];
    total = t.memOff.empty() ? 0 : t.memOff[n];
    if (len < total || (total > 0 && !members)) { err = "buffer too small"; return false; }
    for (int k = 0; k < total; ++k) members[k] = t.memIdx[k];
    return true;
}

}  // namespace xpbd
