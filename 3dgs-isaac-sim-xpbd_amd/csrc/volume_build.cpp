// volume_build.cpp — host construction of volume-constraint clusters.
//
// The reference rule (rebuildVolumeClusters, forward.cu:9814-10240, with
// farthest-point member selection and no cover maximum, which is what the C
// ABI selects): each leader's cluster is its k-hop graph ball (BFS in CSR
// order, at most 8192 nodes), reduced to `cap` members by farthest-point
// sampling on rest positions when larger.  Leaders are every node when the
// minimum leader hop r is 1 (built in parallel); for r >= 2 a sequential
// greedy cover picks seeds in index order, suppresses their (r-1)-hop balls,
// then promotes any node still covered fewer than min(6, r) times.  The
// reverse CSR lists, for every node, its own cluster first (if it leads one)
// followed by the clusters containing it in leader order.  The output is
// deterministic and identical to the reference's for the same graph.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <thread>
#include <vector>

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

namespace {

struct BallScratch {
    std::vector<int> stamp, frontier, next, collected;
    std::vector<float> md;
    std::vector<char> taken;
};

void ball_members(const BodyInput& in, int i, int k, int cap, BallScratch& w, std::vector<int>& out)
{
    const int n = in.n;
    const int kHardCap = 8192;
    std::vector<int>& col = w.collected;
    col.clear();
    if (k <= 1) {
        for (int e = in.rowOff[i]; e < in.rowOff[i + 1]; ++e) {
            const int j = in.nbr[e];
            if (j >= 0 && j < n && j != i) col.push_back(j);
        }
    } else {
        w.frontier.assign(1, i);
        w.stamp[i] = i;
        for (int d = 0; d < k && !w.frontier.empty() && (int)col.size() < kHardCap; ++d) {
            w.next.clear();
            bool capped = false;
            for (int f : w.frontier) {
                for (int e = in.rowOff[f]; e < in.rowOff[f + 1]; ++e) {
                    const int j = in.nbr[e];
                    if (j < 0 || j >= n || w.stamp[j] == i) continue;
                    w.stamp[j] = i;
                    col.push_back(j);
                    w.next.push_back(j);
                    if ((int)col.size() >= kHardCap) { capped = true; break; }
                }
                if (capped) break;
            }
            w.frontier.swap(w.next);
            if (capped) break;
        }
    }
    if ((int)col.size() <= cap) {
        out.assign(col.begin(), col.end());
        return;
    }
    // farthest-point sampling from the leader, squared float distances
    const float* P = in.pos.data();
    auto d2 = [&](int a, int b) {
        const float dx = P[3 * a] - P[3 * b], dy = P[3 * a + 1] - P[3 * b + 1], dz = P[3 * a + 2] - P[3 * b + 2];
        return dx * dx + dy * dy + dz * dz;
    };
    const int m = (int)col.size();
    w.md.resize(m);
    w.taken.assign(m, 0);
    int cur = 0;
    float best = -1.0f;
    for (int q = 0; q < m; ++q) {
        const float dd = d2(col[q], i);
        w.md[q] = dd;
        if (dd > best) { best = dd; cur = q; }
    }
    out.clear();
    for (int s = 0; s < cap; ++s) {
        w.taken[cur] = 1;
        out.push_back(col[cur]);
        int nxt = -1;
        float nbest = -1.0f;
        for (int q = 0; q < m; ++q) {
            if (w.taken[q]) continue;
            const float dd = d2(col[q], col[cur]);
            if (dd < w.md[q]) w.md[q] = dd;
            if (w.md[q] > nbest) { nbest = w.md[q]; nxt = q; }
        }
        if (nxt < 0) break;
        cur = nxt;
    }
}

}  // namespace

void VolumeTopo::finish(int n)
{
    touch();
    listsOnDevice = false;
    revOff.assign(n + 1, 0);
    std::vector<int> cnt(n, 0);
    for (int i = 0; i < n; ++i) {
        if (isLeader[i]) ++cnt[i];
        for (int k = memOff[i]; k < memOff[i + 1]; ++k) ++cnt[memIdx[k]];
    }
    for (int i = 0; i < n; ++i) revOff[i + 1] = revOff[i] + cnt[i];
    revIdx.resize(revOff[n]);
    std::vector<int> cur(revOff.begin(), revOff.end() - 1);
    for (int i = 0; i < n; ++i)
        if (isLeader[i]) revIdx[cur[i]++] = i;
    for (int i = 0; i < n; ++i)
        for (int k = memOff[i]; k < memOff[i + 1]; ++k) revIdx[cur[memIdx[k]]++] = i;
}

void VolumeTopo::touch()
{
    static std::atomic<long long> versions{0};   // engines may build on several threads
    version = ++versions;
}

namespace {

// Leader selection of the greedy cover (leaderR >= 2) over every node's ball
// (row i of ballOff/ballIdx): the rules of build_volume_reference below.
void greedy_cover(const BodyInput& in, int leaderR, const std::vector<int>& ballOff, const int* ballIdx,
                  std::vector<char>& isLeader)
{
    const int n = in.n;
    const int target = std::max(1, std::min(6, leaderR));
    std::vector<int> cover(n, 0), stampR(n, -1), frontier, next;
    isLeader.assign(n, 0);
    auto lead = [&](int i) {
        isLeader[i] = 1;
        ++cover[i];
        for (int q = ballOff[i]; q < ballOff[i + 1]; ++q) ++cover[ballIdx[q]];
    };
    for (int seed = 0; seed < n; ++seed) {
        if (cover[seed] >= target || stampR[seed] == -2) continue;
        lead(seed);
        frontier.assign(1, seed);
        stampR[seed] = -2;
        for (int d = 0; d < leaderR - 1 && !frontier.empty(); ++d) {
            next.clear();
            for (int f : frontier)
                for (int e = in.rowOff[f]; e < in.rowOff[f + 1]; ++e) {
                    const int j = in.nbr[e];
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
            if (cover[i] >= target || isLeader[i]) continue;
            lead(i);
            progressed = true;
        }
    }
}

template <typename T>
void swap_bufs(DevBuf<T>& a, DevBuf<T>& b)
{
    std::swap(a.p, b.p);
    std::swap(a.n, b.n);
}

}  // namespace

bool device_volume_build(const BodyInput& in, int k, int cap, int leaderR, hipStream_t s, VolumeTopo& t,
                         DevBuf<int>& memIdxD, DevBuf<int>& revIdxD, std::string& err)
{
    const int n = in.n;
    const auto t0 = std::chrono::steady_clock::now();
    auto ms = [&]() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); };
    const bool timing = getenv("XPBD_BUILD_TIMING") != nullptr;
    t.isLeader.clear();
    t.memOff.assign(n + 1, 0);
    t.memIdx.clear();
    t.revOff.assign(n + 1, 0);
    t.revIdx.clear();
    t.minMembers = 4;
    t.listsOnDevice = true;
    t.touch();
    memIdxD.free();
    revIdxD.free();
    if (n <= 0) return true;
    // 1. every node's ball: counts, offsets, lists
    DevBuf<int> rowOff, nbr, count, ballOff, ballIdx;
    DevBuf<float> pos;
    const size_t e = in.nbr.size();
    rowOff.alloc(n + 1); nbr.alloc(std::max<size_t>(e, 1)); pos.alloc(3 * (size_t)n); count.alloc(n);
    ballOff.alloc(n + 1);
    if (!rowOff.p || !nbr.p || !pos.p || !count.p || !ballOff.p) { err = "volume build: hipMalloc failed"; return false; }
    XPBD_HIP(hipMemcpyAsync(rowOff.p, in.rowOff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    if (e) XPBD_HIP(hipMemcpyAsync(nbr.p, in.nbr.data(), sizeof(int) * e, hipMemcpyHostToDevice, s));
    XPBD_HIP(hipMemcpyAsync(pos.p, in.pos.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
    VbArgs a{};
    a.n = n; a.k = k; a.cap = cap;
    a.rowOff = rowOff.p; a.nbr = nbr.p; a.pos = pos.p; a.count = count.p; a.off = ballOff.p; a.out = nullptr;
    launch_volume_balls(s, a, false);
    XPBD_HIP(hipGetLastError());
    std::vector<int> cnt(n);
    XPBD_HIP(hipMemcpyAsync(cnt.data(), count.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    XPBD_HIP(hipStreamSynchronize(s));
    if (timing) fprintf(stderr, "[build] device ball counts: %.1f ms\n", ms());
    // balls too large for the device: the host builds them
    std::vector<std::pair<int, std::vector<int>>> host;
    {
        BallScratch w;
        for (int i = 0; i < n; ++i)
            if (cnt[i] < 0) {
                if (w.stamp.empty()) w.stamp.assign(n, -1);
                host.emplace_back(i, std::vector<int>());
                ball_members(in, i, k, cap, w, host.back().second);
                cnt[i] = (int)host.back().second.size();
            }
    }
    std::vector<int> bOff(n + 1, 0);
    for (int i = 0; i < n; ++i) bOff[i + 1] = bOff[i] + cnt[i];
    if (bOff[n] < 0) { err = "volume build: member lists too large"; return false; }
    const int total = bOff[n];
    ballIdx.alloc(std::max(total, 1));
    if (!ballIdx.p) { err = "volume build: hipMalloc failed"; return false; }
    XPBD_HIP(hipMemcpyAsync(ballOff.p, bOff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    if (total > 0) {
        a.out = ballIdx.p;
        launch_volume_balls(s, a, true);   // rows of host-built balls keep count -1: skipped
        XPBD_HIP(hipGetLastError());
        for (const auto& h : host)
            if (!h.second.empty())
                XPBD_HIP(hipMemcpyAsync(ballIdx.p + bOff[h.first], h.second.data(), sizeof(int) * h.second.size(),
                                        hipMemcpyHostToDevice, s));
    }
    XPBD_HIP(hipStreamSynchronize(s));
    if (timing) fprintf(stderr, "[build] device balls: %.1f ms (%d entries, %zu on the host)\n", ms(), total, host.size());
    // 2. leaders: every node (r <= 1), else the greedy cover over the balls
    DevBuf<int> memOffD, leaderPos;
    int L = n;
    if (leaderR <= 1) {
        t.isLeader.assign(n, 1);
        t.memOff = bOff;
        swap_bufs(memIdxD, ballIdx);
        swap_bufs(memOffD, ballOff);
    } else {
        // every node's ball, unset until the copy lands (no zero fill of a
        // buffer this size)
        std::unique_ptr<int[]> hbBuf(new int[std::max(total, 1)]);
        int* hb = hbBuf.get();
        if (total > 0)
            XPBD_HIP(hipMemcpyAsync(hb, ballIdx.p, sizeof(int) * total, hipMemcpyDeviceToHost, s));
        XPBD_HIP(hipStreamSynchronize(s));
        greedy_cover(in, leaderR, bOff, hb, t.isLeader);
        std::vector<int> lp(n, -1);
        L = 0;
        for (int i = 0; i < n; ++i) {
            t.memOff[i + 1] = t.memOff[i] + (t.isLeader[i] ? cnt[i] : 0);
            if (t.isLeader[i]) lp[i] = L++;
        }
        t.memIdx.resize(t.memOff[n]);
        for (int i = 0; i < n; ++i)
            if (t.isLeader[i]) std::copy(hb + bOff[i], hb + bOff[i + 1], t.memIdx.begin() + t.memOff[i]);
        memIdxD.alloc(std::max<size_t>(t.memIdx.size(), 1));
        memOffD.alloc(n + 1);
        leaderPos.alloc(n);
        if (!memIdxD.p || !memOffD.p || !leaderPos.p) { err = "volume build: hipMalloc failed"; return false; }
        if (!t.memIdx.empty())
            XPBD_HIP(hipMemcpyAsync(memIdxD.p, t.memIdx.data(), sizeof(int) * t.memIdx.size(), hipMemcpyHostToDevice, s));
        XPBD_HIP(hipMemcpyAsync(memOffD.p, t.memOff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
        XPBD_HIP(hipMemcpyAsync(leaderPos.p, lp.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        if (timing) fprintf(stderr, "[build] greedy cover: %.1f ms (%d leaders)\n", ms(), L);
    }
    ballIdx.free();
    ballOff.free();
    rowOff.free(); nbr.free(); pos.free();
    // 3. reverse lists: (node, cluster) pairs sorted stably by node
    const int m = t.memOff[n];
    const long long pairs = (long long)L + m;
    if (pairs >= 0x7fffffffLL) { err = "volume build: reverse lists too large"; return false; }
    DevBuf<unsigned int> keys, keysAlt, hist;
    DevBuf<int> vals, valsAlt, revCnt, revOffD, scanTmp;
    keys.alloc(std::max<long long>(pairs, 1)); keysAlt.alloc(std::max<long long>(pairs, 1));
    vals.alloc(std::max<long long>(pairs, 1)); valsAlt.alloc(std::max<long long>(pairs, 1));
    hist.alloc(radix_sort_hist_size((int)std::max<long long>(pairs, 1)));
    revCnt.alloc(n); revOffD.alloc(n + 1); scanTmp.alloc(exclusive_scan_tmp_size(n));
    if (!keys.p || !keysAlt.p || !vals.p || !valsAlt.p || !hist.p || !revCnt.p || !revOffD.p || !scanTmp.p) {
        err = "volume build: hipMalloc failed";
        return false;
    }
    launch_volume_reverse(s, n, leaderR <= 1 ? nullptr : leaderPos.p, L, memOffD.p, memIdxD.p, m, revCnt.p, keys.p,
                          vals.p);
    XPBD_HIP(hipGetLastError());
    launch_exclusive_scan(s, n, revCnt.p, revOffD.p, scanTmp.p);
    int bits = 0;
    while ((1LL << bits) < (long long)n) ++bits;
    const bool inAlt = launch_radix_sort(s, (int)pairs, bits, keys.p, vals.p, keysAlt.p, valsAlt.p, hist.p, false);
    XPBD_HIP(hipGetLastError());
    if (inAlt) swap_bufs(vals, valsAlt);
    swap_bufs(revIdxD, vals);
    XPBD_HIP(hipMemcpyAsync(t.revOff.data(), revOffD.p, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, s));
    XPBD_HIP(hipStreamSynchronize(s));
    if (t.revOff[n] != pairs) { err = "volume build: reverse lists do not add up"; return false; }
    if (timing) fprintf(stderr, "[build] device reverse lists: %.1f ms (%lld pairs)\n", ms(), pairs);
    return true;
}

void build_volume_reference(const BodyInput& in, int k, int cap, int leaderR, VolumeTopo& t)
{
    const int n = in.n;
    std::vector<std::vector<int>> members(n);
    t.isLeader.assign(n, 0);
    if (leaderR <= 1) {
        std::fill(t.isLeader.begin(), t.isLeader.end(), (char)1);
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::atomic<int> next{0};
        auto worker = [&]() {
            BallScratch w;
            w.stamp.assign(n, -1);
            for (;;) {
                const int s = next.fetch_add(512);
                if (s >= n) break;
                for (int i = s; i < std::min(n, s + 512); ++i) ball_members(in, i, k, cap, w, members[i]);
            }
        };
        if (n < 4096) {
            worker();
        } else {
            std::vector<std::thread> pool;
            for (unsigned q = 0; q < hw; ++q) pool.emplace_back(worker);
            for (auto& th : pool) th.join();
        }
    } else {
        const int target = std::max(1, std::min(6, leaderR));
        std::vector<int> cover(n, 0), stampR(n, -1), frontier, next;
        BallScratch w;
        w.stamp.assign(n, -1);
        for (int seed = 0; seed < n; ++seed) {
            if (cover[seed] >= target || stampR[seed] == -2) continue;
            t.isLeader[seed] = 1;
            ball_members(in, seed, k, cap, w, members[seed]);
            ++cover[seed];
            for (int j : members[seed]) ++cover[j];
            frontier.assign(1, seed);
            stampR[seed] = -2;
            for (int d = 0; d < leaderR - 1 && !frontier.empty(); ++d) {
                next.clear();
                for (int f : frontier)
                    for (int e = in.rowOff[f]; e < in.rowOff[f + 1]; ++e) {
                        const int j = in.nbr[e];
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
                if (cover[i] >= target || t.isLeader[i]) continue;
                t.isLeader[i] = 1;
                ball_members(in, i, k, cap, w, members[i]);
                ++cover[i];
                for (int j : members[i]) ++cover[j];
                progressed = true;
            }
        }
    }
    t.memOff.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) t.memOff[i + 1] = t.memOff[i] + (int)members[i].size();
    t.memIdx.resize(t.memOff[n]);
    for (int i = 0; i < n; ++i) std::copy(members[i].begin(), members[i].end(), t.memIdx.begin() + t.memOff[i]);
    t.minMembers = 4;
    const auto t0 = std::chrono::steady_clock::now();
    t.finish(n);
    if (getenv("XPBD_BUILD_TIMING"))
        fprintf(stderr, "[build] reverse lists: %.1f ms\n",
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
}

}  // namespace xpbd
