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
#include <thread>
#include <vector>

#include "engine.h"

namespace xpbd {

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
    static std::atomic<long long> versions{0};   // engines may build on several threads
    version = ++versions;
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
    t.finish(n);
}

}  // namespace xpbd
