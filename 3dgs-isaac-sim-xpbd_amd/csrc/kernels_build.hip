// kernels_build.hip — device construction of the volume clusters' member
// lists: the k-hop graph ball of every node, reduced to `cap` members by
// farthest-point sampling (rebuildVolumeClusters' computeMembers,
// forward.cu:9862-10032, farthest-point selection).
//
// One wave64 per node at a time (persistent grid, nodes strided by the grid).
// The BFS keeps the host's order exactly: hop by hop, the frontier's rows in
// frontier order, each row in CSR order, a node appended at its first
// occurrence.  The wave walks the concatenated rows of 64 frontier nodes 64
// entries at a time; the visited set is an LDS hash set (linear probing)
// whose slots carry the step that inserted them, so the lanes of one step
// that meet the same new node agree on the lowest lane (LDS atomicMin) and
// only that lane appends it, at its rank among the step's appenders.  Balls
// larger than kVbMaxBall are left to the host (count -1).  Farthest-point
// sampling keeps the host's float expressions and its first-maximum rule
// (ties to the lowest candidate index).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace xpbd {

constexpr int kVbSetBits = 11;
constexpr int kVbSet = 1 << kVbSetBits;   // visited-set slots (load <= 0.76 at kVbMaxBall)
static_assert(kVbMaxBall + 64 < kVbSet * 4 / 5, "visited set too small for the largest device ball");

struct VbShared {
    union {
        int set[kVbSet];    // visited node ids (-1 free)
        float md[kVbSet];   // then the sampling's distances to the picked members
    };
    int meta[kVbSet];    // (step << 7) | lowest lane of that step holding the slot's node
    int col[kVbMaxBall]; // the ball in BFS order
    int pre[64];         // exclusive prefix of the row lengths of 64 frontier nodes
    int rs[64];          // their row starts
};

__device__ __forceinline__ unsigned int vb_hash(int j) { return ((unsigned int)j * 2654435761u) >> (32 - kVbSetBits); }

// Squared distance of rest positions a and b, as the host evaluates it.
__device__ __forceinline__ float vb_d2(const float* __restrict__ P, int a, int b)
{
    const float dx = P[3 * a] - P[3 * b], dy = P[3 * a + 1] - P[3 * b + 1], dz = P[3 * a + 2] - P[3 * b + 2];
    return dx * dx + dy * dy + dz * dz;
}

// Wave-wide first maximum: the largest value, the lowest index among equal ones.
__device__ __forceinline__ void vb_argmax(float& v, int& q)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const float v2 = __shfl_xor(v, o, 64);
        const int q2 = __shfl_xor(q, o, 64);
        if (v2 > v || (v2 == v && q2 >= 0 && (q < 0 || q2 < q))) {
            v = v2;
            q = q2;
        }
    }
}

// The ball of node i into sh.col[0, m); returns m, or -1 past kVbMaxBall.
__device__ int vb_ball(const VbArgs& a, int i, VbShared& sh, int lane)
{
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    if (a.k <= 1) {
        // the node's own row, duplicates kept (forward.cu:9882-9889)
        const int r0 = a.rowOff[i], r1 = a.rowOff[i + 1];
        int m = 0;
        for (int e0 = r0; e0 < r1; e0 += 64) {
            const int e = e0 + lane;
            const int j = e < r1 ? a.nbr[e] : -1;
            const bool ok = j >= 0 && j < a.n && j != i;
            const unsigned long long bal = __ballot(ok);
            const int pos = m + __popcll(bal & lt);
            if (ok && pos < kVbMaxBall) sh.col[pos] = j;
            m += __popcll(bal);
            if (m > kVbMaxBall) return -1;
        }
        __syncthreads();
        return m;
    }
    for (int s = lane; s < kVbSet; s += 64) sh.set[s] = -1;
    __syncthreads();
    if (lane == 0) {
        const unsigned int h = vb_hash(i);   // the set is empty: the home slot is free
        sh.set[h] = i;
        sh.meta[h] = 0;
    }
    __syncthreads();
    int m = 0, lo = 0, hi = 0;   // hop d > 0: the frontier is col[lo, hi)
    int step = 1;
    for (int d = 0; d < a.k; ++d) {
        const int nf = d == 0 ? 1 : hi - lo;
        if (nf <= 0) break;
        const int start = m;
        for (int g = 0; g < nf; g += 64) {
            const int q = g + lane;
            int rs = 0, rl = 0;
            if (q < nf) {
                const int f = d == 0 ? i : sh.col[lo + q];
                rs = a.rowOff[f];
                rl = a.rowOff[f + 1] - rs;
            }
            int x = rl;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(x, o, 64);
                if (lane >= o) x += t;
            }
            const int total = __shfl(x, 63, 64);
            sh.pre[lane] = x - rl;
            sh.rs[lane] = rs;
            __syncthreads();
            for (int base = 0; base < total; base += 64) {
                const int v = base + lane;
                int j = -1;
                if (v < total) {
                    // the last frontier lane whose run starts at or before v
                    // holds v (runs of length 0 share their successor's start)
                    int l0 = 0, l1 = 63;
                    while (l0 < l1) {
                        const int mid = (l0 + l1 + 1) >> 1;
                        if (sh.pre[mid] <= v) l0 = mid;
                        else l1 = mid - 1;
                    }
                    j = a.nbr[sh.rs[l0] + (v - sh.pre[l0])];
                }
                const bool cand = j >= 0 && j < a.n;
                unsigned int h = cand ? vb_hash(j) : 0u;
                bool probing = cand, won = false, seen = false;
                while (probing) {
                    const int old = atomicCAS(&sh.set[h], -1, j);
                    if (old == -1) {
                        won = true;
                        probing = false;
                    } else if (old == j) {
                        seen = true;
                        probing = false;
                    } else {
                        h = (h + 1) & (kVbSet - 1);
                    }
                }
                __syncthreads();
                if (won) sh.meta[h] = (step << 7) | 64;
                __syncthreads();
                const bool isNew = won || (seen && (sh.meta[h] >> 7) == step);
                __syncthreads();
                if (isNew) atomicMin(&sh.meta[h], (step << 7) | lane);
                __syncthreads();
                const bool app = isNew && (sh.meta[h] & 127) == lane;
                const unsigned long long bal = __ballot(app);
                const int pos = m + __popcll(bal & lt);
                if (app && pos < kVbMaxBall) sh.col[pos] = j;
                m += __popcll(bal);
                ++step;
                if (m > kVbMaxBall) return -1;
            }
            __syncthreads();
        }
        lo = start;
        hi = m;
    }
    __syncthreads();
    return m;
}

// Farthest-point sampling of `cap` members out of col[0, m) (m > cap), seeded
// by the candidate farthest from i; writes them to out when given; returns
// the number picked.
__device__ int vb_fps(const VbArgs& a, int i, int m, VbShared& sh, int lane, int* __restrict__ out)
{
    float* md = sh.md;
    float best = -1.0f;
    int cur = -1;
    for (int q = lane; q < m; q += 64) {
        const float dd = vb_d2(a.pos, sh.col[q], i);
        md[q] = dd;
        if (dd > best) {
            best = dd;
            cur = q;
        }
    }
    vb_argmax(best, cur);
    if (cur < 0) cur = 0;
    __syncthreads();
    int picked = 0;
    for (int s = 0; s < a.cap; ++s) {
        const int cj = sh.col[cur];
        if (lane == 0) {
            md[cur] = -INFINITY;   // taken
            if (out) out[s] = cj;
        }
        ++picked;
        __syncthreads();
        float nbest = -1.0f;
        int nxt = -1;
        for (int q = lane; q < m; q += 64) {
            float mq = md[q];
            if (mq == -INFINITY) continue;
            const float dd = vb_d2(a.pos, sh.col[q], cj);
            if (dd < mq) {
                mq = dd;
                md[q] = dd;
            }
            if (mq > nbest) {
                nbest = mq;
                nxt = q;
            }
        }
        vb_argmax(nbest, nxt);
        __syncthreads();
        if (nxt < 0) break;
        cur = nxt;
    }
    return picked;
}

// FILL = false: count[i] = members of node i (-1: the host builds it);
// FILL = true: the members of every device-built node at out[off[i]].
template <bool FILL>
__global__ void __launch_bounds__(64) vb_ball_kernel(VbArgs a)
{
    __shared__ VbShared sh;
    const int lane = threadIdx.x;
    for (int i = blockIdx.x; i < a.n; i += gridDim.x) {
        if (FILL && a.count[i] < 0) continue;   // uniform per block
        const int m = vb_ball(a, i, sh, lane);
        int cnt = m;
        if (m >= 0) {
            if (m <= a.cap) {
                if (FILL) {
                    int* o = a.out + a.off[i];
                    for (int q = lane; q < m; q += 64) o[q] = sh.col[q];
                }
            } else {
                cnt = vb_fps(a, i, m, sh, lane, FILL ? a.out + a.off[i] : nullptr);
            }
        }
        if (!FILL && lane == 0) a.count[i] = cnt;
        __syncthreads();
    }
}

// Reverse lists (caller order).  cnt[i]: clusters containing node i, its own
// (leaders) plus one per member entry naming it.
__global__ void __launch_bounds__(256) vb_rev_count_kernel(int n, const int* __restrict__ leaderPos, int* __restrict__ cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = leaderPos ? (leaderPos[i] >= 0 ? 1 : 0) : 1;
}

__global__ void __launch_bounds__(256) vb_rev_hist_kernel(int m, const int* __restrict__ memIdx, int* __restrict__ cnt)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < m) atomicAdd(&cnt[memIdx[e]], 1);
}

// (node, cluster) pairs in the order the stable sort must keep among equal
// nodes: every leader's own pair first (leaders ascending), then the member
// entries in list order (clusters ascending).  One wave per node.
__global__ void __launch_bounds__(256) vb_rev_pairs_kernel(int n, const int* __restrict__ leaderPos,
                                                          const int* __restrict__ memOff,
                                                          const int* __restrict__ memIdx, int L,
                                                          unsigned int* __restrict__ keys, int* __restrict__ vals)
{
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const int lp = leaderPos ? leaderPos[i] : i;
    if (lane == 0 && lp >= 0) {
        keys[lp] = (unsigned int)i;
        vals[lp] = i;
    }
    const int e0 = memOff[i], e1 = memOff[i + 1];
    for (int e = e0 + lane; e < e1; e += 64) {
        keys[L + e] = (unsigned int)memIdx[e];
        vals[L + e] = i;
    }
}

// Rows of one body in the padded device order: row nb + d is caller row
// ord[d], its entries mapped to device ids nb + inv[x].  One wave per row.
__global__ void __launch_bounds__(256) vb_permute_rows_kernel(int n, int nb, const int* __restrict__ ord,
                                                             const int* __restrict__ inv,
                                                             const int* __restrict__ srcOff,
                                                             const int* __restrict__ srcIdx,
                                                             const int* __restrict__ dstOff, int* __restrict__ dst)
{
    const int d = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (d >= n) return;
    const int i = ord[d];
    const int s0 = srcOff[i], len = srcOff[i + 1] - s0;
    const int o = dstOff[nb + d];
    for (int q = lane; q < len; q += 64) dst[o + q] = nb + inv[srcIdx[s0 + q]];
}

static inline int vb_blocks(long long items, int per) { return (int)((items + per - 1) / per); }

void launch_volume_reverse(hipStream_t s, int n, const int* leaderPos, int L, const int* memOff, const int* memIdx,
                           int m, int* cnt, unsigned int* keys, int* vals)
{
    if (n <= 0) return;
    vb_rev_count_kernel<<<vb_blocks(n, 256), 256, 0, s>>>(n, leaderPos, cnt);
    if (m > 0) vb_rev_hist_kernel<<<vb_blocks(m, 256), 256, 0, s>>>(m, memIdx, cnt);
    vb_rev_pairs_kernel<<<vb_blocks(64LL * n, 256), 256, 0, s>>>(n, leaderPos, memOff, memIdx, L, keys, vals);
}

void launch_permute_rows(hipStream_t s, int n, int nb, const int* ord, const int* inv, const int* srcOff,
                         const int* srcIdx, const int* dstOff, int* dst)
{
    if (n > 0) vb_permute_rows_kernel<<<vb_blocks(64LL * n, 256), 256, 0, s>>>(n, nb, ord, inv, srcOff, srcIdx, dstOff, dst);
}

// ---------------------------------------------------------------------------
// Gauss-Seidel colourings (buildDistanceGSEdges forward.cu:10407-10517,
// buildVolumeGSColors 9745-9812).  The reference colours greedily: items in
// index order, each the smallest colour its earlier neighbours do not hold.
// The same colouring comes out of rounds in which every item whose earlier
// neighbours all hold colours takes the smallest free one: an item decides
// from the same set of colours as in the sequential pass, and the first
// uncoloured item is always ready, so every round makes progress.  A colour
// read while its item is being coloured in the same round is either the
// final value or still -1 (the reader then waits a round).

// Directed CSR entries of one body as (min, max) pairs; invalid entries
// (out of range, self) get n, n and sort last.  Entry e keeps value e, so a
// stable sort by (min, max) keeps the CSR order among duplicates.
__global__ void __launch_bounds__(256) gs_entry_keys_kernel(int n, const int* __restrict__ rowOff,
                                                           const int* __restrict__ nbr, unsigned int* __restrict__ ka,
                                                           unsigned int* __restrict__ kb, int* __restrict__ val)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int e = rowOff[i]; e < rowOff[i + 1]; ++e) {
        const int j = nbr[e];
        const bool ok = j >= 0 && j < n && j != i;
        ka[e] = ok ? (unsigned int)min(i, j) : (unsigned int)n;
        kb[e] = ok ? (unsigned int)max(i, j) : (unsigned int)n;
        val[e] = e;
    }
}

// key2[k] = ka[val[k]]: the major key of the second (stable) pass.
__global__ void __launch_bounds__(256) gs_gather_key_kernel(int m, const unsigned int* __restrict__ ka,
                                                           const int* __restrict__ val, unsigned int* __restrict__ key)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) key[k] = ka[val[k]];
}

// head[k] = 1 where a run of equal (min, max) pairs starts (valid pairs only).
__global__ void __launch_bounds__(256) gs_heads_kernel(int n, int m, const unsigned int* __restrict__ sa,
                                                      const unsigned int* __restrict__ kb,
                                                      const int* __restrict__ val, int* __restrict__ head)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const unsigned int a = sa[k], b = kb[val[k]];
    head[k] = a < (unsigned int)n && (k == 0 || sa[k - 1] != a || kb[val[k - 1]] != b) ? 1 : 0;
}

// One thread per run: the undirected edge with its rest length and stiffness
// summed in the run's order (CSR order) and averaged.
__global__ void __launch_bounds__(256) gs_edges_kernel(int n, int m, const unsigned int* __restrict__ sa,
                                                      const unsigned int* __restrict__ kb, const int* __restrict__ val,
                                                      const int* __restrict__ head, const int* __restrict__ rank,
                                                      const float* __restrict__ nrest, const float* __restrict__ nstiff,
                                                      int* __restrict__ ea, int* __restrict__ eb,
                                                      float* __restrict__ erest, float* __restrict__ estiff,
                                                      int* __restrict__ oneWay)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m || !head[k]) return;
    const unsigned int a = sa[k], b = kb[val[k]];
    float rest = nrest[val[k]], stiff = nstiff[val[k]];
    int dup = 1;
    for (int q = k + 1; q < m && sa[q] == a && kb[val[q]] == b; ++q) {
        rest += nrest[val[q]];
        stiff += nstiff[val[q]];
        ++dup;
    }
    const int u = rank[k];
    ea[u] = (int)a;
    eb[u] = (int)b;
    erest[u] = rest / float(dup);
    estiff[u] = stiff / float(dup);
    if (dup == 1) atomicAdd(oneWay, 1);
}

// Incidence lists: each edge under both of its nodes (order inside a list free).
__global__ void __launch_bounds__(256) gs_degree_kernel(int me, const int* __restrict__ ea, const int* __restrict__ eb,
                                                       int* __restrict__ deg)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= me) return;
    atomicAdd(&deg[ea[u]], 1);
    atomicAdd(&deg[eb[u]], 1);
}

__global__ void __launch_bounds__(256) gs_incidence_kernel(int me, const int* __restrict__ ea,
                                                          const int* __restrict__ eb, int* __restrict__ cursor,
                                                          int* __restrict__ inc)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= me) return;
    inc[atomicAdd(&cursor[ea[u]], 1)] = u;
    inc[atomicAdd(&cursor[eb[u]], 1)] = u;
}

// Colours held by the earlier items listed in rows `rows` of (off, idx)
// (items < u); false when one of them has none yet.
__device__ __forceinline__ bool gs_collect(int u, int r0, int r1, const int* __restrict__ idx,
                                           const int* __restrict__ color, unsigned long long* used)
{
    for (int k = r0; k < r1; ++k) {
        const int v = idx[k];
        if (v >= u) continue;
        const int c = __atomic_load_n(&color[v], __ATOMIC_RELAXED);
        if (c < 0) return false;
        const unsigned long long bit = 1ull << (c & 63);
#pragma unroll
        for (int w = 0; w < kGsWords; ++w) used[w] |= (c >> 6) == w ? bit : 0ull;   // static indices: no scratch
    }
    return true;
}

__device__ __forceinline__ int gs_first_free(const unsigned long long* used)
{
#pragma unroll
    for (int w = 0; w < kGsWords; ++w)
        if (~used[w]) return 64 * w + __builtin_ctzll(~used[w]);
    return -1;
}

// One round over the edges: an edge whose earlier incident edges (at either
// node) hold colours takes the smallest colour none of them holds.
__global__ void __launch_bounds__(256) gs_edge_round_kernel(int me, const int* __restrict__ ea,
                                                           const int* __restrict__ eb, const int* __restrict__ incOff,
                                                           const int* __restrict__ inc, int* __restrict__ color,
                                                           int* __restrict__ flags)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= me || color[u] >= 0) return;
    unsigned long long used[kGsWords] = {};
    const int a = ea[u], b = eb[u];
    if (!gs_collect(u, incOff[a], incOff[a + 1], inc, color, used) ||
        !gs_collect(u, incOff[b], incOff[b + 1], inc, color, used)) {
        flags[0] = 1;   // another round
        return;
    }
    const int c = gs_first_free(used);
    if (c < 0) {
        flags[1] = 1;   // more colours than the rounds track: the host colours instead
        __atomic_store_n(&color[u], 64 * kGsWords, __ATOMIC_RELAXED);
        return;
    }
    __atomic_store_n(&color[u], c, __ATOMIC_RELAXED);
}

// One round over the leaders (nodes with a cluster, in index order): a
// leader whose earlier conflicting leaders (clusters holding the leader or
// one of its members) hold colours takes the smallest colour none holds.
__global__ void __launch_bounds__(256) gs_cluster_round_kernel(int L, const int* __restrict__ leaders,
                                                              const int* __restrict__ memOff,
                                                              const int* __restrict__ memIdx,
                                                              const int* __restrict__ revOff,
                                                              const int* __restrict__ revIdx, int* __restrict__ color,
                                                              int* __restrict__ flags)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= L) return;
    const int l = leaders[q];
    if (color[l] >= 0) return;
    unsigned long long used[kGsWords] = {};
    bool ready = gs_collect(l, revOff[l], revOff[l + 1], revIdx, color, used);
    for (int k = memOff[l]; ready && k < memOff[l + 1]; ++k) {
        const int j = memIdx[k];
        ready = gs_collect(l, revOff[j], revOff[j + 1], revIdx, color, used);
    }
    if (!ready) {
        flags[0] = 1;
        return;
    }
    const int c = gs_first_free(used);
    if (c < 0) {
        flags[1] = 1;
        __atomic_store_n(&color[l], 64 * kGsWords, __ATOMIC_RELAXED);
        return;
    }
    __atomic_store_n(&color[l], c, __ATOMIC_RELAXED);
}

void launch_gs_entry_keys(hipStream_t s, int n, const int* rowOff, const int* nbr, unsigned int* ka, unsigned int* kb,
                          int* val)
{
    if (n > 0) gs_entry_keys_kernel<<<vb_blocks(n, 256), 256, 0, s>>>(n, rowOff, nbr, ka, kb, val);
}
void launch_gs_gather_key(hipStream_t s, int m, const unsigned int* ka, const int* val, unsigned int* key)
{
    if (m > 0) gs_gather_key_kernel<<<vb_blocks(m, 256), 256, 0, s>>>(m, ka, val, key);
}
void launch_gs_heads(hipStream_t s, int n, int m, const unsigned int* sa, const unsigned int* kb, const int* val,
                     int* head)
{
    if (m > 0) gs_heads_kernel<<<vb_blocks(m, 256), 256, 0, s>>>(n, m, sa, kb, val, head);
}
void launch_gs_edges(hipStream_t s, int n, int m, const unsigned int* sa, const unsigned int* kb, const int* val,
                     const int* head, const int* rank, const float* nrest, const float* nstiff, int* ea, int* eb,
                     float* erest, float* estiff, int* oneWay)
{
    if (m > 0)
        gs_edges_kernel<<<vb_blocks(m, 256), 256, 0, s>>>(n, m, sa, kb, val, head, rank, nrest, nstiff, ea, eb, erest,
                                                           estiff, oneWay);
}
void launch_gs_incidence(hipStream_t s, int me, const int* ea, const int* eb, int* deg, int* cursor, int* inc,
                         bool countOnly)
{
    if (me <= 0) return;
    if (countOnly) gs_degree_kernel<<<vb_blocks(me, 256), 256, 0, s>>>(me, ea, eb, deg);
    else gs_incidence_kernel<<<vb_blocks(me, 256), 256, 0, s>>>(me, ea, eb, cursor, inc);
}
void launch_gs_edge_round(hipStream_t s, int me, const int* ea, const int* eb, const int* incOff, const int* inc,
                          int* color, int* flags)
{
    if (me > 0) gs_edge_round_kernel<<<vb_blocks(me, 256), 256, 0, s>>>(me, ea, eb, incOff, inc, color, flags);
}
void launch_gs_cluster_round(hipStream_t s, int L, const int* leaders, const int* memOff, const int* memIdx,
                             const int* revOff, const int* revIdx, int* color, int* flags)
{
    if (L > 0)
        gs_cluster_round_kernel<<<vb_blocks(L, 256), 256, 0, s>>>(L, leaders, memOff, memIdx, revOff, revIdx, color,
                                                                   flags);
}

void launch_volume_balls(hipStream_t s, const VbArgs& a, bool fill)
{
    if (a.n <= 0) return;
    const int grid = a.n < kVbGrid ? a.n : kVbGrid;
    if (fill) vb_ball_kernel<true><<<grid, 64, 0, s>>>(a);
    else vb_ball_kernel<false><<<grid, 64, 0, s>>>(a);
}

}  // namespace xpbd
