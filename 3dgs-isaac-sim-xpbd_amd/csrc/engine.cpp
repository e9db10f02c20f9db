// engine.cpp — host orchestration of the MI355X XPBD solver (see engine.h).
#include "engine.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace xpbd {

#define XPBD_HIP(call)                                                                    \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            err = std::string(#call) + ": " + hipGetErrorString(e_);                      \
            return false;                                                                 \
        }                                                                                 \
    } while (0)

// A captured graph holds raw pointers, so any reallocation must trigger a
// recapture.  Allocations made inside an Engine call bump that engine's own
// counter (an AllocScope names it for the calling thread); any other
// allocation bumps the global counter, which every engine's plan includes.
// Engines that build on one thread while others step on another therefore do
// not force each other's graphs to be recaptured.
static std::atomic<int> g_allocEpoch{0};
static thread_local int* t_allocEpoch = nullptr;
static int device_alloc_epoch() { return g_allocEpoch.load(); }

AllocScope::AllocScope(int* e) : prev(t_allocEpoch) { t_allocEpoch = e; }
AllocScope::~AllocScope() { t_allocEpoch = prev; }

template <typename T>
void DevBuf<T>::alloc(size_t count)
{
    free();
    if (t_allocEpoch) ++*t_allocEpoch;
    else ++g_allocEpoch;
    n = count;
    if (count) {
        if (hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * count) != hipSuccess) {
            p = nullptr;
            n = 0;
        }
    }
}

template <typename T>
void DevBuf<T>::free()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
}

template struct DevBuf<float>;
template struct DevBuf<float3>;
template struct DevBuf<float4>;
template struct DevBuf<int>;
template struct DevBuf<BodyDev>;
template struct DevBuf<ColliderDev>;
template struct DevBuf<float2>;
template struct DevBuf<unsigned int>;
template struct DevBuf<unsigned char>;
template struct DevBuf<double>;
template struct DevBuf<int2>;
template struct DevBuf<int4>;

// wave64-per-cluster volume solve from this mean member count up
static const double kVolWaveMembers = 24.0;
// self-collision search radius and fast-pair threshold in units of d_c
// (forward.cu:230, 262)
static const float kScSearchScale = 1.5f;
static const float kScFastRelScale = 0.0f;
// node window (a multiple of the 64-node slice) sorted by row length
static const int kSortWindow = 8 * kTile;

void build_input(BodyInput& in, int n, const float* pos, const float* scales, const float* rots, const float* opacity,
                 int e, const int* edges, const float* rest, const float* stiff)
{
    in.n = n;
    in.pos.assign(pos, pos + 3 * (size_t)n);
    in.scales.assign(scales, scales + 3 * (size_t)n);
    in.rots.assign(rots, rots + 4 * (size_t)n);
    in.opacity.assign(opacity, opacity + n);
    // counting pass, then fill in input edge order (a->b then b->a per edge)
    std::vector<int> deg(n + 1, 0);
    for (int k = 0; k < e; ++k) {
        const int a = edges[2 * k], b = edges[2 * k + 1];
        if (a < 0 || b < 0 || a >= n || b >= n || a == b) continue;
        deg[a]++;
        deg[b]++;
    }
    in.rowOff.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) in.rowOff[i + 1] = in.rowOff[i] + deg[i];
    const int m = in.rowOff[n];
    in.nbr.resize(m);
    in.nrest.resize(m);
    in.nstiff.resize(m);
    std::vector<int> cur(in.rowOff.begin(), in.rowOff.end() - 1);
    for (int k = 0; k < e; ++k) {
        const int a = edges[2 * k], b = edges[2 * k + 1];
        if (a < 0 || b < 0 || a >= n || b >= n || a == b) continue;
        int s = cur[a]++;
        in.nbr[s] = b; in.nrest[s] = rest[k]; in.nstiff[s] = stiff[k];
        s = cur[b]++;
        in.nbr[s] = a; in.nrest[s] = rest[k]; in.nstiff[s] = stiff[k];
    }
}

// Z-order of a body's rest positions: coordinates quantised to 21 bits on
// the body's bounding cube, bits interleaved x, y, z; ties keep caller order.
static uint64_t spread_bits21(uint64_t v)
{
    v &= 0x1fffffULL;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

// fn(begin, end) over [0, n) on up to 16 host threads (one thread below
// 65536 items); every item is handled once, so the result does not depend
// on the split.
template <typename F>
static void host_parallel(long long n, F fn)
{
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 65536 || hw == 1) {
        fn(0LL, n);
        return;
    }
    std::vector<std::thread> pool;
    const long long chunk = (n + hw - 1) / hw;
    for (unsigned t = 0; t < hw; ++t) {
        const long long b = (long long)t * chunk, e = std::min(n, b + chunk);
        if (b < e) pool.emplace_back(fn, b, e);
    }
    for (std::thread& th : pool) th.join();
}

// fn(k) for k in [0, count), tasks spread over up to 16 host threads.
template <typename F>
static void host_tasks(int count, F fn)
{
    const int hw = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (count <= 1 || hw == 1) {
        for (int k = 0; k < count; ++k) fn(k);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < std::min(hw, count); ++t)
        pool.emplace_back([&]() {
            for (int k = next++; k < count; k = next++) fn(k);
        });
    for (std::thread& th : pool) th.join();
}

static void morton_order(const BodyInput& in, std::vector<int>& ord)
{
    const int n = in.n;
    ord.resize(n);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            const double v = in.pos[3 * i + a];
            if (!std::isfinite(v)) continue;
            lo[a] = std::min(lo[a], v);
            hi[a] = std::max(hi[a], v);
        }
    double ext = 0.0;
    for (int a = 0; a < 3; ++a) ext = std::max(ext, hi[a] - lo[a]);
    const double scale = ext > 0.0 ? 2097151.0 / ext : 0.0;
    std::vector<std::pair<uint64_t, int>> key(n);
    host_parallel(n, [&](long long i0, long long i1) {
        for (long long i = i0; i < i1; ++i) {
            uint64_t c = 0;
            for (int a = 0; a < 3; ++a) {
                double q = (in.pos[3 * i + a] - lo[a]) * scale;
                if (!(q >= 0.0)) q = 0.0;
                if (q > 2097151.0) q = 2097151.0;
                c |= spread_bits21((uint64_t)q) << a;
            }
            key[i] = {c, (int)i};
        }
    });
    // (key, index) pairs are distinct: sorted runs merged give std::sort's order
    const int runs = n >= 65536 ? 16 : 1;
    std::vector<long long> cut(runs + 1);
    for (int r = 0; r <= runs; ++r) cut[r] = (long long)n * r / runs;
    host_tasks(runs, [&](int r) { std::sort(key.begin() + cut[r], key.begin() + cut[r + 1]); });
    for (int width = 1; width < runs; width *= 2) {
        const int pairs = (runs + 2 * width - 1) / (2 * width);
        host_tasks(pairs, [&](int q) {
            const int a = q * 2 * width, mid = std::min(runs, a + width), e = std::min(runs, a + 2 * width);
            if (mid < e) std::inplace_merge(key.begin() + cut[a], key.begin() + cut[mid], key.begin() + cut[e]);
        });
    }
    for (int i = 0; i < n; ++i) ord[i] = key[i].second;
}

// A node listed more than once takes its last target, as a serial loop over
// the list leaves it: earlier repeats become -1 (skipped by the kernel), so
// the parallel attach writes each node once.
void keep_last_of_repeats(std::vector<int>& ids)
{
    std::vector<std::pair<int, int>> s;
    s.reserve(ids.size());
    for (size_t k = 0; k < ids.size(); ++k)
        if (ids[k] >= 0) s.emplace_back(ids[k], (int)k);
    std::sort(s.begin(), s.end());
    for (size_t k = 0; k + 1 < s.size(); ++k)
        if (s[k].first == s[k + 1].first) ids[s[k].second] = -1;
}

// Caller-order copy of body b's committed positions after a host-side change.
void Engine::refreshCallerPositions(int b)
{
    if (!permuted_) return;
    launch_scatter_caller(stream_, bodyBegin_[b], bodyBegin_[b] + bodyCount_[b], callerPos_.p, x_.p, xOut_.p);
}

Engine::Engine()
{
    (void)hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking);
}

Engine::~Engine()
{
    if (graphExec_) (void)hipGraphExecDestroy(graphExec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
    for (StageSlot& sl : stage_) {
        if (sl.pin) (void)hipHostFree(sl.pin);
        if (sl.done) (void)hipEventDestroy(sl.done);
    }
}

bool Engine::build(const std::vector<const BodyInput*>& inputs, std::string& err)
{
    const auto tBuild = std::chrono::steady_clock::now();
    struct BuildTimer {
        std::chrono::steady_clock::time_point t0;
        ~BuildTimer()
        {
            if (getenv("XPBD_BUILD_TIMING"))
                fprintf(stderr, "[build] engine build %.1f ms\n",
                        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
    } buildTimer{tBuild};
    XPBD_ALLOC_SCOPE;
    const int B = (int)inputs.size();
    if (B <= 0) { err = "engine: no bodies"; return false; }
    inputs_ = inputs;
    bodyBegin_.resize(B);
    bodyCount_.resize(B);
    long long base = 0, m = 0;
    for (int b = 0; b < B; ++b) {
        bodyBegin_[b] = (int)base;
        bodyCount_[b] = inputs[b]->n;
        base += ((long long)inputs[b]->n + kTile - 1) / kTile * kTile;
        m += inputs[b]->rowOff[inputs[b]->n];
    }
    if (base > 0x7fffffffLL - kTile || m > 0x7fffffffLL) { err = "engine: problem too large for int32 indexing"; return false; }
    nPad_ = (int)base;
    m_ = m;
    uniformStiff_.assign(B, 1);
    uniformStiffValue_.assign(B, 1.0f);
    for (int b = 0; b < B; ++b) {
        const std::vector<float>& st = inputs[b]->nstiff;
        if (st.empty()) continue;
        uniformStiffValue_[b] = st[0];
        for (float v : st)
            if (std::memcmp(&v, &st[0], sizeof(float)) != 0) { uniformStiff_[b] = 0; break; }
    }
    set_.assign(B, Settings());
    press_.assign(B, PressState());
    attActive_.assign(B, false);
    colHost_.assign((size_t)B * kMaxColliders, ColliderDev{});
    colCount_.assign(B, 0);
    colMargin_.assign(B, 0.0f);
    colFriction_.assign(B, 0.0f);
    colDirty_ = true;
    colStatsValid_ = false;
    volTopo_.assign(B, VolumeTopo());
    volDev_.clear();
    for (int b = 0; b < B; ++b) volDev_.emplace_back(new VolDevLists());
    volDeviceValid_ = false;
    volActive_ = 0;
    detRest_.free(); Vrest_.free(); restLen_.free(); matType_.free(); lamVol_.free();
    bbC_.free(); bbS_.free(); bbS2_.free(); volJ_.free();
    gsScale_.free(); gsRot_.free(); gsOpacity_.free(); outScale_.free(); outQuat_.free(); shapeCtr_.free();
    cM_.free(); cRmax_.free(); cRg_.free();
    contactOn_.assign(B, 0); contactColliders_.assign(B, 0); contactGround_.assign(B, 0);
    contactTau_.assign(B, 0.2f); contactCap_.assign(B, 0.0f);
    objTopo_.assign(B, ObjTopo());
    objBase_.assign(B, 0);
    objShiftValid_.assign(B, 0);
    objTotal_ = 0;
    objDeviceValid_ = false;
    scSpacing_.assign(B, 0.0f); scCellOverride_.assign(B, 0.0f);
    scTable_.assign(B, 0); scTableBase_.assign(B, 0); scTableOverride_.assign(B, 0);
    objSingle_.assign(B, 0);
    fineTable_.assign(B, 0); fineBase_.assign(B, 0);
    fineTotal_ = 0;
    scTableTotal_ = 0; scSortBits_ = 0;
    scDeviceValid_ = false;
    scSortedIdx_ = nullptr;
    scSortedKeys_ = nullptr;
    dgs_.assign(B, DistanceGSTopo());
    vgs_.assign(B, VolumeGSTopo());
    dgsIn_.clear(); vgsIn_.clear();
    dgsColorOff_.clear(); vgsColorOff_.clear();
    dgsDevValid_ = false; vgsDevValid_ = false;
    dgsAlphaKey_.clear();
    shapeClusterOf_.assign(B, std::vector<int>());
    shcDevValid_ = true;
    angTopo_.assign(B, AngleTopo());
    angBodies_.assign(B, 0);
    angDevValid_ = false;
    shcK_ = 0;

    auto phase = [&](const char* what) {
        if (getenv("XPBD_BUILD_TIMING"))
            fprintf(stderr, "[build]   %s at %.1f ms\n", what,
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tBuild).count());
    };
    phase("settings");
    // device node order of every body (caller order or Morton order)
    ord_.assign(B, std::vector<int>());
    inv_.assign(B, std::vector<int>());
    permuted_ = false;
    for (int b = 0; b < B; ++b) {
        const BodyInput& in = *inputs[b];
        std::vector<int>& o = ord_[b];
        if (nodeOrderMode_ == 1) {
            morton_order(in, o);
            // longest rows first inside windows of 8 slices: the slices of the
            // sliced-ELL distance entries then pad far less (SELL-C-sigma)
            const int windows = (in.n + kSortWindow - 1) / kSortWindow;
            host_tasks(windows, [&](int w) {
                const int w0 = w * kSortWindow, w1 = std::min(in.n, w0 + kSortWindow);
                std::stable_sort(o.begin() + w0, o.begin() + w1, [&](int x, int y) {
                    return in.rowOff[x + 1] - in.rowOff[x] > in.rowOff[y + 1] - in.rowOff[y];
                });
            });
        } else {
            o.resize(in.n);
            for (int i = 0; i < in.n; ++i) o[i] = i;
        }
        inv_[b].assign(in.n, 0);
        for (int d = 0; d < in.n; ++d) {
            inv_[b][o[d]] = d;
            permuted_ |= o[d] != d;
        }
    }

    phase("node order");
    // host staging of the padded node space (device order); every row keeps
    // the caller's entry order, neighbours are renumbered
    std::vector<float3> hx(nPad_, make_float3(0, 0, 0));
    std::vector<int> hrow(nPad_ + 1, 0), htile(nPad_ / kTile, 0), hcaller(nPad_);
    std::vector<int> hnbr(m);
    std::vector<float> hrest(m), hstiff(m);
    for (int i = 0; i < nPad_; ++i) hcaller[i] = i;
    long long eo = 0;
    for (int b = 0; b < B; ++b) {
        const BodyInput& in = *inputs[b];
        const int nb = bodyBegin_[b];
        const std::vector<int>& o = ord_[b];
        const std::vector<int>& iv = inv_[b];
        const int padEnd = (b + 1 < B) ? bodyBegin_[b + 1] : nPad_;
        for (int t = nb / kTile; t < padEnd / kTile; ++t) htile[t] = b;
        // row starts first, then the rows on several threads
        for (int d = 0; d < in.n; ++d) {
            const int c = o[d];
            hrow[nb + d] = (int)eo;
            eo += in.rowOff[c + 1] - in.rowOff[c];
        }
        host_parallel(in.n, [&](long long d0, long long d1) {
            for (long long d = d0; d < d1; ++d) {
                const int c = o[d];
                hx[nb + d] = make_float3(in.pos[3 * c], in.pos[3 * c + 1], in.pos[3 * c + 2]);
                hcaller[nb + d] = nb + c;
                long long w = hrow[nb + d];
                for (int k = in.rowOff[c]; k < in.rowOff[c + 1]; ++k, ++w) {
                    hnbr[w] = nb + iv[in.nbr[k]];
                    hrest[w] = in.nrest[k];
                    hstiff[w] = in.nstiff[k];
                }
            }
        });
        for (int i = nb + in.n; i <= padEnd && i <= nPad_; ++i) hrow[i] = (int)eo;
    }
    hrow[nPad_] = (int)eo;

    phase("renumbered rows");
    x_.alloc(nPad_); v_.alloc(nPad_); rest_.alloc(nPad_);
    p0_.alloc(nPad_); p1_.alloc(nPad_);
    invMass_.alloc(nPad_); invMassSaved_.alloc(nPad_); objW_.alloc(nPad_);
    rowOff_.alloc(nPad_ + 1); nbr_.alloc(std::max<long long>(m, 1));
    callerPos_.alloc(nPad_);
    devOfCaller_.alloc(nPad_);
    if (permuted_) xOut_.alloc(nPad_);
    else xOut_.free();
    nstiff_.alloc(std::max<long long>(m, 1));
    tileBody_.alloc(nPad_ / kTile);
    bodies_.alloc(B);
    cols_.alloc((size_t)B * kMaxColliders);
    colHits_.alloc((size_t)B * kMaxColliders);
    colPush_.alloc((size_t)B * kMaxColliders * kColAcc);
    if (!x_.p || !v_.p || !rest_.p || !p0_.p || !p1_.p || !invMass_.p || !invMassSaved_.p || !objW_.p || !rowOff_.p ||
        !nbr_.p || !nstiff_.p || !tileBody_.p || !bodies_.p || !cols_.p || !callerPos_.p || !devOfCaller_.p || (permuted_ && !xOut_.p) ||
        !colHits_.p || !colPush_.p) {
        err = "engine: hipMalloc failed (out of device memory?)";
        return false;
    }
    XPBD_HIP(hipMemcpyAsync(x_.p, hx.data(), sizeof(float3) * nPad_, hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipMemcpyAsync(rest_.p, hx.data(), sizeof(float3) * nPad_, hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipMemsetAsync(v_.p, 0, sizeof(float3) * nPad_, stream_));
    XPBD_HIP(hipMemsetAsync(p0_.p, 0, sizeof(float4) * nPad_, stream_));
    XPBD_HIP(hipMemsetAsync(p1_.p, 0, sizeof(float4) * nPad_, stream_));
    XPBD_HIP(hipMemcpyAsync(rowOff_.p, hrow.data(), sizeof(int) * (nPad_ + 1), hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipMemcpyAsync(callerPos_.p, hcaller.data(), sizeof(int) * nPad_, hipMemcpyHostToDevice, stream_));
    {
        std::vector<int> hdev(nPad_);
        for (int i = 0; i < nPad_; ++i) hdev[hcaller[i]] = i;
        XPBD_HIP(hipMemcpyAsync(devOfCaller_.p, hdev.data(), sizeof(int) * nPad_, hipMemcpyHostToDevice, stream_));
        XPBD_HIP(hipStreamSynchronize(stream_));   // hdev leaves scope
    }
    if (permuted_) XPBD_HIP(hipMemsetAsync(xOut_.p, 0, sizeof(float3) * nPad_, stream_));
    if (m > 0) {
        XPBD_HIP(hipMemcpyAsync(nbr_.p, hnbr.data(), sizeof(int) * m, hipMemcpyHostToDevice, stream_));
        XPBD_HIP(hipMemcpyAsync(nstiff_.p, hstiff.data(), sizeof(float) * m, hipMemcpyHostToDevice, stream_));
    }
    phase("uploads");
    // sliced-ELL copy of the directed entries for the distance pass: slice t
    // = node tile t, width = its longest row, entry k of lane l at
    // sliceOff[t] + 64 k + l; rows keep their order, padding has nbr -1
    {
        const int tiles = nPad_ / kTile;
        std::vector<int> hOff(tiles), hW(tiles);
        long long total = 0;
        for (int t = 0; t < tiles; ++t) {
            int w = 0;
            for (int l = 0; l < kTile; ++l) w = std::max(w, hrow[t * kTile + l + 1] - hrow[t * kTile + l]);
            hOff[t] = (int)total;
            hW[t] = w;
            total += (long long)w * kTile;
            if (total > 0x7fffffffLL) { err = "engine: sliced distance entries exceed int32 indexing"; return false; }
        }
        std::vector<int> hSellNbr(std::max<long long>(total, 1), -1);
        std::vector<float> hSellRest(std::max<long long>(total, 1), 0.0f);
        host_parallel(nPad_, [&](long long i0, long long i1) {
            for (long long i = i0; i < i1; ++i) {
                const int t = (int)(i / kTile), l = (int)(i % kTile);
                for (int e = hrow[i]; e < hrow[i + 1]; ++e) {
                    const size_t slot = (size_t)hOff[t] + (size_t)(e - hrow[i]) * kTile + l;
                    hSellNbr[slot] = hnbr[e];
                    hSellRest[slot] = hrest[e];
                }
            }
        });
        sellEntries_ = total;
        sliceOff_.alloc(std::max(tiles, 1)); sliceW_.alloc(std::max(tiles, 1));
        // per-edge alpha tilde only for bodies whose edges differ in stiffness
        bool anyVarying = false;
        for (int b = 0; b < B; ++b) anyVarying |= !uniformStiff_[b];
        sellNbr_.alloc(hSellNbr.size()); sellRest_.alloc(hSellRest.size()); sellLam_.alloc(hSellNbr.size());
        sellAlpha_.alloc(anyVarying ? hSellNbr.size() : 1);
        if (!sliceOff_.p || !sliceW_.p || !sellNbr_.p || !sellRest_.p || !sellAlpha_.p || !sellLam_.p) {
            err = "engine: hipMalloc failed (out of device memory?)";
            return false;
        }
        XPBD_HIP(hipMemcpyAsync(sliceOff_.p, hOff.data(), sizeof(int) * tiles, hipMemcpyHostToDevice, stream_));
        XPBD_HIP(hipMemcpyAsync(sliceW_.p, hW.data(), sizeof(int) * tiles, hipMemcpyHostToDevice, stream_));
        XPBD_HIP(hipMemcpyAsync(sellNbr_.p, hSellNbr.data(), sizeof(int) * hSellNbr.size(), hipMemcpyHostToDevice,
                                stream_));
        XPBD_HIP(hipMemcpyAsync(sellRest_.p, hSellRest.data(), sizeof(float) * hSellRest.size(), hipMemcpyHostToDevice,
                                stream_));
        XPBD_HIP(hipStreamSynchronize(stream_));   // host staging goes out of scope
    }
    phase("sliced entries");
    XPBD_HIP(hipMemcpyAsync(tileBody_.p, htile.data(), sizeof(int) * htile.size(), hipMemcpyHostToDevice, stream_));
    launch_fill_float(stream_, nPad_, invMass_.p, 1.0f);
    launch_fill_float(stream_, nPad_, objW_.p, 1.0f);
    launch_fill_int(stream_, B * kMaxColliders, colHits_.p, 0);
    launch_fill_float(stream_, B * kMaxColliders * kColAcc, colPush_.p, 0.0f);
    bodyHost_.assign(B, BodyDev{});
    for (int b = 0; b < B; ++b) refreshCallerPositions(b);
    lastBodies_.clear();
    alphaKey_.clear();
    XPBD_HIP(hipStreamSynchronize(stream_));
    XPBD_HIP(hipGetLastError());
    return true;
}

// forward.cu:4478-4500: same list length keeps the previous pose for friction.
void Engine::setColliders(int b, int count, const int* types, const float* poses, const float* dims, float margin,
                          float friction)
{
    colCount_[b] = collider_list_update(&colHost_[(size_t)b * kMaxColliders], colCount_[b], count, types, poses, dims);
    colMargin_[b] = positive_or_zero(margin);
    colFriction_[b] = positive_or_zero(friction);
    colDirty_ = true;
}

void Engine::setColliderArray(int b, const ColliderDev* cols, int count, float margin, float friction)
{
    count = std::max(0, std::min(count, kMaxColliders));
    for (int c = 0; c < count; ++c) colHost_[(size_t)b * kMaxColliders + c] = cols[c];
    colCount_[b] = count;
    colMargin_[b] = margin;
    colFriction_[b] = friction;
    colDirty_ = true;
}

// forward.cu:4502-4519
int Engine::colliderStats(int b, int* hits, float* push, float* centroid, int maxCount)
{
    const int n = std::min(maxCount, colCount_[b]);
    if (n <= 0 || !colStatsValid_) return 0;
    std::vector<int> h(n);
    std::vector<float> acc((size_t)kColAcc * n);
    const size_t c0 = (size_t)b * kMaxColliders;
    if (hipMemcpyAsync(h.data(), colHits_.p + c0, sizeof(int) * n, hipMemcpyDeviceToHost, stream_) != hipSuccess) return 0;
    if (hipMemcpyAsync(acc.data(), colPush_.p + c0 * kColAcc, sizeof(float) * kColAcc * n, hipMemcpyDeviceToHost,
                       stream_) != hipSuccess)
        return 0;
    if (hipStreamSynchronize(stream_) != hipSuccess) return 0;
    for (int c = 0; c < n; ++c) {
        if (hits) hits[c] = h[c];
        const float* a = &acc[(size_t)kColAcc * c];
        for (int k = 0; k < 3; ++k) {
            if (push) push[3 * c + k] = a[k];
            if (centroid)
                centroid[3 * c + k] = (a[6] > 0.0f) ? a[3 + k] / a[6] : colHost_[c0 + c].t[k];
        }
    }
    return n;
}

// setAttachedParticles (forward.cu:6759-6788)
bool Engine::attach(int b, int count, const int* idx, const float* pos, float weight)
{
    XPBD_ALLOC_SCOPE;
    const int nb = bodyBegin_[b], nn = bodyCount_[b];
    if (attActive_[b])
        (void)hipMemcpyAsync(invMass_.p + nb, invMassSaved_.p + nb, sizeof(float) * nn, hipMemcpyDeviceToDevice, stream_);
    if (count <= 0 || !idx || !pos) {
        attActive_[b] = false;
        return hipStreamSynchronize(stream_) == hipSuccess;
    }
    if (!attActive_[b])
        (void)hipMemcpyAsync(invMassSaved_.p + nb, invMass_.p + nb, sizeof(float) * nn, hipMemcpyDeviceToDevice, stream_);
    launch_fill_float(stream_, nn, objW_.p + nb, 1.0f);
    if ((size_t)count > attIdx_.n) {
        attIdx_.alloc(count);
        attPos_.alloc((size_t)3 * count);
    }
    // caller ids -> device ids (out-of-range ids stay out of range and are skipped)
    std::vector<int> dIdx(idx, idx + count);
    for (int& k : dIdx)
        if (k >= 0 && k < nn) k = inv_[b][k];
    keep_last_of_repeats(dIdx);
    (void)hipMemcpyAsync(attIdx_.p, dIdx.data(), sizeof(int) * count, hipMemcpyHostToDevice, stream_);
    (void)hipMemcpyAsync(attPos_.p, pos, sizeof(float) * 3 * count, hipMemcpyHostToDevice, stream_);
    launch_attach(stream_, count, attIdx_.p, attPos_.p, nb, nb + nn, x_.p, v_.p, invMass_.p, objW_.p,
                  std::fmax(weight, 1.0f));
    refreshCallerPositions(b);
    attActive_[b] = true;
    return hipStreamSynchronize(stream_) == hipSuccess;
}

void Engine::addRigidVelocity(int b, const float* dv, const float* dw, const float* c)
{
    const int nb = bodyBegin_[b];
    launch_rigid_velocity(stream_, nb, nb + bodyCount_[b], x_.p, v_.p, invMass_.p, make_float3(dv[0], dv[1], dv[2]),
                          make_float3(dw[0], dw[1], dw[2]), make_float3(c[0], c[1], c[2]));
    (void)hipStreamSynchronize(stream_);
}

// squashReset (forward.cu:4411-4427)
void Engine::resetToRest(int b)
{
    const int nb = bodyBegin_[b], nn = bodyCount_[b];
    (void)hipMemcpyAsync(x_.p + nb, rest_.p + nb, sizeof(float3) * nn, hipMemcpyDeviceToDevice, stream_);
    (void)hipMemsetAsync(v_.p + nb, 0, sizeof(float3) * nn, stream_);
    launch_fill_float(stream_, nn, invMass_.p + nb, 1.0f);
    refreshCallerPositions(b);
    (void)hipStreamSynchronize(stream_);
    PressState& ps = press_[b];
    ps.active = false; ps.curDisp = 0.0f; ps.ckptIdx = 0; ps.dwell = 0;
}

void Engine::releaseInvMass(int b)
{
    launch_fill_float(stream_, bodyCount_[b], invMass_.p + bodyBegin_[b], 1.0f);
    (void)hipStreamSynchronize(stream_);
}

// groundLaunch (forward.cu:4551-4610): rigid tilt about the rest centre of
// mass plus a rigid velocity field, computed on the host like the reference.
void Engine::launchFromRest(int b, const float* lin, const float* ang, const float* tilt)
{
    const BodyInput& in = *inputs_[b];
    const int n = in.n;
    double cx = 0, cy = 0, cz = 0;
    for (int i = 0; i < n; ++i) { cx += in.pos[3 * i]; cy += in.pos[3 * i + 1]; cz += in.pos[3 * i + 2]; }
    const float comX = float(cx / n), comY = float(cy / n), comZ = float(cz / n);
    const float ax = tilt ? tilt[0] : 0.0f, ay = tilt ? tilt[1] : 0.0f, az = tilt ? tilt[2] : 0.0f;
    const float th = std::sqrt(ax * ax + ay * ay + az * az);
    float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (th > 1e-8f) {
        const float kx = ax / th, ky = ay / th, kz = az / th;
        const float s = std::sin(th), c = std::cos(th), t = 1.0f - c;
        R[0] = c + t * kx * kx;      R[1] = t * kx * ky - s * kz; R[2] = t * kx * kz + s * ky;
        R[3] = t * ky * kx + s * kz; R[4] = c + t * ky * ky;      R[5] = t * ky * kz - s * kx;
        R[6] = t * kz * kx - s * ky; R[7] = t * kz * ky + s * kx; R[8] = c + t * kz * kz;
    }
    const float vx = lin ? lin[0] : 0.0f, vy = lin ? lin[1] : 0.0f, vz = lin ? lin[2] : 0.0f;
    const float wx = ang ? ang[0] : 0.0f, wy = ang ? ang[1] : 0.0f, wz = ang ? ang[2] : 0.0f;
    std::vector<float3> hp(n), hv(n);
    for (int i = 0; i < n; ++i) {
        const float px = in.pos[3 * i] - comX, py = in.pos[3 * i + 1] - comY, pz = in.pos[3 * i + 2] - comZ;
        const float rx = R[0] * px + R[1] * py + R[2] * pz;
        const float ry = R[3] * px + R[4] * py + R[5] * pz;
        const float rz = R[6] * px + R[7] * py + R[8] * pz;
        hp[i] = make_float3(comX + rx, comY + ry, comZ + rz);
        hv[i] = make_float3(vx + (wy * rz - wz * ry), vy + (wz * rx - wx * rz), vz + (wx * ry - wy * rx));
    }
    const int nb = bodyBegin_[b];
    std::vector<float3> dp(n), dv(n);
    toDevice(b, hp.data(), dp.data(), 1);
    toDevice(b, hv.data(), dv.data(), 1);
    (void)hipMemcpyAsync(x_.p + nb, dp.data(), sizeof(float3) * n, hipMemcpyHostToDevice, stream_);
    (void)hipMemcpyAsync(v_.p + nb, dv.data(), sizeof(float3) * n, hipMemcpyHostToDevice, stream_);
    launch_fill_float(stream_, n, invMass_.p + nb, 1.0f);
    refreshCallerPositions(b);
    (void)hipStreamSynchronize(stream_);
    PressState& ps = press_[b];
    ps.active = false; ps.curDisp = 0.0f; ps.ckptIdx = 0; ps.dwell = 0;
}

// Press ramp / checkpoint dwell of one step (forward.cu:12095-12112).
void Engine::pressAdvance(int b)
{
    PressState& ps = press_[b];
    if (!ps.active) return;
    if (ps.autoLog && ps.dwell > 0) {
        if (--ps.dwell == 0) ++ps.ckptIdx;
    } else {
        ps.curDisp = std::fmin(ps.curDisp + ps.ramp * set_[b].dt, ps.maxDisp);
        if (ps.autoLog && ps.ckptIdx < 4) {
            static const float ck[4] = {0.25f, 0.50f, 0.75f, 1.00f};
            const float target = ck[ps.ckptIdx] * ps.maxDisp;
            if (ps.curDisp >= target) { ps.curDisp = target; ps.dwell = ps.dwellFrames; }
        }
    }
}

void Engine::fillBodyDev(int b, BodyDev& d) const
{
    const Settings& s = set_[b];
    const PressState& ps = press_[b];
    std::memset(&d, 0, sizeof(d));
    d.nodeBegin = bodyBegin_[b];
    d.nodeEnd = bodyBegin_[b] + bodyCount_[b];
    // distance: the coloured Gauss-Seidel pass replaces Jacobi when enabled
    // (forward.cu:12065, 12251-12286)
    d.distGSOn = (s.useDistance && s.distGS && dgsDevValid_ && b < (int)dgsIn_.size() && dgsIn_[b]) ? 1 : 0;
    d.distanceOn = (s.useDistance && !s.distGS) ? 1 : 0;
    d.iters = s.iters;
    d.dt = s.dt;
    d.underRelax = s.underRelax;
    d.velDamping = s.velDamping;
    d.invMassScale = s.invMassScale;
    d.distCompliance = s.distCompliance;
    d.stiffnessScale = s.stiffnessScale;
    // the per-entry expression of the alpha kernel, once for the body
    d.distAlphaUniform = uniformStiff_[b] ? 1 : 0;
    d.distAlphaU = (s.distCompliance / std::fmax(uniformStiffValue_[b] * s.stiffnessScale, 1e-6f)) /
                   std::fmax(s.dt * s.dt, 1e-8f);
    // stepPhysicsOnly (forward.cu:13258-13262): gravity along -up, off during press
    const float g = (s.groundGravity > 0.0f && !ps.active) ? s.groundGravity : 0.0f;
    d.grav[0] = g > 0.0f ? -g * s.groundN[0] : 0.0f;
    d.grav[1] = g > 0.0f ? -g * s.groundN[1] : 0.0f;
    d.grav[2] = g > 0.0f ? -g * s.groundN[2] : 0.0f;
    d.groundOn = (s.groundEnabled && !ps.active) ? 1 : 0;
    for (int k = 0; k < 3; ++k) d.gn[k] = s.groundN[k];
    d.planeD = s.groundHeight + s.groundRadius;
    d.slop = s.groundSlop;
    d.friction = s.groundFriction;
    d.restitution = s.groundRestitution;
    d.restingSpeed = 2.0f * s.groundGravity * s.dt;
    d.colBegin = b * kMaxColliders;
    d.colCount = colCount_[b];
    d.colMargin = colMargin_[b];
    d.colFriction = colFriction_[b];
    d.pressOn = ps.active ? 1 : 0;
    d.pressAxis = ps.axis;
    d.pressLo = ps.lo + (ps.fromLo ? ps.curDisp : 0.0f);
    d.pressHi = ps.hi - (ps.fromLo ? 0.0f : ps.curDisp);
    d.contactColliders = 0;
    d.contactGround = 0;
    // volume: the coloured pass replaces the gather pass unless its colouring
    // was skipped (forward.cu:12066, 12310-12324)
    d.volGSOn = (s.useVolume && s.volGS && vgsDevValid_ && b < (int)vgsIn_.size() && vgsIn_[b]) ? 1 : 0;
    d.volumeOn = (s.useVolume && !d.volGSOn) ? 1 : 0;
    d.angleOn = (s.useAngle && angDevValid_) ? 1 : 0;
    d.angleCompliance = s.angleCompliance;
    d.angleBlend = s.angleBlend;
    d.volCompliance = s.volCompliance;
    d.volMinMembers = volTopo_[b].minMembers;
    d.shapeOn = s.useShape ? 1 : 0;
    d.shapeClusters = (shcDevValid_ && !shapeClusterOf_[b].empty()) ? 1 : 0;
    d.shapeRobust = s.shapeRobust ? 1 : 0;
    d.shapeCompliance = s.shapeCompliance;
    d.shapeBlend = s.shapeBlend;
    // contact ellipsoids act once packed (forward.cu:12219-12226)
    d.contactColliders = (cM_.p && contactColliders_[b]) ? 1 : 0;
    d.contactGround = (cM_.p && contactGround_[b] && s.groundEnabled) ? 1 : 0;
    // object shape after predict (forward.cu:12124-12166), off while pressing
    const int K = objDeviceValid_ ? objTopo_[b].K : 0;
    d.objOn = (s.objShapeStiffness > 0.0f && !ps.active && K > 0) ? 1 : 0;
    d.objBase = objDeviceValid_ ? objBase_[b] : 0;
    d.objK = K;
    // the attachment-weighted fit is the device path's (forward.cu:11671-11702);
    // with the device path off, or past 64 objects, the reference fits on the
    // host without weights (11573-11668, chosen at 12164): same sums, unweighted
    d.objWeighted = (attActive_[b] && s.objShapeGPU && K <= 64) ? 1 : 0;
    d.objStiffness = s.objShapeStiffness;
    d.objBoxes = (d.objOn && s.scEnabled) ? 1 : 0;
    // active set only when the previous shift is valid (forward.cu:6959-7030)
    d.objActiveAllowed = (d.objBoxes && !s.scWithinBody && objShiftValid_[b] && K >= 1 && K <= 64) ? 1 : 0;
    // self-collision build / iterations (forward.cu:11796-11849)
    const float dc = scDeviceValid_ ? s.scRadiusScale * scSpacing_[b] : 0.0f;
    d.scOn = (s.scEnabled && !ps.active && dc > 0.0f) ? 1 : 0;
    d.scWithin = s.scWithinBody ? 1 : 0;
    d.scTableBase = scDeviceValid_ ? scTableBase_[b] : 0;
    d.scTableMask = scDeviceValid_ ? scTable_[b] - 1 : 0;
    d.scDc = dc;
    d.scSearch = kScSearchScale * dc;
    const float excl = s.scExcludeScale * dc;
    d.scExcl2 = excl * excl;
    d.scFastRel = kScFastRelScale * dc;
    d.scCellOverride = scCellOverride_[b];
    // the predicted-position grid serves bodies whose candidates are all
    // same-object ones (self-contact on, every node in one object); its scan
    // order key packs the coarse sorted position into 27 bits
    const bool fine = scDeviceValid_ && fineTable_[b] > 0 && s.scWithinBody && b < (int)objSingle_.size() &&
                      objSingle_[b] && scFineOn_ && nPad_ <= (1 << 27);
    d.scFine = fine ? 1 : 0;
    d.scFineBase = fine ? fineBase_[b] : 0;
    d.scFineMask = fine ? fineTable_[b] - 1 : 0;
    d.scFineInv = 1.0f / (d.scSearch * 1.01f + 1e-6f);
}

// Queue an upload of `bytes` from host memory through the next pinned slot.
// The slot's previous copy must have left it before it is overwritten: the
// launches of several substeps may still be queued.
bool Engine::stageUpload(void* dst, const void* src, size_t bytes, std::string& err)
{
    StageSlot& sl = stage_[stageNext_];
    stageNext_ = (stageNext_ + 1) % kStageSlots;
    if (sl.pending) {
        XPBD_HIP(hipEventSynchronize(sl.done));
        sl.pending = false;
    }
    if (sl.cap < bytes) {
        if (sl.pin) (void)hipHostFree(sl.pin);
        sl.pin = nullptr;
        sl.cap = 0;
        XPBD_HIP(hipHostMalloc(&sl.pin, bytes, 0));
        sl.cap = bytes;
    }
    if (!sl.done) XPBD_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    std::memcpy(sl.pin, src, bytes);
    XPBD_HIP(hipMemcpyAsync(dst, sl.pin, bytes, hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipEventRecord(sl.done, stream_));
    sl.pending = true;
    return true;
}

bool Engine::uploadParams(std::string& err)
{
    const int B = numBodies();
    for (int b = 0; b < B; ++b) fillBodyDev(b, bodyHost_[b]);
    bool changed = (int)lastBodies_.size() != B ||
                   std::memcmp(lastBodies_.data(), bodyHost_.data(), sizeof(BodyDev) * B) != 0;
    if (changed) {
        if (!stageUpload(bodies_.p, bodyHost_.data(), sizeof(BodyDev) * B, err)) return false;
        lastBodies_ = bodyHost_;
    }
    if (colDirty_) {
        if (!stageUpload(cols_.p, colHost_.data(), sizeof(ColliderDev) * colHost_.size(), err)) return false;
        colDirty_ = false;
    }
    // alpha tilde depends on (dt, compliance, stiffness scale) of each body
    std::vector<float> key(3 * (size_t)B);
    for (int b = 0; b < B; ++b) {
        key[3 * b] = set_[b].dt;
        key[3 * b + 1] = set_[b].distCompliance;
        key[3 * b + 2] = set_[b].stiffnessScale;
    }
    if (key != alphaKey_ && sellAlpha_.n > 1) {
        NodeArgs na{nPad_, tileBody_.p, bodies_.p};
        launch_distance_alpha(stream_, na, rowOff_.p, nstiff_.p, sellArgs());
        alphaKey_ = key;
    }
    if (dgsDevValid_ && dgsColorOff_.size() > 1 && dgsColorOff_.back() > 0 && key != dgsAlphaKey_) {
        launch_dist_gs_alpha(stream_, dgsColorOff_.back(), dgsAB_.p, dgsStiff_.p, tileBody_.p, bodies_.p, dgsAlpha_.p);
        dgsAlphaKey_ = key;
    }
    return true;
}

SellArgs Engine::sellArgs() const
{
    return SellArgs{sliceOff_.p, sliceW_.p, sellNbr_.p, sellRest_.p, sellAlpha_.p, sellLam_.p};
}

// Everything that changes the launch sequence (not the parameter values).
std::vector<int> Engine::planSignature() const
{
    const int B = numBodies();
    int iters = set_[0].iters;
    int anyDist = 0, anyCol = 0, anyGround = 0, anyPress = 0, anyVol = 0, anyShape = 0, anyCGround = 0;
    int anyObj = 0, anyBoxes = 0, anySc = 0, anyDistGS = 0, anyVolGS = 0, anyFine = 0, anyAngle = 0;
    for (int b = 0; b < B; ++b) {
        const BodyDev& d = bodyHost_[b];
        anyObj |= d.objOn;
        anyBoxes |= d.objBoxes;
        anySc |= d.scOn;
        anyFine |= d.scOn && d.scFine;
        anyDist |= d.distanceOn;
        anyDistGS |= d.distGSOn;
        anyVolGS |= d.volGSOn;
        anyCol |= colCount_[b] > 0 ? 1 : 0;
        anyGround |= set_[b].groundEnabled ? 1 : 0;
        anyPress |= press_[b].active ? 1 : 0;
        anyVol |= d.volumeOn;
        anyShape |= set_[b].useShape ? 1 : 0;
        anyAngle |= d.angleOn;
        anyCGround |= (cM_.p && contactGround_[b] && set_[b].groundEnabled) ? 1 : 0;
    }
    const int volWave = anyVol ? (tune_.volMode >= 0 ? (tune_.volMode ? 1 : 0) : (volAvgMembers_ >= kVolWaveMembers ? 1 : 0))
                               : 0;
    // fine candidate scan: the automatic choice is the lane scan
    const int groupLanes = tune_.scGroupLanes < 0 ? 1 : tune_.scGroupLanes;
    const int nDistGS = (anyDistGS && dgsDevValid_) ? (int)dgsColorOff_.size() - 1 : 0;
    const int nVolGS = (anyVolGS && vgsDevValid_) ? (int)vgsColorOff_.size() - 1 : 0;
    return {iters, anyDist, anyCol, anyGround, anyPress, nPad_, anyVol, volWave, anyVol ? volActive_ : 0, anyShape,
            anyCGround, anyObj, anyBoxes, anySc, anyObj ? objTotal_ : 0, anySc ? scTableTotal_ : 0,
            anySc ? scSortBits_ : 0, nDistGS, nVolGS, device_alloc_epoch(), (anyShape && shcDevValid_) ? shcK_ : 0,
            allocEpoch_, tune_.jacobiTiles, tune_.volMode, tune_.shapeClusterWaves, tune_.scBuildWaves,
            tune_.scBuildNodes, tune_.scBuildXcd, anyFine, anyAngle, groupLanes};
}

// Event pairs around kernel launches (profile mode only).
struct Engine::Timer {
    hipStream_t s;
    std::vector<std::string> label;
    std::vector<hipEvent_t> ev;
    void begin(const char* l)
    {
        hipEvent_t a;
        (void)hipEventCreate(&a);
        (void)hipEventRecord(a, s);
        label.emplace_back(l);
        ev.push_back(a);
    }
    void end()
    {
        hipEvent_t b;
        (void)hipEventCreate(&b);
        (void)hipEventRecord(b, s);
        ev.push_back(b);
    }
    ~Timer()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

#define XPBD_T_BEGIN(l) do { if (t) t->begin(l); } while (0)
#define XPBD_T_END() do { if (t) t->end(); } while (0)

// Record the kernels of one substep on stream_ (under graph capture when t is
// null, eagerly with timing events otherwise).
void Engine::recordSubstep(Timer* t)
{
    const std::vector<int> sig = planSignature();
    const int iters = sig[0];
    const bool anyDist = sig[1], anyCol = sig[2], anyGround = sig[3], anyPress = sig[4];
    const NodeArgs na{nPad_, tileBody_.p, bodies_.p};
    EpiArgs ea{x_.p, v_.p, cols_.p, colHits_.p, colPush_.p, cM_.p, cRmax_.p, cRg_.p, callerPos_.p,
               permuted_ ? xOut_.p : nullptr};
    const bool shapePass = sig[9] != 0;
    if (sig[10]) {
        XPBD_T_BEGIN("contact_ground_radius");
        launch_contact_ground(stream_, na, cM_.p, cRg_.p);
        XPBD_T_END();
    }

    XPBD_T_BEGIN("predict");
    launch_predict(stream_, na, x_.p, v_.p, invMass_.p, p0_.p);
    XPBD_T_END();
    const bool objPass = sig[11] != 0, scPass = sig[13] != 0;
    const SelfColArgs sa = scArgs();
    if (objPass) {
        XPBD_T_BEGIN("object_shape");
        launch_object_shape(stream_, na, sa, objTotal_, p0_.p, rest_.p, sig[12] != 0);
        XPBD_T_END();
    }
    if (scPass) {
        XPBD_T_BEGIN("selfcol_build");
        const int* sorted = nullptr;
        // profile mode: the build's phases get labels of their own
        void (*mark)(void*, const char*) = nullptr;
        if (t) mark = [](void* c, const char* l) { static_cast<Timer*>(c)->end(); static_cast<Timer*>(c)->begin(l); };
        Tuning tu = tune_;
        tu.scGroupLanes = sig[30];
        const bool inAlt = launch_selfcol_build(stream_, na, numBodies(), sa, x_.p, p0_.p, rest_.p, scSortBits_,
                                                scKeysAlt_.p, scIdxAlt_.p, scHist_.p, &sorted, tu, mark, t);
        XPBD_T_END();
        scSortedIdx_ = sorted;
        scSortedKeys_ = inAlt ? scKeysAlt_.p : scKeys_.p;
    }
    if (anyCol) {
        const int B = numBodies();
        XPBD_T_BEGIN("collider_stats_clear");
        launch_fill_int(stream_, B * kMaxColliders, colHits_.p, 0);
        launch_fill_float(stream_, B * kMaxColliders * kColAcc, colPush_.p, 0.0f);
        XPBD_T_END();
    }
    float4* pin = p0_.p;
    float4* pout = p1_.p;
    for (int it = 0; it < iters; ++it) {
        const bool last = it == iters - 1;
        unsigned epi = 0;
        if (anyPress) epi |= kEpiPress;
        if (anyCol) epi |= kEpiColliders | (last ? kEpiColliderHits : 0u);
        if (anyGround) epi |= kEpiGround;
        // with self-collision the velocity commit moves behind the contact push
        if (last && !scPass) epi |= kEpiVelocity;
        const bool volPass = sig[6] != 0;
        // constraint passes in order: distance GS, distance Jacobi, volume GS,
        // volume gather, shape matching (forward.cu:12250-12470).  The
        // epilogue runs on the last of them when that pass visits every node;
        // after a Gauss-Seidel pass (which visits only its colours' nodes) it
        // runs as its own pass.
        const int nDistGS = sig[17], nVolGS = sig[18];
        enum { kNone, kDistGS, kJacobi, kVolGS, kVolGather, kShape, kAngle };
        const bool anglePass = sig[29] != 0;
        int lastPass = kNone;
        if (nDistGS) lastPass = kDistGS;
        if (anyDist) lastPass = kJacobi;
        if (nVolGS) lastPass = kVolGS;
        if (volPass) lastPass = kVolGather;
        if (shapePass) lastPass = kShape;
        if (anglePass) lastPass = kAngle;
        // the press slab precedes self-collision; colliders and ground follow it
        // (forward.cu:12474-12560)
        const unsigned preEpi = scPass ? (lastPass != kNone ? (epi & kEpiPress) : 0u) : epi;
        const unsigned scEpi = lastPass != kNone ? (epi & ~kEpiPress) : epi;
        if (nDistGS) {
            XPBD_T_BEGIN("distance_gs");
            for (int c = 0; c < nDistGS; ++c)
                launch_dist_gs(stream_, it == 0, last, dgsColorOff_[c], dgsColorOff_[c + 1] - dgsColorOff_[c], dgsAB_.p,
                               dgsRest_.p, dgsAlpha_.p, dgsLam_.p, pin, tileBody_.p, bodies_.p);
            XPBD_T_END();
        }
        if (anyDist) {
            XPBD_T_BEGIN(it == 0 ? (last ? "distance_jacobi_only" : "distance_jacobi_first")
                                 : (last ? "distance_jacobi_last" : "distance_jacobi_mid"));
            launch_distance_jacobi(stream_, na, it == 0, last, pin, pout, sellArgs(),
                                   lastPass == kJacobi ? preEpi : 0u, ea, tune_);
            XPBD_T_END();
            std::swap(pin, pout);
        }
        if (nVolGS) {
            const VolumeArgs va = volArgs();
            XPBD_T_BEGIN("volume_gs");
            for (int c = 0; c < nVolGS; ++c)
                launch_vol_gs(stream_, it == 0, last, vgsColorOff_[c], vgsColorOff_[c + 1] - vgsColorOff_[c],
                              vgsOrder_.p, pin, va, tileBody_.p, bodies_.p);
            XPBD_T_END();
        }
        if (volPass) {
            const VolumeArgs va = volArgs();
            XPBD_T_BEGIN(sig[7] ? "volume_solve_wave" : "volume_solve_thread");
            launch_vol_solve(stream_, na, sig[7] != 0, it == 0, last, pin, va);
            XPBD_T_END();
            XPBD_T_BEGIN("volume_apply");
            launch_vol_apply(stream_, na, pin, pout, va, lastPass == kVolGather ? preEpi : 0u, ea);
            XPBD_T_END();
            std::swap(pin, pout);
        }
        if (shapePass) {
            XPBD_T_BEGIN("shape_match");
            ShapeClusterArgs sca = shcArgs();
            if (!sig[20]) sca.K = 0;
            launch_shape_match(stream_, na, pin, pout, rest_.p, rowOff_.p, nbr_.p, sca,
                               lastPass == kShape ? preEpi : 0u, ea, tune_);
            XPBD_T_END();
            std::swap(pin, pout);
        }
        // umbrella angle constraint after shape matching (forward.cu:12486-12517)
        if (anglePass) {
            XPBD_T_BEGIN("angle");
            launch_angle(stream_, na, pin, pout, angArgs(), preEpi, ea);
            XPBD_T_END();
            std::swap(pin, pout);
        }
        if ((lastPass == kDistGS || lastPass == kVolGS) && preEpi) {
            XPBD_T_BEGIN("epilogue");
            launch_epilogue(stream_, na, pin, preEpi, ea);
            XPBD_T_END();
        }
        if (scPass) {
            XPBD_T_BEGIN("selfcol_iteration");
            launch_selfcol_iteration(stream_, na, sa, pin, pout, x_.p, rest_.p, scEpi, ea);
            XPBD_T_END();
            std::swap(pin, pout);
        } else if (lastPass == kNone && epi) {
            XPBD_T_BEGIN("epilogue");
            launch_epilogue(stream_, na, pin, epi, ea);
            XPBD_T_END();
        }
    }
    if (scPass) {
        // contact-average push, ground re-projection and velocity commit
        // (forward.cu:11858-11890, 12588-12596)
        XPBD_T_BEGIN("contact_push_commit");
        launch_contact_push(stream_, na, sa, objPass ? objTotal_ : 0, pin, ea);
        XPBD_T_END();
    }
}

bool Engine::rebuildGraph(std::string& err)
{
    if (graphExec_) { (void)hipGraphExecDestroy(graphExec_); graphExec_ = nullptr; }
    if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
    XPBD_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    recordSubstep(nullptr);
    hipError_t ce = hipStreamEndCapture(stream_, &graph_);
    if (ce != hipSuccess) { err = std::string("hipStreamEndCapture: ") + hipGetErrorString(ce); return false; }
    XPBD_HIP(hipGraphInstantiate(&graphExec_, graph_, nullptr, nullptr, 0));
    graphSig_ = planSignature();
    return true;
}

bool Engine::prepare(std::string& err, bool advance)
{
    XPBD_ALLOC_SCOPE;
    if (advance)
        for (int b = 0; b < numBodies(); ++b) pressAdvance(b);
    bool anyObj = false, anySc = false;
    for (int b = 0; b < numBodies(); ++b) {
        anyObj |= set_[b].objShapeStiffness > 0.0f || set_[b].scEnabled;
        anySc |= set_[b].scEnabled;
    }
    if (anyObj && !ensureObjects(err)) return false;
    if (anySc && !ensureSelfCol(err)) return false;
    bool anyVol = false, anyVolGS = false, anyDistGS = false;
    for (int b = 0; b < numBodies(); ++b) {
        anyVol |= set_[b].useVolume;
        anyVolGS |= set_[b].useVolume && set_[b].volGS;
        anyDistGS |= set_[b].useDistance && set_[b].distGS;
    }
    if (anyVol && !ensureVolume(false, err)) return false;
    if ((anyVolGS || vgsDevValid_) && !ensureVolumeGS(err)) return false;
    if ((anyDistGS || dgsDevValid_) && !ensureDistanceGS(err)) return false;
    if (!ensureShapeClusters(err)) return false;
    // angle arrays cover exactly the bodies with the constraint on
    std::vector<char> angOn(numBodies(), 0);
    bool anyAngle = false;
    for (int b = 0; b < numBodies(); ++b) { angOn[b] = set_[b].useAngle ? 1 : 0; anyAngle |= set_[b].useAngle; }
    if (angOn != angBodies_) { angBodies_ = angOn; angDevValid_ = false; }
    if (anyAngle && !ensureAngle(err)) return false;
    return uploadParams(err);
}

// The object shift of a body that ran object shape matching is a valid
// reference for the next substep's active set.
void Engine::afterSubstep()
{
    for (int b = 0; b < numBodies(); ++b) {
        if (bodyHost_[b].objOn) objShiftValid_[b] = 1;
        if (bodyHost_[b].scOn && b < (int)scBuilt_.size()) scBuilt_[b] = 1;
    }
}

bool Engine::warm(std::string& err)
{
    XPBD_ALLOC_SCOPE;
    for (int b = 1; b < numBodies(); ++b)
        if (set_[b].iters != set_[0].iters) { err = "engine: every body of one batch needs the same iteration count"; return false; }
    if (!prepare(err, false)) return false;
    if (!graphExec_ || planSignature() != graphSig_) {
        if (!rebuildGraph(err)) return false;
    }
    return true;
}

bool Engine::step(int substeps, std::string& err)
{
    XPBD_ALLOC_SCOPE;
    for (int b = 1; b < numBodies(); ++b)
        if (set_[b].iters != set_[0].iters) { err = "engine: every body of one batch needs the same iteration count"; return false; }
    const bool timing = getenv("XPBD_BUILD_TIMING") != nullptr;
    for (int s = 0; s < substeps; ++s) {
        const auto t0 = std::chrono::steady_clock::now();
        if (!prepare(err)) return false;
        const auto t1 = std::chrono::steady_clock::now();
        if (!graphExec_ || planSignature() != graphSig_) {
            if (!rebuildGraph(err)) return false;
            if (timing)
                fprintf(stderr, "[build] prepare %.1f ms, graph capture %.1f ms\n",
                        std::chrono::duration<double, std::milli>(t1 - t0).count(),
                        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count());
        }
        XPBD_HIP(hipGraphLaunch(graphExec_, stream_));
        afterSubstep();
    }
    colStatsValid_ = true;
    return true;
}

bool Engine::profile(int substeps, std::string& json, std::string& err)
{
    XPBD_ALLOC_SCOPE;
    Timer t;
    t.s = stream_;
    for (int s = 0; s < substeps; ++s) {
        if (!prepare(err)) return false;
        recordSubstep(&t);
        afterSubstep();
    }
    colStatsValid_ = true;
    XPBD_HIP(hipStreamSynchronize(stream_));
    XPBD_HIP(hipGetLastError());
    std::vector<std::string> keys;
    std::vector<double> tot;
    std::vector<int> cnt;
    for (size_t k = 0; k < t.label.size(); ++k) {
        float ms = 0.0f;
        XPBD_HIP(hipEventElapsedTime(&ms, t.ev[2 * k], t.ev[2 * k + 1]));
        size_t j = 0;
        while (j < keys.size() && keys[j] != t.label[k]) ++j;
        if (j == keys.size()) { keys.push_back(t.label[k]); tot.push_back(0.0); cnt.push_back(0); }
        tot[j] += ms;
        cnt[j] += 1;
    }
    json = "{";
    for (size_t j = 0; j < keys.size(); ++j) {
        char buf[256];
        std::snprintf(buf, sizeof(buf), "%s\"%s\": [%.6f, %d]", j ? ", " : "", keys[j].c_str(), tot[j], cnt[j]);
        json += buf;
    }
    json += "}";
    return true;
}

bool Engine::sync(std::string& err)
{
    XPBD_HIP(hipStreamSynchronize(stream_));
    XPBD_HIP(hipGetLastError());
    return true;
}

bool Engine::getState(int b, float* pos, float* vel)
{
    const int nb = bodyBegin_[b], n = bodyCount_[b];
    std::vector<float> hp(pos ? 3 * (size_t)n : 0), hv(vel ? 3 * (size_t)n : 0);
    if (pos && hipMemcpyAsync(hp.data(), x_.p + nb, sizeof(float3) * n, hipMemcpyDeviceToHost, stream_) != hipSuccess)
        return false;
    if (vel && hipMemcpyAsync(hv.data(), v_.p + nb, sizeof(float3) * n, hipMemcpyDeviceToHost, stream_) != hipSuccess)
        return false;
    if (hipStreamSynchronize(stream_) != hipSuccess) return false;
    if (pos) toCaller(b, hp.data(), pos, 3);
    if (vel) toCaller(b, hv.data(), vel, 3);
    return true;
}

bool Engine::setState(int b, const float* pos, const float* vel)
{
    const int nb = bodyBegin_[b], n = bodyCount_[b];
    std::vector<float> hp(pos ? 3 * (size_t)n : 0), hv(vel ? 3 * (size_t)n : 0);
    if (pos) toDevice(b, pos, hp.data(), 3);
    if (vel) toDevice(b, vel, hv.data(), 3);
    if (pos && hipMemcpyAsync(x_.p + nb, hp.data(), sizeof(float3) * n, hipMemcpyHostToDevice, stream_) != hipSuccess)
        return false;
    if (vel && hipMemcpyAsync(v_.p + nb, hv.data(), sizeof(float3) * n, hipMemcpyHostToDevice, stream_) != hipSuccess)
        return false;
    if (pos) refreshCallerPositions(b);
    return hipStreamSynchronize(stream_) == hipSuccess;
}

bool Engine::getInvMass(int b, float* w)
{
    const int nb = bodyBegin_[b], n = bodyCount_[b];
    std::vector<float> h(n);
    if (hipMemcpyAsync(h.data(), invMass_.p + nb, sizeof(float) * n, hipMemcpyDeviceToHost, stream_) != hipSuccess)
        return false;
    if (hipStreamSynchronize(stream_) != hipSuccess) return false;
    toCaller(b, h.data(), w, 1);
    return true;
}

}  // namespace xpbd
