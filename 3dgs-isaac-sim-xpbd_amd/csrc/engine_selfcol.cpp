// engine_selfcol.cpp — object components, object shape matching state and the
// self-collision hash tables of the Engine.
//
// Objects are the connected components of each body's graph with at least
// `objMinComponent` nodes, numbered in node order (forward.cu:11309-11374);
// their rest centroids and rest offsets are cached in fp64 like the reference
// (11376-11439).  Each body owns one hash table of nextpow2(2N) (min 1024)
// cells, or a fixed override, inside one shared cell-range array; keys carry
// the table base so bodies never see each other, and one radix sort orders
// every body at once.  The node spacing d_c is the median nearest-neighbour
// rest length (11765-11792).
#include <algorithm>
#include <cfloat>
#include <cmath>
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

namespace {

int find_root(std::vector<int>& parent, int x)
{
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

}  // namespace

SelfColArgs Engine::scArgs() const
{
    SelfColArgs a;
    a.tileBody = tileBody_.p;
    a.bodies = bodies_.p;
    a.objId = objId_.p;
    a.objBody = objBody_.p;
    a.objCount = objCount_.p;
    a.objC0 = objC0_.p;
    a.objSumQ = objSumQ_.p;
    a.objShift = objShift_.p;
    a.objRestRel = objRestRel_.p;
    a.objW = objW_.p;
    a.objSums = objSums_.p;
    a.objParams = objParams_.p;
    a.objDisp = objDisp_.p;
    a.objBoxesArr = objBoxes_.p;
    a.scMax2 = scMax2_.p;
    a.scCounters = scCounters_.p;
    a.scUseActive = scUseActive_.p;
    a.scExact = scExact_.p;
    a.scSortGate = scSortGate_.p;
    a.devOfCaller = devOfCaller_.p;
    a.scBodyDisp = scBodyDisp_.p;
    a.scInvCell = scInvCell_.p;
    a.scActive = scActive_.p;
    a.keysDev = scKeysDev_.p;
    a.keys = scKeys_.p;
    a.idx = scIdx_.p;
    a.sentinel = (unsigned int)scTableTotal_;
    a.tableTotal = scTableTotal_;
    a.cellStart = cellStart_.p;
    a.cellEnd = cellEnd_.p;
    a.contacts = contacts_.p;
    a.counts = counts_.p;
    a.overflow = scOver_.p;
    a.accum = scAccum_.p;
    a.pushSums = pushSums_.p;
    a.pushT = pushT_.p;
    a.callerPos = callerPos_.p;
    a.sortP = scSortP_.p;
    a.sortX = scSortX_.p;
    a.sortR = scSortR_.p;
    a.sortedIdx = nullptr;
    bool fine = false;
    for (int b = 0; b < numBodies() && !fine; ++b) fine = bodyHost_[b].scFine != 0;
    a.coarseRank = fine ? coarseRank_.p : nullptr;
    a.fineTotal = fine ? fineTotal_ : 0;
    a.fkeys = fkeys_.p;
    a.frank = frank_.p;
    a.fcount = fcount_.p;
    a.fstart = fstart_.p;
    a.fscan = fscan_.p;
    a.finePos = finePos_.p;
    a.fineE = fineE_.p;
    a.fineC = fineC_.p;
    a.laneN = laneN_.p;
    a.scLong = scLong_.p;
    a.laneKey = laneKey_.p;
    a.accCount = fine ? accCount_.p : nullptr;
    return a;
}

// Components and rest caches for every body (built once per minimum size).
bool Engine::ensureObjects(std::string& err)
{
    const int B = numBodies();
    bool changed = !objDeviceValid_;
    for (int b = 0; b < B; ++b) {
        ObjTopo& t = objTopo_[b];
        const int minComp = std::max(1, set_[b].objMinComponent);
        if (t.minComp == minComp) continue;
        const BodyInput& in = *inputs_[b];
        const int n = in.n;
        std::vector<int> parent(n);
        for (int i = 0; i < n; ++i) parent[i] = i;
        for (int i = 0; i < n; ++i)
            for (int e = in.rowOff[i]; e < in.rowOff[i + 1]; ++e) {
                const int ri = find_root(parent, i), rj = find_root(parent, in.nbr[e]);
                if (ri != rj) parent[ri] = rj;
            }
        std::vector<int> size(n, 0), rootId(n, -1);
        for (int i = 0; i < n; ++i) ++size[find_root(parent, i)];
        t.id.assign(n, -1);
        int K = 0;
        for (int i = 0; i < n; ++i) {
            const int r = find_root(parent, i);
            if (rootId[r] == -1) rootId[r] = (size[r] >= minComp) ? K++ : -2;
            if (rootId[r] >= 0) t.id[i] = rootId[r];
        }
        t.K = K;
        t.count.assign(K, 0);
        for (int i = 0; i < n; ++i)
            if (t.id[i] >= 0) ++t.count[t.id[i]];
        std::vector<double> sum(3 * (size_t)K, 0.0);
        for (int i = 0; i < n; ++i) {
            const int id = t.id[i];
            if (id < 0) continue;
            for (int a = 0; a < 3; ++a) sum[3 * id + a] += in.pos[3 * i + a];
        }
        t.c0.assign(3 * (size_t)K, 0.0f);
        for (int k = 0; k < K; ++k)
            for (int a = 0; a < 3; ++a) t.c0[3 * k + a] = (float)(sum[3 * k + a] / std::max(t.count[k], 1));
        t.restRel.assign(3 * (size_t)n, 0.0);
        t.sumQ.assign(3 * (size_t)K, 0.0);
        for (int i = 0; i < n; ++i) {
            const int id = t.id[i];
            if (id < 0) continue;
            for (int a = 0; a < 3; ++a) {
                t.restRel[3 * (size_t)i + a] = (double)(in.pos[3 * i + a] - t.c0[3 * id + a]);
                t.sumQ[3 * id + a] += t.restRel[3 * (size_t)i + a];
            }
        }
        t.minComp = minComp;
        objShiftValid_[b] = 0;
        changed = true;
    }
    if (!changed) return true;
    objSingle_.assign(B, 0);
    for (int b = 0; b < B; ++b) {
        const std::vector<int>& id = objTopo_[b].id;
        bool same = true;
        for (size_t i = 1; i < id.size() && same; ++i) same = id[i] == id[0];
        objSingle_[b] = same ? 1 : 0;
    }
    XPBD_HIP(hipStreamSynchronize(stream_));   // queued substeps may still use the old buffers
    objBase_.assign(B, 0);
    int K = 0;
    for (int b = 0; b < B; ++b) { objBase_[b] = K; K += objTopo_[b].K; }
    objTotal_ = K;
    const int Kc = std::max(K, 1);
    std::vector<int> hId(nPad_, -1), hBody(Kc, 0), hCount(Kc, 0);
    std::vector<float> hC0(3 * (size_t)Kc, 0.0f);
    std::vector<double> hSumQ(3 * (size_t)Kc, 0.0), hShift(3 * (size_t)Kc, 0.0), hRel(3 * (size_t)nPad_, 0.0);
    for (int b = 0; b < B; ++b) {
        const ObjTopo& t = objTopo_[b];
        const int nb = bodyBegin_[b], ob = objBase_[b];
        for (int i = 0; i < bodyCount_[b]; ++i) {   // caller node i -> device node nb + d
            const size_t d = (size_t)nb + inv_[b][i];
            hId[d] = t.id[i] >= 0 ? ob + t.id[i] : -1;
            for (int a = 0; a < 3; ++a) hRel[3 * d + a] = t.restRel[3 * (size_t)i + a];
        }
        for (int k = 0; k < t.K; ++k) {
            hBody[ob + k] = b;
            hCount[ob + k] = t.count[k];
            for (int a = 0; a < 3; ++a) {
                hC0[3 * (ob + k) + a] = t.c0[3 * k + a];
                hSumQ[3 * (ob + k) + a] = t.sumQ[3 * k + a];
                hShift[3 * (ob + k) + a] = (double)t.c0[3 * k + a];
            }
        }
    }
    objId_.alloc(nPad_);
    objBody_.alloc(Kc); objCount_.alloc(Kc);
    objC0_.alloc(3 * (size_t)Kc); objParams_.alloc(16 * (size_t)Kc); objDisp_.alloc(Kc); objBoxes_.alloc(6 * (size_t)Kc);
    pushSums_.alloc(4 * (size_t)Kc); pushT_.alloc(4 * (size_t)Kc);
    objSumQ_.alloc(3 * (size_t)Kc); objShift_.alloc(3 * (size_t)Kc); objSums_.alloc(16 * (size_t)Kc);
    objRestRel_.alloc(3 * (size_t)nPad_);
    if (!objId_.p || !objBody_.p || !objCount_.p || !objC0_.p || !objParams_.p || !objDisp_.p || !objBoxes_.p ||
        !pushSums_.p || !pushT_.p || !objSumQ_.p || !objShift_.p || !objSums_.p || !objRestRel_.p) {
        err = "objects: hipMalloc failed";
        return false;
    }
    XPBD_HIP(hipMemcpyAsync(objId_.p, hId.data(), sizeof(int) * nPad_, hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipMemcpyAsync(objBody_.p, hBody.data(), sizeof(int) * Kc, hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipMemcpyAsync(objCount_.p, hCount.data(), sizeof(int) * Kc, hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipMemcpyAsync(objC0_.p, hC0.data(), sizeof(float) * 3 * Kc, hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipMemcpyAsync(objSumQ_.p, hSumQ.data(), sizeof(double) * 3 * Kc, hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipMemcpyAsync(objShift_.p, hShift.data(), sizeof(double) * 3 * Kc, hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipMemcpyAsync(objRestRel_.p, hRel.data(), sizeof(double) * 3 * nPad_, hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipMemsetAsync(objParams_.p, 0, sizeof(float) * 16 * Kc, stream_));
    XPBD_HIP(hipMemsetAsync(objDisp_.p, 0, sizeof(float) * Kc, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    objShiftValid_.assign(B, 0);   // every shift restarts at the rest centroid
    objDeviceValid_ = true;
    return true;
}

void Engine::setSelfCollisionGrid(int b, float cellSize, int tableSize)
{
    scCellOverride_[b] = (std::isfinite(cellSize) && cellSize > 0.0f) ? cellSize : 0.0f;
    scTableOverride_[b] = tableSize;
    scDeviceValid_ = false;
}

// Spacing, per-body tables and the shared build buffers.
bool Engine::ensureSelfCol(std::string& err)
{
    if (scDeviceValid_) return true;
    XPBD_HIP(hipStreamSynchronize(stream_));   // queued substeps may still use the old buffers
    const int B = numBodies();
    for (int b = 0; b < B; ++b) {
        if (scSpacing_[b] > 0.0f) continue;
        const BodyInput& in = *inputs_[b];
        std::vector<float> mins;
        mins.reserve(in.n);
        for (int i = 0; i < in.n; ++i) {
            float m = FLT_MAX;
            for (int e = in.rowOff[i]; e < in.rowOff[i + 1]; ++e) {
                const float d = in.nrest[e];
                if (d > 0.0f && std::isfinite(d) && d < m) m = d;
            }
            if (m < FLT_MAX) mins.push_back(m);
        }
        if (mins.empty()) continue;
        std::nth_element(mins.begin(), mins.begin() + mins.size() / 2, mins.end());
        scSpacing_[b] = mins[mins.size() / 2];
    }
    long long total = 0;
    for (int b = 0; b < B; ++b) {
        int t = 1024;
        while (t < 2 * bodyCount_[b]) t <<= 1;
        if (scTableOverride_[b] > 0) t = scTableOverride_[b];
        scTable_[b] = t;
        scTableBase_[b] = (int)total;
        total += t;
    }
    if (total >= 0x7fffffffLL) { err = "self-collision: hash tables too large"; return false; }
    scTableTotal_ = (int)total;
    int bits = 0;
    while ((1LL << bits) <= total) ++bits;   // keys in [0, total], the sentinel included
    scSortBits_ = bits;
    const size_t n = nPad_;
    scMax2_.alloc(B); scCounters_.alloc(3 * (size_t)B); scUseActive_.alloc(B); scExact_.alloc(B); scBodyDisp_.alloc(B);
    scInvCell_.alloc(B);
    scKeysDev_.alloc(n); scKeys_.alloc(n); scKeysAlt_.alloc(n); scIdx_.alloc(n); scIdxAlt_.alloc(n);
    scHist_.alloc(radix_sort_hist_size(nPad_));
    cellStart_.alloc(scTableTotal_); cellEnd_.alloc(scTableTotal_);
    contacts_.alloc(n * kSelfColCap); counts_.alloc(n); scOver_.alloc(n);
    scActive_.alloc(n); scAccum_.alloc(n);
    scSortP_.alloc(n); scSortX_.alloc(n); scSortR_.alloc(n);
    scSortGate_.alloc(1);
    if (!scSortGate_.p || !scMax2_.p || !scCounters_.p || !scUseActive_.p || !scExact_.p || !scBodyDisp_.p || !scInvCell_.p || !scKeysDev_.p || !scKeys_.p ||
        !scKeysAlt_.p || !scIdx_.p || !scIdxAlt_.p || !scHist_.p || !cellStart_.p || !cellEnd_.p || !contacts_.p ||
        !counts_.p || !scOver_.p || !scActive_.p || !scAccum_.p || !scSortP_.p ||
        !scSortX_.p || !scSortR_.p) {
        err = "self-collision: hipMalloc failed";
        return false;
    }
    // predicted-position grid: one table of nextpow2(2N) cells per body
    fineTable_.assign(B, 0);
    fineBase_.assign(B, 0);
    long long fineTotal = 0;
    for (int b = 0; b < B; ++b) {
        int t = 1024;
        while (t < 2 * bodyCount_[b]) t <<= 1;
        fineTable_[b] = t;
        fineBase_[b] = (int)fineTotal;
        fineTotal += t;
    }
    if (fineTotal >= 0x7fffffffLL) { err = "self-collision: hash tables too large"; return false; }
    fineTotal_ = (int)fineTotal;
    fkeys_.alloc(n); frank_.alloc(n); finePos_.alloc(n);
    fcount_.alloc(fineTotal_); fstart_.alloc((size_t)fineTotal_ + 1);
    fscan_.alloc(exclusive_scan_tmp_size(fineTotal_));
    fineE_.alloc(2 * n); fineC_.alloc(n); coarseRank_.alloc(n);
    // the lane fine scan's unsorted lists: one contiguous list of kLaneList
    // entries per fine position
    laneN_.alloc(n); laneKey_.alloc(n * kLaneList); scLong_.alloc(n + 1);
    // per node: what its candidate search still needs
    accCount_.alloc(n);
    if (!fkeys_.p || !frank_.p || !finePos_.p || !fcount_.p || !fstart_.p || !fscan_.p || !fineE_.p ||
        !fineC_.p || !coarseRank_.p || !laneN_.p || !laneKey_.p || !scLong_.p || !accCount_.p) {
        err = "self-collision: hipMalloc failed";
        return false;
    }
    XPBD_HIP(hipMemsetAsync(counts_.p, 0, sizeof(int) * n, stream_));
    XPBD_HIP(hipMemsetAsync(scOver_.p, 0, sizeof(int) * n, stream_));
    XPBD_HIP(hipMemsetAsync(scActive_.p, 0, n, stream_));
    XPBD_HIP(hipMemsetAsync(scCounters_.p, 0, sizeof(int) * 3 * B, stream_));
    XPBD_HIP(hipMemsetAsync(scUseActive_.p, 0, sizeof(int) * B, stream_));
    XPBD_HIP(hipMemsetAsync(scSortGate_.p, 0, sizeof(int), stream_));
    XPBD_HIP(hipMemsetAsync(scAccum_.p, 0, sizeof(float3) * n, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    scSortedIdx_ = nullptr;
    scSortedKeys_ = nullptr;
    scDeviceValid_ = true;
    scBuilt_.assign(numBodies(), 0);
    return true;
}

static const float kScSearchScaleHost = 1.5f;   // forward.cu:230

bool Engine::selfCollisionGridInfo(int b, float info[3], std::string& err)
{
    XPBD_ALLOC_SCOPE;
    info[0] = info[1] = info[2] = 0.0f;
    if (!scDeviceValid_) return true;
    float inv = 0.0f;
    XPBD_HIP(hipMemcpyAsync(&inv, scInvCell_.p + b, sizeof(float), hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    info[0] = inv > 0.0f ? 1.0f / inv : 0.0f;
    info[1] = set_[b].scRadiusScale * scSpacing_[b];
    info[2] = kScSearchScaleHost * info[1];
    return true;
}

bool Engine::selfCollisionPaths(int b, int out[4], std::string& err)
{
    XPBD_ALLOC_SCOPE;
    out[0] = out[1] = out[2] = out[3] = 0;
    if (!scDeviceValid_ || b >= (int)scBuilt_.size() || !scBuilt_[b]) return true;
    const int nb = bodyBegin_[b], n = bodyCount_[b];
    if (!bodyHost_[b].scFine) {
        out[2] = n;   // every node takes the coarse walk
        return true;
    }
    std::vector<int> acc(n), cnt(n);
    XPBD_HIP(hipMemcpyAsync(acc.data(), accCount_.p + nb, sizeof(int) * n, hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipMemcpyAsync(cnt.data(), counts_.p + nb, sizeof(int) * n, hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    for (int i = 0; i < n; ++i) {
        const int c = acc[i];
        if (c == kScAccLane) ++out[0];
        else if (c == kScAccRule) ++out[1];
        else if (c == kScAccCoarse) ++out[2];
        else ++out[3];
    }
    return true;
}

bool Engine::selfCollisionStats(int b, int out[3], std::string& err)
{
    XPBD_ALLOC_SCOPE;
    out[0] = out[1] = out[2] = 0;
    if (!scDeviceValid_ || b >= (int)scBuilt_.size() || !scBuilt_[b]) return true;
    if (!scStats_.p) {
        scStats_.alloc(3);
        if (!scStats_.p) { err = "self-collision: hipMalloc failed"; return false; }
    }
    // totals over the body's nodes: list lengths, overflow, active mask (all
    // ones when the active restriction is off)
    launch_selfcol_stats(stream_, bodyBegin_[b], bodyBegin_[b] + bodyCount_[b], scArgs(), scStats_.p);
    int c[3] = {0, 0, 0};
    XPBD_HIP(hipMemcpyAsync(c, scStats_.p, sizeof(int) * 3, hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    XPBD_HIP(hipGetLastError());
    out[0] = c[0];
    out[1] = c[2];
    out[2] = c[1];
    return true;
}

// Keys (body-local table, sentinel = table size), the body's nodes in sorted
// order (table range first, then sentinel-keyed nodes in index order), counts
// and contact lists [n x 32] with local ids.
bool Engine::selfCollisionLists(int b, unsigned int* keys, int* sorted, int* counts, int* contacts, std::string& err)
{
    XPBD_ALLOC_SCOPE;
    if (!scDeviceValid_ || !scSortedIdx_) { err = "no self-collision build yet"; return false; }
    const int nb = bodyBegin_[b], n = bodyCount_[b];
    std::vector<unsigned int> hk(nPad_), hks(nPad_);
    // lists are slot-major per 64-node tile: slot k of tile-local node l at
    // tile * 64 * 32 + 64 k + l (the body starts on a tile boundary)
    const size_t tiles = ((size_t)n + kTile - 1) / kTile;
    std::vector<int> hs(nPad_), hc(n), hl(tiles * kTile * kSelfColCap);
    // a build whose nodes all took the fine scans left its keys unsorted:
    // sort them now (once; the sort lands where the build's would have)
    int gate = 1;
    XPBD_HIP(hipMemcpyAsync(&gate, scSortGate_.p, sizeof(int), hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    if (gate == 0) {
        launch_sc_caller_order(stream_, nPad_, scKeysDev_.p, devOfCaller_.p, scKeys_.p, scIdx_.p, nullptr);
        launch_radix_sort(stream_, nPad_, scSortBits_, scKeys_.p, scIdx_.p, scKeysAlt_.p, scIdxAlt_.p, scHist_.p, false);
        const int sorted2 = 2;
        XPBD_HIP(hipMemcpyAsync(scSortGate_.p, &sorted2, sizeof(int), hipMemcpyHostToDevice, stream_));
        XPBD_HIP(hipStreamSynchronize(stream_));
    }
    XPBD_HIP(hipMemcpyAsync(hks.data(), scSortedKeys_, sizeof(unsigned int) * nPad_, hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipMemcpyAsync(hs.data(), scSortedIdx_, sizeof(int) * nPad_, hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipMemcpyAsync(hc.data(), counts_.p + nb, sizeof(int) * n, hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipMemcpyAsync(hl.data(), contacts_.p + (size_t)nb * kSelfColCap, sizeof(int) * hl.size(),
                            hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    // the sort consumed the key buffer: invert the sorted permutation
    for (int t = 0; t < nPad_; ++t) {
        if (hs[t] < 0 || hs[t] >= nPad_) { err = "self-collision: sorted order is not a permutation"; return false; }
        hk[hs[t]] = hks[t];
    }
    const unsigned int base = (unsigned int)scTableBase_[b], T = (unsigned int)scTable_[b];
    const std::vector<int>& ord = ord_[b];
    const std::vector<int>& inv = inv_[b];
    auto inBody = [&](int t) { return hs[t] >= nb && hs[t] < nb + n; };
    auto inTable = [&](int t) { return hks[t] >= base && hks[t] < base + T; };
    if (keys)
        for (int i = 0; i < n; ++i) {
            const unsigned int k = hk[nb + inv[i]];
            keys[i] = (k >= base && k < base + T) ? k - base : T;
        }
    if (sorted) {
        int w = 0;
        for (int t = 0; t < nPad_; ++t)
            if (inBody(t) && inTable(t)) sorted[w++] = ord[hs[t] - nb];
        for (int t = 0; t < nPad_; ++t)
            if (inBody(t) && !inTable(t)) sorted[w++] = ord[hs[t] - nb];
    }
    if (counts)
        for (int i = 0; i < n; ++i) counts[i] = hc[inv[i]];
    if (contacts)
        for (int i = 0; i < n; ++i) {
            const int d = inv[i];
            for (int k = 0; k < kSelfColCap; ++k) {
                const int code = hl[(size_t)(d / kTile) * kTile * kSelfColCap + (size_t)k * kTile + d % kTile];
                int out = 0;   // slots past the count hold earlier builds' entries
                if (k < hc[d]) out = code >= 0 ? ord[code - nb] : -(ord[-code - 1 - nb] + 1);
                contacts[(size_t)i * kSelfColCap + k] = out;
            }
        }
    return true;
}

}  // namespace xpbd
