// engine_batch.cpp — Engine services used by batches of many bodies: state
// hand-over between engines, grouped attachment, all-body shape updates and
// the launch-shape autotuner.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

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

bool Engine::copyBodyFrom(int b, Engine& src, int sb, std::string& err)
{
    XPBD_ALLOC_SCOPE;
    if (b < 0 || b >= numBodies() || sb < 0 || sb >= src.numBodies()) { err = "copy: body out of range"; return false; }
    const int nn = bodyCount_[b];
    if (src.bodyCount_[sb] != nn || src.inputs_[sb]->rowOff != inputs_[b]->rowOff || src.ord_[sb] != ord_[b]) {
        err = "copy: the two bodies differ in size, graph or node order";
        return false;
    }
    // settings, topology choices and host-side state first: the preparation
    // below builds the device buffers the copies land in
    set_[b] = src.set_[sb];
    press_[b] = src.press_[sb];
    attActive_[b] = src.attActive_[sb];
    for (int c = 0; c < kMaxColliders; ++c)
        colHost_[(size_t)b * kMaxColliders + c] = src.colHost_[(size_t)sb * kMaxColliders + c];
    colCount_[b] = src.colCount_[sb];
    colMargin_[b] = src.colMargin_[sb];
    colFriction_[b] = src.colFriction_[sb];
    colDirty_ = true;
    scCellOverride_[b] = src.scCellOverride_[sb];
    if (scTableOverride_[b] != src.scTableOverride_[sb]) {
        scTableOverride_[b] = src.scTableOverride_[sb];
        scDeviceValid_ = false;
    }
    if (scSpacing_[b] <= 0.0f) scSpacing_[b] = src.scSpacing_[sb];
    if (shapeClusterOf_[b] != src.shapeClusterOf_[sb]) {
        shapeClusterOf_[b] = src.shapeClusterOf_[sb];
        shcDevValid_ = false;
    }
    // device-built lists are compared and handed over through host copies
    if (!src.volumeHostLists(sb, err) || !volumeHostLists(b, err)) return false;
    const VolumeTopo& sv = src.volTopo_[sb];
    VolumeTopo& dv = volTopo_[b];
    if (sv.version != 0 &&
        (dv.explicitClusters != sv.explicitClusters || dv.minMembers != sv.minMembers || dv.k != sv.k ||
         dv.cap != sv.cap || dv.r != sv.r || dv.isLeader != sv.isLeader || dv.memOff != sv.memOff ||
         dv.memIdx != sv.memIdx)) {
        dv = sv;
        volDeviceValid_ = false;
    }
    contactOn_[b] = src.contactOn_[sb];
    contactColliders_[b] = src.contactColliders_[sb];
    contactGround_[b] = src.contactGround_[sb];
    contactTau_[b] = src.contactTau_[sb];
    contactCap_[b] = src.contactCap_[sb];
    if (!prepare(err, false)) return false;

    XPBD_HIP(hipStreamSynchronize(src.stream_));
    const size_t nb = bodyBegin_[b], snb = src.bodyBegin_[sb];
    auto copy = [&](void* dst, const void* from, size_t bytes) -> bool {
        return hipMemcpyAsync(dst, from, bytes, hipMemcpyDeviceToDevice, stream_) == hipSuccess;
    };
    bool ok = copy(x_.p + nb, src.x_.p + snb, sizeof(float3) * nn) && copy(v_.p + nb, src.v_.p + snb, sizeof(float3) * nn) &&
              copy(invMass_.p + nb, src.invMass_.p + snb, sizeof(float) * nn) &&
              copy(invMassSaved_.p + nb, src.invMassSaved_.p + snb, sizeof(float) * nn) &&
              copy(objW_.p + nb, src.objW_.p + snb, sizeof(float) * nn);
    if (ok && contactOn_[b] && src.cM_.p) {
        std::string e2;
        if (!cM_.p && !setContactShape(b, 1 | (contactGround_[b] ? 2 : 0), contactTau_[b], contactCap_[b], e2)) {
            err = "copy: " + e2;
            return false;
        }
        ok = copy(cM_.p + 6 * nb, src.cM_.p + 6 * snb, sizeof(float) * 6 * nn) &&
             copy(cRmax_.p + nb, src.cRmax_.p + snb, sizeof(float) * nn);
    }
    // object-shape shifts continue only where both engines hold the same components
    objShiftValid_[b] = 0;
    if (ok && src.objShiftValid_[sb] && objDeviceValid_ && src.objDeviceValid_ && objTopo_[b].K == src.objTopo_[sb].K &&
        objTopo_[b].K > 0) {
        ok = copy(objShift_.p + 3 * (size_t)objBase_[b], src.objShift_.p + 3 * (size_t)src.objBase_[sb],
                  sizeof(double) * 3 * objTopo_[b].K);
        objShiftValid_[b] = 1;
    }
    if (!ok) { err = "copy: device copy failed"; return false; }
    refreshCallerPositions(b);
    XPBD_HIP(hipStreamSynchronize(stream_));
    XPBD_HIP(hipGetLastError());
    return true;
}

bool Engine::attachMany(int count, const int* bodies, const int* off, const int* idx, const float* pos, float weight,
                        std::string& err)
{
    XPBD_ALLOC_SCOPE;
    std::vector<int> dIdx;
    std::vector<float> dPos;
    bool any = false;
    for (int r = 0; r < count; ++r) {
        const int b = bodies[r];
        if (b < 0 || b >= numBodies()) { err = "attach: body out of range"; return false; }
        const int nb = bodyBegin_[b], nn = bodyCount_[b];
        const int k0 = off[r], k1 = off[r + 1];
        if (attActive_[b])
            XPBD_HIP(hipMemcpyAsync(invMass_.p + nb, invMassSaved_.p + nb, sizeof(float) * nn, hipMemcpyDeviceToDevice,
                                    stream_));
        if (k1 <= k0 || !idx || !pos) {
            attActive_[b] = false;
            continue;
        }
        if (!attActive_[b])
            XPBD_HIP(hipMemcpyAsync(invMassSaved_.p + nb, invMass_.p + nb, sizeof(float) * nn, hipMemcpyDeviceToDevice,
                                    stream_));
        launch_fill_float(stream_, nn, objW_.p + nb, 1.0f);
        for (int k = k0; k < k1; ++k) {
            const int c = idx[k];
            dIdx.push_back((c >= 0 && c < nn) ? nb + inv_[b][c] : -1);
            dPos.insert(dPos.end(), pos + 3 * (size_t)k, pos + 3 * (size_t)k + 3);
        }
        attActive_[b] = true;
        any = true;
    }
    if (any) {
        keep_last_of_repeats(dIdx);   // global device ids: a repeat is the same node
        const size_t m = dIdx.size();
        if (m > attIdx_.n) {
            attIdx_.alloc(m);
            attPos_.alloc(3 * m);
            if (!attIdx_.p || !attPos_.p) { err = "attach: hipMalloc failed"; return false; }
        }
        XPBD_HIP(hipMemcpyAsync(attIdx_.p, dIdx.data(), sizeof(int) * m, hipMemcpyHostToDevice, stream_));
        XPBD_HIP(hipMemcpyAsync(attPos_.p, dPos.data(), sizeof(float) * 3 * m, hipMemcpyHostToDevice, stream_));
        launch_attach(stream_, (int)m, attIdx_.p, attPos_.p, 0, nPad_, x_.p, v_.p, invMass_.p, objW_.p,
                      std::fmax(weight, 1.0f));
        if (permuted_) launch_scatter_caller(stream_, 0, nPad_, callerPos_.p, x_.p, xOut_.p);
    }
    XPBD_HIP(hipStreamSynchronize(stream_));   // host staging goes out of scope
    XPBD_HIP(hipGetLastError());
    return true;
}

bool Engine::computeShapesAll(float deformEps, std::string& err)
{
    XPBD_ALLOC_SCOPE;
    if (!ensureShapeBuffers(err)) return false;
    // padding nodes have no neighbours and are never counted as deformed
    launch_deformed_shapes(stream_, 0, nPad_, rest_.p, x_.p, rowOff_.p, nbr_.p, gsScale_.p, gsRot_.p, deformEps,
                           outScale_.p, outQuat_.p, shapeCtr_.p);
    for (int b = 0; b < numBodies(); ++b) {
        if (!contactOn_[b] || !cM_.p) continue;
        const size_t nb = bodyBegin_[b];
        const float cap = (std::isfinite(contactCap_[b]) && contactCap_[b] > 0.0f) ? contactCap_[b] : 0.0f;
        launch_contact_pack(stream_, (int)nb, bodyCount_[b], outScale_.p + 3 * nb, outQuat_.p + 4 * nb,
                            gsOpacity_.p + nb, contactTau_[b], cap, cM_.p, cRmax_.p);
    }
    XPBD_HIP(hipGetLastError());
    return true;
}

// ---------------------------------------------------------------------------
// Autotuner.  Coordinate descent over the launch shapes of kernels.h Tuning:
// each candidate runs `substeps` substeps from one saved state (the state is
// put back after every trial and at the end), timed with events on the engine
// stream; the fastest of two repetitions counts.

namespace {

struct Snapshot {
    DevBuf<float3> x, v, xOut;
    DevBuf<float> w, wSaved, objW;
    DevBuf<double> shift;
    DevBuf<float> cM, cRmax;
};

}  // namespace

bool Engine::autotune(int substeps, int flags, std::string& report, std::string& err)
{
    XPBD_ALLOC_SCOPE;
    if (substeps < 1) substeps = 1;
    if (!warm(err)) return false;
    XPBD_HIP(hipStreamSynchronize(stream_));
    const size_t n = nPad_;
    Snapshot s;
    s.x.alloc(n); s.v.alloc(n); s.w.alloc(n); s.wSaved.alloc(n); s.objW.alloc(n);
    if (permuted_) s.xOut.alloc(n);
    if (objShift_.p) s.shift.alloc(objShift_.n);
    if (cM_.p) { s.cM.alloc(cM_.n); s.cRmax.alloc(cRmax_.n); }
    if (!s.x.p || !s.v.p || !s.w.p || !s.wSaved.p || !s.objW.p) { err = "autotune: hipMalloc failed"; return false; }
    const std::vector<PressState> press = press_;
    const std::vector<char> shiftValid = objShiftValid_;
    auto save = [&](bool back) -> bool {
        auto cp = [&](void* a, void* b2, size_t bytes) {
            return hipMemcpyAsync(back ? a : b2, back ? b2 : a, bytes, hipMemcpyDeviceToDevice, stream_) == hipSuccess;
        };
        bool ok = cp(x_.p, s.x.p, sizeof(float3) * n) && cp(v_.p, s.v.p, sizeof(float3) * n) &&
                  cp(invMass_.p, s.w.p, sizeof(float) * n) && cp(invMassSaved_.p, s.wSaved.p, sizeof(float) * n) &&
                  cp(objW_.p, s.objW.p, sizeof(float) * n);
        if (ok && s.xOut.p) ok = cp(xOut_.p, s.xOut.p, sizeof(float3) * n);
        if (ok && s.shift.p && objShift_.p && s.shift.n == objShift_.n) ok = cp(objShift_.p, s.shift.p, sizeof(double) * s.shift.n);
        if (ok && s.cM.p && cM_.p) ok = cp(cM_.p, s.cM.p, sizeof(float) * s.cM.n) && cp(cRmax_.p, s.cRmax.p, sizeof(float) * s.cRmax.n);
        if (back) {
            press_ = press;
            objShiftValid_ = shiftValid;
        }
        return ok && hipStreamSynchronize(stream_) == hipSuccess;
    };
    if (!save(false)) { err = "autotune: snapshot failed"; return false; }
    hipEvent_t e0, e1;
    XPBD_HIP(hipEventCreate(&e0));
    XPBD_HIP(hipEventCreate(&e1));
    auto trial = [&](float& best) -> bool {
        best = 1e30f;
        if (!warm(err)) return false;          // capture outside the timed region
        for (int rep = 0; rep < 2; ++rep) {
            if (!save(true)) { err = "autotune: restore failed"; return false; }
            XPBD_HIP(hipEventRecord(e0, stream_));
            if (!step(substeps, err)) return false;
            XPBD_HIP(hipEventRecord(e1, stream_));
            XPBD_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            XPBD_HIP(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms);
        }
        return true;
    };
    bool anyVol = false, anySc = false, anyClusters = false;
    for (int b = 0; b < numBodies(); ++b) {
        anyVol |= set_[b].useVolume;
        anySc |= set_[b].scEnabled;
        anyClusters |= set_[b].useShape && !shapeClusterOf_[b].empty();
    }
    // A knob whose first value is automatic (-1) keeps it unless a fixed
    // value is faster by more than kFixedMargin: the automatic rule follows
    // later changes of the batch (merges, bodies that stop self-colliding)
    // that a fixed value cannot.
    const float kFixedMargin = 0.03f;
    struct Knob {
        const char* name;
        int* field;
        std::vector<int> values;
        bool on;
    };
    std::vector<Knob> knobs = {
        {"jacobi_tiles_per_block", &tune_.jacobiTiles, {4, 2, 1}, true},
        {"volume_mode", &tune_.volMode, {0, 1}, anyVol},
        {"shape_cluster_waves_per_block", &tune_.shapeClusterWaves, {4, 2, 1}, anyClusters},
        {"selfcol_nodes_per_wave", &tune_.scBuildNodes, {32, 16, 64, 8, 4}, anySc},
        {"selfcol_waves_per_block", &tune_.scBuildWaves, {4, 2, 1}, anySc},
        {"selfcol_xcd_runs", &tune_.scBuildXcd, {16, 1, 0, 64}, anySc},
        {"selfcol_fine_scan", &tune_.scGroupLanes, {-1, 0}, anySc},
    };
    std::string rep = "{";
    float base = 0.0f;
    bool ok = trial(base);
    char buf[96];
    std::snprintf(buf, sizeof(buf), "\"start_ms\": %.4f", base);
    rep += buf;
    for (Knob& k : knobs) {
        if (!ok || !k.on) continue;
        const int keep = *k.field;
        int bestV = keep;
        float bestMs = 1e30f;
        rep += std::string(", \"") + k.name + "\": {";
        for (size_t j = 0; j < k.values.size() && ok; ++j) {
            *k.field = k.values[j];
            float ms = 0.0f;
            ok = trial(ms);
            const bool autoFirst = k.values[0] == -1;
            const float need = (autoFirst && j > 0 && bestV == -1) ? bestMs * (1.0f - kFixedMargin) : bestMs;
            if (ok && ms < need) { bestMs = ms; bestV = k.values[j]; }
            std::snprintf(buf, sizeof(buf), "%s\"%d\": %.4f", j ? ", " : "", k.values[j], ms);
            rep += buf;
        }
        *k.field = ok ? bestV : keep;
        std::snprintf(buf, sizeof(buf), ", \"chosen\": %d}", *k.field);
        rep += buf;
    }
    // hash-table size (flag 1): every self-colliding body's table at 1, 4 or
    // 16 times the reference size nextpow2(2N) (at most 2^24 cells).  Unlike
    // the launch shapes this can reorder candidate lists that overflow the
    // 32 slots, so it is only tuned on request.
    if (ok && anySc && (flags & 1)) {
        const std::vector<int> keepT = scTableOverride_;
        int bestF = 0;
        float bestMs = 1e30f;
        rep += ", \"hash_table_factor\": {";
        const int factors[3] = {1, 4, 16};
        for (int j = 0; j < 3 && ok; ++j) {
            for (int b = 0; b < numBodies(); ++b) {
                if (!set_[b].scEnabled) continue;
                long long t = 1024;
                while (t < 2LL * bodyCount_[b]) t <<= 1;
                scTableOverride_[b] = (int)std::min<long long>(t * factors[j], 1LL << 24);
            }
            scDeviceValid_ = false;
            float ms = 0.0f;
            ok = trial(ms);
            if (ok && ms < bestMs) { bestMs = ms; bestF = factors[j]; }
            std::snprintf(buf, sizeof(buf), "%s\"%d\": %.4f", j ? ", " : "", factors[j], ms);
            rep += buf;
        }
        if (ok) {
            for (int b = 0; b < numBodies(); ++b) {
                if (!set_[b].scEnabled) continue;
                long long t = 1024;
                while (t < 2LL * bodyCount_[b]) t <<= 1;
                scTableOverride_[b] = (int)std::min<long long>(t * bestF, 1LL << 24);
            }
        } else {
            scTableOverride_ = keepT;
        }
        scDeviceValid_ = false;
        std::snprintf(buf, sizeof(buf), ", \"chosen\": %d}", bestF);
        rep += buf;
    }
    rep += "}";
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!save(true)) { if (ok) err = "autotune: restore failed"; return false; }
    if (!ok) return false;
    report = rep;
    return warm(err);
}

}  // namespace xpbd
