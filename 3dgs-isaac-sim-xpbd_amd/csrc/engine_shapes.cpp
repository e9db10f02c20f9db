// engine_shapes.cpp — deformed Gaussian shapes and contact ellipsoids of the
// Engine (forward.cu:13230-13247 computeDeformedShapes; 6733-6757
// updateContactShapes; native/xpbd_capi.cu:179-191, 248-260).
#include <cstdint>
#include <algorithm>
#include <cmath>
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

// Scales, rotations and opacity of every body in node space (uploaded once).
bool Engine::ensureShapeBuffers(std::string& err)
{
    if (gsScale_.p) return true;
    gsScale_.alloc((size_t)3 * nPad_);
    gsRot_.alloc((size_t)4 * nPad_);
    gsOpacity_.alloc(nPad_);
    outScale_.alloc((size_t)3 * nPad_);
    outQuat_.alloc((size_t)4 * nPad_);
    shapeCtr_.alloc(deformed_shape_blocks(nPad_) + 1);
    if (!gsScale_.p || !gsRot_.p || !gsOpacity_.p || !outScale_.p || !outQuat_.p || !shapeCtr_.p) {
        err = "shapes: hipMalloc failed";
        return false;
    }
    for (int b = 0; b < numBodies(); ++b) {
        const BodyInput& in = *inputs_[b];
        const size_t nb = bodyBegin_[b], n = in.n;
        std::vector<float> sc(3 * n), rt(4 * n), op(n);
        toDevice(b, in.scales.data(), sc.data(), 3);
        toDevice(b, in.rots.data(), rt.data(), 4);
        toDevice(b, in.opacity.data(), op.data(), 1);
        XPBD_HIP(hipMemcpyAsync(gsScale_.p + 3 * nb, sc.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice, stream_));
        XPBD_HIP(hipMemcpyAsync(gsRot_.p + 4 * nb, rt.data(), sizeof(float) * 4 * n, hipMemcpyHostToDevice, stream_));
        XPBD_HIP(hipMemcpyAsync(gsOpacity_.p + nb, op.data(), sizeof(float) * n, hipMemcpyHostToDevice, stream_));
        XPBD_HIP(hipStreamSynchronize(stream_));   // staging vectors go out of scope
    }
    return true;
}

static float cap_or_zero(float c) { return (std::isfinite(c) && c > 0.0f) ? c : 0.0f; }

bool Engine::setContactShape(int b, int enabled, float tau, float radiusCap, std::string& err)
{
    XPBD_ALLOC_SCOPE;
    if (!ensureShapeBuffers(err)) return false;
    contactOn_[b] = enabled != 0;
    contactTau_[b] = tau;
    contactCap_[b] = radiusCap;
    if (contactOn_[b]) {
        if (!cM_.p) {
            cM_.alloc((size_t)6 * nPad_);
            cRmax_.alloc(nPad_);
            cRg_.alloc(nPad_);
            if (!cM_.p || !cRmax_.p || !cRg_.p) { err = "contact: hipMalloc failed"; return false; }
            XPBD_HIP(hipMemsetAsync(cM_.p, 0, sizeof(float) * 6 * nPad_, stream_));
            XPBD_HIP(hipMemsetAsync(cRmax_.p, 0, sizeof(float) * nPad_, stream_));
            XPBD_HIP(hipMemsetAsync(cRg_.p, 0, sizeof(float) * nPad_, stream_));
        }
        const size_t nb = bodyBegin_[b];
        launch_contact_pack(stream_, (int)nb, bodyCount_[b], gsScale_.p + 3 * nb, gsRot_.p + 4 * nb, gsOpacity_.p + nb,
                            tau, cap_or_zero(radiusCap), cM_.p, cRmax_.p);
    }
    contactColliders_[b] = (enabled & 1) != 0;
    contactGround_[b] = (enabled & 2) != 0;
    XPBD_HIP(hipStreamSynchronize(stream_));
    XPBD_HIP(hipGetLastError());
    return true;
}

bool Engine::computeShapes(int b, float deformEps, float* hostScales, float* hostQuats, int& deformed, std::string& err)
{
    XPBD_ALLOC_SCOPE;
    if (!ensureShapeBuffers(err)) return false;
    const size_t nb = bodyBegin_[b];
    const int n = bodyCount_[b];
    launch_deformed_shapes(stream_, (int)nb, n, rest_.p, x_.p, rowOff_.p, nbr_.p, gsScale_.p + 3 * nb, gsRot_.p + 4 * nb,
                           deformEps, outScale_.p + 3 * nb, outQuat_.p + 4 * nb, shapeCtr_.p);
    if (contactOn_[b] && cM_.p)
        launch_contact_pack(stream_, (int)nb, n, outScale_.p + 3 * nb, outQuat_.p + 4 * nb, gsOpacity_.p + nb,
                            contactTau_[b], cap_or_zero(contactCap_[b]), cM_.p, cRmax_.p);
    unsigned int c = 0;
    std::vector<float> sc(hostScales ? 3 * (size_t)n : 0), qt(hostQuats ? 4 * (size_t)n : 0);
    XPBD_HIP(hipMemcpyAsync(&c, shapeCtr_.p + deformed_shape_blocks(n), sizeof(unsigned int), hipMemcpyDeviceToHost,
                            stream_));
    if (hostScales)
        XPBD_HIP(hipMemcpyAsync(sc.data(), outScale_.p + 3 * nb, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, stream_));
    if (hostQuats)
        XPBD_HIP(hipMemcpyAsync(qt.data(), outQuat_.p + 4 * nb, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    XPBD_HIP(hipGetLastError());
    if (hostScales) toCaller(b, sc.data(), hostScales, 3);
    if (hostQuats) toCaller(b, qt.data(), hostQuats, 4);
    deformed = (int)c;
    return true;
}

bool Engine::frameOutputs(int b, float* dPos, float* dScale, float* dQuat, bool xyzw, float deformEps, float* dPrev,
                          float moveEps, double host4[4], int& deformed, std::string& err)
{
    XPBD_ALLOC_SCOPE;
    const int nb = (int)bodyBegin_[b];
    const int n = bodyCount_[b];
    // caller index of device node i: callerPos[i] - callerBase
    int callerBase = 0;
    for (int k = 0; k < b; ++k) callerBase += bodyCount_[k];
    deformed = 0;
    unsigned int c = 0;
    const bool shapes = dScale || dQuat;
    if (shapes) {
        if (!ensureShapeBuffers(err)) return false;
        launch_deformed_shapes(stream_, nb, n, rest_.p, x_.p, rowOff_.p, nbr_.p, gsScale_.p + 3 * (size_t)nb,
                               gsRot_.p + 4 * (size_t)nb, deformEps, outScale_.p + 3 * (size_t)nb,
                               outQuat_.p + 4 * (size_t)nb, shapeCtr_.p);
        if (contactOn_[b] && cM_.p)
            launch_contact_pack(stream_, nb, n, outScale_.p + 3 * (size_t)nb, outQuat_.p + 4 * (size_t)nb,
                                gsOpacity_.p + nb, contactTau_[b], cap_or_zero(contactCap_[b]), cM_.p, cRmax_.p);
        launch_shapes_to_caller(stream_, nb, n, callerBase, callerPos_.p, outScale_.p + 3 * (size_t)nb,
                                outQuat_.p + 4 * (size_t)nb, dScale, dQuat, xyzw);
        XPBD_HIP(hipMemcpyAsync(&c, shapeCtr_.p + deformed_shape_blocks(n), sizeof(unsigned int),
                                hipMemcpyDeviceToHost, stream_));
    }
    const size_t need = 4 * (size_t)frame_output_blocks(n) + 4;
    if (frameScratch_.n < need) {
        frameScratch_.alloc(need);
        if (!frameScratch_.p) { err = "hipMalloc failed"; return false; }
    }
    // the positions in caller order (the caller-order copy, or the device
    // array itself when the nodes keep the caller's order)
    const float3* pos = permuted_ ? xOut_.p + callerBase : x_.p + nb;
    double* tot = frameScratch_.p + need - 4;
    launch_frame_positions(stream_, n, pos, dPos, dPrev, moveEps, frameScratch_.p, tot);
    XPBD_HIP(hipMemcpyAsync(host4, tot, 4 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    XPBD_HIP(hipGetLastError());
    deformed = shapes ? (int)c : 0;
    return true;
}

}  // namespace xpbd

namespace xpbd {

// Shape-matching clusters (extension).  Members of a cluster are listed in
// increasing caller order and its rest centroid is summed in that order, so
// the result is independent of the device node order.
bool Engine::setShapeClusters(int b, const int* clusterOf, std::string& err)
{
    const int n = bodyCount_[b];
    std::vector<int> of;
    if (clusterOf) {
        of.assign(clusterOf, clusterOf + n);
        for (int v : of)
            if (v < -1 || v >= n) { err = "shape clusters: cluster id out of range"; return false; }
    }
    shapeClusterOf_[b] = std::move(of);
    shcDevValid_ = false;
    return true;
}

ShapeClusterArgs Engine::shcArgs() const
{
    return ShapeClusterArgs{shcK_, shcOff_.p, shcMem_.p, shcRest_.p, shcOf_.p, shcParams_.p};
}

bool Engine::ensureShapeClusters(std::string& err)
{
    if (shcDevValid_) return true;
    XPBD_HIP(hipStreamSynchronize(stream_));   // queued substeps may still use the old buffers
    std::vector<int> hOff(1, 0), hMem, hOf(nPad_, -1);
    std::vector<float3> hRest;
    for (int b = 0; b < numBodies(); ++b) {
        const std::vector<int>& of = shapeClusterOf_[b];
        if (of.empty()) continue;
        const BodyInput& in = *inputs_[b];
        const int n = bodyCount_[b], nb = bodyBegin_[b];
        int K = 0;
        for (int v : of) K = std::max(K, v + 1);
        std::vector<int> cnt(K, 0), off(K + 1, 0);
        for (int v : of)
            if (v >= 0) ++cnt[v];
        for (int k = 0; k < K; ++k) off[k + 1] = off[k] + cnt[k];
        std::vector<int> mem(off[K]), fill(off.begin(), off.end() - 1);
        for (int i = 0; i < n; ++i)
            if (of[i] >= 0) mem[fill[of[i]]++] = i;
        const int g0 = (int)hOff.size() - 1;
        // clusters numbered by their first member in the device order, so
        // consecutive clusters gather from nearby memory (each cluster's
        // result does not depend on its number)
        std::vector<int> first(K, INT32_MAX), order(K), rank(K);
        for (int i = 0; i < n; ++i)
            if (of[i] >= 0) first[of[i]] = std::min(first[of[i]], inv_[b][i]);
        for (int k = 0; k < K; ++k) order[k] = k;
        std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return first[u] < first[v]; });
        for (int r = 0; r < K; ++r) rank[order[r]] = r;
        for (int r = 0; r < K; ++r) {
            const int k = order[r];
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            for (int t = off[k]; t < off[k + 1]; ++t) {
                const int i = mem[t];
                sx = sx + in.pos[3 * i]; sy = sy + in.pos[3 * i + 1]; sz = sz + in.pos[3 * i + 2];
                hMem.push_back(nb + inv_[b][i]);
            }
            const int c = off[k + 1] - off[k];
            const float ic = c > 0 ? 1.0f / (float)c : 0.0f;
            hRest.push_back(make_float3(sx * ic, sy * ic, sz * ic));
            hOff.push_back((int)hMem.size());
        }
        for (int i = 0; i < n; ++i)
            if (of[i] >= 0) hOf[nb + inv_[b][i]] = g0 + rank[of[i]];
    }
    shcK_ = (int)hOff.size() - 1;
    const size_t K = std::max(shcK_, 1);
    shcOff_.alloc(K + 1); shcMem_.alloc(std::max<size_t>(hMem.size(), 1)); shcOf_.alloc(nPad_);
    shcRest_.alloc(K); shcParams_.alloc(16 * K);
    if (!shcOff_.p || !shcMem_.p || !shcOf_.p || !shcRest_.p || !shcParams_.p) {
        err = "shape clusters: hipMalloc failed";
        return false;
    }
    XPBD_HIP(hipMemcpyAsync(shcOff_.p, hOff.data(), sizeof(int) * hOff.size(), hipMemcpyHostToDevice, stream_));
    if (!hMem.empty())
        XPBD_HIP(hipMemcpyAsync(shcMem_.p, hMem.data(), sizeof(int) * hMem.size(), hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipMemcpyAsync(shcOf_.p, hOf.data(), sizeof(int) * nPad_, hipMemcpyHostToDevice, stream_));
    if (!hRest.empty())
        XPBD_HIP(hipMemcpyAsync(shcRest_.p, hRest.data(), sizeof(float3) * hRest.size(), hipMemcpyHostToDevice, stream_));
    XPBD_HIP(hipStreamSynchronize(stream_));
    shcDevValid_ = true;
    return true;
}

}  // namespace xpbd
