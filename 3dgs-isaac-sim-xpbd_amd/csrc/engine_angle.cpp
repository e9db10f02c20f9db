// engine_angle.cpp — device arrays of the umbrella angle constraint.
//
// Per directed entry of the device CSR: the paired neighbour (device id) and
// the rest cosine; per node, the list of corrections it receives, ordered as
// a serial loop over centres in caller order, their entries, and the roles
// centre -> first neighbour -> paired neighbour (forward.cu:4990-5096 adds
// them in that order when run on one thread).
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

AngleArgs Engine::angArgs() const
{
    AngleArgs a;
    a.rowOff = rowOff_.p;
    a.nbr = nbr_.p;
    a.pair = angPair_.p;
    a.restCos = angRestCos_.p;
    a.dp = angDp_.p;
    a.revOff = angRevOff_.p;
    a.revEnt = angRevEnt_.p;
    return a;
}

bool Engine::ensureAngle(std::string& err)
{
    if (angDevValid_) return true;
    XPBD_HIP(hipStreamSynchronize(stream_));   // queued substeps may still use the old buffers
    const int B = numBodies();
    if ((int)angTopo_.size() != B) angTopo_.assign(B, AngleTopo());
    // device CSR row starts (device order) are needed to place the entries
    std::vector<int> hRow(nPad_ + 1);
    XPBD_HIP(hipMemcpy(hRow.data(), rowOff_.p, sizeof(int) * (nPad_ + 1), hipMemcpyDeviceToHost));
    const size_t M = (size_t)hRow[nPad_];
    std::vector<int> hPair(std::max<size_t>(M, 1), -1);
    std::vector<float> hCos(std::max<size_t>(M, 1), 2.0f);
    std::vector<int> revCnt(nPad_ + 1, 0);
    std::vector<int> revOff(nPad_ + 1, 0), revEnt;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            for (int j = 0; j < nPad_; ++j) revOff[j + 1] = revOff[j] + revCnt[j];
            revEnt.assign(std::max(revOff[nPad_], 1), 0);
            std::fill(revCnt.begin(), revCnt.end(), 0);
        }
        for (int b = 0; b < B; ++b) {
            if (!set_[b].useAngle) continue;
            AngleTopo& t = angTopo_[b];
            const BodyInput& in = *inputs_[b];
            if (!t.built) build_angle_pairs(in, t);
            const int nb = bodyBegin_[b];
            const std::vector<int>& inv = inv_[b];
            for (int c = 0; c < in.n; ++c) {
                const int dc = nb + inv[c];
                for (int k = 0; k < in.rowOff[c + 1] - in.rowOff[c]; ++k) {
                    const int e = in.rowOff[c] + k;          // caller entry
                    const int ed = hRow[dc] + k;              // device entry (rows keep their order)
                    const int p2 = t.pairNext[e];
                    const float rc = t.restCos[e];
                    if (pass == 0) {
                        hPair[ed] = p2 >= 0 ? nb + inv[p2] : -1;
                        hCos[ed] = rc;
                    }
                    if (p2 < 0 || p2 == in.nbr[e] || !(rc >= -1.0f && rc <= 1.0f)) continue;
                    const int who[3] = {dc, nb + inv[in.nbr[e]], nb + inv[p2]};
                    for (int role = 0; role < 3; ++role) {
                        const int j = who[role];
                        if (pass == 0) ++revCnt[j];
                        else revEnt[revOff[j] + revCnt[j]++] = 3 * ed + role;
                    }
                }
            }
        }
    }
    angPair_.alloc(hPair.size());
    angRestCos_.alloc(hCos.size());
    angDp_.alloc(3 * hPair.size());
    angRevOff_.alloc(nPad_ + 1);
    angRevEnt_.alloc(revEnt.size());
    if (!angPair_.p || !angRestCos_.p || !angDp_.p || !angRevOff_.p || !angRevEnt_.p) {
        err = "angle: hipMalloc failed";
        return false;
    }
    XPBD_HIP(hipMemcpy(angPair_.p, hPair.data(), sizeof(int) * hPair.size(), hipMemcpyHostToDevice));
    XPBD_HIP(hipMemcpy(angRestCos_.p, hCos.data(), sizeof(float) * hCos.size(), hipMemcpyHostToDevice));
    XPBD_HIP(hipMemcpy(angRevOff_.p, revOff.data(), sizeof(int) * (nPad_ + 1), hipMemcpyHostToDevice));
    XPBD_HIP(hipMemcpy(angRevEnt_.p, revEnt.data(), sizeof(int) * revEnt.size(), hipMemcpyHostToDevice));
    angDevValid_ = true;
    return true;
}

bool Engine::anglePairs(int b, int* pairNext, float* restCos, std::string& err)
{
    if ((int)angTopo_.size() != numBodies()) angTopo_.assign(numBodies(), AngleTopo());
    AngleTopo& t = angTopo_[b];
    if (!t.built) build_angle_pairs(*inputs_[b], t);
    const size_t m = t.pairNext.size();
    if (pairNext && m) std::memcpy(pairNext, t.pairNext.data(), sizeof(int) * m);
    if (restCos && m) std::memcpy(restCos, t.restCos.data(), sizeof(float) * m);
    (void)err;
    return true;
}

}  // namespace xpbd
