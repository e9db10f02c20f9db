// engine.h — host orchestration of the MI355X XPBD solver.
//
// An Engine owns the device state of one or more bodies (independent
// simulated instances) laid out back to back in one padded node space, and
// steps all of them with one launch sequence per substep.  The C ABI
// singleton (capi.cpp) drives an Engine with one body; the batched API drives
// one with many.  Every substep's launch sequence is captured once in a
// hipGraph and replayed; per-body settings live in device memory so changing
// them does not invalidate the graph.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "kernels.h"
#include "xpbd_device.h"

namespace xpbd {

// Solver settings of one body, with the reference defaults (forward.cu:41-612).
struct Settings {
    int iters = 20;
    float dt = 1.0f / 60.0f, underRelax = 0.6f, velDamping = 0.59f;
    float stiffnessScale = 1.0f, invMassScale = 1.0f, distCompliance = 1e-3f;
    float shapeCompliance = 0.0f, shapeBlend = 0.1f, angleCompliance = 0.0f, angleBlend = 0.2f;
    bool useDistance = true, useShape = false, useAngle = false, useVolume = false;
    bool shapeRobust = true;
    bool groundEnabled = false;
    float groundN[3] = {0, 0, 1};
    float groundHeight = 0, groundFriction = 0.5f, groundRestitution = 0.3f;
    float groundRadius = 0, groundSlop = 0, groundGravity = 0;
    bool contactColliders = false, contactGround = false;
    float contactTau = 0.2f, contactCap = 0.0f;
    float volCompliance = 1e-6f, volAniso = 0.05f;
    int volRingK = 1, volMaxMembers = 64, volLeaderMinHop = 1, volMemberSelect = 1;
    float objShapeStiffness = 0.0f;
    int objMinComponent = 50;
    bool objShapeGPU = true;
    bool scEnabled = false, scWithinBody = false;
    float scRadiusScale = 1.5f, scExcludeScale = 2.0f;
    bool distGS = false, volGS = false;   // forward.cu:310-313 (FORWARD::set*GSEnabled)
};

// Setter clamps of the reference (settings.cpp), shared by both C APIs.
void settings_solver(Settings& s, int iters, float dt, float underRelax, float velDamping, float stiffnessScale,
                     float invMassScale, float distCompliance, float shapeCompliance, float shapeBlend,
                     float angleCompliance, float angleBlend);
void settings_constraints(Settings& s, int distance, int shape, int angle, int volume);
void settings_volume(Settings& s, float compliance, int ringK, int maxMembers, int leaderMinHop);
void settings_ground(Settings& s, int enabled, float ux, float uy, float uz, float height, float friction,
                     float restitution, float contactRadius, float contactSlop, float gravity);
void settings_object_shape(Settings& s, float stiffness);
void settings_self_collision(Settings& s, int enabled, float radiusScale, float excludeScale, int withinBody);
// Updates cols[0 .. count) in place; returns the clamped count.
int collider_list_update(ColliderDev* cols, int prevCount, int count, const int* types, const float* poses,
                         const float* dims);
float positive_or_zero(float v);
// Attachment lists: earlier repeats of a node id become -1 (the last wins).
void keep_last_of_repeats(std::vector<int>& ids);

// One body's input graph in the reference's loadGraph form (local ids).
struct BodyInput {
    int n = 0;
    std::vector<float> pos, scales, rots, opacity;
    std::vector<int> rowOff, nbr;
    std::vector<float> nrest, nstiff;
};

// Build the symmetric CSR of ChainMail::loadGraph (forward.cu:2702-2779).
void build_input(BodyInput& in, int n, const float* pos, const float* scales, const float* rots, const float* opacity,
                 int e, const int* edges, const float* rest, const float* stiff);

// Volume clusters of one body (local ids): leader i owns {i} + members
// memIdx[memOff[i] .. memOff[i+1]); revOff/revIdx list the clusters containing
// each node (own cluster first).
struct VolumeTopo {
    std::vector<char> isLeader;
    std::vector<int> memOff, memIdx, revOff, revIdx;
    int minMembers = 4;       // rest-precompute gate (forward.cu:7745)
    int k = -1, cap = -1, r = -1;
    bool explicitClusters = false;
    // memIdx and revIdx of a topology built on the GPU stay in the engine's
    // device copy (Engine::volumeHostLists fetches them); isLeader, memOff
    // and revOff are always here
    bool listsOnDevice = false;
    long long version = 0;    // changes with every finish() or device build: colourings follow it
    void finish(int n);       // builds the reverse CSR from isLeader / members
    void touch();             // a new version
};

// Gauss-Seidel colourings of one body in caller ids.
// Distance (forward.cu:10407-10517): undirected edges a < b, colour-major.
struct DistanceGSTopo {
    bool built = false;
    std::vector<int> a, b, colorOff;
    std::vector<float> rest, stiff;
    int oneWay = 0;
};
// Volume (forward.cu:9745-9812): leaders in colour-major order.
struct VolumeGSTopo {
    bool built = false, skipped = false;
    long long version = -1;
    int lowerBound = 0;
    std::vector<int> order, colorOff;
};
// Umbrella angle pairing of one body (forward.cu:10708-10819), per directed
// CSR entry in caller order: paired neighbour (-1 none) and rest cosine (2 none).
struct AngleTopo {
    bool built = false;
    std::vector<int> pairNext;
    std::vector<float> restCos;
};
void build_angle_pairs(const BodyInput& in, AngleTopo& t);

void build_distance_gs(const BodyInput& in, DistanceGSTopo& t);
void build_volume_gs(const VolumeTopo& v, int n, VolumeGSTopo& t);

// k-ring clusters by the reference rule (forward.cu:9814-10240).
void build_volume_reference(const BodyInput& in, int k, int cap, int leaderR, VolumeTopo& t);

// Press-plate state of one body (forward.cu:535-575 squash globals, press mode).
struct PressState {
    bool active = false, fromLo = false, autoLog = true;
    int axis = 1, ckptIdx = 0, dwell = 0, dwellFrames = 90;
    float ramp = 0.15f, maxDisp = 0.30f, curDisp = 0.0f, lo = 0.0f, hi = 0.0f;
};

// While alive, device allocations made by this thread bump *e (an engine's
// allocation counter) instead of the process-wide one.
struct AllocScope {
    int* prev;
    explicit AllocScope(int* e);
    ~AllocScope();
    AllocScope(const AllocScope&) = delete;
    AllocScope& operator=(const AllocScope&) = delete;
};
#define XPBD_ALLOC_SCOPE ::xpbd::AllocScope allocScope_(&allocEpoch_)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count);
    void free();
    ~DevBuf() { free(); }
};

// The same topology as build_volume_reference, built on the GPU: every
// node's k-hop ball and its farthest-point sample (balls past the device
// limit on the host), the greedy leader cover on the host (r >= 2) over the
// device balls, the reverse lists by a stable device sort.  memIdx and revIdx
// are left in the device buffers (t.listsOnDevice); identical lists.
bool device_volume_build(const BodyInput& in, int k, int cap, int leaderR, hipStream_t s, VolumeTopo& t,
                         DevBuf<int>& memIdx, DevBuf<int>& revIdx, std::string& err);

class Engine {
public:
    Engine();
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // Node order in device memory, applied by the next build: 0 keeps the
    // caller's order, 1 (default) sorts each body's nodes along a Morton
    // (Z-order) curve of the rest positions so that graph neighbours share
    // cache lines.  Every host-facing array stays in the caller's order.
    void setNodeOrder(int mode) { nodeOrderMode_ = mode; }
    // Where the k-ring volume clusters' member lists and the Gauss-Seidel
    // colourings are built: 0 on the host, 1 on the GPU (the same lists and
    // colours), -1 (default) on the GPU for bodies of at least
    // kDeviceBuildMin nodes.  A change rebuilds the k-ring topologies at the
    // next step.
    void setDeviceBuilders(int mode);
    static constexpr int kDeviceBuildMin = 16384;
    bool permuted() const { return permuted_; }

    // Allocate and upload: x = rest, v = 0, invMass = 1 for every body.
    bool build(const std::vector<const BodyInput*>& inputs, std::string& err);

    int numBodies() const { return (int)bodyBegin_.size(); }
    int nodeBegin(int b) const { return bodyBegin_[b]; }
    int nodeCount(int b) const { return bodyCount_[b]; }
    int padNodes() const { return nPad_; }
    long long directedEntries() const { return m_; }
    long long slicedEntries() const { return sellEntries_; }

    Settings& settings(int b) { return set_[b]; }
    PressState& press(int b) { return press_[b]; }

    void setColliders(int b, int count, const int* types, const float* poses, const float* dims, float margin,
                      float friction);
    // Install a collider list whose previous poses were already tracked by the caller.
    void setColliderArray(int b, const ColliderDev* cols, int count, float margin, float friction);
    int colliderStats(int b, int* hits, float* push, float* centroid, int maxCount);
    bool attach(int b, int count, const int* idx, const float* pos, float weight);
    void addRigidVelocity(int b, const float* dv, const float* dw, const float* c);
    void resetToRest(int b);
    void launchFromRest(int b, const float* lin, const float* ang, const float* tilt);
    void releaseInvMass(int b);   // invMass = 1 on the body (squashStop)

    // Advance every body by `substeps` substeps (one graph replay each).  The
    // launches are queued on the engine stream; sync() waits for them.  Every
    // body of one engine runs the same iteration count.
    bool step(int substeps, std::string& err);
    bool sync(std::string& err);
    // All host-side set-up of the next substep (topologies, colourings, device
    // buffers, parameter upload, graph capture) without launching it.
    bool warm(std::string& err);

    // Copy the dynamic state of body sb of `src` into body b of this engine
    // (both built from the same input with the same node order): positions,
    // velocities, inverse masses (live and saved), attachment weights and
    // flags, object-shape shifts, collider poses, press state and settings.
    // A body continues bit for bit where it left off.
    bool copyBodyFrom(int b, Engine& src, int sb, std::string& err);

    // Attach several bodies at once: request r pins idx[off[r] .. off[r+1])
    // (caller ids of body bodies[r]) to pos; a request with no nodes releases.
    bool attachMany(int count, const int* bodies, const int* off, const int* idx, const float* pos, float weight,
                    std::string& err);

    // Deformed shapes of every body into device memory (device node order;
    // callerPosDevice maps a device node to its caller-order global index).
    bool computeShapesAll(float deformEps, std::string& err);
    const float* shapeScalesDevice() const { return outScale_.p; }
    const float* shapeQuatsDevice() const { return outQuat_.p; }
    const int* callerPosDevice() const { return callerPos_.p; }

    // Launch shapes (kernels.h Tuning).  autotune() times each candidate on
    // `substeps` substeps from the current state, restores the state after
    // each trial and keeps the fastest; results do not depend on the choice.
    Tuning& tuning() { return tune_; }
    // flags bit 0 also tunes the self-collision hash-table sizes.
    bool autotune(int substeps, int flags, std::string& report, std::string& err);
    // Run `substeps` substeps eagerly with HIP events around every kernel on the
    // engine stream; returns JSON {"label": [total_ms, launches], ...}.
    bool profile(int substeps, std::string& json, std::string& err);

    // Committed positions in the caller's node order (device memory).
    float3* positionsDevice() { return permuted_ ? xOut_.p : x_.p; }
    hipStream_t stream() const { return stream_; }
    bool getState(int b, float* pos, float* vel);
    bool setState(int b, const float* pos, const float* vel);
    bool getInvMass(int b, float* invMass);
    const BodyInput* input(int b) const { return inputs_[b]; }

    // Volume clusters: explicit per body (local ids, off == nullptr returns to
    // the k-ring rule).  Statistics evaluate J on the committed positions.
    bool setVolumeClusters(int b, const int* off, const int* members, int minMembers, std::string& err);
    bool volumeStats(int b, float& jvw, float& jvwStd, std::string& err);
    bool volumeTopology(int b, int* leaders, int* memberCount, int* revCount, int* matType, float* detRest,
                        std::string& err);
    double volumeAvgMembers() const { return volAvgMembers_; }
    // Members of every cluster of body b (local ids) in leader order.
    bool volumeMembers(int b, int* members, int len, int& total, std::string& err);

    // Deformed Gaussian shapes of body b from the committed positions
    // (forward.cu:13230-13247); host outputs may be null.  Returns the
    // deformed count in `deformed`.  Repacks the contact ellipsoids when on.
    bool computeShapes(int b, float deformEps, float* hostScales, float* hostQuats, int& deformed, std::string& err);
    // Frame outputs of body b into caller-owned device buffers (any may be
    // null); host4 = sums of x, y, z and the largest change against prev,
    // which takes the new positions when that change exceeds moveEps.
    bool frameOutputs(int b, float* dPos, float* dScale, float* dQuat, bool xyzw, float deformEps, float* dPrev,
                      float moveEps, double host4[4], int& deformed, std::string& err);
    // Contact ellipsoids (native/xpbd_capi.cu:179-191): bit 1 colliders, bit 2 ground.
    bool setContactShape(int b, int enabled, float tau, float radiusCap, std::string& err);

    // Self-collision (native/xpbd_capi.cu:203-212): statistics of the last build
    // (candidates, active particles, overflow), an optional fixed grid
    // (cell > 0 replaces the displacement rule; table a power of two >= 1024),
    // and the build's artefacts in body-local ids.
    bool selfCollisionStats(int b, int out[3], std::string& err);
    void setSelfCollisionGrid(int b, float cellSize, int tableSize);
    // Candidate search of one-object bodies through the predicted-position
    // grid (on by default; off = the coarse scan for every body).  The lists
    // are the same either way.
    void setSelfCollisionFine(bool on) { scFineOn_ = on; }
    bool selfCollisionGridInfo(int b, float info[3], std::string& err);   // cell, d_c, search
    // node counts per candidate-search path of the last build: lane scan (<= 32),
    // lane scan + slot rule, coarse walk, outside the grid
    bool selfCollisionPaths(int b, int out[4], std::string& err);
    bool selfCollisionLists(int b, unsigned int* keys, int* sorted, int* counts, int* contacts, std::string& err);

    // Gauss-Seidel colourings (built on demand, caller ids): info = distance
    // edges, distance colours, one-way edges, volume clusters, volume colours,
    // volume colouring skipped.
    bool gsInfo(int b, int info[6], std::string& err);
    bool distanceGSEdges(int b, int* ab, float* rest, float* stiff, int* colorOff, std::string& err);
    bool volumeGSOrder(int b, int* order, int* colorOff, std::string& err);

    // Shape matching over node clusters (caller ids, -1 = none; null = 1-ring).
    bool setShapeClusters(int b, const int* clusterOf, std::string& err);

    // Umbrella angle pairing of body b (directed entries in caller CSR order).
    bool anglePairs(int b, int* pairNext, float* restCos, std::string& err);

private:
    // angle constraint (forward.cu:4990-5130): pairing per body, device
    // entry arrays, and per node the contributions it receives in the
    // serial order (centre caller index, entry, role)
    std::vector<AngleTopo> angTopo_;
    std::vector<char> angBodies_;           // bodies the device arrays were built for
    bool angDevValid_ = false;
    DevBuf<int> angPair_, angRevOff_, angRevEnt_;
    DevBuf<float> angRestCos_;
    DevBuf<float4> angDp_;
    bool ensureAngle(std::string& err);
    AngleArgs angArgs() const;
    std::vector<std::vector<int>> shapeClusterOf_;
    bool shcDevValid_ = true;
    int shcK_ = 0;
    DevBuf<int> shcOff_, shcMem_, shcOf_;
    DevBuf<float3> shcRest_;
    DevBuf<float> shcParams_;
    bool ensureShapeClusters(std::string& err);
    ShapeClusterArgs shcArgs() const;
    bool ensureDistanceGS(std::string& err);
    bool ensureVolumeGS(std::string& err);
    std::vector<DistanceGSTopo> dgs_;
    std::vector<VolumeGSTopo> vgs_;
    std::vector<char> dgsIn_, vgsIn_;       // bodies merged into the device colourings
    std::vector<int> dgsColorOff_, vgsColorOff_;
    bool dgsDevValid_ = false, vgsDevValid_ = false;
    std::vector<float> dgsAlphaKey_;
    DevBuf<int2> dgsAB_;
    DevBuf<float> dgsRest_, dgsStiff_, dgsAlpha_, dgsLam_;
    DevBuf<int> vgsOrder_;
    bool ensureShapeBuffers(std::string& err);
    bool ensureObjects(std::string& err);
    bool ensureSelfCol(std::string& err);
    SelfColArgs scArgs() const;
    bool ensureVolume(bool needAll, std::string& err);
    bool prepare(std::string& err, bool advance = true);   // per-substep host work before launching
    Tuning tune_;
    int allocEpoch_ = 0;              // device allocations made by this engine
    void afterSubstep();              // host bookkeeping after a substep is queued
    void pressAdvance(int b);
    void fillBodyDev(int b, BodyDev& d) const;
    bool uploadParams(std::string& err);
    std::vector<int> planSignature() const;
    struct Timer;
    void recordSubstep(Timer* t);
    bool rebuildGraph(std::string& err);

    // caller <-> device node order: device local d of body b holds the
    // caller's node ord_[b][d]; inv_[b] is the inverse
    std::vector<std::vector<int>> ord_, inv_;
    int nodeOrderMode_ = 1;
    int deviceBuilders_ = -1;
    bool permuted_ = false;
    DevBuf<int> callerPos_;      // [nPad] caller-order global index of each device node
    DevBuf<int> devOfCaller_;    // [nPad] its inverse: device node of each caller-order index
    DevBuf<float3> xOut_;        // committed positions in caller order (permuted only)
    void refreshCallerPositions(int b);
    template <typename T>
    void toDevice(int b, const T* src, T* dst, int comps) const
    {
        const std::vector<int>& o = ord_[b];
        for (size_t d = 0; d < o.size(); ++d)
            for (int c = 0; c < comps; ++c) dst[d * comps + c] = src[(size_t)o[d] * comps + c];
    }
    template <typename T>
    void toCaller(int b, const T* src, T* dst, int comps) const
    {
        const std::vector<int>& o = ord_[b];
        for (size_t d = 0; d < o.size(); ++d)
            for (int c = 0; c < comps; ++c) dst[(size_t)o[d] * comps + c] = src[d * comps + c];
    }

    hipStream_t stream_ = nullptr;
    int nPad_ = 0;
    long long m_ = 0;
    std::vector<int> bodyBegin_, bodyCount_;
    std::vector<const BodyInput*> inputs_;
    std::vector<Settings> set_;
    std::vector<PressState> press_;
    std::vector<bool> attActive_;

    // colliders (host mirror keeps the previous pose for friction)
    std::vector<ColliderDev> colHost_;
    std::vector<int> colCount_;
    std::vector<float> colMargin_, colFriction_;
    bool colDirty_ = true;
    bool colStatsValid_ = false;

    // device state
    DevBuf<float3> x_, v_, rest_;
    DevBuf<float4> p0_, p1_;
    DevBuf<float> invMass_, invMassSaved_, objW_;
    DevBuf<int> rowOff_, nbr_;
    DevBuf<float> nstiff_;
    // distance entries in sliced-ELL form (kernels.h SellArgs)
    DevBuf<int> sliceOff_, sliceW_, sellNbr_;
    DevBuf<float> sellRest_, sellAlpha_, sellLam_;
    std::vector<char> uniformStiff_;        // every edge of the body has one stiffness
    std::vector<float> uniformStiffValue_;
    long long sellEntries_ = 0;
    SellArgs sellArgs() const;
    DevBuf<int> tileBody_;
    DevBuf<BodyDev> bodies_;
    DevBuf<ColliderDev> cols_;
    DevBuf<int> colHits_, attIdx_;
    DevBuf<float> colPush_, attPos_;

    // parameter staging: bodyHost_ holds the parameter blocks of the next
    // substep; uploads go through pinned slots, and a slot is written again
    // only after the copy that last read it has completed (its event), so
    // substeps queued back to back each see their own parameters
    struct StageSlot {
        void* pin = nullptr;
        size_t cap = 0;
        hipEvent_t done = nullptr;
        bool pending = false;
    };
    static constexpr int kStageSlots = 4;
    StageSlot stage_[kStageSlots];
    int stageNext_ = 0;
    bool stageUpload(void* dst, const void* src, size_t bytes, std::string& err);
    std::vector<BodyDev> bodyHost_;
    std::vector<BodyDev> lastBodies_;
    std::vector<float> alphaKey_;

    // volume constraint
    std::vector<VolumeTopo> volTopo_;
    struct VolDevLists {
        DevBuf<int> memIdx, revIdx;   // caller order, of topologies built on the GPU
    };
    std::vector<std::unique_ptr<VolDevLists>> volDev_;   // per body
    // Host copies of a device-built topology's member and reverse lists.
    bool volumeHostLists(int b, std::string& err);
    // Gauss-Seidel colourings of body b, on the GPU or the host (the same
    // colours; see setDeviceBuilders).
    bool buildDistanceGS(int b, std::string& err);
    bool buildVolumeGS(int b, std::string& err);
    bool volDeviceValid_ = false;
    int volActive_ = 0;
    double volAvgMembers_ = 0.0;
    float volAniso_ = 0.05f;
    DevBuf<int> volOff_, volIdx_, revOff_, revIdx_, matType_, volActiveList_;
    DevBuf<float> detRest_, Vrest_, restLen_, lamVol_, volJ_;
    DevBuf<float4> bbC_, bbS_;
    DevBuf<float2> bbS2_;
    std::vector<int> hMatType_;
    std::vector<float> hVrest_;
    VolumeArgs volArgs() const;

    // Gaussian attributes, deformed shapes, contact ellipsoids (node space)
    DevBuf<float> gsScale_, gsRot_, gsOpacity_, outScale_, outQuat_;
    DevBuf<float> cM_, cRmax_, cRg_;
    DevBuf<unsigned int> shapeCtr_;
    DevBuf<double> frameScratch_;   // frame-output partials, then the four totals
    std::vector<char> contactOn_, contactColliders_, contactGround_;
    std::vector<float> contactTau_, contactCap_;

    // object components (forward.cu:11303-11439) and object shape matching
    struct ObjTopo {
        int K = 0, minComp = -1;
        std::vector<int> id, count;
        std::vector<float> c0;
        std::vector<double> restRel, sumQ;
    };
    std::vector<ObjTopo> objTopo_;
    std::vector<int> objBase_;
    std::vector<char> objShiftValid_;
    int objTotal_ = 0;
    bool objDeviceValid_ = false;
    DevBuf<int> objId_, objBody_, objCount_;
    DevBuf<float> objC0_, objParams_, objDisp_, objBoxes_, pushSums_, pushT_;
    DevBuf<double> objSumQ_, objShift_, objRestRel_, objSums_;

    // self-collision: one hash table per body inside shared cell arrays
    // candidate search through a second grid on the predicted positions for
    // bodies whose nodes form one object (same lists as the coarse scan)
    std::vector<char> objSingle_;           // every node of the body in one object
    std::vector<int> fineTable_, fineBase_;
    int fineTotal_ = 0;
    bool scFineOn_ = true;
    DevBuf<unsigned int> fkeys_;
    DevBuf<int> frank_, fcount_, fstart_, fscan_, finePos_, laneN_, scLong_;
    DevBuf<int> coarseRank_, accCount_;
    DevBuf<float4> fineE_;
    DevBuf<int4> fineC_;
    DevBuf<int2> laneKey_;
    std::vector<float> scSpacing_, scCellOverride_;
    std::vector<int> scTable_, scTableBase_, scTableOverride_;
    int scTableTotal_ = 0, scSortBits_ = 0;
    bool scDeviceValid_ = false;
    std::vector<char> scBuilt_;             // a build has filled the body's lists since the arrays were made
    const int* scSortedIdx_ = nullptr;
    const unsigned int* scSortedKeys_ = nullptr;
    DevBuf<unsigned int> scMax2_, scKeysDev_, scKeys_, scKeysAlt_, scHist_;
    DevBuf<int> scCounters_, scUseActive_, scExact_, scIdx_, scIdxAlt_, cellStart_, cellEnd_, contacts_, counts_;
    DevBuf<int> scOver_, scStats_;   // per-node overflow; statistics read-back
    DevBuf<int> scSortGate_;          // 0: the last build skipped the coarse sort (2: sorted since, on demand)
    DevBuf<float> scBodyDisp_, scInvCell_;
    DevBuf<unsigned char> scActive_;
    DevBuf<float3> scAccum_;
    DevBuf<float4> scSortP_, scSortX_, scSortR_;

    // graph
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graphExec_ = nullptr;
    std::vector<int> graphSig_;
};

}  // namespace xpbd
