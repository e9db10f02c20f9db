// kernels.h — launcher declarations shared by the host orchestration.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "xpbd_device.h"

namespace xpbd {

// Grid size for kernels that remap blocks per XCD (a multiple of 8).
inline int xcd_grid(int blocks) { return (blocks + 7) & ~7; }

// Launch shapes the autotuner may change.  None of them changes a result:
// each only regroups independent waves into blocks or picks between two
// kernels that sum in the same order.
struct Tuning {
    int jacobiTiles = 4;        // 64-node tiles (one wave each) per distance Jacobi block: 1, 2, 4
    int volMode = -1;           // volume cluster solve: -1 by mean member count, 0 thread, 1 wave per cluster
    int shapeClusterWaves = 4;  // waves (4 clusters each) per cluster shape-matching block: 1, 2, 4
    int scBuildWaves = 4;       // waves per self-collision candidate-build block: 1, 2, 4
    int scBuildNodes = 32;      // nodes per wave of the candidate build: 4, 8, 16, 32, 64
    int scBuildXcd = 16;        // candidate build: runs of that many blocks dealt round-robin over the XCDs
                                // (2-256), one contiguous run per XCD (1), plain block order (0)
    int scGroupLanes = -1;      // fine candidate scan: 1 one lane per node in the predicted-position grid's
                                // own order (the nodes it hands over take the coarse walk); 0 every node
                                // takes the coarse walk; -1 automatic (1)
};

// Node-space launch description shared by every per-node kernel.
struct NodeArgs {
    int nPad;                 // padded node count (multiple of kTile)
    const int* tileBody;      // [nPad / kTile]
    const BodyDev* bodies;    // [B]
};

// Directed distance entries in sliced-ELL form: slice t (= node tile t) holds
// sliceW[t] columns of 64 lanes from entry sliceOff[t]; padding has nbr -1.
struct SellArgs {
    const int* sliceOff;      // [tiles]
    const int* sliceW;        // [tiles]
    const int* nbr;           // [entries]
    const float* rest;        // [entries] rest length
    float* alpha;             // [entries] alpha tilde (bodies with per-edge stiffness)
    float* lam;               // [entries]
};

// Buffers the per-node epilogue touches.
struct EpiArgs {
    float3* x;                // committed positions (friction reference, commit)
    float3* v;
    const ColliderDev* cols;
    int* colHits;
    float* colPush;
    const float* cM;          // contact ellipsoids k^2 Sigma [N x 6] or null
    const float* cRmax;       // [N] or null
    const float* rg;          // [N] ground support radius or null
    const int* callerPos;     // [N] caller-order index of each node
    float3* xCaller;          // committed positions in caller order, or null
};

// Volume-constraint buffers (node space; leaders own clusters).
struct VolumeArgs {
    int nActive;              // leaders with matType 0
    const int* active;        // [nActive]
    const int* volOff;        // [nPad + 1] members CSR by leader
    const int* volIdx;
    const int* revOff;        // [nPad + 1] clusters containing each node
    const int* revIdx;
    float* detRest;
    float* Vrest;
    float* restLen;
    int* matType;
    float* lam;               // [nPad] per leader
    float4* bbC;              // [nPad] centroid + coef
    float4* bbS;              // [nPad] S^-1 00 11 22 10
    float2* bbS2;             // [nPad] S^-1 20 21
};

void launch_vol_precompute(hipStream_t s, const NodeArgs& na, const float3* rest, const int* rowOff, const int* nbr,
                           const VolumeArgs& va, float aniso);
void launch_vol_solve(hipStream_t s, const NodeArgs& na, bool wave, bool first, bool last, const float4* P,
                      const VolumeArgs& va);
void launch_vol_apply(hipStream_t s, const NodeArgs& na, const float4* Pin, float4* Pout, const VolumeArgs& va,
                      unsigned epi, const EpiArgs& ea);
void launch_vol_J(hipStream_t s, const NodeArgs& na, const float3* X, const VolumeArgs& va, float* outJ);

// Umbrella angle constraint over the device CSR: entry e of centre i pairs
// nbr[e] with pair[e]; dp[3e + role] holds the contribution to the centre
// (role 0), nbr[e] (1) and pair[e] (2) with w = 1 when it counts; node j
// sums the entries revEnt[revOff[j] .. revOff[j+1]) in that order.
struct AngleArgs {
    const int* rowOff;        // [nPad + 1]
    const int* nbr;           // [M]
    const int* pair;          // [M] device id or -1
    const float* restCos;     // [M]
    float4* dp;               // [3M]
    const int* revOff;        // [nPad + 1]
    const int* revEnt;        // [contributions] 3e + role
};
void launch_angle(hipStream_t s, const NodeArgs& na, const float4* Pin, float4* Pout, const AngleArgs& aa,
                  unsigned epi, const EpiArgs& ea);

// Graph-coloured Gauss-Seidel: one launch per colour, in place on P.
void launch_dist_gs_alpha(hipStream_t s, int m, const int2* ab, const float* stiff, const int* tileBody,
                          const BodyDev* bodies, float* alpha);
void launch_dist_gs(hipStream_t s, bool first, bool last, int begin, int count, const int2* ab, const float* rest,
                    const float* alpha, float* lam, float4* P, const int* tileBody, const BodyDev* bodies);
void launch_vol_gs(hipStream_t s, bool first, bool last, int begin, int count, const int* order, float4* P,
                   const VolumeArgs& va, const int* tileBody, const BodyDev* bodies);

constexpr int kSelfColCap = 32;   // forward.cu:223 SELFCOL_MAX_CONTACTS
// per-node candidate-search path codes (accCount of fine-grid bodies)
constexpr int kScAccCoarse = -2;   // the coarse walk (the reference's own)
constexpr int kScAccLane = -5;     // the lane fine scan
constexpr int kScAccRule = -6;     // the lane fine scan, more than 32 candidates (slot rule)
constexpr int kScAccLong = -8;     // the lane fine scan, list ranked by the group kernel (transient)
constexpr int kLaneList = 160;     // candidates the lane fine scan keeps per node (more: the coarse walk)

// Object-shape and self-collision buffers.  Objects are numbered globally
// (objBase of each body + local component id).
struct SelfColArgs {
    const int* tileBody;
    const BodyDev* bodies;
    // objects
    const int* objId;          // [nPad] global object id or -1
    const int* objBody;        // [K]
    const int* objCount;       // [K]
    const float* objC0;        // [3K]
    const double* objSumQ;     // [3K]
    double* objShift;          // [3K]
    const double* objRestRel;  // [3 nPad]
    const float* objW;         // [nPad] fit weights while attached
    double* objSums;           // [16K]
    float* objParams;          // [16K]
    float* objDisp;            // [K]
    float* objBoxesArr;        // [6K]
    // self-collision
    unsigned int* scMax2;      // [B] max |p - x|^2 bits
    int* scCounters;           // [3B] candidates, overflow, active
    int* scUseActive;          // [B]
    int* scExact;              // [B] the cell edge is below search + 2 max move (capped or fixed):
                               //     candidates may lie outside the 27 cells around x
    int* scSortGate;           // [1] 1 when some node takes the coarse walk this build: the coarse
                               //     sort, cell ranges, sorted copies and walk run only then
    const int* devOfCaller;    // [nPad] device node of each caller-order index
    float* scBodyDisp;         // [B]
    float* scInvCell;          // [B]
    unsigned char* scActive;   // [nPad]
    unsigned int* keysDev;     // [nPad] coarse key of each node (node order)
    unsigned int* keys;        // [nPad] the coarse sort's input: keys in caller order,
    int* idx;                  // [nPad]   with their nodes (written when the sort runs)
    unsigned int sentinel;     // = tableTotal
    int tableTotal;
    int* cellStart;            // [tableTotal]
    int* cellEnd;
    int* contacts;             // [nPad * kSelfColCap]
    int* counts;               // [nPad]
    int* overflow;             // [nPad] accepted candidates beyond the 32 slots
    float3* accum;
    float* pushSums;           // [4K]
    float* pushT;              // [4K]
    const int* callerPos;      // [nPad] caller-order index (sort position of equal keys)
    float4* sortP;             // [nPad] sorted order: p, object id bits
    float4* sortX;             // [nPad] sorted order: x, node id bits
    float4* sortR;             // [nPad] sorted order: rest
    int* coarseRank;           // [nPad] sorted position of each node
    const int* sortedIdx;      // [nPad] node at each sorted position (set for the coarse scan)
    // second grid on the predicted positions (cell edge >= search), used for
    // bodies whose nodes all belong to one object; ordered by a counting sort
    // (the order inside a cell does not matter: lists are ranked by scan key)
    int fineTotal;             // cells of all fine tables
    unsigned int* fkeys;       // [nPad] fine key of each node (fineTotal = none)
    int* frank;                // [nPad] rank of the node inside its fine cell
    int* fcount;               // [fineTotal] nodes per fine cell
    int* fstart;               // [fineTotal + 1] first fine position of each cell (exclusive prefix sum)
    int* fscan;                // block sums of the prefix sum
    int* finePos;              // [nPad] fine position of each node, -1 when not in the grid
    float4* fineE;             // [2 nPad] fine order, two float4 per entry: (p.x, p.y, p.z, rest.x),
                               //   (rest.y, rest.z, caller-order position | code << 27, node id |
                               //   (code >> 5) << 27), code = low 3 bits of each coarse cell coordinate
                               //   of x (x, y, z from bit 0); both fields below 2^27;
                               //   one 32 B entry per candidate test
    int4* fineC;               // [nPad] fine order: coarse cell of x, node id
    // lists of the lane fine scan, one contiguous list per fine position
    // (entry k of position t at t kLaneList + k)
    int* laneN;                // [nPad] fine order: candidates of the lane scan, -1 handed over
    int* scLong;               // [1 + nPad] this build's long-list queue: count, then the nodes whose
                               //   lane-scan lists exceed 32 (any order; null without the fine grid)
    int2* laneKey;             // [nPad * kLaneList] (scan key, node id) of each candidate,
                               //   unsorted; (node id, priority bits) in rank order for lists that take
                               //   the slot rule
    int* accCount;             // [nPad] per node: what its candidate search needs (kScLane, kScNeedCoarse,
                               //        kScRule, or 0: nothing to scan)
};

void launch_object_shape(hipStream_t s, const NodeArgs& na, const SelfColArgs& a, int K, float4* P, const float3* rest,
                         bool boxes);
// Returns true when the sorted (keys, idx) ended in the alternate buffers.
bool launch_selfcol_build(hipStream_t s, const NodeArgs& na, int numBodies, const SelfColArgs& a, const float3* X,
                          const float4* P, const float3* rest, int sortBits, unsigned int* keysAlt, int* idxAlt,
                          unsigned int* hist, const int** sortedIdxOut, const Tuning& tu,
                          void (*mark)(void*, const char*) = nullptr, void* markCtx = nullptr);
void launch_selfcol_iteration(hipStream_t s, const NodeArgs& na, const SelfColArgs& a, const float4* Pin,
                              float4* Pout, const float3* X, const float3* rest, unsigned epi, const EpiArgs& ea);
void launch_contact_push(hipStream_t s, const NodeArgs& na, const SelfColArgs& a, int K, float4* P, const EpiArgs& ea);
// out[3] = candidates, overflow, active nodes of the last build over [nodeBegin, nodeEnd).
void launch_selfcol_stats(hipStream_t s, int nodeBegin, int nodeEnd, const SelfColArgs& a, int* out);
int radix_sort_tiles(int n);
size_t radix_sort_hist_size(int n);   // entries of the sort's histogram buffer
// gate (device int, may be null): when it reads 0 the sort's kernels return
// at once and the buffers keep their contents
bool launch_radix_sort(hipStream_t s, int n, int bits, unsigned int* k0, int* v0, unsigned int* k1, int* v1,
                       unsigned int* hist, bool iota, const int* gate = nullptr);
// The coarse sort's input from the node-order keys: keys[pos] =
// keysDev[devOfCaller[pos]], idx[pos] = devOfCaller[pos] (gate as above).
void launch_sc_caller_order(hipStream_t s, int nPad, const unsigned int* keysDev, const int* devOfCaller,
                            unsigned int* keys, int* idx, const int* gate);
// Volume-cluster member lists on the device (one body, caller order): the
// k-hop ball of every node, farthest-point sampled to cap.  Balls of more
// than kVbMaxBall nodes are left to the host (count -1).
constexpr int kVbMaxBall = 1536;
constexpr int kVbGrid = 4096;   // waves of the persistent ball kernel
struct VbArgs {
    int n, k, cap;
    const int* rowOff;   // [n + 1] symmetric CSR
    const int* nbr;
    const float* pos;    // [3n] rest positions
    int* count;          // [n] members per node (-1: host)
    const int* off;      // [n] first member of each device-built node (fill pass)
    int* out;            // members
};
void launch_volume_balls(hipStream_t s, const VbArgs& a, bool fill);
// Reverse lists of a member CSR (caller order): cnt[n] = clusters containing
// each node; (keys, vals)[L + m] = the (node, cluster) pairs, own cluster
// first, for a stable sort by node.  leaderPos: rank of each leader among the
// L leaders, -1 otherwise (null: every node leads, L = n).
void launch_volume_reverse(hipStream_t s, int n, const int* leaderPos, int L, const int* memOff, const int* memIdx,
                           int m, int* cnt, unsigned int* keys, int* vals);
// dst rows nb + d (offsets dstOff, padded device order) = caller rows ord[d]
// of (srcOff, srcIdx) with entries mapped to nb + inv[x].
void launch_permute_rows(hipStream_t s, int n, int nb, const int* ord, const int* inv, const int* srcOff,
                         const int* srcIdx, const int* dstOff, int* dst);

// Gauss-Seidel colourings on the device (the reference's greedy colourings,
// computed in rounds; colours up to 64 kGsWords - 1, more: flags[1] set).
constexpr int kGsWords = 4;
void launch_gs_entry_keys(hipStream_t s, int n, const int* rowOff, const int* nbr, unsigned int* ka, unsigned int* kb,
                          int* val);
void launch_gs_gather_key(hipStream_t s, int m, const unsigned int* ka, const int* val, unsigned int* key);
void launch_gs_heads(hipStream_t s, int n, int m, const unsigned int* sa, const unsigned int* kb, const int* val,
                     int* head);
void launch_gs_edges(hipStream_t s, int n, int m, const unsigned int* sa, const unsigned int* kb, const int* val,
                     const int* head, const int* rank, const float* nrest, const float* nstiff, int* ea, int* eb,
                     float* erest, float* estiff, int* oneWay);
void launch_gs_incidence(hipStream_t s, int me, const int* ea, const int* eb, int* deg, int* cursor, int* inc,
                         bool countOnly);
void launch_gs_edge_round(hipStream_t s, int me, const int* ea, const int* eb, const int* incOff, const int* inc,
                          int* color, int* flags);
void launch_gs_cluster_round(hipStream_t s, int L, const int* leaders, const int* memOff, const int* memIdx,
                             const int* revOff, const int* revIdx, int* color, int* flags);

// Exclusive prefix sum of in[0, n) into out[0, n] (out[n] = total); tmp holds
// exclusive_scan_tmp_size(n) ints.
size_t exclusive_scan_tmp_size(int n);
void launch_exclusive_scan(hipStream_t s, int n, const int* in, int* out, int* tmp);
void launch_fill_double(hipStream_t s, int n, double* a);

// Shape-matching node clusters (K clusters; members in increasing caller
// order, device ids); params per cluster: centroid, rotation, valid flag.
struct ShapeClusterArgs {
    int K;
    const int* off;           // [K + 1]
    const int* mem;           // [members]
    const float3* comRest;    // [K]
    const int* clusterOf;     // [nPad] cluster of each node or -1
    float* params;            // [16 K]
};

void launch_shape_match(hipStream_t s, const NodeArgs& na, const float4* Pin, float4* Pout, const float3* rest,
                        const int* rowOff, const int* nbr, const ShapeClusterArgs& sca, unsigned epi,
                        const EpiArgs& ea, const Tuning& tu);
// scales / rots / outScale / outQuat are the body's own arrays (local index);
// node arrays are indexed from nodeBegin.  counts holds deformed_shape_blocks(n)
// per-block counts followed by their total.
int deformed_shape_blocks(int n);
void launch_deformed_shapes(hipStream_t s, int nodeBegin, int n, const float3* rest, const float3* cur,
                            const int* rowOff, const int* nbr, const float* scales, const float* rots, float deformEps,
                            float* outScale, float* outQuat, unsigned int* counts);
// The file-scope count array of the allocation-free deformed-shape call on
// the current device, and the most Gaussians one call may take.
bool deformed_shape_scratch(unsigned int** out, std::string& err);
int deformed_shape_scratch_max();
void launch_contact_pack(hipStream_t s, int nodeBegin, int n, const float* sc, const float* qt, const float* op,
                         float tau, float rcap, float* M, float* rmax);
void launch_contact_ground(hipStream_t s, const NodeArgs& na, const float* M, float* rg);
// Frame outputs (caller-owned device buffers): positions copied, compared with
// prev (overwritten when the largest change exceeds moveEps) and summed; part
// holds frame_output_blocks(n) * 4 doubles, out receives sum x, y, z and the
// largest change.
int frame_output_blocks(int n);
void launch_frame_positions(hipStream_t s, int n, const float3* pos, float* dPos, float* dPrev, float moveEps,
                            double* part, double* out);
void launch_shapes_to_caller(hipStream_t s, int nodeBegin, int n, int callerBase, const int* callerPos,
                             const float* scale, const float* quat, float* dScale, float* dQuat, bool xyzw);

void launch_predict(hipStream_t s, const NodeArgs& na, const float3* X, float3* V, const float* invMass, float4* P);
void launch_distance_jacobi(hipStream_t s, const NodeArgs& na, bool first, bool last, const float4* Pin, float4* Pout,
                            const SellArgs& sa, unsigned epi, const EpiArgs& ea, const Tuning& tu);
void launch_epilogue(hipStream_t s, const NodeArgs& na, float4* P, unsigned epi, const EpiArgs& ea);
void launch_distance_alpha(hipStream_t s, const NodeArgs& na, const int* rowOff, const float* nstiff, const SellArgs& sa);
void launch_attach(hipStream_t s, int count, const int* idx, const float* pos, int nodeBegin, int nodeEnd, float3* X,
                   float3* V, float* invMass, float* objW, float weight);
void launch_rigid_velocity(hipStream_t s, int nodeBegin, int nodeEnd, const float3* X, float3* V, const float* invMass,
                           float3 dv, float3 dw, float3 c);
// xCaller[callerPos[i]] = X[i] over nodes [nodeBegin, nodeEnd).
void launch_scatter_caller(hipStream_t s, int nodeBegin, int nodeEnd, const int* callerPos, const float3* X,
                           float3* xCaller);
void launch_fill_float(hipStream_t s, int n, float* a, float v);
void launch_fill_int(hipStream_t s, int n, int* a, int v);

}  // namespace xpbd
