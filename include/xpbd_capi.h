/*
 * xpbd_capi.h — C ABI of libxpbd_isaac.so, the MI355X XPBD Gaussian-graph solver.
 *
 * Part 1 is the drop-in surface: the 29 symbols the reference caller binds with
 * ctypes (src/xpbd_bridge.py:58-112), with the reference's argument meaning,
 * return codes (0 ok / -1 error, message in xpbd_last_error) and call-order
 * rules.  Each declaration cites the reference entry point it replaces in
 * native/xpbd_capi.cu.  State is one library-global instance, as in the
 * reference (native/xpbd_capi.cu:6, 17-29); solver settings persist across
 * xpbd_create / xpbd_destroy like the reference's forward.cu globals.
 *
 * Part 2 (xpbd_ext_*) has no counterpart in the reference: state access for
 * parity checks, the explicit-cluster / hash-override recipe API, and
 * diagnostics.  Part 3 (xpbd_batch_*) is the batched many-object API.
 *
 * All arrays are C-contiguous, little-endian, float32 / int32, in the caller's
 * units and ordering.  No function here takes a device pointer except where
 * stated.
 */
#ifndef XPBD_CAPI_H
#define XPBD_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Part 1: drop-in symbols (native/xpbd_capi.cu:66-310)
 * ------------------------------------------------------------------------- */

/* native/xpbd_capi.cu:66 — last error string, owned by the library. */
const char* xpbd_last_error(void);

/* native/xpbd_capi.cu:71-102 — n Gaussians, e undirected edges (i,j).
 * pos[n*3], scales[n*3] (activated), rots[n*4] (w,x,y,z), opacity[n],
 * edges[e*2], rest[e], stiff[e].  Frees any previous instance.  Invalid or
 * self edges are dropped (forward.cu:2729-2731). */
int xpbd_create(int n, const float* pos, const float* scales, const float* rots, const float* opacity,
                int e, const int* edges, const float* rest, const float* stiff);

/* native/xpbd_capi.cu:104-109 — clears colliders and attachments, frees. */
void xpbd_destroy(void);

/* native/xpbd_capi.cu:111 */
int xpbd_count(void);

/* native/xpbd_capi.cu:113-119 -> forward.cu:4896-4921 (clamps as there). */
void xpbd_set_solver(int iters, float dt, float underRelax, float velDamping, float stiffnessScale,
                     float invMassScale, float distCompliance, float shapeCompliance, float shapeBlend,
                     float angleCompliance, float angleBlend);

/* native/xpbd_capi.cu:121-127 */
void xpbd_set_constraints(int distance, int shape, int angle, int volume);

/* native/xpbd_capi.cu:130-137 -> forward.cu:3849-3853, 3934-3944, 3953-3965 */
void xpbd_set_volume(float compliance, int ringK, int maxMembers, int leaderMinHop);

/* native/xpbd_capi.cu:140-145 -> forward.cu:4430-4446 */
void xpbd_set_ground(int enabled, float ux, float uy, float uz, float height, float friction,
                     float restitution, float contactRadius, float contactSlop, float gravity);

/* native/xpbd_capi.cu:148 -> forward.cu:4464-4476 (bodyContact off, minComponent 50) */
void xpbd_set_object_shape(float stiffness);

/* native/xpbd_capi.cu:151 -> forward.cu:4521 */
void xpbd_set_object_shape_gpu(int enabled);

/* native/xpbd_capi.cu:154-158 -> forward.cu:4478-4500.  types: 0 sphere,
 * 1 box, 2 capsule; poses[count*12] = R (row-major 3x3, local->input) + t;
 * dims[count*3].  At most 32 colliders are used. */
void xpbd_set_colliders(int count, const int* types, const float* poses, const float* dims, float margin,
                        float friction);

/* native/xpbd_capi.cu:161-164 -> forward.cu:4502-4519 */
int xpbd_collider_stats(int* hits, float* push, float* centroid, int maxCount);

/* native/xpbd_capi.cu:167-172 -> forward.cu:6790-6796 */
int xpbd_add_rigid_velocity(const float* dv, const float* dw, const float* center);

/* native/xpbd_capi.cu:179-191 -> forward.cu:6733-6757 */
int xpbd_set_contact_shape(int enabled, float tau, float radiusCap);

/* native/xpbd_capi.cu:195-200 -> forward.cu:6759-6788 */
int xpbd_set_attached(int count, const int* idx, const float* pos, float weight);

/* native/xpbd_capi.cu:203-206 -> forward.cu:4532-4540 */
void xpbd_set_self_collision(int enabled, float radiusScale, float excludeScale, int withinBody);

/* native/xpbd_capi.cu:209-212: out[0] candidates, out[1] active, out[2] overflow */
void xpbd_self_collision_stats(int* out);

/* native/xpbd_capi.cu:215 -> forward.cu:3738 */
void xpbd_set_shape_robust(int enabled);

/* native/xpbd_capi.cu:218-225 — one substep; returns after device completion. */
int xpbd_step(void);

/* native/xpbd_capi.cu:228-232 — device pointer to float3[n] committed positions;
 * 0 before the first step. */
unsigned long long xpbd_positions_device(void);

/* native/xpbd_capi.cu:235-243 */
int xpbd_positions_host(float* out, int n);

/* native/xpbd_capi.cu:248-260 — returns the deformed count or -1. */
int xpbd_compute_shapes(float deformEps, float* hostScales, float* hostQuats, int n);

/* native/xpbd_capi.cu:263-269 */
int xpbd_reset(void);

/* native/xpbd_capi.cu:272-276 -> forward.cu:4551-4610 */
int xpbd_launch(const float* lin, const float* ang, const float* tilt);

/* native/xpbd_capi.cu:281-288 */
int xpbd_press_start(int axis, float rampPerSec, float maxDispPct, int fromLo);

/* native/xpbd_capi.cu:291 */
void xpbd_press_autolog(int enabled);

/* native/xpbd_capi.cu:294 */
void xpbd_press_stop(void);

/* native/xpbd_capi.cu:297-303 */
int xpbd_press_state(float* lo, float* hi, float* curDisp, float* maxDisp);

/* native/xpbd_capi.cu:306-310 */
void xpbd_volume_stats(int enable, float* jvw, float* jvwStd);

/* ---------------------------------------------------------------------------
 * Part 2: extensions (no counterpart in the reference)
 * ------------------------------------------------------------------------- */

/* Copy committed positions x[n*3] and velocities v[n*3] of the instance
 * (after the first step).  Either pointer may be null. */
int xpbd_ext_get_state(float* pos, float* vel, int n);

/* Overwrite committed positions and velocities (after the first step), so two
 * solvers can be stepped from an identical state.  Either pointer may be null. */
int xpbd_ext_set_state(const float* pos, const float* vel, int n);

/* Inverse masses invMass[n] (after the first step). */
int xpbd_ext_get_inv_mass(float* invMass, int n);

/* Step `substeps` substeps eagerly (no graph) with HIP events around every
 * kernel on the library's stream; writes JSON {"kernel": [total_ms, launches]}
 * into json[len].  The state advances exactly as with xpbd_step. */
int xpbd_ext_profile(int substeps, char* json, int len);

/* Launch-shape autotuner of the instance (after the first step), as
 * xpbd_batch_autotune: the state is restored and results do not depend on
 * the launch shapes it picks; flags bit 0 also sizes the hash table. */
int xpbd_ext_autotune(int substeps, int flags, char* json, int len);
/* Fine candidate scan of the instance: 1 one lane per node in the
 * predicted-position grid's order (the nodes it hands over take the
 * one-wave-per-node scan); 0 sends every node to the one-wave-per-node scan;
 * -1, the default, is 1.  Candidate lists do not depend on it. */
int xpbd_ext_set_group_lanes(int lanes);

/* Volume clusters given explicitly instead of by the k-ring rule: node i leads
 * the cluster {i} + members[off[i] .. off[i+1]) (an empty range = no cluster).
 * minMembers is the rest-precompute gate (the reference rule rejects clusters
 * with fewer than 4 members; 3 admits node + 3 nearest neighbours, a
 * tetrahedron, whose covariance J equals |V/V0|).  off == NULL returns to the
 * k-ring rule.  Call after xpbd_create. */
int xpbd_ext_set_volume_clusters(const int* off, const int* members, int minMembers);

/* Volume topology diagnostics per node (after the first step): leader flag,
 * member count, number of clusters containing the node, material type
 * (0 volume, 1 surface, 2 fibre, 3 too few members, 4 degenerate) and rest
 * determinant.  Any pointer may be null. */
int xpbd_ext_volume_topology(int* leaders, int* memberCount, int* revCount, int* matType, float* detRest, int n);

/* Members of every volume cluster, concatenated in leader order (caller ids,
 * each cluster in the builder's order; its size is memberCount of
 * xpbd_ext_volume_topology).  Returns the total, or -1 when len is smaller. */
int xpbd_ext_volume_members(int* members, int len);

/* Self-collision grid: cellSize > 0 fixes the hash-cell edge (instead of
 * search radius + 2 x max displacement, capped at search + 16 d_c); tableSize
 * 0 keeps nextpow2(max(1024, 2n)) cells, else a power of two >= 1024.  Applies
 * to later builds; persists like the solver settings. */
int xpbd_ext_set_self_collision_grid(float cellSize, int tableSize);

/* How candidate lists are found for a body whose nodes all belong to one
 * object with self-contact on: 1 (default) through a second grid on the
 * predicted positions (cells of edge >= the search radius) with the results
 * put in the reference's scan order, 0 by the reference's own walk of the 27
 * frame-start cells.  The lists, counts and statistics are identical; only
 * the work differs.  The oracle accepts and ignores it. */
int xpbd_ext_set_self_collision_fine(int enabled);

/* Umbrella angle pairing (forward.cu:10708-10819) of the instance, per
 * directed graph entry in the reference's CSR order (both directions of each
 * edge, input order): the paired neighbour pairNext[m] (-1 none) and the rest
 * cosine restCos[m] (2 none); m = 2 x the number of valid edges. */
int xpbd_ext_angle_pairs(int* pairNext, float* restCos, int m);

/* Artefacts of the last self-collision build (after the first step): cell key
 * of every node (table size = inactive / non-finite), nodes in stable key
 * order, candidate counts counts[n] and candidate lists contacts[n*32]
 * (j, or -(j+1) for a fast cross-object pair).  Any pointer may be null. */
int xpbd_ext_self_collision_lists(unsigned int* keys, int* sorted, int* counts, int* contacts, int n);

/* Device memory layout of the next instance (applied at its first step):
 * 1 (default) stores the nodes sorted along a Morton curve of the rest
 * positions, so graph neighbours share cache lines; 0 keeps the caller's
 * order.  Results are identical either way; every array crossing this ABI
 * (including xpbd_positions_device) stays in the caller's order. */
int xpbd_ext_set_node_order(int mode);

/* Grid of the last self-collision build (after the first step): info[0] cell
 * edge, info[1] contact distance d_c, info[2] search radius. */
int xpbd_ext_self_collision_grid_info(float* info);

/* Candidate-search paths of the last self-collision build (after the first
 * step), node counts: out[0] lists of at most 32 from the lane fine scan,
 * out[1] longer lane-scan lists (slot rule), out[2] the coarse walk (the
 * reference's walk; every node of bodies without the fine grid), out[3]
 * nodes outside the grid (inactive or non-finite). */
int xpbd_ext_self_collision_paths(int* out);

/* Where the volume clusters of the k-ring rule (xpbd_set_volume) get their
 * member lists and the Gauss-Seidel passes (xpbd_ext_set_gauss_seidel) their
 * edges and colourings: 0 the host, 1 the GPU (each node's k-hop ball and
 * its farthest-point sample, the reverse lists, the deduplicated distance
 * edges and both greedy colourings on the device; balls of more than 1536
 * nodes and the greedy leader cover stay on the host), -1 (default) the GPU
 * for bodies of at least 16384 Gaussians.  Lists, edges and colours are
 * identical either way.  A change takes effect at the next step.  No
 * counterpart in the reference, whose rebuildVolumeClusters,
 * buildVolumeGSColors and buildDistanceGSEdges (forward.cu:9745-10517) run
 * on the host. */
int xpbd_ext_set_device_builders(int mode);

/* Shape matching over node clusters (clusterOf[n]: cluster id, -1 = none;
 * e.g. k-means clusters of ~64 nodes) instead of the graph 1-ring, with the
 * same compliance / blend projection (forward.cu:4972-4988).  Clusters of fewer
 * than 4 nodes do not act.  NULL returns to the 1-ring.  Call after
 * xpbd_create. */
int xpbd_ext_set_shape_clusters(const int* clusterOf, int n);

/* Graph-coloured Gauss-Seidel instead of the Jacobi distance pass and the
 * gather volume pass (the reference's FORWARD::setDistanceGSEnabled /
 * setVolumeGSEnabled, forward.cu:3780-3810, off by default).  Settings persist
 * like the solver settings. */
int xpbd_ext_set_gauss_seidel(int distance, int volume);

/* Colouring sizes (after the first step): info[0] distance edges, [1]
 * distance colours, [2] one-way edges, [3] volume clusters, [4] volume
 * colours, [5] 1 when the volume colouring was skipped as too large. */
int xpbd_ext_gs_info(int* info);

/* The distance colouring in colour-major order: edge ends ab[2E] (a < b,
 * caller ids), rest[E], stiff[E], colour offsets colorOff[C+1].  Any pointer
 * may be null. */
int xpbd_ext_distance_gs_edges(int* ab, float* rest, float* stiff, int* colorOff);

/* The volume colouring: leaders in colour-major order order[L], colour
 * offsets colorOff[C+1].  Either pointer may be null. */
int xpbd_ext_volume_gs_order(int* order, int* colorOff);

/* Library identity: "hip-gfx950" for the GPU library, "oracle-cpu" for the oracle. */
const char* xpbd_ext_backend(void);

/* The persistent solver settings after the setters' clamps, as floats in this
 * order (booleans 0/1): 0 iters, 1 dt, 2 underRelax, 3 velDamping,
 * 4 stiffnessScale, 5 invMassScale, 6 distCompliance, 7 shapeCompliance,
 * 8 shapeBlend, 9 angleCompliance, 10 angleBlend, 11-14 distance / shape /
 * angle / volume enabled, 15 volume compliance, 16 ring k, 17 member cap,
 * 18 leader min hop, 19 ground enabled, 20-22 ground normal, 23 height,
 * 24 friction, 25 restitution, 26 contact radius, 27 contact slop,
 * 28 gravity, 29 object-shape stiffness, 30 minimum component,
 * 31 object shape on the device, 32 self-collision enabled, 33 radius scale,
 * 34 exclude scale, 35 within body, 36 robust polar, 37 collider count,
 * 38 collider margin, 39 collider friction, 40 press autolog, 41 J statistics
 * enabled.  Returns the count written (42), or -1 when len is smaller. */
int xpbd_ext_get_settings(float* out, int len);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the deformed Gaussian shapes of xpbd_compute_shapes
 * (forward.cu:13230-13247) from caller-owned device arrays, with no device
 * allocation.  n Gaussians; restDev and posDev float[n*3] (for example the
 * array of xpbd_positions_device), the symmetric CSR rowOffDev int[n+1] and
 * nbrDev (each edge (i,j) as i->j and j->i, in input edge order, as
 * xpbd_create builds it), scalesDev float[n*3] and rotsDev float[n*4]
 * (w,x,y,z) of the rest shapes; writes outScalesDev float[n*3] and
 * outQuatsDev float[n*4].  Runs on the null stream and returns the deformed
 * count (or -1) after the work has completed.  The per-block counts go to a
 * file-scope device array sized for n <= 16,777,216; larger n is refused.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * race.  The CSR must index [0, n). */
int xpbd_ext_compute_shapes_noalloc(int n, const float* restDev, const float* posDev, const int* rowOffDev,
                                    const int* nbrDev, const float* scalesDev, const float* rotsDev, float deformEps,
                                    float* outScalesDev, float* outQuatsDev);

/* Frame outputs in device memory: what a scene writer consumes every frame,
 * without the host copies of the reference's runtime (runtime_loop.py:387-402
 * reads positions and shapes to the host; 938-946 then takes their mean and
 * the largest change to decide whether to write).  All pointers are device
 * memory owned by the caller, n the instance size:
 *   posDev   float[n*3]  positions, caller order (may be null);
 *   scaleDev float[n*3], quatDev float[n*4]: the deformed shapes of
 *            xpbd_compute_shapes (either may be null; both null skips the
 *            shape update); quaternions (w,x,y,z), or (x,y,z,w) when quatXYZW;
 *   prevDev  float[n*3]  the positions of the last frame written (may be
 *            null): host4[3] receives the largest |coordinate change| against
 *            it, and it takes the new positions when that change exceeds
 *            moveEps (every call for moveEps < 0), decided on the device;
 *   host4    double[4]   host memory: the sums of x, y, z over the positions
 *            (double, fixed order) and the largest change (0 without prevDev).
 * The caller's own work on these buffers must have completed before the call;
 * the library's work on them has completed when it returns.  Returns the
 * deformed count (0 without shapes) or -1. */
int xpbd_ext_frame_outputs(float* posDev, float* scaleDev, float* quatDev, int quatXYZW, float deformEps,
                           float* prevDev, float moveEps, double* host4, int n);

/* ---------------------------------------------------------------------------
 * Part 3: batched objects (no counterpart in the reference, whose state is one
 * global instance, native/xpbd_capi.cu:6, 17-29).  A batch holds many
 * independent objects in one device node space and steps all of them with one
 * launch sequence per substep, captured in a HIP graph.  Each object keeps
 * its own settings, colliders, attachments, clusters and hash table, and
 * advances exactly as it would alone (same launches per node, same summation
 * orders).  Objects of one batch share the iteration count.  Functions return
 * 0 / -1 (message in xpbd_batch_last_error) unless stated; `obj` < 0 in a
 * setter means every object.  Handles are independent: different batches may
 * be used from different threads at once; one batch from one thread at a time.
 * ------------------------------------------------------------------------- */
typedef struct xpbd_batch xpbd_batch;

/* count objects back to back: object k owns nodes nodeOff[k] .. nodeOff[k+1]
 * of pos[*3], scales[*3], rots[*4], opacity[*] and edges edgeOff[k] ..
 * edgeOff[k+1] of edges[*2], rest[*], stiff[*], whose ends are node ids local
 * to object k.  Every object starts at rest with the reference defaults.
 * Returns null on failure (message in xpbd_batch_last_error(null)). */
xpbd_batch* xpbd_batch_create(int count, const int* nodeOff, const float* pos, const float* scales, const float* rots,
                              const float* opacity, const int* edgeOff, const int* edges, const float* rest,
                              const float* stiff);
void xpbd_batch_destroy(xpbd_batch* b);
/* Last error of the batch; for null, the last failed xpbd_batch_create of the calling thread. */
const char* xpbd_batch_last_error(const xpbd_batch* b);
int xpbd_batch_objects(const xpbd_batch* b);
/* Nodes of object obj (obj < 0: of the whole batch). */
int xpbd_batch_count(const xpbd_batch* b, int obj);

/* Per-object settings with the clamps of the Part 1 setters. */
int xpbd_batch_set_solver(xpbd_batch* b, int obj, int iters, float dt, float underRelax, float velDamping,
                          float stiffnessScale, float invMassScale, float distCompliance, float shapeCompliance,
                          float shapeBlend, float angleCompliance, float angleBlend);
int xpbd_batch_set_constraints(xpbd_batch* b, int obj, int distance, int shape, int angle, int volume);
int xpbd_batch_set_volume(xpbd_batch* b, int obj, float compliance, int ringK, int maxMembers, int leaderMinHop);
int xpbd_batch_set_ground(xpbd_batch* b, int obj, int enabled, float ux, float uy, float uz, float height,
                          float friction, float restitution, float contactRadius, float contactSlop, float gravity);
int xpbd_batch_set_object_shape(xpbd_batch* b, int obj, float stiffness);
int xpbd_batch_set_self_collision(xpbd_batch* b, int obj, int enabled, float radiusScale, float excludeScale,
                                  int withinBody);
int xpbd_batch_set_self_collision_grid(xpbd_batch* b, int obj, float cellSize, int tableSize);
int xpbd_batch_set_gauss_seidel(xpbd_batch* b, int obj, int distance, int volume);
int xpbd_batch_set_shape_robust(xpbd_batch* b, int obj, int enabled);
/* Explicit volume clusters / shape clusters of one object (as the Part 2
 * functions, local ids; null returns to the reference rule). */
int xpbd_batch_set_volume_clusters(xpbd_batch* b, int obj, const int* off, const int* members, int minMembers);
int xpbd_batch_set_shape_clusters(xpbd_batch* b, int obj, const int* clusterOf, int n);
/* Colliders of one object (as xpbd_set_colliders). */
int xpbd_batch_set_colliders(xpbd_batch* b, int obj, int count, const int* types, const float* poses,
                             const float* dims, float margin, float friction);
/* Attachments of several objects in one call: request r pins nodes
 * idx[off[r] .. off[r+1]) (local ids of object objs[r]) to pos[3*k]; a request
 * with no nodes releases that object (as xpbd_set_attached). */
int xpbd_batch_set_attached(xpbd_batch* b, int requests, const int* objs, const int* off, const int* idx,
                            const float* pos, float weight);
int xpbd_batch_add_rigid_velocity(xpbd_batch* b, int obj, const float* dv, const float* dw, const float* center);
int xpbd_batch_reset(xpbd_batch* b, int obj);

/* Host-side set-up of the next substep (topologies, colourings, buffers, graph
 * capture) without stepping; safe to run on another thread than the one
 * stepping other batches. */
int xpbd_batch_prepare(xpbd_batch* b);
/* `substeps` substeps of every object; xpbd_batch_step returns after device
 * completion, xpbd_batch_step_async after queueing (xpbd_batch_sync waits). */
int xpbd_batch_step(xpbd_batch* b, int substeps);
int xpbd_batch_step_async(xpbd_batch* b, int substeps);
int xpbd_batch_sync(xpbd_batch* b);

/* Device pointer to float3 committed positions of every object in caller
 * order (object k from node nodeOff[k]); valid until destroy. */
unsigned long long xpbd_batch_positions_device(xpbd_batch* b);
int xpbd_batch_positions_host(xpbd_batch* b, int obj, float* out, int n);
int xpbd_batch_get_state(xpbd_batch* b, int obj, float* pos, float* vel, int n);
int xpbd_batch_set_state(xpbd_batch* b, int obj, const float* pos, const float* vel, int n);
/* Deformed shapes (as xpbd_compute_shapes) of object obj; obj < 0 updates
 * every object in one launch and leaves the results in device memory
 * (returns 0, see xpbd_batch_shapes_device). */
int xpbd_batch_compute_shapes(xpbd_batch* b, int obj, float deformEps, float* hostScales, float* hostQuats, int n);
/* ptrs[0] scales float[3*P], ptrs[1] quaternions float[4*P] (w,x,y,z), both in
 * device node order, ptrs[2] int[P]: the caller-order node index (nodeOff[k] +
 * local id) held by each device node, P = padded node count returned. */
int xpbd_batch_shapes_device(xpbd_batch* b, unsigned long long* ptrs);
int xpbd_batch_self_collision_stats(xpbd_batch* b, int obj, int* out);
int xpbd_batch_collider_stats(xpbd_batch* b, int obj, int* hits, float* push, float* centroid, int maxCount);

/* Continue object srcObj of batch src as object dstObj of batch dst (both
 * created from the same object data): settings, colliders, attachments,
 * clusters and dynamic state move over and the object advances bit for bit as
 * if it had stayed.  Used to regroup live objects into new batches. */
int xpbd_batch_copy_object(xpbd_batch* dst, int dstObj, xpbd_batch* src, int srcObj);

/* Launch-shape autotuner: times candidates of each tunable over `substeps`
 * substeps from the current state (restored afterwards) and keeps the
 * fastest; writes a JSON report.  Results do not depend on these choices.
 * tuning[7]: Jacobi tiles per block, volume mode (-1 auto, 0 thread, 1 wave),
 * cluster shape waves per block, self-collision build waves per block (1, 2,
 * 4), nodes per wave (4, 8, 16, 32, 64), XCD mapping (0 plain block order,
 * 1 one contiguous node run per XCD, 2-256 runs of that many blocks dealt
 * round-robin over the XCDs) and the fine candidate scan (1 one lane per
 * node, 0 one wave per node, -1 the default: 1).
 * flags bit 0 also picks the self-collision hash-table size
 * (1, 4 or 16 x nextpow2(2N) cells per object), which can reorder candidate
 * lists that overflow their 32 slots. */
int xpbd_batch_autotune(xpbd_batch* b, int substeps, int flags, char* json, int len);
int xpbd_batch_get_tuning(xpbd_batch* b, int* tuning);
int xpbd_batch_set_tuning(xpbd_batch* b, const int* tuning);
/* Per-kernel HIP-event timing of `substeps` eager substeps (JSON). */
int xpbd_batch_profile(xpbd_batch* b, int substeps, char* json, int len);

#ifdef __cplusplus
}
#endif

#endif /* XPBD_CAPI_H */
