/*
 * mcgpu -- MI355X intersection front end for MCUT (C ABI).
 *
 * This library replaces MCUT's intersection front end: the broad phase of
 * source/bvh.cpp and the narrow phase at the head of dispatch() in
 * source/kernel.cpp.  Given a source mesh and a cut mesh exactly as
 * mcDispatch receives them (include/mcut/mcut.h:1008-1020), it computes on the
 * GPU:
 *   - the coordinate shift of preproc.cpp:2124-2290 (vertices moved into the
 *     positive quadrant; all outputs are in that shifted space, as in MCUT);
 *   - the polygon-soup numbering of edges and halfedges (hmesh.cpp:406-733,
 *     kernel.cpp:1593-1732);
 *   - the candidate face pairs whose AABBs overlap (bvh.cpp:219-783);
 *   - the edge-face tests after expansion and culling (kernel.cpp:1779-2177);
 *   - the plane of every candidate face (kernel.cpp:2184-2356);
 *   - the edge-face narrow phase with exact orient3d signs and the
 *     intersection registry with its points (kernel.cpp:2415-2878), sorted by
 *     (edge, face);
 *   - the dispatch status the reference derives from these
 *     (kernel.cpp:2237-2331, 2532-2604, 3211-3229, 3439-3452).
 *
 * Plain C types only.  Every function returns an mcgpu_result; no exception
 * crosses this boundary.  A context owns one HIP stream and its device
 * buffers; it must not be used from two threads at once.
 *
 * The front end is a superset of what MCUT computes in one place: its broad
 * phase runs on the unshifted coordinates with a rounding bound, overlapped
 * with the coordinate sums, and an exact test on the shifted boxes then leaves
 * exactly MCUT's candidate pairs.  Inputs whose bound is not finite
 * (coordinates near the double range) return MCGPU_ERROR_UNSUPPORTED.
 */
#ifndef MCGPU_H_
#define MCGPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define MCGPU_API __attribute__((visibility("default")))
#else
#define MCGPU_API
#endif

typedef struct mcgpu_context_t* mcgpu_context;

/* Return codes of every entry point (the ABI's McResult, mcut.h:202-208). */
typedef enum {
    MCGPU_SUCCESS = 0,
    MCGPU_ERROR_INVALID_VALUE = -1, /* bad argument: null pointer, size 0, face size < 3, index out of range */
    MCGPU_ERROR_DEVICE = -2,        /* HIP runtime error */
    MCGPU_ERROR_OUT_OF_MEMORY = -3,
    MCGPU_ERROR_UNSUPPORTED = -4,   /* candidate polygon with more than 65535 vertices; coordinates too large to shift */
    MCGPU_ERROR_NO_RESULT = -5      /* query before a successful dispatch */
} mcgpu_result;

/* Front-end status: the values of MCUT's status_t (kernel.h:51-86). */
typedef enum {
    MCGPU_STATUS_SUCCESS = 0,
    MCGPU_STATUS_INVALID_SRC_MESH = -1,
    MCGPU_STATUS_INVALID_CUT_MESH = -2,
    MCGPU_STATUS_INVALID_MESH_INTERSECTION = -3,
    MCGPU_STATUS_GENERAL_POSITION_VIOLATION = -4
} mcgpu_status;

/* MCUT dispatch flags the front end reads (mcut.h:379-441). */
#define MCGPU_DISPATCH_REQUIRE_THROUGH_CUTS (1u << 2)
#define MCGPU_DISPATCH_ENFORCE_GENERAL_POSITION (1u << 15)
#define MCGPU_DISPATCH_ENFORCE_GENERAL_POSITION_ABSOLUTE (1u << 16)
/* libmcgpu's own: refuse (MCGPU_ERROR_INVALID_VALUE) an input mesh whose
 * vertex graph has more than one connected component, as MCUT's
 * check_input_mesh refuses it with MC_INVALID_VALUE (preproc.cpp:505-578).
 * MCUT makes that check on the host before its front end, so a caller inside
 * MCUT has it already; a standalone caller asks for it with this flag. */
#define MCGPU_DISPATCH_CHECK_COMPONENTS (1u << 30)

/* A mesh as mcDispatch takes it: xyz doubles per vertex, concatenated face
 * vertex indices, and face sizes (NULL = every face is a triangle,
 * preproc.cpp:206).  Pointers are host memory unless
 * mcgpu_dispatch_desc.input_on_device is set. */
typedef struct {
    const double* vertices;
    const uint32_t* face_indices;
    const uint32_t* face_sizes;
    uint32_t num_vertices;
    uint32_t num_faces;
} mcgpu_mesh;

typedef struct {
    uint32_t flags;              /* MCGPU_DISPATCH_* bits, same values as MC_DISPATCH_* */
    double gp_constant;          /* MC_CONTEXT_GENERAL_POSITION_ENFORCEMENT_CONSTANT (default 1e-4) */
    uint32_t enforcement_count;  /* perturbation attempts made so far (input_t, kernel.h:125-172) */
    uint32_t input_on_device;    /* 1: mesh pointers are device pointers already resident in HBM */
    uint32_t keep_tests;         /* 1: keep per-test signs/types for mcgpu_get_tests */
    uint32_t reserved;
} mcgpu_dispatch_desc;

typedef struct {
    int32_t status;              /* mcgpu_status */
    uint32_t gp_violation;       /* narrow phase saw a touching or edge/vertex hit */
    uint64_t num_candidate_pairs;
    uint64_t num_candidate_faces;
    uint64_t num_edges;          /* polygon-soup edges */
    uint64_t num_tests;          /* edge-face tests after expansion, dedup and culling */
    uint64_t num_escalated;      /* tests whose orient3d signs needed exact arithmetic */
    uint64_t num_registry;       /* intersecting (edge, face) pairs */
    uint64_t num_raw_pairs;      /* pairs of the raw-coordinate broad phase before the exact shifted-box test */
    uint64_t num_decided_f32;    /* tests whose two orient3d signs the FP32 pre-filter settled */
    uint64_t num_decided_f64;    /* tests settled by the FP64 semi-static filter (num_escalated: by exact arithmetic) */
    double com[3];               /* shift: x' = (x - com) + translation */
    double translation[3];
    double eps;                  /* cut-face AABB enlargement */
    float device_ms;             /* GPU time of the whole front end */
    uint32_t partial_cut;        /* 1: the front end succeeded and an intersection lies on a cut-mesh border
                                    edge (kernel.cpp:2646-2650): a partial cut.  With
                                    MCGPU_DISPATCH_REQUIRE_THROUGH_CUTS the reference ends the dispatch there
                                    with success and no cut (kernel.cpp:3269-3278) */
} mcgpu_frontend_info;

/* A batch of independent (source, cut) pairs, each dispatched as
 * mcgpu_dispatch_frontend would dispatch it alone.  `src` holds every pair's
 * source mesh concatenated in pair order and `cut` every cut mesh; face
 * indices stay local to their own mesh (as each pair's mcDispatch would take
 * them), num_vertices / num_faces are the totals.  The offset arrays are
 * host memory with num_pairs + 1 entries: pair p's source vertices are
 * [src_vertex_offsets[p], src_vertex_offsets[p + 1]), and likewise for faces
 * and the cut meshes.  Every mesh needs at least 3 vertices and one face. */
typedef struct {
    uint32_t num_pairs;
    mcgpu_mesh src;
    mcgpu_mesh cut;
    const uint32_t* src_vertex_offsets;
    const uint32_t* src_face_offsets;
    const uint32_t* cut_vertex_offsets;
    const uint32_t* cut_face_offsets;
} mcgpu_batch;

/* --- context (one per MCUT context / API thread) ---------------------------- */

/* mcCreateContext (mcut.h:567) counterpart: binds a device and creates the stream. */
MCGPU_API mcgpu_result mcgpu_create_context(int device, mcgpu_context* out);
/* mcReleaseContext (mcut.h:1377) counterpart. */
MCGPU_API mcgpu_result mcgpu_release_context(mcgpu_context ctx);
MCGPU_API const char* mcgpu_result_string(mcgpu_result r);

/* --- front end ---------------------------------------------------------------
 * Replaces, in one call, preproc.cpp:2304-2902 (shift, BVH builds
 * bvh.cpp:219, traversal bvh.cpp:638) and kernel.cpp:1779-3229 (edge-face
 * pairs, cull, face planes, narrow phase, registry).  Returns after the
 * front end has finished; results stay on the device for the getters below. */
MCGPU_API mcgpu_result mcgpu_dispatch_frontend(mcgpu_context ctx, const mcgpu_mesh* src, const mcgpu_mesh* cut,
                                               const mcgpu_dispatch_desc* desc, mcgpu_frontend_info* info);

/* The front-end block of dispatch() alone (kernel.cpp:1779-3229), on the
 * inputs preproc hands it (input_t, kernel.h:125-172): the two meshes with the
 * coordinates dispatch() holds (already shifted, and perturbed in a
 * general-position retry; they are used as given), face lists in input order,
 * the face AABBs (6 doubles per face, min xyz then max xyz, cut faces
 * enlarged: source_hmesh_face_aabb_array / cut_hmesh_face_aabb_array) and the
 * candidate pairs (`pairs`: num_pairs (source face, cut face) pairs, cut faces
 * numbered from 0: ps_face_to_potentially_intersecting_others).  It numbers
 * the polygon-soup edges, expands and culls the edge-face tests, computes the
 * candidate-face planes and runs the narrow phase into the sorted registry,
 * as mcgpu_dispatch_frontend does after its broad phase; the getters then read
 * its results (com and translation are zero).  `topology` (optional) is the
 * soup's edge numbering as the caller's halfedge mesh has it (hmesh_t edge ids
 * and halfedge-0 directions, which a mesh edited in place -- preproc's
 * floating-polygon partitioning -- no longer creates in face order); without
 * it the edges are numbered as hmesh_t::add_face numbers them for faces added
 * in order.  Host arrays only; desc->flags and enforcement_count select the
 * status rules. */
typedef struct {
    /* per mesh, per corner of its face lists as given: the soup edge from that
     * corner's vertex to the next one of its face */
    const uint32_t* corner_edges[2];
    /* per soup edge: source and target of halfedge 0 (soup vertex ids: source
     * vertices, then cut vertices), faces of halfedges 0 and 1 (-1: none) */
    const uint32_t* edge_v0;
    const uint32_t* edge_v1;
    const int32_t* edge_f0;
    const int32_t* edge_f1;
    uint32_t num_edges;
    uint32_t reserved;
} mcgpu_soup_topology;

MCGPU_API mcgpu_result mcgpu_dispatch_narrow(mcgpu_context ctx, const mcgpu_mesh* src, const mcgpu_mesh* cut,
                                             const double* src_boxes, const double* cut_boxes, const uint32_t* pairs,
                                             uint64_t num_pairs, const mcgpu_soup_topology* topology,
                                             const mcgpu_dispatch_desc* desc, mcgpu_frontend_info* info);

/* Re-runs the narrow phase with the cut mesh moved by `perturbation` (3
 * doubles), keeping the candidate pairs and cut-face AABBs of the last
 * mcgpu_dispatch_frontend, as the reference's general-position retry does
 * (preproc.cpp:2590-2698, 2826). */
MCGPU_API mcgpu_result mcgpu_redispatch_perturbed(mcgpu_context ctx, const double* perturbation,
                                                  uint32_t enforcement_count, mcgpu_frontend_info* info);

/* mcDispatch's general-position loop (preproc.cpp:2562-2631) around the front
 * end: dispatch; then, while the status is a general-position violation, an
 * enforce flag is set and fewer than `attempts` perturbations were made
 * (the context setting MC_CONTEXT_GENERAL_POSITION_ENFORCEMENT_ATTEMPTS,
 * default 4), draw the next perturbation exactly as the reference draws it
 * (per context, the sequence continuing across dispatches like the
 * reference's thread-local generator) and rerun the narrow phase with the
 * cut mesh moved by it.  A violation left after `attempts` perturbations is
 * where the reference throws "max perturbation iterations reached".
 * perturbation (3 doubles) and perturbations (count) are optional outputs. */
MCGPU_API mcgpu_result mcgpu_dispatch_enforced(mcgpu_context ctx, const mcgpu_mesh* src, const mcgpu_mesh* cut,
                                               const mcgpu_dispatch_desc* desc, uint32_t attempts,
                                               mcgpu_frontend_info* info, double* perturbation, uint32_t* perturbations);

/* Broad phase alone, at the seam of the reference's build_oibvh (bvh.cpp:219)
 * and intersectOIBVHs (bvh.cpp:638): given the face AABBs of both meshes
 * (6 doubles per face: min xyz, max xyz, as build_oibvh's face_bboxes), finds
 * every (source face, cut face) pair whose boxes overlap, with the inclusive
 * test of intersect_bounding_boxes (math.h:931-941).  The set is the one
 * intersectOIBVHs reports; read it with mcgpu_get_candidate_pairs.  Boxes are
 * device pointers when input_on_device is 1.  After this call only the
 * candidate pairs are available from ctx. */
MCGPU_API mcgpu_result mcgpu_broad_phase(mcgpu_context ctx, const double* src_boxes, uint32_t num_src,
                                         const double* cut_boxes, uint32_t num_cut, uint32_t input_on_device,
                                         uint64_t* num_pairs);

/* Broad phase without device allocation.  No counterpart in the reference:
 * it is mcgpu_broad_phase for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc / hipFree: every array it needs lives in
 * a static device array of MCGPU_NOALLOC_SCRATCH_BYTES (50.1 MiB), which each
 * device that runs it holds until the process ends.  Limits: at most
 * MCGPU_NOALLOC_MAX_FACES faces per mesh (larger inputs return
 * MCGPU_ERROR_INVALID_VALUE before any device work) and at most
 * MCGPU_NOALLOC_MAX_PAIRS pairs of the raw-coordinate broad phase (more
 * return MCGPU_ERROR_OUT_OF_MEMORY with *num_pairs = that count).
 * src_boxes / cut_boxes: device pointers, 6 doubles per face as in
 * mcgpu_broad_phase.  pairs: device array of `capacity` entries that receives
 * the overlapping pairs, unsorted, each (source face << 32) | cut face; a
 * capacity below the pair count returns MCGPU_ERROR_OUT_OF_MEMORY with
 * *num_pairs = the count.  stream: a hipStream_t (NULL: the null stream); the
 * call returns once the pairs are written.  All calls on one device share its
 * array, so calls that overlap in time on one device race. */
#define MCGPU_NOALLOC_MAX_FACES 65536u
#define MCGPU_NOALLOC_MAX_PAIRS (1u << 20)
#define MCGPU_NOALLOC_SCRATCH_BYTES 52566528ull
MCGPU_API mcgpu_result mcgpu_broad_phase_noalloc(int device, const double* src_boxes, uint32_t num_src,
                                                 const double* cut_boxes, uint32_t num_cut, uint64_t* pairs,
                                                 uint64_t capacity, uint64_t* num_pairs, void* stream);

/* Batched front end: every pair of `batch` through the front end in one
 * pass of the pipeline (each stage one launch over all pairs), with results
 * bit-identical to dispatching each pair alone.  This has no single
 * counterpart in the reference, whose mcDispatch (mcut.h:988-1020) takes
 * one pair per call: it replaces a loop of those calls' front ends.  The
 * general-position retry is per pair with mcgpu_dispatch_frontend and
 * mcgpu_redispatch_perturbed.  info holds totals over the batch, its status
 * is the first pair's status that is not SUCCESS (com, translation and eps are
 * zero); per-pair results come from the getters below. */
MCGPU_API mcgpu_result mcgpu_dispatch_batch(mcgpu_context ctx, const mcgpu_batch* batch, const mcgpu_dispatch_desc* desc,
                                            mcgpu_frontend_info* info);
/* After mcgpu_dispatch_batch: each pair's status (mcgpu_status) and its
 * entries [registry_offsets[p], registry_offsets[p + 1]) of the batch
 * registry.  capacity >= num_pairs; registry_offsets needs num_pairs + 1. */
MCGPU_API mcgpu_result mcgpu_get_batch_status(mcgpu_context ctx, int32_t* status, uint64_t* registry_offsets,
                                              uint64_t capacity);
/* After mcgpu_dispatch_batch: each pair's mcgpu_frontend_info.partial_cut
 * (capacity >= num_pairs). */
MCGPU_API mcgpu_result mcgpu_get_batch_partial_cuts(mcgpu_context ctx, uint32_t* partial_cut, uint64_t capacity);
/* After mcgpu_dispatch_batch: candidate pairs as (pair, source face, cut
 * face) triples, face ids local to the pair, sorted. */
MCGPU_API mcgpu_result mcgpu_get_batch_candidate_pairs(mcgpu_context ctx, uint32_t* triples, uint64_t capacity);

/* --- serving ------------------------------------------------------------------
 * One resident source mesh cut by requests as they arrive.  The server runs
 * the source through the front end once and keeps every part of it that no
 * request changes -- coordinate sums, raw box, polygon-soup edges, raw-space
 * leaf boxes and BVH -- resident in HBM.  Each request is then only its cut
 * mesh: its sums, shift, cut-mesh edges and BVH, the traversal against the
 * resident source tree, and the narrow phase.  Requests run in slots (one
 * HIP stream and buffers sized once for the largest request each), so
 * requests that arrive while others run overlap on the device.  No reference
 * counterpart: it replaces a host loop of mcDispatch calls with one source
 * mesh; every request's status and registry are those of mcgpu_dispatch_frontend
 * on (source, cut). */
typedef struct mcgpu_server_t* mcgpu_server;

typedef struct {
    uint32_t max_cut_vertices;   /* largest request a slot holds */
    uint32_t max_cut_faces;
    uint32_t max_cut_corners;    /* sum of face sizes (0: 3 * max_cut_faces) */
    uint32_t max_slots;          /* requests in flight at most */
    uint64_t hbm_budget;         /* bytes the in-flight requests' footprints may add up to */
} mcgpu_server_limits;

typedef struct {
    double arrival_s;            /* from the schedule */
    double start_s;              /* admitted to a slot (seconds since the run began) */
    double done_s;               /* results read back */
    int32_t status;              /* mcgpu_status, or a negative mcgpu_result */
    uint32_t partial_cut;        /* as mcgpu_frontend_info.partial_cut */
    uint64_t num_registry;
    uint64_t num_candidate_pairs;
} mcgpu_request_result;

/* desc: flags, gp_constant and input_on_device for the source and every
 * request (device pointers must outlive the server / the run). */
MCGPU_API mcgpu_result mcgpu_server_create(int device, const mcgpu_mesh* src, const mcgpu_dispatch_desc* desc,
                                           const mcgpu_server_limits* limits, mcgpu_server* out);
MCGPU_API mcgpu_result mcgpu_server_release(mcgpu_server srv);
MCGPU_API mcgpu_result mcgpu_server_info(mcgpu_server srv, uint32_t* slots, uint64_t* slot_bytes, uint32_t* source_edges);
/* Runs a schedule to completion: request i (cuts[i]) arrives arrival_s[i]
 * seconds after the call.  Arrived requests are admitted oldest first while a
 * slot is free and the sum of in-flight footprints (an estimate from each
 * request's face, corner and vertex counts) stays within hbm_budget; an idle
 * server always admits.  Completion is polled; out[i] gets its times, status
 * and counts.  keep_registries: copy each registry to the host for
 * mcgpu_server_get_registry. */
MCGPU_API mcgpu_result mcgpu_server_run(mcgpu_server srv, const mcgpu_mesh* cuts, const double* arrival_s, uint32_t n,
                                        uint32_t keep_registries, mcgpu_request_result* out);
MCGPU_API mcgpu_result mcgpu_server_get_registry(mcgpu_server srv, uint32_t request, uint32_t* edges, uint32_t* faces,
                                                 double* points, uint64_t capacity);

/* --- results of the last dispatch (host copies; capacity in elements) -----
 * After mcgpu_dispatch_batch, mcgpu_get_registry and mcgpu_registry_device
 * give the batch registry: pairs in order, each pair's entries with the
 * edge and face ids of that pair's own polygon soup; the other getters
 * describe the batch's concatenated soup (every source mesh, then every cut
 * mesh). */

/* Intersection registry sorted by (edge, face): m0_ivtx_to_intersection_registry_entry
 * and the intersection points (kernel.cpp:2603-2615).  points: 3 doubles per entry. */
MCGPU_API mcgpu_result mcgpu_get_registry(mcgpu_context ctx, uint32_t* edges, uint32_t* faces, double* points,
                                          uint64_t capacity);
/* Candidate (src face, cut face) pairs, sorted: the contents of
 * ps_face_to_potentially_intersecting_others (bvh.cpp:709-716) without the mirror. */
MCGPU_API mcgpu_result mcgpu_get_candidate_pairs(mcgpu_context ctx, uint32_t* pairs, uint64_t capacity);
/* Polygon-soup edges: h0 = v0 -> v1, face of h0 / h1 (-1 = border). */
MCGPU_API mcgpu_result mcgpu_get_edges(mcgpu_context ctx, uint32_t* v0, uint32_t* v1, int32_t* f0, int32_t* f1,
                                       uint64_t capacity);
/* Edge-face tests sorted by (edge, face) with orient3d signs of q and r, the
 * segment/plane type ('0','1','p','q','r') and the last point-in-polygon
 * result ('-' when none).  Needs desc.keep_tests. */
MCGPU_API mcgpu_result mcgpu_get_tests(mcgpu_context ctx, uint32_t* edges, uint32_t* faces, int8_t* qsign,
                                       int8_t* rsign, char* type, char* pip, uint64_t capacity);
/* Planes of the candidate faces sorted by face: unit normal, d, largest
 * component, the three vertices used by orient3d (kernel.cpp:2188-2191). */
MCGPU_API mcgpu_result mcgpu_get_face_planes(mcgpu_context ctx, uint32_t* faces, double* normals, double* d,
                                             int32_t* max_comp, int32_t* triple, uint64_t capacity);
/* Shifted polygon-soup vertex coordinates (source vertices, then cut vertices). */
MCGPU_API mcgpu_result mcgpu_get_ps_vertices(mcgpu_context ctx, double* xyz, uint64_t capacity);

/* Device pointers to the sorted registry (valid until the next dispatch on ctx),
 * for consumers that stay on the GPU. */
MCGPU_API mcgpu_result mcgpu_registry_device(mcgpu_context ctx, const uint32_t** edges, const uint32_t** faces,
                                             const double** points, uint64_t* count);

/* --- measurement -------------------------------------------------------------- */

/* The predicate stage alone, over the edge-face tests of the last
 * mcgpu_dispatch_frontend: both orient3d signs of every test (its edge's
 * endpoints against its face's plane triplet, or the face's first three
 * vertices where the plane has none) through the staged filter -- the FP32
 * pre-filter (error bound 16 * 2^-24 * permanent), the FP64 semi-static
 * filter ((7 + 56 eps) eps * permanent, shewchuk.c:424) and exact expansion
 * arithmetic.  counts[6] = decided in FP32, in FP64, exactly; signs negative,
 * zero, positive (all over 2 * num_tests predicates).  qsign / rsign
 * (optional) get the signs in mcgpu_get_tests order.  device_ms: the
 * kernel's time (this call replaces the stage times of the dispatch). */
MCGPU_API mcgpu_result mcgpu_predicate_stage(mcgpu_context ctx, uint64_t* counts, int8_t* qsign, int8_t* rsign,
                                             uint64_t capacity, float* device_ms);

/* Per-stage device times of the last dispatch in ms (stage names via
 * mcgpu_stage_name); returns the number of stages written, 0 when that
 * dispatch ran without stage timing. */
MCGPU_API int mcgpu_stage_times(mcgpu_context ctx, float* ms, int capacity);

/* Record the per-stage events in this context's later dispatches (off by
 * default: thirteen recorded events cost about 50 us a dispatch).  enable = 2
 * records only the traversal kernel's own dispatch timestamps (stage
 * "traversal_kernel"; the other stages read 0), cheap enough for timed runs.
 * No counterpart in the reference; for profiling callers. */
MCGPU_API mcgpu_result mcgpu_set_stage_timing(mcgpu_context ctx, int enable);
MCGPU_API const char* mcgpu_stage_name(int stage);

/* The dispatch of mcgpu_dispatch_frontend, enqueued without waiting for it
 * (mcEnqueueDispatch's contract, frontend.cpp:690-706: the call returns once
 * the work is queued and the caller waits for its event).  stream (a
 * hipStream_t, or NULL): the dispatch starts after the work queued so far on
 * it, so device inputs written there are ready.  Host (and device) inputs
 * must stay valid until mcgpu_dispatch_wait returns: a dispatch whose lists
 * overflow is re-run from them there.  Until then, the context takes no other
 * dispatch (MCGPU_ERROR_INVALID_VALUE) and its getters have no result. */
MCGPU_API mcgpu_result mcgpu_dispatch_frontend_async(mcgpu_context ctx, const mcgpu_mesh* src, const mcgpu_mesh* cut,
                                                     const mcgpu_dispatch_desc* desc, void* stream);
/* 1 when the enqueued dispatch has finished on the device (or none is
 * pending), 0 while it runs; never blocks.  A list that overflowed is only
 * seen, and the dispatch re-run with grown lists, by mcgpu_dispatch_wait. */
MCGPU_API int mcgpu_dispatch_query(mcgpu_context ctx);
/* Waits for the enqueued dispatch and completes it as mcgpu_dispatch_frontend
 * would (info filled, results readable).  stream (or NULL): made to wait for
 * the dispatch's end, so work queued there afterwards sees the registry.
 * MCGPU_ERROR_NO_RESULT when nothing was enqueued. */
MCGPU_API mcgpu_result mcgpu_dispatch_wait(mcgpu_context ctx, mcgpu_frontend_info* info, void* stream);

/* Cut-path edges of the last successful mcgpu_dispatch_frontend or
 * mcgpu_dispatch_narrow (kernel.cpp:3427-3615, the first consumer of the
 * registry): the edges between intersection points along which each
 * (source face, cut face) pair cuts, in the order MCUT creates them.
 * order: NULL (registry order) or a permutation of the registry giving the
 * caller's intersection-point order (order[k] = registry index of point k;
 * mcDispatch numbers its points in its own order); points and edges are then
 * named by rank k.  short_group = 1 when a group holds a single point: MCUT
 * then stops the dispatch with a general-position violation when it enforces
 * general position, else with an invalid mesh intersection
 * (kernel.cpp:3445-3459). */
typedef struct mcgpu_cutpath_info {
    uint64_t num_edges;      /* cut-path edges made */
    uint64_t num_groups;     /* (source face, cut face) groups */
    uint64_t num_candidates; /* consecutive point pairs examined */
    uint32_t short_group;    /* a group of one point */
    uint32_t unsupported;    /* a midpoint test named a face without a plane */
    float device_ms;
    float reserved;
} mcgpu_cutpath_info;
MCGPU_API mcgpu_result mcgpu_cutpath_edges(mcgpu_context ctx, const uint32_t* order, mcgpu_cutpath_info* info);
/* The edges in creation order: source and target point (ranks), the two faces
 * the edge is filed under (a two-point group's source and cut face, else the
 * two faces whose area holds its midpoint; polygon-soup face ids), the group
 * (in (source face, cut face) order) and kind (0: group of two points, 1:
 * consecutive points of a longer group).  Any output may be NULL. */
MCGPU_API mcgpu_result mcgpu_get_cutpath_edges(mcgpu_context ctx, uint32_t* src, uint32_t* tgt, uint32_t* faces,
                                               uint32_t* groups, uint8_t* kinds, uint64_t capacity);

/* Graph replay of dispatches.  No counterpart in the reference (a CPU
 * pipeline); for callers that dispatch the same shape repeatedly (the same
 * sizes and input buffers, new coordinates or connectivity behind them).
 * enable != 0: mcgpu_dispatch_frontend and mcgpu_dispatch_batch record their
 * launches once per dispatch shape as HIP graphs, one per stream segment
 * between the cross-stream event records and waits, and replay them (each on
 * the stream, and CU mask, it was captured from) while the shape stays the
 * same; a different shape (sizes, capacities, flags, traversal shape, buffer
 * addresses) records anew.  Dispatches with stage timing on run eagerly.
 * MCGPU_ERROR_UNSUPPORTED on a serving slot's context. */
MCGPU_API mcgpu_result mcgpu_set_graphs(mcgpu_context ctx, int enable);
/* programs recorded and replayed so far, and the graphs of the current one */
MCGPU_API mcgpu_result mcgpu_graph_stats(mcgpu_context ctx, uint64_t* recordings, uint64_t* replays,
                                         uint32_t* graph_launches);

/* Launch shape of the broad-phase traversal of a dispatch.  No counterpart
 * in the reference, whose broad phase is a recursive CPU walk
 * (bvh.cpp:638-783): these are the knobs of the GPU traversal that the
 * autotuner picks per workload class (single pairs or batches, by the bit
 * lengths of the two face counts). */
typedef struct mcgpu_tuning {
    uint32_t wave_leaves;       /* source leaves per traversal block: 8, 16 or 32 */
    uint32_t tree_stride;       /* cut-tree levels per breadth-first step: 2, 3 or 4 */
    uint32_t queues;            /* 0: one wave per block; 1: the same, each XCD's blocks in runs of
                                   `chunk` Morton-consecutive ones; 8: persistent waves over 8 per-XCD
                                   work queues */
    uint32_t chunk;             /* blocks a persistent wave takes per queue pop; the run length with
                                   queues 1 */
    uint32_t broad_share_pct;   /* percent of the CUs outside the sums' reservation that the
                                   broad-phase stream holds (reported; 0 in mcgpu_set_tuning) */
    uint32_t candidates;        /* traversal shapes the autotuner timed (0: defaults, not tuned) */
    float traversal_ms;         /* kernel time of the chosen traversal on the tuned dispatch */
    float default_traversal_ms; /* kernel time of the default traversal on it */
    float dispatch_ms;          /* device time of the dispatch with the chosen shape */
    float default_dispatch_ms;  /* device time of the dispatch with the default shape */
} mcgpu_tuning;

/* Tunes the traversal shape for the workload class of the last successful
 * mcgpu_dispatch_frontend or mcgpu_dispatch_batch (its inputs must still be
 * valid): 36 shapes timed alone on its trees, three runs each, each required
 * to find the same raw pairs; then the dispatch re-run twice with the default
 * and twice with the chosen shape for the report.  Later dispatches of the
 * class use the choice.  The last dispatch's outputs stay valid. */
MCGPU_API mcgpu_result mcgpu_autotune(mcgpu_context ctx, mcgpu_tuning* out);
/* The CU split of a context's batch streams: broad_share_pct percent (10-90)
 * of the CUs outside the sums' reservation for the broad phase, the rest for
 * the edge numbering (default 58).  The streams are made with their masks by
 * the context's first batch, so this is accepted only before it
 * (MCGPU_ERROR_INVALID_VALUE after). */
MCGPU_API mcgpu_result mcgpu_set_batch_split(mcgpu_context ctx, uint32_t broad_share_pct);
/* enable != 0: the first dispatch of each untuned class tunes it (above) */
MCGPU_API mcgpu_result mcgpu_set_autotune(mcgpu_context ctx, int enable);
/* Forces a traversal shape for every later dispatch (t = NULL: back to the
 * tuned or default choice).  MCGPU_ERROR_INVALID_VALUE for a shape outside
 * the lists above or a nonzero broad_share_pct. */
MCGPU_API mcgpu_result mcgpu_set_tuning(mcgpu_context ctx, const mcgpu_tuning* t);
/* the shape the last dispatch ran with (and its tuning report) */
MCGPU_API mcgpu_result mcgpu_get_tuning(mcgpu_context ctx, mcgpu_tuning* out);

#ifdef __cplusplus
}
#endif
#endif /* MCGPU_H_ */
