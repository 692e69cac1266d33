/* CPU oracle for the MCUT intersection front end.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * algorithm, used as the checker for the HIP implementation.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never links or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the MCUT source tree).
 */
#ifndef MCUT_AMD_FRONTEND_ORACLE_H
#define MCUT_AMD_FRONTEND_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    const double* v;    /* xyz per vertex */
    const uint32_t* f;  /* concatenated face vertex indices */
    const uint32_t* sz; /* face sizes, NULL = all triangles */
    uint32_t nv, nf;
} om_mesh;

typedef struct {
    double gp_constant;     /* context general-position constant (frontend.h:596, default 1e-4) */
    int eps_absolute;       /* MC_DISPATCH_ENFORCE_GENERAL_POSITION_ABSOLUTE (preproc.cpp:2669) */
    int enforce_gp;         /* MC_DISPATCH_ENFORCE_GENERAL_POSITION[_ABSOLUTE] */
    int enforcement_count;  /* input_t::general_position_enforcement_count (kernel.cpp:3213) */
    int require_through_cuts;
    int has_perturbation;   /* cut mesh perturbed by `perturbation` (preproc.cpp:2650-2665) */
    double perturbation[3];
} om_options;

/* status_t values (kernel.h:51-86) */
enum {
    OM_SUCCESS = 0,
    OM_INVALID_SRC_MESH = -1,
    OM_INVALID_CUT_MESH = -2,
    OM_INVALID_MESH_INTERSECTION = -3,
    OM_GENERAL_POSITION_VIOLATION = -4,
    OM_UNSUPPORTED = -100 /* candidate face whose triplet choice cannot be replayed */
};

enum { OM_N_PAIRS = 0, OM_N_TESTS, OM_N_REGISTRY, OM_N_EDGES, OM_N_CAND_FACES };

typedef struct om_result om_result;

om_result* om_frontend(const om_mesh* src, const om_mesh* cut, const om_options* opt);
void om_free(om_result* r);
int om_status(const om_result* r);
int om_stage(const om_result* r); /* 0 faces, 1 narrow phase, 2 empty registry, 3 pair check, 4 done */
/* 1: the front end succeeded and its registry holds an intersection on a
 * cut-mesh border edge (kernel.cpp:2646-2650): a partial cut, which ends the
 * dispatch early under MC_DISPATCH_REQUIRE_THROUGH_CUTS (kernel.cpp:3269-3278) */
int om_partial(const om_result* r);
void om_shift(const om_result* r, double com[3], double t[3], double* eps);
int64_t om_count(const om_result* r, int what);
void om_copy_pairs(const om_result* r, uint32_t* pairs);
void om_copy_tests(const om_result* r, uint32_t* edge, uint32_t* face, int8_t* qsign, int8_t* rsign, char* type,
                   char* pip);
void om_copy_registry(const om_result* r, uint32_t* edge, uint32_t* face, double* points);
void om_copy_edges(const om_result* r, uint32_t* v0, uint32_t* v1, int32_t* f0, int32_t* f1);
void om_copy_planes(const om_result* r, uint32_t* face, double* normal, double* d, int32_t* k, int32_t* triple);
void om_copy_ps_vertices(const om_result* r, double* xyz);
double om_time_ms(const om_result* r);

/* building blocks, exported for direct checks */
double om_orient2d(const double* a, const double* b, const double* c);
int om_orient3d_sign(const double* a, const double* b, const double* c, const double* d);
int om_plane(const double* verts, int n, double* normal, double* d);
int om_best_triple(const double* verts, int n, const double* normal, int k, int* ijk);
/* libstdc++'s std::sort replayed over keys with the triplet ranking's
 * comparator (greater key first): order[i] = original position of the i-th */
void om_trip_sort_keys(const double* keys, long n, int* order);
char om_segment_type(const double* q, const double* r, const double* verts, int n, const int* ijk, int8_t* qs,
                     int8_t* rs);
void om_segment_point(const double* normal, double d, const double* q, const double* r, double* p);
char om_point_in_polygon(const double* p, const double* verts, int n, const double* normal, int k);
/* cut-path edges of the registry in the given order (kernel.cpp:3427-3615) */
int64_t om_cutpath(const om_result* r, const uint32_t* order, uint32_t* src, uint32_t* tgt, uint32_t* faces,
                   uint32_t* groups, uint8_t* kinds, int64_t cap, int* short_group);

#ifdef __cplusplus
}
#endif
#endif
