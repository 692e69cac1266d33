// The GPU intersection front end as it sits inside MCUT's dispatch(): the
// build's copy of source/kernel.cpp includes this header at file scope and
// frontend_block.inc in place of its front-end block (kernel.cpp:1779-3231,
// "Calculate polygon intersection points" through the empty-registry check).
#pragma once

#include <cstdint>
#include <map>
#include <unordered_map>
#include <utility>
#include <vector>

#include "mcut/internal/hmesh.h"
#include "mcut/internal/kernel.h"
#include "mcut/internal/math.h"
#include "mcut/internal/utils.h"

// Runs the front end on libmcgpu with what preproc handed dispatch() (the two
// meshes, their face AABBs and the candidate face pairs, input_t) and the
// polygon soup `ps` dispatch() built from them, and fills the containers the
// rest of dispatch() reads, as the replaced block fills them:
//   - the candidate faces' planes and vertices (kernel.cpp:2184-2356);
//   - the intersection points, appended to m0 in registry order, and the
//     registry with its side maps (kernel.cpp:2603-2654).
// Returns true when dispatch() must return at once with output.status set:
// a zero-area candidate face (kernel.cpp:2237-2331), a general-position
// violation in the narrow phase (kernel.cpp:2532-2604), or an empty registry
// (kernel.cpp:3211-3229).  A libmcgpu failure throws std::runtime_error.
bool mcgpu_seam_frontend(const input_t& input, output_t& output, const hmesh_t& sm, const hmesh_t& cs,
                         const hmesh_t& ps, hmesh_t& m0,
                         std::unordered_map<fd_t, vec3>& face_normal,
                         std::unordered_map<fd_t, scalar_t>& face_d,
                         std::unordered_map<fd_t, int>& face_max_comp,
                         std::unordered_map<fd_t, std::vector<vec3>>& face_vertices,
                         std::vector<std::pair<ed_t, fd_t>>& registry,
                         std::vector<vd_t>& cm_border_reentrant,
                         std::unordered_map<ed_t, std::vector<vd_t>>& intersecting_edges,
                         std::map<pair<fd_t>, std::vector<vd_t>>& cutpath_edges,
                         std::unordered_map<fd_t, std::vector<vd_t>>& face_ivertices,
                         bool& partial_cut);

// The cut-path edges libmcgpu made for the last mcgpu_seam_frontend call of
// this thread (kernel.cpp:3427-3615), in creation order: points as ranks of
// the registry (vertex ps_vtx_cnt + rank of m0), the two faces each edge is
// filed under, and its kind (0: a group of two points, whose faces' neighbours
// are updated too; 1: consecutive points of a longer group).  short_group: a
// group held a single point.
struct McgpuSeamCutpath {
    bool short_group = false;
    std::vector<uint32_t> src, tgt, fa, fb;
    std::vector<uint8_t> kind;
};
const McgpuSeamCutpath& mcgpu_seam_cutpath_result();
