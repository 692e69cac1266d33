// Broad-phase seam for the MCUT library: the two functions preproc() calls
// for its broad phase, build_oibvh (bvh.cpp:219) and intersectOIBVHs
// (bvh.cpp:638), implemented over libmcgpu's GPU broad phase.
//
// build_oibvh keeps computing the per-face AABBs the rest of MCUT reads
// (kernel.cpp uses them for the edge-face cull), with the reference's own
// bounding_box_t::expand/enlarge, and hands them to intersectOIBVHs in place
// of a tree, behind the tree's root box (their union).  intersectOIBVHs sends the two sets of face boxes to
// mcgpu_broad_phase and fills ps_face_to_potentially_intersecting_others the
// way bvh.cpp:709-716 does: each pair in both directions, cut faces offset by
// the source face count.  The order inside each list differs from the CPU
// traversal's; MCUT sorts those lists before use (kernel.cpp:1801-1815).
//
// Build: compiled with MCUT's sources, bvh.cpp among them with these two
// functions renamed on its command line, and linked against libmcgpu.
#include "mcut/internal/bvh.h"

#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "mcgpu.h"

namespace {

// One libmcgpu context for the process.  MCUT runs every context's dispatches
// on that context's own API thread, so broad phases from several MCUT
// contexts take turns on it (a libmcgpu context is not thread-safe).
std::mutex seam_mutex;

mcgpu_context seam_context()
{
    static mcgpu_context ctx = nullptr;
    if (!ctx) {
        const mcgpu_result r = mcgpu_create_context(0, &ctx);
        if (r != MCGPU_SUCCESS) throw std::runtime_error(std::string("mcgpu_create_context: ") + mcgpu_result_string(r));
    }
    return ctx;
}

// boxes[first..] as six doubles each (minimum, then maximum)
void flatten(const std::vector<bounding_box_t<vec3_<double>>>& boxes, std::vector<double>& out, size_t first)
{
    const size_t n = boxes.size() > first ? boxes.size() - first : 0;
    out.resize(6 * n);
    for (size_t i = 0; i < n; ++i) {
        for (int j = 0; j < 3; ++j) {
            out[6 * i + j] = boxes[first + i].minimum()[j];
            out[6 * i + 3 + j] = boxes[first + i].maximum()[j];
        }
    }
}

} // namespace

void build_oibvh(
#if defined(MCUT_WITH_COMPUTE_HELPER_THREADPOOL)
    thread_pool& /*pool*/,
#endif
    const hmesh_t& mesh, std::vector<bounding_box_t<vec3_<double>>>& bvhAABBs, std::vector<fd_t>& bvhLeafNodeFaces,
    std::vector<bounding_box_t<vec3_<double>>>& face_bboxes, const double& slightEnlargmentEps, const double /*multiplier*/)
{
    const int nf = static_cast<int>(mesh.number_of_faces());
    face_bboxes.assign(nf, bounding_box_t<vec3_<double>>());
    for (face_array_iterator_t f = mesh.faces_begin(); f != mesh.faces_end(); ++f) {
        const int fi = static_cast<int>(*f);
        const std::vector<vd_t> vertices_on_face = mesh.get_vertices_around_face(*f);
        for (const vd_t& v : vertices_on_face) face_bboxes[fi].expand(mesh.vertex(v));
        if (slightEnlargmentEps > 0.0) face_bboxes[fi].enlarge(slightEnlargmentEps);
    }
    // Entry 0 is the root box, the union of the face boxes (what bvh.cpp's
    // internal levels leave there and preproc.cpp:2896 and :3721 read as the
    // mesh's box for the intersection type); the face boxes follow, which is
    // all the GPU broad phase needs.
    bvhAABBs.assign(1, bounding_box_t<vec3_<double>>());
    for (const auto& b : face_bboxes) bvhAABBs[0].expand(b);
    bvhAABBs.insert(bvhAABBs.end(), face_bboxes.begin(), face_bboxes.end());
    bvhLeafNodeFaces.resize(nf);
    for (int i = 0; i < nf; ++i) bvhLeafNodeFaces[i] = fd_t(i);
}

void intersectOIBVHs(std::map<fd_t, std::vector<fd_t>>& ps_face_to_potentially_intersecting_others,
                     const std::vector<bounding_box_t<vec3_<double>>>& srcMeshBvhAABBs,
                     const std::vector<fd_t>& srcMeshBvhLeafNodeFaces,
                     const std::vector<bounding_box_t<vec3_<double>>>& cutMeshBvhAABBs,
                     const std::vector<fd_t>& cutMeshBvhLeafNodeFaces)
{
    const uint32_t ns = static_cast<uint32_t>(srcMeshBvhLeafNodeFaces.size());
    const uint32_t nc = static_cast<uint32_t>(cutMeshBvhLeafNodeFaces.size());
    std::vector<double> a, b;
    flatten(srcMeshBvhAABBs, a, 1); // the face boxes, past the root box
    flatten(cutMeshBvhAABBs, b, 1);
    std::lock_guard<std::mutex> lock(seam_mutex);
    mcgpu_context ctx = seam_context();
    uint64_t n = 0;
    mcgpu_result r = mcgpu_broad_phase(ctx, a.data(), ns, b.data(), nc, 0, &n);
    if (r != MCGPU_SUCCESS) throw std::runtime_error(std::string("mcgpu_broad_phase: ") + mcgpu_result_string(r));
    std::vector<uint32_t> pairs(2 * n);
    if (n && (r = mcgpu_get_candidate_pairs(ctx, pairs.data(), n)) != MCGPU_SUCCESS)
        throw std::runtime_error(std::string("mcgpu_get_candidate_pairs: ") + mcgpu_result_string(r));
    for (uint64_t i = 0; i < n; ++i) {
        const fd_t sf(pairs[2 * i]);
        const fd_t cf(pairs[2 * i + 1] + ns);
        ps_face_to_potentially_intersecting_others[sf].push_back(cf);
        ps_face_to_potentially_intersecting_others[cf].push_back(sf);
    }
}
