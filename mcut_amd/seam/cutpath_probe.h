// Instrumentation of the MCUT builds under mcut_amd/seam: the cut-path edges
// a dispatch made (kernel.cpp:3427-3615), recorded right after MCUT makes them
// so that two builds (MCUT unchanged, MCUT with those edges from libmcgpu) can
// be compared edge for edge.  The copy of kernel.cpp calls the recorder; the
// tests read the record through the C functions of cutpath_probe.cpp.
#pragma once

#include <unordered_map>
#include <utility>
#include <vector>

#include "mcut/internal/hmesh.h"

void mcut_probe_record_cutpath(const hmesh_t& m0, int ps_vtx_cnt, const std::vector<std::pair<ed_t, fd_t>>& registry,
                               const std::vector<ed_t>& cutpath_edges,
                               const std::unordered_map<fd_t, std::vector<ed_t>>& face_edges);
