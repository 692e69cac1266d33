"""Build step of the seam builds: writes patched copies of two of MCUT's
source files into mcut_amd/seam/_build/src/ (build products, kept out of
version control and off the GPU runs; the sources under the reference tree are
only read).

kernel_gpu.cpp  source/kernel.cpp with its intersection front end -- from the
                declaration of ps_edge_face_intersection_pairs ("Calculate
                polygon intersection points", kernel.cpp:1779) through the
                TIMESTACK_POP() after the empty-registry check (kernel.cpp:3231)
                -- replaced by `#include "frontend_block.inc"`, its cut-path
                edge loop (kernel.cpp:3427-3615) by
                `#include "cutpath_block.inc"`, the cut-path probe's recorder
                called after that loop, and frontend_seam.h and
                cutpath_probe.h included after its own includes.
kernel_probe.cpp  source/kernel.cpp with only the recorder call and its
                include: MCUT unchanged, its cut-path edges recorded.
preproc_quiet.cpp  source/preproc.cpp with `kernel_input.verbose = true;`
                (preproc.cpp:2371, which makes every dispatch write its
                intermediate meshes as .off files) set to false.

Every replaced span is located by its text and checked to be where the
reference has it; any mismatch stops the build.

    python mcut_amd/seam/patch_sources.py /root/reference mcut_amd/seam/_build/src
"""
import os
import sys


def lines_of(path):
    with open(path) as fh:
        return fh.read().split("\n")


def find_one(lines, text, start=0, stop=None):
    hits = [i for i in range(start, len(lines) if stop is None else stop) if lines[i].strip() == text]
    if len(hits) != 1:
        sys.exit(f"patch_sources: expected one line '{text}' in [{start}, {stop}), found {len(hits)}")
    return hits[0]


RECORD = ("    mcut_probe_record_cutpath(m0, ps_vtx_cnt, m0_ivtx_to_intersection_registry_entry, m0_cutpath_edges, "
          "ps_iface_to_m0_edge_list);")


def cutpath_spans(L):
    """The cut-path edge loop [c0, c1] and the line the recorder goes before."""
    dbg = find_one(L, "#if 0 // used for debugging colinearity bug, which occur when we have poly with eg. > 3")
    loop = ("for (std::map<pair<fd_t>, std::vector<vd_t>>::const_iterator cutpath_edge_creation_info_iter = "
            "cutpath_edge_creation_info.cbegin();")
    hits = [i for i in range(dbg, len(L)) if L[i].strip() == loop]
    if len(hits) != 2:  # the disabled debugging copy of the loop, then the loop
        sys.exit(f"patch_sources: expected the cut-path loop twice after the debugging block, found {len(hits)}")
    c0 = hits[1]
    tail = find_one(L, "} // else if (new_ivertices_count > 2) {", c0)
    c1 = tail + 1
    if L[c1].strip() != "}":
        sys.exit("patch_sources: the cut-path loop does not close where expected")
    note = find_one(L, '// intersecting faces in the polygon-soup ("ps").', c1)
    rec = note + 2
    if L[rec].strip() != "if (input.verbose) {" or "m0.v.e" not in L[rec + 1]:
        sys.exit("patch_sources: the m0.v.e dump does not follow the cut-path loop")
    if (c0 + 1, c1 + 1, rec + 1) != (3427, 3615, 3622):
        sys.exit(f"patch_sources: cut-path loop at kernel.cpp:{c0 + 1}-{c1 + 1}, dump at {rec + 1}; "
                 "expected 3427-3615, 3622")
    return c0, c1, rec


def kernel_gpu(ref, out):
    L = lines_of(os.path.join(ref, "source", "kernel.cpp"))
    head = find_one(L, '#include "mcut/internal/utils.h"')
    banner = find_one(L, "// Calculate polygon intersection points")
    a = find_one(L, "std::unordered_map<ed_t, std::vector<fd_t>> ps_edge_face_intersection_pairs;", banner, banner + 6)
    empty = find_one(L, "if (m0_ivtx_to_intersection_registry_entry.empty()) {", a)
    b = find_one(L, "TIMESTACK_POP();", empty, empty + 25)
    if (a + 1, b + 1) != (1779, 3231):
        sys.exit(f"patch_sources: front-end block at kernel.cpp:{a + 1}-{b + 1}, expected 1779-3231")
    c0, c1, rec = cutpath_spans(L)
    new = L[:head + 1] + ['#include "frontend_seam.h"', '#include "cutpath_probe.h"'] + L[head + 1:a] + \
        ['#include "frontend_block.inc"'] + L[b + 1:c0] + ['#include "cutpath_block.inc"'] + L[c1 + 1:rec] + \
        [RECORD] + L[rec:]
    with open(os.path.join(out, "kernel_gpu.cpp"), "w") as fh:
        fh.write("\n".join(new))


def kernel_probe(ref, out):
    L = lines_of(os.path.join(ref, "source", "kernel.cpp"))
    head = find_one(L, '#include "mcut/internal/utils.h"')
    c0, c1, rec = cutpath_spans(L)
    new = L[:head + 1] + ['#include "cutpath_probe.h"'] + L[head + 1:rec] + [RECORD] + L[rec:]
    with open(os.path.join(out, "kernel_probe.cpp"), "w") as fh:
        fh.write("\n".join(new))


def preproc_quiet(ref, out):
    L = lines_of(os.path.join(ref, "source", "preproc.cpp"))
    i = find_one(L, "kernel_input.verbose = true;")
    if i + 1 != 2371:
        sys.exit(f"patch_sources: verbose switch at preproc.cpp:{i + 1}, expected 2371")
    L[i] = L[i].replace("kernel_input.verbose = true;", "kernel_input.verbose = false;")
    with open(os.path.join(out, "preproc_quiet.cpp"), "w") as fh:
        fh.write("\n".join(L))


if __name__ == "__main__":
    ref, out = sys.argv[1], sys.argv[2]
    os.makedirs(out, exist_ok=True)
    kernel_gpu(ref, out)
    kernel_probe(ref, out)
    preproc_quiet(ref, out)
