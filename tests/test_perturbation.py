"""The general-position retry's perturbation draw (preproc.cpp:2621-2631):
default_random_engine(1) seeds an mt19937 whose draws go through
uniform_real_distribution(-1, 1), times the cut mesh's diag * constant (the
dispatch's eps).  libmcgpu's host code draws with the same standard-library
generators; this compiles those three lines with the host compiler and pins
the first vector to the one the reference drew on config 1 (golden fixture),
bit for bit."""
import os
import subprocess

import numpy as np

import meshes

PROGRAM = r"""
#include <cstdio>
#include <random>
int main() {
    std::default_random_engine e(1);
    std::mt19937 m(e());
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (int i = 0; i < 6; ++i) std::printf("%a\n", u(m));
    return 0;
}
"""


def test_first_draw_is_the_reference_perturbation(tmp_path):
    from oracle import oracle as om
    from mcut_amd import meshgen as mg
    src = tmp_path / "draw.cpp"
    exe = tmp_path / "draw"
    src.write_text(PROGRAM)
    subprocess.run(["g++", "-O2", str(src), "-o", str(exe)], check=True)
    draws = [float.fromhex(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.split()]
    eps = om.frontend(*mg.config1(), detail=False)["eps"]
    want = np.load(os.path.join(meshes.GOLDEN, "ref_cases.npz"))["config1_enforce/perturbation"]
    got = np.array(draws[:3]) * eps
    assert np.array_equal(got.view(np.uint64), np.asarray(want, np.float64).view(np.uint64))
    assert all(-1.0 <= d < 1.0 for d in draws)
