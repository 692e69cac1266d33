"""Mesh inputs shared by the tests."""
import os

import numpy as np

from mcut_amd import meshgen as mg

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bench_pair(i):
    s = os.path.join(GOLDEN, "meshes", f"src-mesh{i:03d}.off")
    c = os.path.join(GOLDEN, "meshes", f"cut-mesh{i:03d}.off")
    return mg.read_off(s), mg.read_off(c)


def config1_rotated():
    return mg.config1(phi0=np.pi / 32)


def nfaces(m):
    V, F, S = m
    return F.size // 3 if S is None else S.size


def icosphere_pair(level, amp=0.02, offset=0.5):
    """Config-2 style pair at a chosen subdivision level."""
    V, F, _ = mg.icosphere(level)
    A = mg.radial_displace(V, amp, seed=1)
    B = mg.radial_displace(V, amp, seed=2)
    R = mg.random_rotation(seed=3)
    B = B @ R.T
    B[:, 0] += offset
    return (A, F, None), (B.copy(), F.copy(), None)
