"""Shared test inputs: seeded workloads sliced by the oracle (CPU)."""
import functools
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import es_nufft, geometry, matching  # noqa: E402
from spinifel_amd import synth  # noqa: E402


def oracle_slicer(eps=1e-12):
    def f(ac, quats, ppr, E):
        return matching.model_slices(ac, quats, ppr, E, 1, None, eps)
    return f


@functools.lru_cache(maxsize=None)
def workload(name="m41_1k_px", seed=0, n_ref=None, n_img=None):
    return synth.make_workload(name, oracle_slicer(), seed=seed, n_ref=n_ref, n_img=n_img)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def normalized_points(quats, ppr, E, oversampling=1):
    return geometry.gen_nonuniform_normalized_positions(quats, ppr, E, oversampling)


def nudft_real(ac, H, K, L):
    return es_nufft.nudft3d2(H, K, L, np.asarray(ac, np.float64)).real
