"""Writes tests/golden/m21_16x16.npz: a small orientation-matching case and the
oracle's results for it (CPU only; run from the repository root):

    python tests/golden/make_golden.py

Inputs: M=21 Gaussian-atom autocorrelation (60 atoms), one 16 x 16 Ewald
panel, 96 reference quaternions, 24 Poisson images (0.5 photons/px).
Outputs (oracle, FP64, NUFFT eps 1e-12): argmin index per image, the std
scaling ratio, the minimum distances, and a 2000-value subsample of the
model slices (with its flat indices).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import matching  # noqa: E402
from spinifel_amd import synth  # noqa: E402


def main():
    seed = 21
    def slicer(ac, quats, ppr, E):
        return matching.model_slices(ac, quats, ppr, E, 1, None, 1e-12)
    w = synth.make_workload("m41_1k_px", slicer, seed=seed, n_ref=96, n_img=24, M=21, atoms=60,
                            rows=16, cols=16)
    r = matching.slicing_and_match(w.ac, w.images, w.ppr, w.pdr, w.ref_quats, return_details=True)
    ms = r["model_slices"] / r["alpha"]
    sub = np.random.default_rng(0).choice(ms.size, 2000, replace=False)
    np.savez_compressed(os.path.join(HERE, "m21_16x16.npz"), ac=w.ac, ppr=w.ppr, pdr=w.pdr,
                        ref_quats=w.ref_quats, images=w.images, index=r["index"],
                        alpha=np.float64(r["alpha"]), min_dist=r["min_dist"],
                        slice_flat_idx=sub, slice_values=ms.ravel()[sub])


if __name__ == "__main__":
    main()
