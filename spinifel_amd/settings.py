"""Settings shim: the fields of spinifel's settings singleton that the
orientation-matching path reads.

Reference: ``SpinifelSettings`` (spinifel/settings.py:167-848); the path
reads ``N_orientations``, ``N_batch_size``, ``reduced_det_shape``
(spinifel/sequential/orientation_matching.py:204-211) and ``oversampling``
(:234-236), and the merge reads ``M``, ``M_ups``, ``solve_ac_maxiter``
(spinifel/sequential/autocorrelation.py:44-53), and phasing reads ``Mquat``,
``beta``, ``nER``, ``nHIO``, ``N_phase_loops`` (spinifel/sequential/phasing.py:249-296),
and the data prep ``N_clipping``, ``N_binning`` (spinifel/prep.py:116-134).
The rest of the reference's TOML/CLI system is out of scope.

``use_single_prec`` (spinifel/settings.py:303-307, default False) selects the
distance precision as in the reference (f_type float64 unless set,
spinifel/sequential/orientation_matching.py:32-37): with it False the FP32
form rescores its best candidates in FP64, so the argmin is the FP64 one;
with it True the FP32 form is the plain FP32 argmin.

Fields added for this implementation: ``nufft_eps`` (the reference hard-codes
1e-12 for finufft, spinifel/autocorrelation.py:59; FP32 slices cannot carry
more than ~1e-7 so the default here is 1e-6), ``precision`` of the cross-term
GEMM, ``topk`` candidates and ``rescore`` for the precision forms, and
``device``.
"""
from dataclasses import dataclass


@dataclass
class SpinifelSettings:
    N_orientations: int = 10000
    N_batch_size: int = 1000
    reduced_det_shape: tuple = (1, 128, 128)
    oversampling: int = 1
    nufft_eps: float = 1e-6
    solve_ac_maxiter: int = 100     # CG iterations of the merge (settings.py:383-386)
    beta: float = 0.9               # HIO feedback (settings.py:388-391)
    nER: int = 10                   # ER iterations per block (settings.py:398-401)
    nHIO: int = 5                   # HIO iterations per block (settings.py:403-406)
    N_phase_loops: int = 5          # ER/HIO/ER/shrink-wrap loops (settings.py:408-411)
    N_clipping: int = 1             # data prep clipping factor 2^n (settings.py:413-416)
    N_binning: int = 4              # data prep binning factor 2^n (settings.py:418-421)
    use_single_prec: bool = False   # FP32 distances only (settings.py:303-307)
    precision: str = "fp32"     # 'fp32' | 'bf16' | 'fp8'
    topk: int = 0               # 0: 8 for bf16 / fp8; fp32: 1 with use_single_prec, else 4
    rescore: bool = None        # None: FP64 rescoring unless fp32 with use_single_prec
    device: int = 0
    verbose: bool = False

    @property
    def Mquat(self):
        """int(oversampling * 20) (spinifel/settings.py:805-810)."""
        return int(self.oversampling * 20)

    @property
    def M(self):
        """4 * Mquat + 1 (spinifel/settings.py:813-818)."""
        return 4 * self.Mquat + 1

    @property
    def M_ups(self):
        """2 * M, the grid of the merge's convolution kernel (spinifel/settings.py:821-825)."""
        return 2 * self.M


settings = SpinifelSettings()
