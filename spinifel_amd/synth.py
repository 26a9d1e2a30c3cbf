"""Seeded synthetic workloads for orientation matching.

Every array comes from ``numpy.random.default_rng`` with a recorded seed
per stream (atoms, reference quaternions, true quaternions, Poisson noise,
arrivals).  Nothing here is timed.

Model: ``n_atoms`` Gaussian atoms (sigma 1 voxel) with centres uniform in a
ball of radius ``0.3 M`` about the grid centre; the model handed to
matching is the autocorrelation in the form spinifel's phasing step returns
(spinifel/sequential/phasing.py:305-313):
``fftshift(|ifftn(|fftn(rho)|^2)|)``, float32, centrosymmetric.

Detector: flat panels stacked into a square about the beam, pixel
q-vectors on the Ewald sphere ``q = (r_hat - z_hat) / lambda``; the
returned ``pixel_distance_reciprocal`` is ``|q| / edge_frac`` so that
``reciprocal_extent = max(pdr)`` puts the corner pixel at
``|h| = edge_frac * pi`` (spinifel/sequential/orientation_matching.py:221,
:234-236).

Images: the model slice at a random true orientation, clipped at 0,
scaled globally to a mean photon count per pixel, Poisson-sampled, float32.
The slicer is passed in by the caller (GPU path or a CPU checker).
"""
from dataclasses import dataclass, field

import numpy as np

STREAM_ATOMS, STREAM_REF, STREAM_TRUE, STREAM_POISSON, STREAM_ARRIVAL = 0, 1, 2, 3, 4


def rng_for(seed, stream):
    return np.random.default_rng([int(seed), int(stream)])


def density(M, n_atoms, rng, sigma=1.0, radius_frac=0.3):
    """Sum of unit-amplitude isotropic Gaussians on an M^3 grid (float64)."""
    c0 = (M - 1) / 2.0
    r = radius_frac * M
    # uniform in a ball: direction * r * u^(1/3)
    d = rng.standard_normal((n_atoms, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rad = r * rng.uniform(0.0, 1.0, n_atoms) ** (1.0 / 3.0)
    centres = c0 + d * rad[:, None]
    ax = np.arange(M, dtype=np.float64)
    g = [np.exp(-(ax[None, :] - centres[:, i:i + 1]) ** 2 / (2.0 * sigma ** 2)) for i in range(3)]
    gxy = (g[0][:, :, None] * g[1][:, None, :]).reshape(n_atoms, M * M)
    rho = (gxy.T @ g[2]).reshape(M, M, M)
    return rho


def autocorrelation(rho):
    """``fftshift(|ifftn(|fftn(rho)|^2)|)`` symmetrised exactly, float32."""
    F = np.fft.fftn(rho)
    ac = np.fft.fftshift(np.abs(np.fft.ifftn(np.abs(F) ** 2)))
    M = ac.shape[0]
    if M % 2 == 1:
        ac = 0.5 * (ac + ac[::-1, ::-1, ::-1])
    return ac.astype(np.float32)


def detector(panels, rows, cols, edge_frac=0.9, two_theta_corner_deg=30.0, wavelength=1.0):
    """Flat panels stacked along the row axis into one square about the beam.

    Returns ``(ppr (3, panels, rows, cols), pdr (panels, rows, cols))`` float64.
    """
    R = panels * rows
    C = cols
    gy = (np.arange(R) - (R - 1) / 2.0).reshape(panels, rows)[:, :, None] * np.ones((1, 1, C))
    gx = (np.arange(C) - (C - 1) / 2.0)[None, None, :] * np.ones((panels, rows, 1))
    corner = np.hypot((R - 1) / 2.0, (C - 1) / 2.0)
    dist = corner / np.tan(np.radians(two_theta_corner_deg))
    r = np.stack([gx, gy, np.full_like(gx, dist)])
    rhat = r / np.linalg.norm(r, axis=0, keepdims=True)
    q = rhat.copy()
    q[2] -= 1.0
    q /= wavelength
    qn = np.linalg.norm(q, axis=0)
    return q, qn / edge_frac


def random_quaternions(n, rng):
    """Uniform on S^3 (normalised 4-D Gaussians), canonical w >= 0."""
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[q[:, 0] < 0] *= -1.0
    return q


def poisson_images(slices, mean_photons, rng):
    """Clip at 0, scale globally to ``mean_photons`` per pixel, Poisson-sample."""
    s = np.maximum(np.asarray(slices, np.float64), 0.0)
    c = mean_photons / max(s.mean(), 1e-300)
    return rng.poisson(s * c).astype(np.float32)


@dataclass
class Workload:
    name: str
    M: int
    ac: np.ndarray
    ppr: np.ndarray
    pdr: np.ndarray
    ref_quats: np.ndarray
    true_quats: np.ndarray
    images: np.ndarray          # (N_i,) + det_shape float32
    det_shape: tuple
    seed: int
    meta: dict = field(default_factory=dict)

    @property
    def N_pixels(self):
        return int(np.prod(self.det_shape))

    @property
    def reciprocal_extent(self):
        return float(self.pdr.max())


# Shapes of the standard workloads (name -> parameters).
WORKLOADS = {
    # CPU-runnable plain case: M=41, 1 panel 32x32, 2000 refs, 500 images
    "m41_1k_px": dict(M=41, atoms=300, panels=1, rows=32, cols=32,
                      n_ref=2000, n_img=500, mean_photons=0.5),
    # quickstart scale: M=81, 4 x 32 x 128 px, 10k refs, 10k images
    "m81_16k_px": dict(M=81, atoms=800, panels=4, rows=32, cols=128,
                       n_ref=10000, n_img=10000, mean_photons=1.0),
}


def make_model(M, atoms, seed):
    rho = density(M, atoms, rng_for(seed, STREAM_ATOMS))
    return autocorrelation(rho)


def make_workload(name, slicer, seed=0, n_ref=None, n_img=None, **over):
    """Build a named workload.  ``slicer(ac, quats, ppr, E) -> (N, N_p)`` array."""
    p = dict(WORKLOADS[name])
    p.update(over)
    if n_ref is not None:
        p["n_ref"] = n_ref
    if n_img is not None:
        p["n_img"] = n_img
    ac = make_model(p["M"], p["atoms"], seed)
    ppr, pdr = detector(p["panels"], p["rows"], p["cols"])
    det_shape = (p["panels"], p["rows"], p["cols"])
    ref = random_quaternions(p["n_ref"], rng_for(seed, STREAM_REF))
    true = random_quaternions(p["n_img"], rng_for(seed, STREAM_TRUE))
    E = float(pdr.max())
    sl = np.asarray(slicer(ac, true, ppr, E), np.float64).reshape(p["n_img"], -1)
    imgs = poisson_images(sl, p["mean_photons"], rng_for(seed, STREAM_POISSON))
    return Workload(name=name, M=p["M"], ac=ac, ppr=ppr, pdr=pdr, ref_quats=ref,
                    true_quats=true, images=imgs.reshape((p["n_img"],) + det_shape),
                    det_shape=det_shape, seed=seed, meta=p)
