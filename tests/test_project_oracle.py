"""Template projection: the FP64 oracle (oracle/gisspa_oracle.cpp,
or_project), pinned by closed forms on the CPU:
  * the identity view is the sum of the volume along z, exactly;
  * alt = 90 degrees is the sum along y (volume zero on its faces);
  * phi = 90 degrees turns the alt/az projection by a quarter turn, pixel
    for pixel;
  * a volume of Gaussian blobs projects to the analytic sum of 2D Gaussians
    (gisspa_amd.synth.project_templates, the templates the picker tests use)
    at any view, to the accuracy of trilinear sampling.
The GPU projector is compared with this oracle in test_gpu_project.py."""
import numpy as np

from gisspa_amd import synth
from oracle import oracle as orc


def _vol(n, seed):
    rng = np.random.default_rng(seed)
    v = np.zeros((n, n, n), np.float32)
    v[1:-1, 1:-1, 1:-1] = rng.standard_normal((n - 2, n - 2, n - 2))
    return v


def test_identity_view_sums_along_z():
    v = _vol(12, 0)
    p = orc.project(v, [(0, 0, 0)])[0]
    assert np.allclose(p, v.astype(np.float64).sum(axis=0), atol=1e-12)


def test_tilt_90_sums_along_y():
    v = _vol(12, 1)
    p = orc.project(v, [(90, 0, 0)])[0]
    # sample (u, v, w) -> (x, y, z) = (u, -w, v) + c: row = z, col = x
    assert np.allclose(p, v.astype(np.float64).sum(axis=1), atol=1e-9)


def test_phi_quarter_turn():
    v = _vol(14, 2)
    n, c = 14, 7
    p0, p90 = orc.project(v, [(30, 40, 0), (30, 40, 90)])
    # image_90(u, v) = image_0(-v, u): row r, col k <- row k, col 2c - r
    # (row 0 would need column 2c = n of image_0, outside its array)
    exp = np.zeros_like(p0)
    for r in range(1, n):
        exp[r, :] = p0[:, 2 * c - r]
    assert np.allclose(p90[1:], exp[1:], atol=1e-9)


def test_blob_volume_projects_to_analytic_templates():
    n = 48
    dens = synth.blob_density(12, n, seed=3, sigma_range=(2.5, 4.0))
    vol = synth.render_volume(dens, n)
    eul = synth.random_eulers(6, seed=4)
    got = orc.project(vol, [(a, z, 0.0) for a, z in eul])
    ref = synth.project_templates(dens, n, eul).astype(np.float64)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("max relative deviation from the analytic projection", err)
    assert err < 0.02
