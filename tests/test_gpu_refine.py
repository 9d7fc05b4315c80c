"""GPU refinement: the objective, its hand-written adjoint and the local
refinement, against the FP64 oracle (oracle/gisspa_oracle.cpp,
or_refine_score).

  * objective: GPU value within 1e-4 * max(|oracle|, 1) of the oracle's;
  * gradient: against FP64 central differences of the oracle objective on
    12 and on 256 candidates, within 1e-3 relative (of the component, or of 5 % of the largest
    component for a near-zero one).  Steps: 1e-3 pixel for the shift,
    1e-3 A (1e-7 um) for the defocus and 1e-6 degree for psi: the objective
    is piecewise smooth in psi (bilinear cells) and defocus (|CTF| at its
    zeros), and a larger step would straddle those kinks;
  * the gradient also drives a torch.autograd.Function (test-only wrapper);
  * refinement never lowers the objective and its final values are the
    oracle's at the final parameters.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import _cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import gisspa_amd
    c = _cases.tiny(seed=0, l=128)
    op = orc.params(c.l, **c.params)
    pk = gisspa_amd.Picker(0, **c.params)
    pk.set_templates(c.templates, c.defocus)
    pk.set_micrograph(c.image, c.defocus)
    picks = pk.pick()
    order = np.argsort(-picks["score"])
    top = picks[order[:12]]
    white = orc.template_whitened(op, c.templates)
    tiles = orc.micrograph_prep(op, orc.ctf(c.defocus), c.image)
    ny, nx = c.image.shape
    ox = orc.tile_offsets(nx, c.l, c.params["overlap"])
    oy = orc.tile_offsets(ny, c.l, c.params["overlap"])
    yield c, op, pk, top, white, tiles, (ox, oy)
    pk.close()


def _perturbed(c, top, seed):
    import gisspa_amd
    base = gisspa_amd.peaks_from_picks(top, c.params["phistep"], c.defocus)
    rng = np.random.default_rng(seed)
    p = base.copy()
    p["psi"] += rng.uniform(-2.5, 2.5, len(p)).astype(np.float32)
    p["dx"] += rng.uniform(-1, 1, len(p)).astype(np.float32)
    p["dy"] += rng.uniform(-1, 1, len(p)).astype(np.float32)
    p["defocus"] += rng.uniform(-0.02, 0.02, len(p)).astype(np.float32)
    return base, p


def _oracle_args(c, op, white, tiles, offs, p):
    ox, oy = offs
    bxn = len(ox)
    x0 = float(p["cx"]) - ox[int(p["tile"]) % bxn]
    y0 = float(p["cy"]) - oy[int(p["tile"]) // bxn]
    theta = (float(p["psi"]), float(p["dx"]), float(p["dy"]), float(p["defocus"]))
    return (op, orc.ctf(c.defocus), white[int(p["view"])], tiles[int(p["tile"])], x0, y0), theta


FD_STEPS = (1e-6, 1e-3, 1e-3, 1e-7)


def _fd(args, theta):
    g = np.zeros(4)
    for k, h in enumerate(FD_STEPS):
        tp = np.array(theta, dtype=np.float64)
        tm = tp.copy()
        tp[k] += h
        tm[k] -= h
        g[k] = (orc.refine_score(*args, tp)[0] - orc.refine_score(*args, tm)[0]) / (2 * h)
    return g


def _check_grad(g, fd, what):
    tol = 1e-3 * np.maximum(np.abs(fd), 0.05 * np.abs(fd).max())
    err = np.abs(g - fd)
    print(what, "grad", g, "fd", fd, "err/tol", err / tol)
    assert np.all(err <= tol), (what, g, fd, tol)


def test_objective_and_gradient_match_oracle(gpu, setup):
    c, op, pk, top, white, tiles, offs = setup
    _, peaks = _perturbed(c, top, 1)
    got = pk.adjoint(peaks)
    for i, p in enumerate(peaks):
        args, theta = _oracle_args(c, op, white, tiles, offs, p)
        ref = orc.refine_score(*args, theta)[0]
        print("peak", i, "score gpu", got["score"][i], "oracle", ref)
        assert abs(got["score"][i] - ref) <= 1e-4 * max(abs(ref), 1.0)
        g = np.array([got["d_psi"][i], got["d_dx"][i], got["d_dy"][i], got["d_defocus"][i]],
                     dtype=np.float64)
        _check_grad(g, _fd(args, theta), f"peak {i}")


def test_gradient_on_256_candidates(gpu, setup):
    """Gradients of 256 candidates (the top picks under 22 perturbations each,
    +-2.5 degree, +-1 px, +-200 A) against FP64 central differences."""
    c, op, pk, top, white, tiles, offs = setup
    rows = []
    seed = 10
    while sum(len(r) for r in rows) < 256:
        rows.append(_perturbed(c, top, seed)[1])
        seed += 1
    peaks = np.concatenate(rows)[:256]
    got = pk.adjoint(peaks)
    worst = 0.0
    for i, p in enumerate(peaks):
        args, theta = _oracle_args(c, op, white, tiles, offs, p)
        ref = orc.refine_score(*args, theta)[0]
        assert abs(got["score"][i] - ref) <= 1e-4 * max(abs(ref), 1.0)
        g = np.array([got["d_psi"][i], got["d_dx"][i], got["d_dy"][i], got["d_defocus"][i]],
                     dtype=np.float64)
        fd = _fd(args, theta)
        tol = 1e-3 * np.maximum(np.abs(fd), 0.05 * np.abs(fd).max())
        worst = max(worst, float((np.abs(g - fd) / tol).max()))
        assert np.all(np.abs(g - fd) <= tol), (i, g, fd)
    print(f"256 candidates: worst gradient error {worst:.3f} of the tolerance")


def test_autograd_function_gradient(gpu, setup):
    """torch.autograd.Function whose backward is the library's adjoint."""
    import torch
    c, op, pk, top, white, tiles, offs = setup
    _, peaks = _perturbed(c, top[:4], 2)

    class RefineScore(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta):
            p = peaks.copy()
            t = theta.detach().cpu().numpy()
            p["psi"], p["dx"], p["dy"], p["defocus"] = t[:, 0], t[:, 1], t[:, 2], t[:, 3]
            g = pk.adjoint(p)
            jac = np.stack([g["d_psi"], g["d_dx"], g["d_dy"], g["d_defocus"]], 1)
            ctx.save_for_backward(torch.from_numpy(jac.astype(np.float64)))
            return torch.from_numpy(g["score"].astype(np.float64))

        @staticmethod
        def backward(ctx, gout):
            (jac,) = ctx.saved_tensors
            return gout[:, None] * jac

    theta0 = np.stack([peaks["psi"], peaks["dx"], peaks["dy"], peaks["defocus"]], 1)
    theta = torch.tensor(theta0.astype(np.float64), requires_grad=True)
    weights = torch.tensor([1.0, -0.5, 2.0, 0.25], dtype=torch.float64)
    (RefineScore.apply(theta) * weights).sum().backward()
    for i, p in enumerate(peaks):
        args, th = _oracle_args(c, op, white, tiles, offs, p)
        _check_grad(theta.grad[i].numpy(), _fd(args, th) * float(weights[i]), f"autograd {i}")


def test_refinement_raises_the_objective(gpu, setup):
    c, op, pk, top, white, tiles, offs = setup
    base, peaks = _perturbed(c, top, 3)
    start = pk.adjoint(peaks)["score"]
    at_grid = pk.adjoint(base)["score"]
    refined, final = pk.refine(peaks, iters=20)
    print("start", start, "\nfinal", final["score"], "\ngrid", at_grid)
    assert np.all(final["score"] >= start)
    # the perturbation is undone: back to (at least) the grid optimum's level
    assert np.median(final["score"] / at_grid) >= 0.999
    for i, p in enumerate(refined):
        args, theta = _oracle_args(c, op, white, tiles, offs, p)
        ref = orc.refine_score(*args, theta)[0]
        assert abs(final["score"][i] - ref) <= 1e-4 * max(abs(ref), 1.0)
