"""Independent pins of the oracle stages that the stage-level tests in
test_oracle.py leave open: norm_type 1 picking, the phase_flip = 1
micrograph whitening / weighting / band pass with the per-tile amplitude
normaliser, and the SNR/|CTF| weighting of the templates.

Each check is a restatement that does not share code with the oracle:
numpy's FFT instead of the oracle's, the textbook CTF instead of CTF_AST,
the weighting written in its algebraically simplified closed form instead
of the reference's sequence of if-branches, or a brute-force search over
numpy arrays.  Inputs are chosen so that a dropped term, a wrong sign or a
wrong index fails (a flat input spectrum turns the whitening into a known
constant; a CTF zero placed on a pixel ring; kk = 0 and kk > 0; a window
whose half is odd, which negates the correlation map).
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc

RNG = np.random.default_rng(8642)


def textbook_ctf(s2, df_um, cs_mm, kv, amp):
    """Phase-contrast CTF at squared spatial frequency s2 (1/A^2), underfocus
    df_um micrometres: sqrt(1-A^2) sin(g) + A cos(g),
    g = pi lambda df s^2 - pi/2 Cs lambda^3 s^4 (relativistic lambda in A)."""
    lam = 12.2639 / math.sqrt(kv * 1000.0 + 0.97845 * kv * kv)
    g = math.pi * lam * df_um * 1e4 * s2 - 0.5 * math.pi * cs_mm * 1e7 * lam ** 3 * s2 * s2
    return math.sqrt(1 - amp * amp) * np.sin(g) + amp * np.cos(g)


def raised_cosine_band(r, lo, hi, width=8.0):
    """Band pass over integer ring index r: 1 on [lo, hi), a half-cosine fall
    over [hi, hi + width) and a half-cosine rise over [lo - width, lo)
    (first matching interval wins), 0 elsewhere and for r < 0."""
    r = np.asarray(r, dtype=np.float64)
    conds = [(r < hi) & (r >= lo),
             (r >= hi) & (r < hi + width),
             (r >= lo - width) & (r < lo) & (r >= 0)]
    vals = [np.ones_like(r),
            0.5 * np.cos(np.pi * (r - hi) / (2 * width)) + 0.5,
            0.5 * np.cos(np.pi * (lo - r) / (2 * width)) + 0.5]
    return np.select(conds, vals, 0.0)


def ring_index(nx, ny):
    """Pixel ring of each spectrum pixel (x, y): nearest integer radius of the
    wrapped frequency (min(x, nx-x), min(y, ny-y)), minus one (DC -> -1)."""
    yy, xx = np.mgrid[0:ny, 0:nx]
    dx = np.minimum(xx, nx - xx).astype(np.float64)
    dy = np.minimum(yy, ny - yy).astype(np.float64)
    return xx, yy, dx, dy, (np.floor(np.hypot(dx, dy) + 0.5) - 1).astype(np.int64)


def snr_curves(p, ss):
    """Signal and noise curves of the weighting (exponentials in ss)."""
    S = np.exp(p.bfactor * ss * ss + p.bfactor2 * ss + p.bfactor3)
    N = np.exp(p.a * ss * ss + p.b * ss + p.b2)
    return S, N


def half_plane_norm(t, l):
    """sqrt(sum |t|^2) / count over the half plane x <= l/2 inside ring l/2."""
    xx, yy, dx, dy, r = ring_index(l, l)
    sel = (xx <= l // 2) & (r >= 0) & (r < l // 2)
    return math.sqrt(float(np.sum(np.abs(t[sel]) ** 2))) / sel.sum()


# ---------------------------------------------------------------------------
# phase_flip = 1 micrograph preparation (GPU_func.cu:155-205, 785-796;
# main.cu:548-586 whole-image path, main.cu:617-653 per-tile normaliser)
# ---------------------------------------------------------------------------
def _mg_params(l, **kw):
    base = dict(angpix=1.0, phistep=30, kk=3, energy=300, cs=2.7, Highres=4, Lowres=20,
                diameter=12, ampconst=0.1, overlap=16, phase_flip=1)
    base.update(kw)
    return orc.params(l, **base)


def flat_spectrum_image(nx, ny, scale, seed):
    """A real image whose spectrum has modulus exactly `scale` at every pixel
    (random phases with Hermitian symmetry); returns (image, unit spectrum)."""
    rng = np.random.default_rng(seed)
    U = np.fft.fft2(rng.standard_normal((ny, nx)))
    U /= np.abs(U)
    img = np.real(np.fft.ifft2(U)) * scale
    return img, U


def micrograph_weight_closed_form(p, c, nx, ny):
    """Real factor multiplying each spectrum pixel of a micrograph whose ring
    power is 1, in closed form: with S, N the signal and noise curves at
    s = |wrapped frequency|, v = |CTF|,
        w = sqrt((S v^2 + N) / (N (kk+1) + kk S v^2)) * damp * band
    inside ring L/2 (L = max(nx, ny)), damp = exp(-100 s^2) above ring
    L apix / 6; outside ring L/2 only the band applies."""
    L = max(nx, ny)
    xx, yy, dx, dy, r = ring_index(nx, ny)
    s2 = ((dx / nx) ** 2 + (dy / ny) ** 2) / (p.apix * p.apix)
    ss = np.sqrt(s2)
    v = np.abs(textbook_ctf(s2, c.defocus_um, p.cs, p.energy, p.ampconst))
    S, N = snr_curves(p, ss)
    kk = p.kk
    inner = (r >= 0) & (r < L // 2)
    w = np.sqrt((S * v * v + N) / (N * (kk + 1) + kk * S * v * v))
    w = np.where(r > L * p.apix / 6, w * np.exp(-100 * s2), w)
    band = raised_cosine_band(r, L * p.apix / p.lowres, L * p.apix / p.highres)
    return np.where(inner, w, 1.0) * band


def phase_flip_expected_tiles(p, c, U, l, scale_out=1.0):
    """Tile spectra of the phase_flip = 1 path for an input whose spectrum is
    scale * U with |U| = 1: the whitened spectrum is U itself, so the
    processed micrograph has spectrum sign(CTF) * U * w (DC zeroed)."""
    ny, nx = U.shape
    xx, yy, dx, dy, r = ring_index(nx, ny)
    s2 = ((dx / nx) ** 2 + (dy / ny) ** 2) / (p.apix * p.apix)
    sgn = np.where(textbook_ctf(s2, c.defocus_um, p.cs, p.energy, p.ampconst) > 0, 1.0, -1.0)
    G = U * sgn * micrograph_weight_closed_form(p, c, nx, ny)
    G[0, 0] = 0
    img = np.real(np.fft.ifft2(G)) * (nx * ny)
    out = []
    for y0 in orc.tile_offsets(ny, l, p.overlap):
        for x0 in orc.tile_offsets(nx, l, p.overlap):
            t = np.fft.fft2(img[y0:y0 + l, x0:x0 + l]) / (l * l)
            out.append(t / half_plane_norm(t, l))
    return np.array(out), G


@pytest.mark.parametrize("kk,nx,ny,invert", [(3, 96, 80, 0), (0, 96, 80, 0), (2.5, 80, 96, 1)])
def test_phase_flip_prep_flat_spectrum_closed_form(kk, nx, ny, invert):
    l = 48
    p = _mg_params(l, kk=kk, invert=invert)
    c = orc.ctf(1.3)
    img, U = flat_spectrum_image(nx, ny, 100.0, seed=int(10 * kk) + nx)
    # no pixel is touched by the hot-pixel clamp (bounds at +-6 sd, int-truncated)
    x = img.astype(np.float32).astype(np.float64)
    avg = x.mean()
    sd = math.sqrt((x * x).mean() - avg * avg)
    assert np.all((x <= int(avg + 6 * sd)) & (x >= int(avg - 6 * sd)))
    got = orc.micrograph_prep(p, c, img.astype(np.float32))
    # the float32 input changes the spectrum's modulus by ~1e-7 relative:
    # the expected tiles use the spectrum of the float32 image, normalised
    F32 = np.fft.fft2(x)
    Uf = F32 / np.abs(F32)
    Uf[0, 0] = 0
    if invert:
        Uf = -Uf
    exp_tiles, G = phase_flip_expected_tiles(p, c, Uf, l)
    assert got.shape == exp_tiles.shape
    err = np.abs(got - exp_tiles).max() / np.abs(exp_tiles).max()
    assert err < 2e-6, err
    # invariants: sqrt(sum |F|^2) / count = 1 on every tile; zero at DC
    for t in got:
        assert half_plane_norm(t, l) == pytest.approx(1.0, rel=1e-12)
    # the processed micrograph's spectrum is zero outside the band and at DC
    xx, yy, dx, dy, r = ring_index(nx, ny)
    L = max(nx, ny)
    outside = (r >= L * p.apix / p.highres + 8) | (r < 0)
    assert np.all(G[outside] == 0)


def test_phase_flip_sign_follows_the_textbook_ctf():
    """Without the sign flip the closed form misses: the oracle's output
    differs from the unflipped expectation on the pixels where the CTF is
    negative (and only there)."""
    l, nx, ny = 48, 96, 80
    p = _mg_params(l)
    c = orc.ctf(2.1)
    img, _ = flat_spectrum_image(nx, ny, 100.0, seed=3)
    x = img.astype(np.float32).astype(np.float64)
    got = orc.micrograph_prep(p, c, img.astype(np.float32))
    F32 = np.fft.fft2(x)
    Uf = F32 / np.abs(F32)
    Uf[0, 0] = 0
    flipped, _ = phase_flip_expected_tiles(p, c, Uf, l)
    assert np.abs(got - flipped).max() < 2e-6 * np.abs(flipped).max()
    # the CTF at 2.1 um has both signs inside the pass band, so ignoring the
    # flip must fail
    xx, yy, dx, dy, r = ring_index(nx, ny)
    s2 = ((dx / nx) ** 2 + (dy / ny) ** 2) / (p.apix * p.apix)
    v = textbook_ctf(s2, 2.1, p.cs, p.energy, p.ampconst)
    band = micrograph_weight_closed_form(p, c, nx, ny) > 0
    assert np.any(v[band] > 0) and np.any(v[band] < 0)
    sgn = np.where(v > 0, 1.0, -1.0)
    unflipped, _ = phase_flip_expected_tiles(p, c, Uf * sgn, l)
    assert np.abs(got - unflipped).max() > 1e-2 * np.abs(flipped).max()


def test_phase_flip_whitening_cancels_radial_filters():
    """The whitening divides each pixel ring by its RMS over the half plane
    x <= nx/2, so a micrograph filtered by any positive function of the ring
    index prepares to the same tiles as the unfiltered one."""
    l, nx, ny = 48, 96, 80
    p = _mg_params(l)
    c = orc.ctf(1.7)
    x0 = RNG.standard_normal((ny, nx)) * 50
    F = np.fft.fft2(x0)
    xx, yy, dx, dy, r = ring_index(nx, ny)
    h = 0.4 + RNG.random(int(r.max()) + 2)
    x1 = np.real(np.fft.ifft2(F * h[np.maximum(r, 0)]))
    # the filtered image keeps the DC term (r = -1 maps to h[0]), the
    # whitening never sees it (DC is zeroed)
    t0 = orc.micrograph_prep(p, c, x0.astype(np.float32))
    t1 = orc.micrograph_prep(p, c, x1.astype(np.float32))
    assert np.abs(t0 - t1).max() < 2e-5 * np.abs(t0).max()
    # a filter that is not constant on the pixel rings does change the output
    h2 = 1.0 + 0.5 * np.cos(np.arctan2(yy - ny / 2, xx - nx / 2))
    x2 = np.real(np.fft.ifft2(F * h2))
    t2 = orc.micrograph_prep(p, c, x2.astype(np.float32))
    assert np.abs(t0 - t2).max() > 1e-2 * np.abs(t0).max()


# ---------------------------------------------------------------------------
# template SNR / |CTF| weighting (GPU_func.cu:249-290) and the amplitude
# normaliser (main.cu:444-452)
# ---------------------------------------------------------------------------
def _tp(l, **kw):
    base = dict(angpix=1.0, phistep=10, kk=3, energy=300, cs=2.7, Highres=4, Lowres=100,
                diameter=0.75 * l / 2, ampconst=0.1)
    base.update(kw)
    return orc.params(l, **base)


def template_weight_closed_form(p, c, l, l_config=None):
    """|CTF| * sqrt(S / (N (kk+1) + kk S |CTF|^2)) * low pass, on rings
    0 <= r < l/2, with S and N at s = r / (apix * l_config) (the weighting's
    frequency step is fixed from the configured window) and the CTF at the
    pixel's own frequency; rings outside [0, l/2) keep only the low pass."""
    l_config = l_config or l
    xx, yy, dx, dy, r = ring_index(l, l)
    rr = np.hypot(dx, dy)
    ss = rr / (p.apix * l_config)
    s2_pix = (dx * dx + dy * dy) / (l * l) / (p.apix * p.apix)
    v = np.abs(textbook_ctf(s2_pix, c.defocus_um, p.cs, p.energy, p.ampconst))
    S, N = snr_curves(p, ss)
    kk = p.kk
    hi, lo = l * p.apix / p.highres, l * p.apix / p.lowres
    band = np.select([(r >= 0) & (r < hi), (r >= hi) & (r < hi + 8),
                      (r >= lo - 8) & (r < lo) & (r >= 0)],
                     [np.ones(r.shape), 0.5 * np.cos(np.pi * (r - hi) / 16) + 0.5,
                      0.5 * np.cos(np.pi * (lo - r) / 16) + 0.5], 0.0)
    inner = (r >= 0) & (r < l // 2)
    w = v * np.sqrt(S / (N * (kk + 1) + kk * S * v * v))
    return np.where(inner, w, 1.0) * band


def _template_case(l, p, c, n=2, box=None):
    box = box or l // 2
    tm = RNG.standard_normal((n, box, box)).astype(np.float32)
    B = orc.template_whitened(p, tm)
    out = orc.template_prep(p, c, tm)
    O = np.fft.fft2(out) / (l * l)    # spectrum of the prepared template
    return B, O


@pytest.mark.parametrize("kk,l,lcfg", [(3, 64, None), (0, 64, None), (1.5, 72, None),
                                       (3, 96, 80)])
def test_template_weight_closed_form(kk, l, lcfg):
    kw = dict(kk=kk)
    if lcfg:
        kw["window_size"] = lcfg
    p = _tp(l, **kw)
    c = orc.ctf(1.4)
    B, O = _template_case(l, p, c)
    w = template_weight_closed_form(p, c, l, lcfg)
    for b, o in zip(B, O):
        A = b * w
        m = half_plane_norm(A, l)
        # 2e-8: the reference's PI (3.141592654) inside its CTF against
        # math.pi here moves the CTF phase by ~1e-10 relative
        assert np.abs(o - A / m).max() < 2e-8 * np.abs(A / m).max()


def test_template_weight_kk0_is_ctf_times_root_snr():
    """kk = 0: the weight is |CTF| sqrt(S/N) inside the pass band."""
    l = 64
    p = _tp(l, kk=0)
    c = orc.ctf(1.9)
    B, O = _template_case(l, p, c, n=1)
    xx, yy, dx, dy, r = ring_index(l, l)
    ss = np.hypot(dx, dy) / (p.apix * l)
    S, N = snr_curves(p, ss)
    v = np.abs(textbook_ctf(ss * ss, 1.9, p.cs, p.energy, p.ampconst))
    sel = (r >= 0) & (r < l * p.apix / p.highres)
    ratio = O[0][sel] / B[0][sel]
    want = v[sel] * np.sqrt(S[sel] / N[sel])
    # the ratio is w / m with one m for the whole template (absolute
    # tolerance: next to the CTF zeros the PI difference is relatively larger)
    k = np.median(ratio.real / want)
    assert np.abs(ratio.imag).max() < 1e-9 * np.abs(ratio.real).max()
    assert np.abs(ratio.real - k * want).max() < 2e-8 * np.abs(ratio.real).max()


def test_template_weight_ring_ratios():
    """Between two rings the prepared-to-whitened ratio is the ratio of the
    closed-form weights (the normaliser cancels), for pixels of different
    rings and both CTF signs."""
    l = 80
    p = _tp(l, kk=2)
    c = orc.ctf(2.4)
    B, O = _template_case(l, p, c, n=1)
    w = template_weight_closed_form(p, c, l)
    pix = [(3, 0), (0, 7), (5, 12), (l - 9, 4), (11, l - 13), (17, 2)]
    ref_x, ref_y = pix[0]
    for x, y in pix[1:]:
        got = (O[0][y, x] / B[0][y, x]) / (O[0][ref_y, ref_x] / B[0][ref_y, ref_x])
        assert got.real == pytest.approx(w[y, x] / w[ref_y, ref_x], rel=2e-8)
        assert abs(got.imag) < 1e-9


def test_template_weight_zero_on_the_ctf_zero_set():
    """With Cs = 0 and no amplitude contrast the CTF is sin(pi lambda df s^2):
    a defocus that puts its first zero exactly on ring radius 10 makes the
    prepared template's spectrum vanish on every pixel at that radius
    ((10, 0), (6, 8), (8, 6), (0, 10) and their mirrors) and nowhere near it."""
    l = 64
    kv = 300
    lam = 12.2639 / math.sqrt(kv * 1000.0 + 0.97845 * kv * kv)
    s2 = 100.0 / (l * l)
    df = 1.0 / (lam * 1e4 * s2)
    p = _tp(l, cs=0.0, ampconst=0.0, kk=1)
    c = orc.ctf(df)
    B, O = _template_case(l, p, c, n=1)
    mag = np.abs(O[0])
    on = [(10, 0), (0, 10), (6, 8), (8, 6), (l - 6, 8), (6, l - 8), (l - 10, 0)]
    for x, y in on:
        assert mag[y, x] < 1e-6 * mag.max(), (x, y)
    # neighbouring rings carry weight
    for x, y in [(9, 0), (11, 0), (0, 9), (7, 7)]:
        assert mag[y, x] > 1e-3 * mag.max(), (x, y)


# ---------------------------------------------------------------------------
# norm_type 1 (main.cu:718-821, GPU_func.cu:618-662, 711-745) by brute force
# ---------------------------------------------------------------------------
def circular_correlation(tile, tmpl):
    """c[u] = sum_x tile[x] tmpl[x - u - l/2] / l^2 (zero lag at (l/2, l/2)),
    through numpy's FFT."""
    l = tile.shape[0]
    corr = np.real(np.fft.ifft2(np.fft.fft2(tile) * np.conj(np.fft.fft2(tmpl))))
    return np.roll(corr, (-(l // 2), -(l // 2)), axis=(0, 1)) / (l * l)


def test_circular_correlation_helper_is_the_direct_sum():
    l = 12
    a = RNG.standard_normal((l, l))
    b = RNG.standard_normal((l, l))
    ref = np.zeros((l, l))
    for uy in range(l):
        for ux in range(l):
            for y in range(l):
                for x in range(l):
                    ref[uy, ux] += a[y, x] * b[(y - uy - l // 2) % l, (x - ux - l // 2) % l]
    assert np.abs(circular_correlation(a, b) - ref / l ** 2).max() < 1e-12


def norm1_brute_force(p, c, tm, img):
    """Records (psi index, view, cx, cy, z) of norm_type 1 by brute force:
    every CC map of every (psi, tile, view) from real-space operands, the
    per-pixel mean and sd of c / l^2 over all views x psi, z = (c/l^2 - mean)
    / sd (0 where sd <= 0), the maximum of z over each run of 1024 pixels of
    the row-major map (first position on ties; the last run is shorter),
    kept when z >= threshold and the centre lies d or more from every edge;
    ordered by psi, tile (y outer, x inner), view, run."""
    l = p.l
    ny, nx = img.shape
    ox = orc.tile_offsets(nx, l, p.overlap)
    oy = orc.tile_offsets(ny, l, p.overlap)
    tiles = orc.micrograph_prep(p, c, img)
    treal = orc.template_prep(p, c, tm)
    P = orc.num_psi(p.phi_step)
    tile_r = np.real(np.fft.ifft2(tiles, axes=(1, 2))) * l * l
    sign = -1.0 if (l // 2) % 2 == 1 else 1.0
    T = len(tiles)
    n = len(tm)
    cc = np.zeros((P, T, n, l, l))
    for ip in range(P):
        spec = orc.rotate_normalize(p, treal, orc.psi_value(p.phi_step, ip))
        rot = np.real(np.fft.ifft2(spec, axes=(1, 2))) * l * l
        for t in range(T):
            for j in range(n):
                cc[ip, t, j] = sign * circular_correlation(tile_r[t], rot[j])
    cur = cc / (l * l)
    total = n * P
    mean = cur.sum(axis=(0, 2)) / total
    sd = np.sqrt(np.maximum((cur * cur).sum(axis=(0, 2)) / total - mean * mean, -1.0))
    with np.errstate(invalid="ignore", divide="ignore"):
        z = np.where(sd[None, :, None] > 0, (cur - mean[None, :, None]) / sd[None, :, None], 0.0)
    recs = []
    nrun = -(-l * l // 1024)
    for ip in range(P):
        for t in range(T):
            iy, ix = divmod(t, len(ox))
            for j in range(n):
                flat = z[ip, t, j].ravel()
                for b in range(nrun):
                    seg = flat[b * 1024:(b + 1) * 1024]
                    k = int(np.argmax(seg))
                    best = seg[k]
                    pos = b * 1024 + k
                    if best >= p.threshold:
                        cx = ox[ix] + pos % l
                        cy = oy[iy] + pos // l
                        d = p.diameter
                        if d <= cx < nx - d and d <= cy < ny - d:
                            recs.append((ip, j, cx, cy, best))
    return recs, z


@pytest.mark.parametrize("l,phistep,size", [(48, 60.0, (80, 64)), (30, 90.0, (56, 50))])
def test_norm1_matches_brute_force(l, phistep, size):
    nx, ny = size
    box = 16 if l == 48 else 12
    tm = (RNG.standard_normal((3, box, box)) + 0.3).astype(np.float32)
    img = RNG.standard_normal((ny, nx)).astype(np.float32)
    p = orc.params(l, angpix=1.0, phistep=phistep, kk=3, energy=300, cs=2.7, Highres=4,
                   Lowres=100, diameter=6, threshold=2.4, overlap=16, ampconst=0.1, norm_type=1)
    c = orc.ctf(1.2)
    want, z = norm1_brute_force(p, c, tm, img)
    got = orc.pick_norm1(p, c, tm, img)
    # no run maximum sits within rounding of the threshold
    assert np.abs(z - p.threshold).min() > 1e-9
    assert len(want) > 5 and len(got) == len(want)
    for g, w in zip(got, want):
        assert (g["psi_index"], g["view"], g["cx"], g["cy"]) == w[:4]
        assert g["score"] == pytest.approx(w[4], rel=1e-10)
    if l == 48:
        # ragged last run: 48^2 = 2 * 1024 + 256
        assert any(int(w[3] - orc.tile_offsets(ny, l, 16)[0]) >= 0 for w in want)


def test_norm1_brute_force_detects_a_planted_error():
    """The brute force is sensitive to the statistics: using the map mean
    over the map's own pixels instead of over views x psi changes the records."""
    l = 48
    tm = (RNG.standard_normal((3, 16, 16)) + 0.3).astype(np.float32)
    img = RNG.standard_normal((64, 80)).astype(np.float32)
    p = orc.params(l, angpix=1.0, phistep=60.0, kk=3, energy=300, cs=2.7, Highres=4,
                   Lowres=100, diameter=6, threshold=2.4, overlap=16, ampconst=0.1, norm_type=1)
    c = orc.ctf(1.2)
    want, z = norm1_brute_force(p, c, tm, img)
    # a per-map z-score (the norm_type 0 flavour) gives a different list
    zz = (z - z.mean(axis=(3, 4), keepdims=True)) / z.std(axis=(3, 4), keepdims=True)
    assert not np.allclose(zz, z)


# ---------------------------------------------------------------------------
# the in-plane step table, the whitened templates, and the norm_type 0 loop
# and selection as compositions of the stages pinned above
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("step,count", [(5, 72), (10, 36), (3, 120), (7, 52), (0.7, 515),
                                        (30, 12), (360, 1)])
def test_psi_table_is_a_float32_accumulator(step, count):
    """main.cu:1230 steps euler3 by += phistep in float: the count is the
    number of float32 partial sums below 360, and psi i is the i-th sum."""
    acc = np.float32(0)
    sums = []
    while acc < np.float32(360):
        sums.append(acc)
        acc = np.float32(acc + np.float32(step))
    assert orc.num_psi(step) == len(sums) == count
    for i in sorted({0, len(sums) // 2, len(sums) - 1}):
        assert orc.psi_value(step, i) == sums[i]


def test_template_whitened_is_ring_normalised_spectrum():
    """Whitened template = FFT / l^2 of the centred, edge-normalised template
    (divided by its sd over l^2), each pixel divided by the RMS of its ring
    over the half plane x <= l/2; rings outside [0, l/2) are left as they are."""
    l, box = 40, 17
    p = _tp(l)
    tm = (RNG.standard_normal((2, box, box)) + 0.5).astype(np.float32)
    got = orc.template_whitened(p, tm)
    xx, yy, dx, dy, r = ring_index(l, l)
    half = (xx <= l // 2) & (r >= 0) & (r < l // 2)
    for t, g in zip(tm, got):
        pad = np.zeros((l, l))
        s = (l - box) // 2
        pad[s:s + box, s:s + box] = t
        pad /= pad.std()
        F = np.fft.fft2(pad) / (l * l)
        pw = np.bincount(r[half], np.abs(F[half]) ** 2, minlength=l)
        cnt = np.bincount(r[half], minlength=l)
        inner = (r >= 0) & (r < l // 2)
        want = F.copy()
        want[inner] /= np.sqrt(pw[r[inner]] / cnt[r[inner]])
        assert np.abs(g - want).max() < 1e-12 * np.abs(want).max()


def _small_norm0_case():
    l = 40
    box = 14
    tm = (RNG.standard_normal((3, box, box)) + 0.2).astype(np.float32)
    img = RNG.standard_normal((70, 90)).astype(np.float32)
    p = orc.params(l, angpix=1.0, phistep=72.0, kk=3, energy=300, cs=2.7, Highres=4,
                   Lowres=100, diameter=10, threshold=3.0, overlap=12, ampconst=0.1)
    return p, orc.ctf(1.6), tm, img


def test_pick_norm0_is_the_composition_of_its_stages():
    """or_pick_norm0 [psi][tile][view] = per psi, the rotated spectra of the
    prepared templates correlated with each tile spectrum (or_correlate_norm0),
    tiles in y-major order."""
    p, c, tm, img = _small_norm0_case()
    st = orc.pick_norm0(p, c, tm, img)
    tiles = orc.micrograph_prep(p, c, img)
    treal = orc.template_prep(p, c, tm)
    P = orc.num_psi(p.phi_step)
    assert st.shape == (P, len(tiles), len(tm))
    for ip in range(P):
        spec = orc.rotate_normalize(p, treal, orc.psi_value(p.phi_step, ip))
        for t in range(len(tiles)):
            for j in range(len(tm)):
                one = orc.correlate_norm0(p, tiles[t], spec[j])
                assert st[ip, t, j]["pos"] == one["pos"]
                assert st[ip, t, j]["score"] == pytest.approx(one["score"], rel=1e-12)
        many = orc.correlate_many(p, tiles, spec)
        assert np.array_equal(many["pos"].reshape(len(tiles), len(tm)), st[ip]["pos"])


def test_select_norm0_by_brute_force():
    """A map gives a pick when its score reaches the threshold and the peak's
    micrograph position (tile offset + pos % l, pos // l) lies d or more from
    every edge; picks in (psi, tile y-major, view) order."""
    p, c, tm, img = _small_norm0_case()
    st = orc.pick_norm0(p, c, tm, img)
    ny, nx = img.shape
    l = p.l
    ox = orc.tile_offsets(nx, l, p.overlap)
    oy = orc.tile_offsets(ny, l, p.overlap)
    d = p.diameter
    want = []
    for ip in range(st.shape[0]):
        for t in range(st.shape[1]):
            iy, ix = divmod(t, len(ox))
            for j in range(st.shape[2]):
                s = st[ip, t, j]
                cx, cy = ox[ix] + s["pos"] % l, oy[iy] + s["pos"] // l
                if s["score"] >= p.threshold and d <= cx < nx - d and d <= cy < ny - d:
                    want.append((ip, j, cx, cy))
    got = orc.select_norm0(p, st, 0, nx, ny)
    assert 3 < len(want) < st.size
    assert [(g["psi_index"], g["view"], g["cx"], g["cy"]) for g in got] == want
