"""Synthetic inputs for the picker: Gaussian-blob densities, analytic
projections, CTF-modulated micrographs with embedded particles, and the
reference's file formats (MRC stacks, micrograph .lst, Euler .lst, config).

Every generator takes an explicit seed and uses numpy's PCG64, so the same
call gives the same bytes on any machine with the same numpy.

Geometry conventions (chosen to match the picker, reference main.cu and
GPU_func.cu:496-560):
  * a template of box b is centred at pixel (b//2, b//2);
  * rotate_subIMG samples out(q) = in(M q) with M = [[c, s], [-s, c]], so an
    in-plane rotation by e degrees moves a point c to M^T c;
  * a particle embedded with view J and rotation e is found by the picker at
    euler3 = e, reported as psi = 360 - e.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import mrc


# ---------------------------------------------------------------------------
# densities and projections
# ---------------------------------------------------------------------------
@dataclass
class BlobDensity:
    centres: np.ndarray  # (n, 3) in pixels, relative to the box centre
    sigmas: np.ndarray   # (n,)
    amps: np.ndarray     # (n,)


def blob_density(n_blobs: int, box: int, seed: int, sigma_range=(2.0, 6.0)) -> BlobDensity:
    """n Gaussian blobs inside a sphere of radius box/4 (so any projection,
    rotated in-plane, stays well inside the box)."""
    rng = np.random.default_rng(seed)
    r_max = box / 4.0
    pts = []
    while len(pts) < n_blobs:
        p = rng.uniform(-r_max, r_max, size=3)
        if np.dot(p, p) <= r_max * r_max:
            pts.append(p)
    sig = rng.uniform(sigma_range[0], sigma_range[1], size=n_blobs)
    amp = rng.uniform(0.5, 1.0, size=n_blobs)
    return BlobDensity(np.array(pts), sig, amp)


def euler_matrix(alt_deg: float, az_deg: float, phi_deg: float = 0.0) -> np.ndarray:
    """EMAN ZXZ rotation: R = Rz(phi) Rx(alt) Rz(az)."""
    a, b, c = (math.radians(v) for v in (az_deg, alt_deg, phi_deg))

    def rz(t):
        return np.array([[math.cos(t), math.sin(t), 0], [-math.sin(t), math.cos(t), 0], [0, 0, 1]])

    def rx(t):
        return np.array([[1, 0, 0], [0, math.cos(t), math.sin(t)], [0, -math.sin(t), math.cos(t)]])

    return rz(c) @ rx(b) @ rz(a)


def projected_blobs(dens: BlobDensity, alt: float, az: float):
    """2D centres, sigmas and integrated amplitudes of the projection along z."""
    R = euler_matrix(alt, az)
    c3 = dens.centres @ R.T
    amps2 = dens.amps * dens.sigmas * math.sqrt(2 * math.pi)
    return c3[:, :2].copy(), dens.sigmas, amps2


def render_blobs(c2: np.ndarray, sig: np.ndarray, amp: np.ndarray, nx: int, ny: int,
                 cx: float, cy: float, out: np.ndarray | None = None) -> np.ndarray:
    """Add sum_b amp_b exp(-|p - (cx,cy) - c_b|^2 / 2 s_b^2) over pixels p of an
    ny x nx image (separable evaluation)."""
    if out is None:
        out = np.zeros((ny, nx), dtype=np.float64)
    xs = np.arange(nx, dtype=np.float64)
    ys = np.arange(ny, dtype=np.float64)
    for (bx, by), s, a in zip(c2, sig, amp):
        gx = np.exp(-((xs - cx - bx) ** 2) / (2 * s * s))
        gy = np.exp(-((ys - cy - by) ** 2) / (2 * s * s))
        x0 = max(0, int(cx + bx - 5 * s) - 1)
        x1 = min(nx, int(cx + bx + 5 * s) + 2)
        y0 = max(0, int(cy + by - 5 * s) - 1)
        y1 = min(ny, int(cy + by + 5 * s) + 2)
        if x0 >= x1 or y0 >= y1:
            continue
        out[y0:y1, x0:x1] += a * np.outer(gy[y0:y1], gx[x0:x1])
    return out


def render_volume(dens: BlobDensity, box: int) -> np.ndarray:
    """The density on a box^3 grid [z][y][x] (float32), voxel (x, y, z) at
    (x, y, z) - box//2 relative to the centre: sum_b amp_b
    exp(-|p - c_b|^2 / 2 s_b^2).  Its projection along z is project_templates'
    image."""
    vol = np.zeros((box, box, box), dtype=np.float64)
    ax = np.arange(box, dtype=np.float64) - box // 2
    for (cx, cy, cz), s, a in zip(dens.centres, dens.sigmas, dens.amps):
        g = [np.exp(-((ax - c) ** 2) / (2 * s * s)) for c in (cx, cy, cz)]
        vol += a * g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :]
    return vol.astype(np.float32)


def project_templates(dens: BlobDensity, box: int, eulers: np.ndarray) -> np.ndarray:
    """Stack (n, box, box) float32 of projections at the given (alt, az)."""
    out = np.zeros((len(eulers), box, box), dtype=np.float32)
    for i, (alt, az) in enumerate(eulers):
        c2, s, a = projected_blobs(dens, alt, az)
        out[i] = render_blobs(c2, s, a, box, box, box // 2, box // 2).astype(np.float32)
    return out


def random_eulers(n: int, seed: int) -> np.ndarray:
    """n random (alt, az) in degrees, uniform on the sphere."""
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1, 1, n)
    alt = np.degrees(np.arccos(z))
    az = rng.uniform(-180, 180, n)
    return np.stack([alt, az], axis=1)


def sphere_eulers(n: int) -> np.ndarray:
    """n near-uniform (alt, az) over the whole sphere (Fibonacci lattice).
    n = 1649 gives the ~5 degree C1 sampling of a 1649-line Euler file."""
    i = np.arange(n, dtype=np.float64) + 0.5
    z = 1 - 2 * i / n
    alt = np.degrees(np.arccos(z))
    golden = math.pi * (3 - math.sqrt(5))
    az = np.degrees((i * golden) % (2 * math.pi)) - 180.0
    return np.stack([alt, az], axis=1)


# ---------------------------------------------------------------------------
# CTF (physical model for the synthetic signal; the picker's own CTF is the
# reference's CTF_AST, implemented in the HIP library)
# ---------------------------------------------------------------------------
def ctf_2d(nx: int, ny: int, apix: float, defocus_um: float, cs_mm: float, kv: float,
           ampconst: float) -> np.ndarray:
    lam = 12.2639 / math.sqrt(kv * 1000.0 + 0.97845 * kv * kv)
    fx = np.fft.fftfreq(nx, d=apix)
    fy = np.fft.fftfreq(ny, d=apix)
    s2 = fx[None, :] ** 2 + fy[:, None] ** 2
    df = defocus_um * 1e4
    cs = cs_mm * 1e7
    gamma = math.pi * lam * df * s2 - 0.5 * math.pi * cs * lam ** 3 * s2 * s2
    return -(math.sqrt(1 - ampconst ** 2) * np.sin(gamma) + ampconst * np.cos(gamma))


@dataclass
class Particle:
    x: float
    y: float
    view: int
    euler3: float


@dataclass
class Micrograph:
    image: np.ndarray            # (ny, nx) float32
    particles: list = field(default_factory=list)


def make_micrograph(nx: int, ny: int, dens: BlobDensity, view_eulers: np.ndarray,
                    n_particles: int, snr: float, seed: int, apix: float, defocus_um: float,
                    cs_mm: float = 2.7, kv: float = 300.0, ampconst: float = 0.1,
                    box: int = 128, psi_step: float = 10.0, margin: int | None = None,
                    phase_flipped: bool = True) -> Micrograph:
    """Gaussian noise (sigma 1) plus n CTF-modulated particles at SNR `snr`
    (signal variance over noise variance inside one box).  Particle views are
    drawn from view_eulers and in-plane angles from the psi grid, so each
    particle is an exact template match up to interpolation.  With
    phase_flipped (the input the picker expects when phase_flip = 0) the
    signal carries |CTF|; otherwise the signed CTF."""
    rng = np.random.default_rng(seed)
    if margin is None:
        margin = box
    sig = np.zeros((ny, nx), dtype=np.float64)
    parts = []
    npsi = int(math.ceil(360.0 / psi_step - 1e-9))
    for _ in range(n_particles):
        x = float(rng.integers(margin, nx - margin))
        y = float(rng.integers(margin, ny - margin))
        v = int(rng.integers(0, len(view_eulers)))
        e = float(rng.integers(0, npsi)) * psi_step
        c2, s, a = projected_blobs(dens, view_eulers[v][0], view_eulers[v][1])
        ce, se = math.cos(math.radians(e)), math.sin(math.radians(e))
        rot = np.array([[ce, -se], [se, ce]])  # M^T
        c2r = c2 @ rot.T
        render_blobs(c2r, s, a, nx, ny, x, y, out=sig)
        parts.append(Particle(x, y, v, e))
    if n_particles:
        C = ctf_2d(nx, ny, apix, defocus_um, cs_mm, kv, ampconst)
        if phase_flipped:
            C = np.abs(C)
        sig = np.real(np.fft.ifft2(np.fft.fft2(sig) * C))
        # scale: variance of one isolated particle inside its box
        c2, s, a = projected_blobs(dens, view_eulers[0][0], view_eulers[0][1])
        one = render_blobs(c2, s, a, 2 * box, 2 * box, box, box)
        Cb = ctf_2d(2 * box, 2 * box, apix, defocus_um, cs_mm, kv, ampconst)
        if phase_flipped:
            Cb = np.abs(Cb)
        one = np.real(np.fft.ifft2(np.fft.fft2(one) * Cb))[box // 2: box // 2 + box, box // 2: box // 2 + box]
        v1 = float(one.var())
        if v1 > 0:
            sig *= math.sqrt(snr / v1)
    noise = rng.standard_normal((ny, nx))
    img = (sig + noise).astype(np.float32)
    return Micrograph(img, parts)


# ---------------------------------------------------------------------------
# reference file formats
# ---------------------------------------------------------------------------
def write_euler_lst(path: str, eulers: np.ndarray) -> None:
    """'alt az 0' per line, as the reference's Euler_file/*.lst."""
    with open(path, "w") as f:
        for alt, az in eulers:
            f.write(f"{alt:.6f}\t{az:.6f}\t0.0\n")


def write_micrograph_lst(path: str, entries) -> None:
    """#LST file: idx<TAB>path<TAB>defocus=..<TAB>dfdiff=..<TAB>dfang=.. ;
    paths are relative to the lst directory (reference main.cu:125-138)."""
    with open(path, "w") as f:
        f.write("#LST\n")
        for idx, p, df, dfdiff, dfang in entries:
            f.write(f"{idx}\t{p}\tdefocus={df:.6f}\tdfdiff={dfdiff:.6f}\tdfang={dfang:.6f}\n")


def write_config(path: str, **kv) -> None:
    with open(path, "w") as f:
        f.write("# GisSPA picker config\n")
        for k, v in kv.items():
            f.write(f"{k} = {v}\n")


# ---------------------------------------------------------------------------
# named workloads
# ---------------------------------------------------------------------------
@dataclass
class Workload:
    name: str
    templates: np.ndarray        # (N, b, b) float32
    eulers: np.ndarray           # (N, 2)
    micrographs: list            # list of (image, defocus_um)
    params: dict
    particles: list


def small_workload(seed: int = 0, n_views: int = 64, box: int = 128, size: int = 1024,
                   n_particles: int = 20, snr: float = 0.1, window: int = 512,
                   overlap: int = 128, phistep: float = 10.0, defocus: float = 1.5,
                   threshold: float = 7.0) -> Workload:
    """Plain method, small: one size^2 micrograph, N views of a 50-blob
    box^3 density at random Euler angles, 300 kV / Cs 2.7 / A 0.1, 1.5 um."""
    dens = blob_density(50, box, seed + 1)
    eul = random_eulers(n_views, seed + 2)
    tm = project_templates(dens, box, eul)
    mg = make_micrograph(size, size, dens, eul, n_particles, snr, seed, apix=1.0,
                         defocus_um=defocus, ampconst=0.1, box=box, psi_step=phistep)
    params = dict(angpix=1.0, phistep=phistep, kk=3, energy=300, cs=2.7, Highres=8, Lowres=100,
                  diameter=0.75 * box, threshold=threshold, window_size=window, overlap=overlap,
                  ampconst=0.1)
    return Workload("small", tm, eul, [(mg.image, defocus)], params, mg.particles)


def full_workload(seed: int = 0, n_views: int = 1649, box: int = 180, nx: int = 5760,
                  ny: int = 4092, n_particles: int = 300, snr: float = 0.05,
                  window: int = 720, overlap: int = 180, phistep: float = 5.0,
                  defocus: float = 1.8, threshold: float = 7.0,
                  eulers: np.ndarray | None = None) -> Workload:
    """Plain method, full size: 5760x4092, 180^2 templates of a 60-blob 180^3
    density over a 1649-view ~5 degree sphere (the given (alt, az) rows, or a
    Fibonacci lattice of n_views), 72 psi, 1.8 um."""
    dens = blob_density(60, box, seed + 1)
    eul = sphere_eulers(n_views) if eulers is None else np.asarray(eulers, np.float64)[:, :2]
    tm = project_templates(dens, box, eul)
    mg = make_micrograph(nx, ny, dens, eul, n_particles, snr, seed, apix=1.0, defocus_um=defocus,
                         ampconst=0.1, box=box, psi_step=phistep)
    params = dict(angpix=1.0, phistep=phistep, kk=3, energy=300, cs=2.7, Highres=8, Lowres=100,
                  diameter=0.75 * box, threshold=threshold, window_size=window, overlap=overlap,
                  ampconst=0.1)
    return Workload("full", tm, eul, [(mg.image, defocus)], params, mg.particles)


def write_workload(wl: Workload, outdir: str, extra: dict | None = None) -> str:
    """Write templates.mrcs, euler.lst, micrographs, micrographs.lst and a
    config; returns the config path."""
    os.makedirs(outdir, exist_ok=True)
    mrc.write_mrc(os.path.join(outdir, "templates.mrcs"), wl.templates)
    write_euler_lst(os.path.join(outdir, "euler.lst"), wl.eulers)
    entries = []
    for i, (img, df) in enumerate(wl.micrographs):
        name = f"mic{i:03d}.mrc"
        mrc.write_mrc(os.path.join(outdir, name), img[None] if img.ndim == 2 else img)
        entries.append((i, name, df, 0.0, 0.0))
    write_micrograph_lst(os.path.join(outdir, "micrographs.lst"), entries)
    cfg = dict(input=os.path.join(outdir, "micrographs.lst"),
               template=os.path.join(outdir, "templates.mrcs"),
               eulerfile=os.path.join(outdir, "euler.lst"),
               output=os.path.join(outdir, "picks.lst"),
               first=0, last=len(wl.micrographs))
    cfg.update(wl.params)
    if extra:
        cfg.update(extra)
    path = os.path.join(outdir, "config")
    write_config(path, **cfg)
    return path
