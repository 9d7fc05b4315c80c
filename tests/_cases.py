"""Small seeded picking cases shared by the CPU and GPU tests."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from gisspa_amd import synth


@dataclass
class Case:
    name: str
    templates: np.ndarray
    image: np.ndarray
    defocus: float
    params: dict  # config-key names

    @property
    def l(self) -> int:
        return int(self.params["window_size"])


def tiny(seed: int = 0, l: int = 128, box: int = 64, n_views: int = 6, size=(320, 288),
         n_particles: int = 3, phistep: float = 30.0, overlap: int = 32, snr: float = 0.5,
         threshold: float = 5.0) -> Case:
    dens = synth.blob_density(20, box, seed + 11, sigma_range=(1.5, 4.0))
    eul = synth.random_eulers(n_views, seed + 12)
    tm = synth.project_templates(dens, box, eul)
    nx, ny = size
    mg = synth.make_micrograph(nx, ny, dens, eul, n_particles, snr, seed + 13, apix=1.0,
                               defocus_um=1.5, ampconst=0.1, box=box, psi_step=phistep,
                               margin=box)
    params = dict(angpix=1.0, phistep=phistep, kk=3, energy=300, cs=2.7, Highres=8, Lowres=100,
                  diameter=0.75 * box, threshold=threshold, window_size=l, overlap=overlap,
                  ampconst=0.1)
    return Case(f"tiny{seed}_l{l}", tm, mg.image, 1.5, params)


def c1_step5_views() -> np.ndarray:
    """(alt, az) of the 1649 views of the reference's euler_C1_step5.lst
    (tests/golden/euler_C1_step5.npy, written by make_euler_fixture.py)."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "euler_C1_step5.npy")
    return np.load(path)[:, :2].astype(np.float64)
