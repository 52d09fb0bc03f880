"""Seeded sets of independent fixed-source problems for the batched sweep.

Each problem draws its own shape: energy groups from {8, 32, 64, 128},
1-30 fine axial intervals per coarse interval, 2D segments per track
uniform in [seg_lo, seg_hi] (5-500 for the benchmark set), flat or
quadratic axial source, and a spread of radial track counts and z-stack
sizes, so that track counts, track lengths and group counts are all
uneven across the batch.  Sources are the generator's U(0,1) fine sources
(fixed), sigma_t U(0.1, 2.0), 2D segment lengths U(0.01, 0.5) cm.
"""
from __future__ import annotations

import numpy as np

from . import Input, Problem

GROUPS = (8, 32, 64, 128)


def shapes(n: int = 64, seed: int = 3, seg_lo: int = 5, seg_hi: int = 500,
           radial_ray_sep=(0.6, 1.6), n_regions: int = 400, groups=GROUPS):
    """Input fields of n problems (list of dicts) and their region-draw seeds."""
    rng = np.random.default_rng(seed)
    out = []
    for p in range(n):
        fai = int(rng.integers(1, 31))
        cai = int(rng.integers(1, 4))
        out.append(dict(
            x_assemblies=1, y_assemblies=1, cai=cai, fai=fai,
            axial_exp=int(rng.choice([0, 2, 2])),
            radial_ray_sep=float(rng.uniform(*radial_ray_sep)),
            axial_z_sep=float(rng.choice([0.25, 0.5, 1.0])),
            n_azimuthal=int(rng.choice([4, 8, 16])), n_polar_angles=int(rng.choice([2, 4, 6])),
            n_egroups=int(groups[p % len(groups)]), decompose=0, decomp_assemblies_ax=1,
            segments_per_track=(seg_lo + seg_hi) // 2, assembly_width=21.42,
            height=float(rng.choice([10.0, 20.0, 40.0])), precision=0.01,
            n_2D_source_regions_per_assembly=n_regions))
    return out, [1000 + 7 * p for p in range(n)]


def problems(n: int = 64, seed: int = 3, seg_lo: int = 5, seg_hi: int = 500, **kw):
    """(problems, seeds): n generated problems of shapes(...)."""
    sh, seeds = shapes(n, seed, seg_lo, seg_hi, **kw)
    out = []
    for p, fields in enumerate(sh):
        inp = Input(**fields)
        inp.derive()
        out.append(Problem.generate(inp, seed=seed * 1000 + p, seg_dist="uniform",
                                    seg_lo=seg_lo, seg_hi=seg_hi))
    return out, seeds
