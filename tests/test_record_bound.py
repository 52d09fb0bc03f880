"""Upper bound on the pieces of a sweep (the batched path sizes its piece
records with it once, for every later sweep).

The oracle's axial trace (a restatement of src/solver.c:366-527, pinned
bit-exact to the reference) gives the actual pieces of every stacked track;
over several sweeps, whose z states drift through the reference's
bookkeeping, no track may exceed the bound of its stack.  The shapes cover
one and two fine intervals, steep and shallow polar angles, long tracks
(hundreds of 2D segments) and z separations off the plane grid.
"""
import numpy as np
import pytest

import simplemoc_amd as sm
from helpers import make_input, oracle_for

CASES = [
    dict(fai=1, cai=3, n_polar_angles=2, segments_per_track=40),
    dict(fai=2, cai=2, n_polar_angles=4, axial_z_sep=0.37),
    dict(fai=5, cai=2, n_polar_angles=8, height=7.3, axial_z_sep=0.29),
    dict(fai=3, cai=6, n_polar_angles=6, height=3.0, axial_z_sep=0.11, axial_exp=0),
    dict(fai=30, cai=1, n_polar_angles=4, height=10.0, axial_z_sep=0.5, decomp_assemblies_ax=2),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_pieces_within_bound_over_sweeps(case):
    kw = dict(n_azimuthal=4, radial_ray_sep=1.5, n_2D_source_regions_per_assembly=50)
    kw.update(CASES[case])
    inp = make_input(**kw)
    seg = dict(seg_dist="uniform", seg_lo=5, seg_hi=200) if case % 2 else {}
    prob = sm.Problem.generate(inp, seed=30 + case, **seg)
    per, total = prob.record_bounds()
    o = oracle_for(prob, psi=False)
    d = prob.dims
    nt2, npol, zs = d["ntracks_2D"], d["n_polar"], d["z_stacked"]
    worst = 0.0
    for sweep in range(3):
        S = 0
        for i in range(nt2):
            for j in range(npol):
                rec, z = o.trace_stack(i, j)
                s0 = (i * npol + j) * zs
                counts = np.bincount(rec["track"] - s0, minlength=zs)
                assert counts.max() <= per[i * npol + j], (sweep, i, j, counts.max(), per[i * npol + j])
                worst = max(worst, counts.max() / per[i * npol + j])
                S += len(rec)
                o.a["z_height"][s0:s0 + zs] = z
        assert S <= total
    print(f"case {case}: bound total {total}, last sweep {S} pieces, worst track at "
          f"{worst:.2f} of its bound")
