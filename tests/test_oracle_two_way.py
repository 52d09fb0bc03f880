"""The oracle's two-way sweep against the reference's own.

src/solver.c:556-891 (two_way_transport_sweep) traces each z-stack forward
with its own interval arithmetic, attenuating into f_psi, then replays
every track's pieces backwards into b_psi with -mu, and resets z.  The
reference harness (oracle/_ref/ref_serial, unmodified sources, serial -O0)
runs it in place of transport_sweep; the oracle in its reference modes
(libc rand() draws, FP32 tally and reductions) must reproduce the run bit
for bit: piece counts, k_eff, residuals, fluxes after the first sweep and
at the end, f_psi, b_psi and z.  Skipped where the harness was not built.
"""
import os

import numpy as np
import pytest

import simplemoc_amd as sm
from helpers import oracle_for
from oracle import pyoracle as po
from test_oracle_refdiff import BASE, write_input

pytestmark = pytest.mark.skipif(not os.path.exists(po.REF_SERIAL),
                                reason="reference harness oracle/_ref/ref_serial not built")

# Every case has a whole-number fine interval height (height / decomp /
# (cai * fai), in cm).  The reference's upward interval, (int)z_height /
# fine_delta_z (src/solver.c:671), is then exact; with a fractional height
# it can lag the true interval by one, the next crossing lies below the
# track and the piece length comes out negative, which sends a negative
# index into the exponential table (src/solver.c:1447): a read before the
# table's start, whose value is whatever memory lies there.  The oracle
# and the GPU clamp that index to the first cell.
VARIANTS = [
    dict(height=10.0),
    dict(height=20.0, axial_exp=0, n_egroups=12),
    dict(n_polar_angles=6, n_egroups=5, cai=3, fai=3, height=9.0),
    dict(decomp_assemblies_ax=3, height=36.0, cai=4, fai=3, n_egroups=24),
    dict(axial_z_sep=0.1, height=4.0, cai=1, fai=4, n_polar_angles=2, segments_per_track=40),
]


@pytest.mark.parametrize("vi", range(len(VARIANTS)))
def test_two_way_oracle_matches_reference(vi, tmp_path, libc):
    d = dict(BASE, **VARIANTS[vi])
    inp_path = tmp_path / "case.in"
    write_input(inp_path, d)
    gen_seed, sweep_seed, iters = 300 + vi, 400 + vi, 2
    ref = po.run_ref_serial(inp_path, gen_seed, sweep_seed, iters, tmp_path / "dump.bin",
                            two_way=True)
    inp = sm.Input.from_file(inp_path)
    inp.derive()
    prob = sm.Problem.generate(inp, seed=gen_seed, mode="reference", alloc_psi=True,
                               alloc_b_psi=True)
    o = oracle_for(prob, qsr_mode=po.QSR_LIBC, tally=po.TALLY_F32, reduce=po.REDUCE_F32)
    libc.srand(sweep_seed)
    segs0 = o.two_way_sweep()
    assert segs0 == ref["segments"][0]
    assert np.array_equal(o["fine_flux"].view(np.uint32), ref["flux_after_sweep0"].view(np.uint32))
    assert np.array_equal(o["z_height"].view(np.uint32), ref["z_after_sweep0"].view(np.uint32))
    o.renormalize_flux()
    r0 = o.update_sources(1.0)
    k0 = o.compute_keff()
    out = [(segs0, k0, r0)] + o.iterate(iters - 1, keff=k0, two_way=True)
    assert [x[0] for x in out] == ref["segments"].tolist()
    assert np.array_equal(np.array([x[1] for x in out], np.float32).view(np.uint32),
                          ref["keff"].view(np.uint32))
    assert np.array_equal(np.array([x[2] for x in out], np.float32).view(np.uint32),
                          ref["residual"].view(np.uint32))
    for k in ("fine_flux", "fine_source", "z_height", "f_psi", "b_psi"):
        assert np.array_equal(o[k].view(np.uint32), ref[k].view(np.uint32)), k
    print(f"variant {vi}: {out[0][0]} counted pieces, keff {out[-1][1]:.6f}")
