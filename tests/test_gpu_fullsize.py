"""GPU parity at the benchmark workloads' full sizes.

* one domain of a full core (17x17 assemblies, 400 cm / 20 axial domains;
  19,386 2D tracks x 10 polar x 80 stacked, G=128): after
  one sweep the angular flux of sampled tracks equals, bit for bit, the
  oracle's attenuation recurrence applied to the oracle's own trace of the
  same stack (psi does not depend on the tally order, so this is exact),
  in both tally modes;
* the same shape, two-way sweep: after one sweep the forward and backward
  angular flux of sampled tracks equal, bit for bit, the oracle's two-way
  pass over the same stack (src/solver.c:632-855), and z is back at its
  starting heights;
* the small core in full (2x2 assemblies, 20 cm; 4,846 2D tracks, G=8): two power iterations
  against the oracle in the same modes.
"""
import ctypes as C
import math

import numpy as np
import pytest

import simplemoc_amd as sm
from helpers import SMALL_CORE, FULL_CORE_DOMAIN, flux_errors, make_input, oracle_for
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tally", ["f32", "f64"])
def test_full_core_domain_sampled_track_psi_bitwise(tally):
    inp = make_input(**FULL_CORE_DOMAIN)
    prob = sm.Problem.generate(inp, seed=2)
    d = prob.dims
    G, zs, npol = d["G"], d["z_stacked"], d["n_polar"]
    o = oracle_for(prob, psi=False)
    z0 = prob.array("z_height").copy()
    azw = prob.array("az_weight")
    polar = prob.array("polar_angle")
    rng = np.random.default_rng(11)
    with sm.Solver(prob, sm.Options(tally=tally)) as s:
        s.transport_sweep()
        checked = 0
        for _ in range(5):
            i = int(rng.integers(d["ntracks_2D"]))
            j = int(rng.integers(npol))
            base = (i * npol + j) * zs
            rec, _ = o.trace_stack(i, j, z0[base:base + zs])
            mu = np.float32(math.cos(float(polar[j])))   # libm cos, as the oracle
            for k in rng.integers(zs, size=3):
                t = base + int(k)
                psi = np.zeros(G, np.float32)
                for pc in rec[rec["track"] == t]:
                    po.lib().orc_attenuate(C.byref(o.p), C.byref(o.opts), t, pc["zh"], pc["ds"],
                                           int(pc["qsr"]), mu, azw[i],
                                           psi.ctypes.data_as(C.POINTER(C.c_float)), None, None)
                got = s.track_psi(t)
                np.testing.assert_array_equal(got, psi)
                checked += 1
    assert checked == 15


def test_full_core_domain_two_way_sampled_tracks_bitwise():
    inp = make_input(**FULL_CORE_DOMAIN)
    prob = sm.Problem.generate(inp, seed=2)
    d = prob.dims
    G, zs, npol = d["G"], d["z_stacked"], d["n_polar"]
    o = oracle_for(prob, psi=False)
    z0 = prob.array("z_height").copy()
    rng = np.random.default_rng(13)
    zero = np.zeros(zs * G, np.float32)
    with sm.Solver(prob, sm.Options(tally="f32")) as s:
        segs = s.two_way_sweep()
        assert segs > 0 and segs % 2 == 0
        checked = 0
        for _ in range(5):
            i = int(rng.integers(d["ntracks_2D"]))
            j = int(rng.integers(npol))
            base = (i * npol + j) * zs
            n, z_end, fpsi, bpsi = o.two_way_stack(i, j, z0[base:base + zs], zero, zero, 0)
            assert n > 0
            np.testing.assert_array_equal(z_end, z0[base:base + zs])
            for k in rng.integers(zs, size=3):
                t = base + int(k)
                np.testing.assert_array_equal(s.track_psi(t), fpsi[k * G:(k + 1) * G])
                np.testing.assert_array_equal(s.track_psi(t, backward=True), bpsi[k * G:(k + 1) * G])
                checked += 1
        st = s.state(psi=False)
    np.testing.assert_array_equal(st["z_height"], z0)
    assert checked == 15


def test_small_core_full_shape_matches_oracle():
    inp = make_input(**SMALL_CORE)
    prob = sm.Problem.generate(inp, seed=1)
    o = oracle_for(prob)
    with sm.Solver(prob, sm.Options(tally="f64")) as s:
        out_g = s.iterate(2)
        st = s.state()
    out_o = o.iterate(2)
    assert [x[0] for x in out_g] == [x[0] for x in out_o]
    np.testing.assert_array_equal(st["z_height"], o["z_height"])
    for (_, kg, _), (_, ko, _) in zip(out_g, out_o):
        assert abs(kg - ko) <= 1e-6 * max(1.0, abs(ko))
    for k in ("fine_flux", "fine_source", "f_psi"):
        norm, _ = flux_errors(st[k], o[k], G=8 if k != "f_psi" else None)
        assert norm <= 1e-6, (k, norm)
