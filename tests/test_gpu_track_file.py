"""Sweeps over 2D tracks loaded from OpenMOC track files
(src/tracks.c:170-323) on the GPU.

* Golden runs (tests/golden/tracks_*.npz: the reference's own loader on
  tests/golden/tiny.tracks, libc draws, FP32): GPU in FP64 tally to the
  oracle's FP64 build within 1e-6, to the reference's FP32 run within
  1e-5 (flux, psi) and k_eff within 1e-5; one-way and two-way sweeps.
* Hash draws on a freshly written track file: GPU to the oracle in the
  same modes within 1e-6.
* Whole programs with -d <file>: the reference's own main
  (oracle/_ref/ref_main_serial), the same main linked against the library
  (oracle/_ref/ref_main_dropin) and the library's C driver print the same
  k_eff (within 1e-5).
"""
import os
import re
import subprocess

import numpy as np
import pytest

import simplemoc_amd as sm
from helpers import GOLDEN, ROOT, flux_errors, make_input, oracle_for
from openmoc_files import write_random_track_file
from oracle import pyoracle as po
from test_oracle_golden import problem_of

pytestmark = pytest.mark.gpu

NEAR = 1e-6
REF_TOL = 1e-5
REF_MAIN = os.path.join(ROOT, "oracle", "_ref", "ref_main_serial")
DROPIN_MAIN = os.path.join(ROOT, "oracle", "_ref", "ref_main_dropin")
DRIVER = os.path.join(ROOT, "simplemoc_amd", "_lib", "simplemoc")


@pytest.mark.parametrize("case", ["tracks_quad", "tracks_two_way"])
def test_loaded_tracks_match_reference_golden_run(case, golden_index, libc):
    meta, prob = problem_of(case, golden_index)
    g = np.load(os.path.join(GOLDEN, f"{case}.npz"))
    two_way = bool(meta["two_way"])
    G = prob.dims["G"]
    iters = meta["iterations"]
    print(f"{case}: {prob.dims['ntracks_2D']} loaded 2D tracks", flush=True)
    with sm.Solver(prob, sm.Options(qsr="libc")) as s:
        libc.srand(meta["sweep_seed"])
        out = s.iterate(iters, two_way=two_way)
        st = s.state(b_psi=two_way)
    o64 = oracle_for(prob, qsr_mode=po.QSR_LIBC, tally=po.TALLY_F64, reduce=po.REDUCE_F64)
    libc.srand(meta["sweep_seed"])
    out64 = o64.iterate(iters, two_way=two_way)
    assert [x[0] for x in out] == g["segments"].tolist()
    np.testing.assert_array_equal(st["z_height"], g["z_height"])
    for (_, k, _), kr, (_, k64, _) in zip(out, g["keff"], out64):
        assert abs(k - kr) <= REF_TOL, (k, kr)
        assert abs(k - k64) <= NEAR * max(1.0, abs(k64)), (k, k64)
    for key in ["fine_flux", "f_psi"] + (["b_psi"] if two_way else []):
        n64, _ = flux_errors(st[key], o64[key], G=G if key == "fine_flux" else None)
        nref, _ = flux_errors(st[key], g[key], G=G if key == "fine_flux" else None)
        print(f"{case} {key}: vs FP64 build {n64:.2e}, vs FP32 reference {nref:.2e}")
        assert n64 <= NEAR, (key, n64)
        assert nref <= REF_TOL, (key, nref)


@pytest.mark.parametrize("two_way,empty_every", [(False, 0), (True, 0), (False, 4), (True, 4)])
def test_loaded_tracks_hash_mode_matches_oracle(two_way, empty_every, tmp_path):
    # empty_every = 4: every fourth 2D track has no segments
    path = tmp_path / "t.tracks"
    write_random_track_file(path, seed=19, n_azim=12, mean_segments=30, empty_every=empty_every)
    tf = sm.TrackFile(path)
    inp = make_input(n_egroups=32, height=10.0)
    inp.apply_track_file(tf)
    prob = sm.Problem.generate(inp, seed=5, alloc_b_psi=two_way, track_file=tf)
    o = oracle_for(prob, seed=1)
    with sm.Solver(prob) as s:
        out_g = s.iterate(2, two_way=two_way)
        st = s.state(b_psi=two_way)
    out_o = o.iterate(2, two_way=two_way)
    assert [x[0] for x in out_g] == [x[0] for x in out_o]
    np.testing.assert_array_equal(st["z_height"], o["z_height"])
    for (_, kg, _), (_, ko, _) in zip(out_g, out_o):
        assert abs(kg - ko) <= NEAR * max(1.0, abs(ko)), (kg, ko)
    for key in ["fine_flux", "fine_source", "f_psi"] + (["b_psi"] if two_way else []):
        norm, _ = flux_errors(st[key], o[key], G=32 if key in ("fine_flux", "fine_source") else None)
        assert norm <= NEAR, (key, norm)


def keff_of(cmd, env=None):
    out = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    m = re.findall(r"keff = ([-+0-9.eEnaNinf]+)", out.stdout)
    assert m, out.stdout[-2000:]
    return float(m[-1])


@pytest.mark.skipif(not (os.path.exists(REF_MAIN) and os.path.exists(DROPIN_MAIN)),
                    reason="reference main harnesses not built")
def test_whole_programs_with_track_file():
    """-d <file> through the reference's main, the relinked main and the C
    driver.  The reference's main seeds rand() with time(NULL) = 1000
    (pinned by oracle/time_shim.c) and runs one iteration; the C driver
    gets the same seed for the reference generator and the same draws."""
    inp = os.path.join(GOLDEN, "tiny_tw.in")
    tracks = os.path.join(GOLDEN, "tiny.tracks")
    k_ref = keff_of([REF_MAIN, "-i", inp, "-d", tracks])
    env = dict(os.environ, SIMPLEMOC_QSR="libc", SIMPLEMOC_DROPIN_OPENMP="0")
    k_dropin = keff_of([DROPIN_MAIN, "-i", inp, "-d", tracks], env)
    k_drv = keff_of([DRIVER, "-i", inp, "-d", tracks, "-g", "reference", "-S", "1000", "-q", "libc",
                     "-n", "1"])
    print(f"-d tiny.tracks: reference {k_ref:.6f}, drop-in {k_dropin:.6f}, driver {k_drv:.6f}")
    for k in (k_dropin, k_drv):
        assert abs(k - k_ref) <= 2e-6 + REF_TOL * abs(k_ref), (k, k_ref)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="C driver not built")
def test_driver_two_way_on_loaded_tracks():
    """The C driver's -w (two-way sweeps) on -d tracks: completes, counts
    2 x pieces for the quadratic source, and its first k_eff equals the
    library's own two-way iteration on the same problem (the C driver seeds
    the reference generator with -S and draws by hash with seed 1)."""
    inp_path = os.path.join(GOLDEN, "tiny_tw.in")
    tracks = os.path.join(GOLDEN, "tiny.tracks")
    out = subprocess.run([DRIVER, "-i", inp_path, "-d", tracks, "-g", "reference", "-S", "77", "-w",
                          "-n", "2"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    ks = [float(x) for x in re.findall(r"keff = ([-+0-9.eEnaNinf]+)", out.stdout)]
    segs = [int(x) for x in re.findall(r"3D segments (\d+)", out.stdout)]
    assert len(ks) == 2 and len(segs) == 2
    tf = sm.TrackFile(tracks)
    inp = sm.Input.from_file(inp_path)
    inp.derive()
    inp.apply_track_file(tf)
    prob = sm.Problem.generate(inp, seed=77, mode="reference", track_file=tf)
    with sm.Solver(prob) as s:
        lib_out = s.iterate(2, two_way=True)
    assert segs == [x[0] for x in lib_out]
    assert segs[0] % 2 == 0
    for k, (_, kl, _) in zip(ks, lib_out):
        assert abs(k - kl) <= 1e-6 * max(1.0, abs(kl)), (k, kl)
