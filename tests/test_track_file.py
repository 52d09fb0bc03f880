"""OpenMOC track files (src/tracks.c:170-323): the library's reader, the
input changes loading makes, and problems built on loaded tracks against
the reference's own loader.

The reference ships no track file; the files here are written by
tests/openmoc_files.py (seeded tracks across an assembly, segments
summing to the chord, random material and region ids).  The reference's
loader runs in oracle/_ref/ref_serial (its build_tracks with load_tracks
set, as its -d option does); those comparisons are skipped where the
harness was not built.
"""
import os

import numpy as np
import pytest

import simplemoc_amd as sm
from helpers import PROBLEM_KEYS, digest, oracle_for
from openmoc_files import write_random_track_file, write_track_file
from oracle import pyoracle as po
from test_oracle_refdiff import BASE, write_input


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("cmfd", [False, True])
def test_reader_returns_what_was_written(cmfd, tmp_path):
    path = tmp_path / "t.tracks"
    per_azim = write_random_track_file(path, seed=3, cmfd=cmfd)
    tf = sm.TrackFile(path, cmfd=cmfd)
    flat = [t for ts in per_azim for t in ts]
    assert tf.geometry == "random tracks, seed 3"
    assert tf.n_azimuthal == len(per_azim) and tf.spacing == 0.1
    assert tf.num_tracks.tolist() == [len(ts) for ts in per_azim]
    assert tf.ntracks_2D == len(flat)
    np.testing.assert_array_equal(tf.endpoints, np.array([t[:5] for t in flat]))
    np.testing.assert_array_equal(tf.azim_index,
                                  np.concatenate([[a] * len(ts) for a, ts in enumerate(per_azim)]))
    counts = [len(t[5]) for t in flat]
    np.testing.assert_array_equal(np.diff(tf.seg_offset), counts)
    segs = [s for t in flat for s in t[5]]
    assert tf.n_seg2d == len(segs)
    # lengths are narrowed to float as the reference stores them (:290)
    np.testing.assert_array_equal(tf.seg_length, np.array([s[0] for s in segs], np.float32))
    np.testing.assert_array_equal(tf.material_id, [s[1] for s in segs])
    np.testing.assert_array_equal(tf.region_id, [s[2] for s in segs])
    tf.close()


def test_empty_tracks_and_empty_geometry(tmp_path):
    path = tmp_path / "e.tracks"
    write_track_file(path, [[(0, 0, 1, 1, 0.5, [])], [(0, 0, 2, 2, 1.0, [(2.5, 1, 7)])]],
                     geometry="")
    tf = sm.TrackFile(path)
    assert tf.geometry == "" and tf.ntracks_2D == 2 and tf.n_seg2d == 1
    assert tf.seg_offset.tolist() == [0, 0, 1]
    assert tf.region_id.tolist() == [7]


def test_truncated_files_are_errors(tmp_path):
    path = tmp_path / "t.tracks"
    write_random_track_file(path, seed=5, n_azim=2, tracks_per_azim=(2, 2), mean_segments=3)
    data = path.read_bytes()
    cut = tmp_path / "cut.tracks"
    # inside the header, the azimuthal tables, a track header, a segment, one byte short
    for n in (2, 10, 30, 70, 100, len(data) - 20, len(data) - 1):
        cut.write_bytes(data[:n])
        with pytest.raises(sm.MOCError, match="ends inside|too short"):
            sm.TrackFile(cut)


@pytest.mark.parametrize("case", ["string", "angles", "tracks", "segments"])
def test_bad_counts_are_errors(case, tmp_path):
    import struct
    path = tmp_path / "b.tracks"
    write_track_file(path, [[(0, 0, 1, 1, 0.5, [(1.0, 0, 0)])]], geometry="g")
    b = bytearray(path.read_bytes())
    # offsets: len(4) 'g'(1) A(4) spacing(8) tracks(4) x(4) y(4) w(8) | track: 5d(40) azim(4) n(4)
    at = {"string": 0, "angles": 5, "tracks": 17, "segments": 37 + 44}[case]
    b[at:at + 4] = struct.pack("<i", -1 if case != "angles" else 0)
    path.write_bytes(bytes(b))
    with pytest.raises(sm.MOCError):
        sm.TrackFile(path)


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(sm.MOCError, match="cannot open"):
        sm.TrackFile(tmp_path / "nope.tracks")


def test_apply_track_file_to_input(tmp_path):
    path = tmp_path / "t.tracks"
    write_random_track_file(path, seed=7)
    tf = sm.TrackFile(path)
    inp = sm.Input(**{k: v for k, v in BASE.items() if k != "papi_event_set"})
    inp.derive()
    zs, npol = inp.z_stacked, inp.n_polar_angles
    inp.apply_track_file(tf)
    assert inp.n_azimuthal == tf.n_azimuthal          # not halved (:186)
    assert inp.radial_ray_sep == np.float32(tf.spacing)
    assert inp.ntracks_2D == tf.ntracks_2D
    assert inp.segments_per_track == tf.n_seg2d // tf.ntracks_2D   # (:311)
    assert inp.ntracks == tf.ntracks_2D * npol * zs
    prob = sm.Problem.generate(inp, seed=1, track_file=tf)
    np.testing.assert_array_equal(prob.array("seg_offset"), tf.seg_offset)
    np.testing.assert_array_equal(prob.array("seg_length"), tf.seg_length)
    # a track file that does not match the input's track count
    inp2 = sm.Input(**{k: v for k, v in BASE.items() if k != "papi_event_set"})
    inp2.derive()
    with pytest.raises(sm.MOCError, match="2D tracks"):
        sm.Problem.generate(inp2, seed=1, track_file=tf)


@pytest.mark.skipif(not os.path.exists(po.REF_SERIAL),
                    reason="reference harness oracle/_ref/ref_serial not built")
@pytest.mark.parametrize("two_way,empty_every", [(False, 0), (True, 0), (False, 3), (True, 3)])
def test_loaded_tracks_match_reference_loader(two_way, empty_every, tmp_path, libc):
    """The reference's build_tracks with its loader and the product's
    reference-mode generator on the same file and seed give the same
    problem bit for bit; the oracle then reproduces the reference's
    iterations on it bit for bit (one-way and two-way sweeps, and a file in
    which every third 2D track has no segments)."""
    tpath = tmp_path / "t.tracks"
    write_random_track_file(tpath, seed=11, empty_every=empty_every)
    d = dict(BASE, height=10.0)        # whole-number fine interval height
    ipath = tmp_path / "case.in"
    write_input(ipath, d)
    gen_seed, sweep_seed, iters = 31, 37, 2
    ref = po.run_ref_serial(ipath, gen_seed, sweep_seed, iters, tmp_path / "dump.bin",
                            two_way=two_way, tracks=str(tpath))
    tf = sm.TrackFile(tpath)
    inp = sm.Input.from_file(ipath)
    inp.derive()
    inp.apply_track_file(tf)
    dims = ref["dims"]
    assert (dims[0], dims[9], dims[11], dims[13]) == (inp.ntracks_2D, inp.ntracks,
                                                      inp.n_azimuthal, inp.segments_per_track)
    assert ref["fdims"][7] == np.float32(inp.radial_ray_sep)
    prob = sm.Problem.generate(inp, seed=gen_seed, mode="reference", alloc_psi=True,
                               alloc_b_psi=two_way, track_file=tf)
    arrs = prob.arrays
    for ours, theirs in PROBLEM_KEYS.items():
        assert digest(arrs[ours]) == digest(ref[theirs]), ours
    o = oracle_for(prob, qsr_mode=po.QSR_LIBC, tally=po.TALLY_F32, reduce=po.REDUCE_F32)
    libc.srand(sweep_seed)
    out = o.iterate(iters, two_way=two_way)
    assert [x[0] for x in out] == ref["segments"].tolist()
    assert bits_equal(np.array([x[1] for x in out], np.float32), ref["keff"])
    keys = ["fine_flux", "fine_source", "z_height", "f_psi"] + (["b_psi"] if two_way else [])
    for k in keys:
        assert bits_equal(o[k], ref[k]), k
