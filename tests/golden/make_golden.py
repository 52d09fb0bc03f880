"""Writes the golden fixtures of tests/golden/ from the reference itself.

Each fixture is one run of oracle/_ref/ref_serial: the unmodified reference
sources (serial, -O0) driven through the reference's per-iteration call
sequence (src/main.c:57-92) with fixed seeds.  Stored per case:
  - sha256 digests of every generated problem array (the product's
    MOC_GEN_REFERENCE generator must reproduce them);
  - per-iteration segments_processed, keff and source residual;
  - fine_flux and z heights after the first sweep;
  - final fine_flux, fine_source, z heights, forward psi and leakage;
  - for the two-way cases (each sweep is the reference's
    two_way_transport_sweep, src/solver.c:556-891), backward psi too.
    Their inputs have a whole-number fine interval height (height /
    decomp / (cai * fai) cm): with a fractional one the reference's
    upward interval (int)z / fine_delta_z (src/solver.c:671) can lag by
    one, giving negative piece lengths and exponential-table reads before
    the table's start (src/solver.c:1447), i.e. undefined values;
  - for the loaded-track cases, 2D tracks from tiny.tracks, an OpenMOC
    track file this script writes (seeded, tests/openmoc_files.py) and the
    reference reads with its own loader (src/tracks.c:170-323).

Usage: python tests/golden/make_golden.py   (needs the reference tree to
build oracle/_ref; the fixtures themselves carry no reference code)
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

sys.path.insert(0, os.path.dirname(HERE))
from oracle import pyoracle as po  # noqa: E402
from openmoc_files import write_random_track_file  # noqa: E402

TRACK_FILE = "tiny.tracks"
TRACK_SEED = 41

# (name, input file, gen_seed, sweep_seed, iterations, two-way sweep, track file)
CASES = [
    ("tiny_quad", "tiny.in", 7, 11, 3, False, None),
    ("tiny_flat", "tiny_flat.in", 5, 13, 3, False, None),
    ("tiny_g16_decomp2", "tiny_g16.in", 9, 17, 2, False, None),
    ("two_way_quad", "tiny_tw.in", 21, 23, 2, True, None),
    ("two_way_flat", "tiny_tw_flat.in", 25, 27, 2, True, None),
    ("tracks_quad", "tiny_tw.in", 43, 47, 2, False, TRACK_FILE),
    ("tracks_two_way", "tiny_tw.in", 53, 59, 2, True, TRACK_FILE),
]

PROBLEM_KEYS = ["az_weight", "seg_offset", "seg_length", "polar_angle", "p_weight",
                "z_height0", "material", "vol", "fine_source0", "sigT", "xs", "scatter", "table"]


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True,
                   stdout=subprocess.DEVNULL)
    write_random_track_file(os.path.join(HERE, TRACK_FILE), seed=TRACK_SEED)
    index = {}
    for name, inp, gs, ss, iters, two_way, tracks in CASES:
        with tempfile.TemporaryDirectory() as td:
            d = po.run_ref_serial(os.path.join(HERE, inp), gs, ss, iters,
                                  os.path.join(td, "dump.bin"), two_way=two_way,
                                  tracks=os.path.join(HERE, tracks) if tracks else None)
        extra = dict(b_psi=d["b_psi"]) if two_way else {}
        np.savez_compressed(
            os.path.join(HERE, f"{name}.npz"),
            segments=d["segments"], keff=d["keff"], residual=d["residual"],
            flux_after_sweep0=d["flux_after_sweep0"], z_after_sweep0=d["z_after_sweep0"],
            fine_flux=d["fine_flux"], fine_source=d["fine_source"], z_height=d["z_height"],
            f_psi=d["f_psi"], leakage=d["leakage"], **extra)
        index[name] = dict(input=inp, gen_seed=gs, sweep_seed=ss, iterations=iters,
                           two_way=two_way, tracks=tracks,
                           problem_sha256={k: digest(d[k]) for k in PROBLEM_KEYS},
                           dims=[int(x) for x in d["dims"]])
        print(name, "segments", d["segments"].tolist(), "keff", d["keff"].tolist())
    with open(os.path.join(HERE, "golden_index.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
