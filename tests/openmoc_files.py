"""Writers of OpenMOC track files (the format load_OpenMOC_tracks reads,
src/tracks.c:170-323) for tests and fixtures.

Layout, native little-endian: i32 geometry string length, the string;
i32 azimuthal angles A, f64 track spacing; i32 tracks[A], i32 num_x[A],
i32 num_y[A], f64 azimuthal weights[A]; per track, angle by angle: f64 x0,
y0, x1, y1, phi, i32 azimuthal index, i32 segment count; per segment f64
length, i32 material, i32 region (+ i32, i32 mesh surfaces with CMFD).
"""
from __future__ import annotations

import math
import struct

import numpy as np


def write_track_file(path, per_azim, spacing=0.05, geometry="Geometry", cmfd=False,
                     num_x=None, num_y=None, azim_weight=None):
    """per_azim[a] = list of tracks of azimuthal angle a; a track is
    (x0, y0, x1, y1, phi, segments) with segments = [(length, material,
    region), ...] (or (length, material, region, fwd, bwd) with cmfd)."""
    A = len(per_azim)
    num_tracks = [len(ts) for ts in per_azim]
    num_x = num_x if num_x is not None else num_tracks
    num_y = num_y if num_y is not None else [0] * A
    azim_weight = azim_weight if azim_weight is not None else [1.0 / max(A, 1)] * A
    g = geometry.encode()
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(g)))
        f.write(g)
        f.write(struct.pack("<id", A, spacing))
        f.write(struct.pack(f"<{A}i", *num_tracks))
        f.write(struct.pack(f"<{A}i", *num_x))
        f.write(struct.pack(f"<{A}i", *num_y))
        f.write(struct.pack(f"<{A}d", *azim_weight))
        for a, ts in enumerate(per_azim):
            for (x0, y0, x1, y1, phi, segs) in ts:
                f.write(struct.pack("<5dii", x0, y0, x1, y1, phi, a, len(segs)))
                for sg in segs:
                    f.write(struct.pack("<dii", sg[0], sg[1], sg[2]))
                    if cmfd:
                        fwd, bwd = (sg[3], sg[4]) if len(sg) > 3 else (-1, -1)
                        f.write(struct.pack("<ii", fwd, bwd))


def random_tracks(seed, n_azim=8, tracks_per_azim=(6, 14), mean_segments=20, width=21.42,
                  n_regions=200, n_materials=8, empty_every=0):
    """Seeded straight-line tracks across a width x width square, each cut
    into segments whose lengths sum to the chord (per_azim for
    write_track_file); with empty_every = k, every k-th track (counting
    from the first) has no segments."""
    rng = np.random.default_rng(seed)
    per_azim = []
    n_made = 0
    for a in range(n_azim):
        phi = math.pi * (a + 0.5) / n_azim
        nt = int(rng.integers(tracks_per_azim[0], tracks_per_azim[1] + 1))
        ts = []
        for _ in range(nt):
            x0, y0 = float(rng.uniform(0, width)), 0.0
            chord = float(rng.uniform(0.1, width * math.sqrt(2)))
            x1, y1 = x0 + chord * math.cos(phi), chord * math.sin(phi)
            ns = max(1, int(rng.poisson(mean_segments)))
            cuts = np.sort(rng.uniform(0, chord, ns - 1))
            lens = np.diff(np.concatenate([[0.0], cuts, [chord]]))
            segs = [(float(L), int(rng.integers(n_materials)), int(rng.integers(n_regions)))
                    for L in lens]
            if empty_every and n_made % empty_every == 0:
                segs = []
            n_made += 1
            ts.append((x0, y0, x1, y1, phi, segs))
        per_azim.append(ts)
    return per_azim


def write_random_track_file(path, seed, cmfd=False, **kw):
    per_azim = random_tracks(seed, **kw)
    if cmfd:
        per_azim = [[(x0, y0, x1, y1, phi, [(L, m, r, r % 4, (r + 1) % 4) for (L, m, r) in segs])
                     for (x0, y0, x1, y1, phi, segs) in ts] for ts in per_azim]
    write_track_file(path, per_azim, spacing=0.1, geometry=f"random tracks, seed {seed}", cmfd=cmfd)
    return per_azim
