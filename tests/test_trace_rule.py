"""The z-stack bookkeeping rule of the device trace, against the serial loop.

src/solver.c:409-525 walks the stacked tracks k = begin..end-1 of a 2D
segment one after another and changes the bounds while it runs: going up,
every exit does end-- before the next k is tested (so the loop stops
early); going down, begin++ only affects later segments.  The device trace
(trace_stack in simplemoc_amd/csrc/moc_sweep.h) traces every active track
of the segment at once and decides afterwards which of them the serial
loop reaches: going up, track k is reached iff k < end - E(k), E(k) the
exits among the active tracks below k, a test that fails monotonically,
so the reached tracks are the prefix [begin, kstop) found with one block
scan and one block minimum; end then drops by the exits of the reached
tracks.  Going down every active track is reached and begin grows by all
exits.  Here both are restated in Python and compared on random exit
patterns (pure host logic; the kernels themselves are checked against the
oracle's trace on the GPU).
"""
import numpy as np


def serial(exits, begin, end, up):
    reached = []
    k = begin
    while k < end:
        reached.append(k)
        if exits[k]:
            if up:
                end -= 1
            else:
                begin += 1
        k += 1
    return reached, begin, end


def scan_rule(exits, begin, end, up):
    active = list(range(begin, end))
    if not up:
        return active, begin + int(sum(exits[k] for k in active)), end
    e = 0
    kstop = None
    for k in active:                       # exclusive prefix of exits
        if k >= end - e and kstop is None:
            kstop = k
        e += exits[k]
    if kstop is None:
        kstop = end
    reached = [k for k in active if k < kstop]
    return reached, begin, end - int(sum(exits[k] for k in reached))


def test_rule_matches_serial_loop_on_random_patterns():
    rng = np.random.default_rng(20)
    for case in range(200_000):
        zs = int(rng.integers(1, 40))
        begin = int(rng.integers(0, zs))
        end = int(rng.integers(begin, zs + 1))
        p = rng.choice([0.02, 0.1, 0.3, 0.7])
        exits = (rng.random(zs) < p).astype(int)
        up = bool(rng.integers(0, 2))
        assert serial(exits, begin, end, up) == scan_rule(exits, begin, end, up), \
            (case, zs, begin, end, up, exits.tolist())


def test_rule_over_whole_stacks():
    """Several segments in a row, bounds carried from one to the next."""
    rng = np.random.default_rng(21)
    for case in range(5_000):
        zs = int(rng.integers(1, 64))
        up = bool(rng.integers(0, 2))
        b1 = b2 = 0
        e1 = e2 = zs
        for seg in range(int(rng.integers(1, 30))):
            exits = (rng.random(zs) < 0.15).astype(int)
            r1, b1, e1 = serial(exits, b1, e1, up)
            r2, b2, e2 = scan_rule(exits, b2, e2, up)
            assert (r1, b1, e1) == (r2, b2, e2), (case, seg)
