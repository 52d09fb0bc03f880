"""The oracle against what the mathematics fixes.

These checks complement the bit-exact golden runs: each one is chosen so
that a plausible slip in the restatement (a dropped term, a wrong sign, a
transposed operand, an off-by-one in the trace) fails it.

  * exponential table: closed form at the nodes, the reference's sign
    quirk below 0.028 (slope -e^{-n dx}), saturation above maxVal;
  * pairwise sum: sequential below 17 values, the n/2 split above;
  * attenuation: exact solution of dpsi/ds = q - sigma*psi for a flat source
    (intrinsic exponential), zero-source decay, exact scaling by 2,
    superposition in (psi, source);
  * axial trace: piece count = plane crossings + 1 and sum of piece lengths
    = the 3D segment length for interior segments, brute force;
  * reductions: constant fields give closed-form normalisation, k_eff and
    source;
  * region hash: range and uniformity.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

f32 = np.float32


# ---------------------------------------------------------------- table
def test_table_closed_forms():
    vals, N, dx, mv = po.build_table(0.01, 10.0)
    assert N == 353
    assert f32(dx) == f32(10.0) / f32(353)
    assert f32(mv) == f32(10.0) - f32(dx)
    for n in (1, 5, 35, 100, 351):
        x = f32(n) * f32(dx)
        F = po.interpolate(vals, N, dx, mv, x)
        assert abs(F - (1 - math.exp(-float(x)))) < 2e-6, n
    # reference sign quirk: cell 0 has slope -1, intercept 0 -> F(x) = -x
    assert po.interpolate(vals, N, dx, mv, f32(0.02)) == pytest.approx(-0.02, abs=1e-7)
    assert po.interpolate(vals, N, dx, mv, f32(1.0)) == pytest.approx(0.625828, abs=1e-6)
    assert po.interpolate(vals, N, dx, mv, f32(9.99)) == 1.0
    # interior of a cell: the line through the node with slope -e^{-n dx}
    n = 35
    x = f32(n * dx + 0.3 * dx)
    e = math.exp(-n * float(f32(dx)))
    assert po.interpolate(vals, N, dx, mv, x) == pytest.approx(-e * float(x) + 1 + (n * float(f32(dx)) - 1) * e, rel=1e-5)


# ---------------------------------------------------------------- pairwise
def _pairwise_model(v):
    if len(v) <= 16:
        s = f32(0)
        for x in v:
            s = f32(s + x)
        return s
    h = len(v) // 2
    return f32(_pairwise_model(v[:h]) + _pairwise_model(v[h:]))


@pytest.mark.parametrize("n", [0, 1, 7, 16, 17, 33, 100, 1000, 4097])
def test_pairwise_sum(n):
    rng = np.random.default_rng(n)
    v = (rng.standard_normal(n) * 10 ** rng.uniform(-3, 3, n)).astype(np.float32)
    got = po.pairwise_sum(v)
    assert f32(got) == _pairwise_model(list(v))
    if 0 < n <= 16:
        assert f32(got) == f32(np.cumsum(v, dtype=np.float32)[-1])


# ---------------------------------------------------------------- attenuation
def _one_region_problem(G=4, fai=5, axial_exp=2, src_rows=None, sig=None, psi=None):
    """One track, one region: enough to call orc_attenuate."""
    rng = np.random.default_rng(0)
    vals, N, dx, mv = po.build_table(0.01, 10.0)
    nsr = 2
    dims = dict(ntracks_2D=1, n_polar=2, z_stacked=1, G=G, nsr=nsr, fai=fai, cai=1, decomp=1,
                axial_exp=axial_exp, height=5.0, axial_z_sep=1.0, assembly_width=1.0,
                domain_height=5.0, ntracks=2, n_mat=1, table_N=N, table_dx=dx, table_maxVal=mv,
                leakage=0.0)
    src = rng.uniform(0.1, 1, (nsr, fai, G)).astype(np.float32) if src_rows is None else src_rows
    arrays = dict(az_weight=np.ones(1, np.float32), seg_offset=np.array([0, 1]),
                  seg_length=np.ones(1, np.float32), polar_angle=np.array([0.7, 2.4], np.float32),
                  p_weight=np.ones(2, np.float32), z_height=np.zeros(2, np.float32),
                  f_psi=np.zeros(2 * G, np.float32) if psi is None else psi,
                  fine_source=src, fine_flux=np.zeros(nsr * fai * G, np.float32),
                  sigT=(rng.uniform(0.1, 2, (nsr, G)) if sig is None else sig).astype(np.float32),
                  vol=np.ones(nsr, np.float32), material=np.zeros(nsr, np.int32),
                  xs=np.ones(3 * G, np.float32), scatter=np.ones(G * G, np.float32), table=vals)
    return dims, arrays


def _attenuate(dims, arrays, zh, ds, mu, psi0, exp_mode=po.EXP_TABLE, qsr=0):
    import ctypes as C
    o = po.Oracle(dims, arrays, exp_mode=exp_mode)
    G = dims["G"]
    psi = np.array(psi0, np.float32).copy()
    acc = np.zeros(dims["nsr"] * dims["fai"] * G, np.float64)
    po.lib().orc_attenuate(C.byref(o.p), C.byref(o.opts), 0, zh, ds, qsr, mu, 1.0,
                           psi.ctypes.data_as(C.POINTER(C.c_float)), None,
                           acc.ctypes.data_as(C.POINTER(C.c_double)))
    return psi, acc


@pytest.mark.parametrize("axial_exp", [0, 2])
def test_flat_source_matches_exact_transport_solution(axial_exp):
    """Equal source rows (quadratic fit degenerates to q0) or the flat mode:
    psi(s) = psi0 e^{-sigma s} + q/sigma (1 - e^{-sigma s}) and the tally is
    the integral of psi over the piece (times mu in flat mode)."""
    G, fai = 4, 5
    src = np.full((2, fai, G), 0.0, np.float32)
    q = np.array([0.3, 1.2, 0.05, 2.0], np.float32)
    src[:, :, :] = q
    sig = np.array([[0.2, 0.9, 1.7, 0.4]] * 2, np.float32)
    dims, arrays = _one_region_problem(G, fai, axial_exp, src_rows=src, sig=sig)
    psi0 = np.array([0.5, 0.0, 3.0, 1.0], np.float32)
    mu = f32(0.6)
    for ds in (0.05, 0.3, 1.7):
        psi, acc = _attenuate(dims, arrays, f32(2.3), f32(ds), mu, psi0, po.EXP_INTRINSIC)
        s64 = sig[0].astype(np.float64)
        E = np.exp(-s64 * ds)
        exact_psi = psi0 * E + q / s64 * (1 - E)
        # integral of psi over [0, ds]
        integ = psi0 * (1 - E) / s64 + q / s64 * (ds - (1 - E) / s64)
        np.testing.assert_allclose(psi, exact_psi, rtol=2e-5, atol=1e-6)
        fine = int(f32(2.3) / (f32(5.0) / f32(fai))) % fai
        tally = acc.reshape(2, fai, G)[0, fine]
        w = float(mu) if axial_exp == 0 else 1.0
        if axial_exp == 0:
            # flat mode tallies the attenuated flux difference (src/solver.c:1108-1111)
            np.testing.assert_allclose(tally, w * (psi0 - exact_psi), rtol=3e-5, atol=1e-6)
        else:
            np.testing.assert_allclose(tally, integ, rtol=3e-5, atol=1e-6)


def test_zero_source_is_pure_decay_and_scaling_is_exact():
    G = 4
    dims, arrays = _one_region_problem(G, src_rows=np.zeros((2, 5, G), np.float32))
    psi0 = np.array([1.0, 2.0, 0.5, 7.0], np.float32)
    psi, acc = _attenuate(dims, arrays, f32(1.1), f32(0.4), f32(0.3), psi0)
    vals = arrays["table"]
    _, N, dx, mv = po.build_table(0.01, 10.0)
    sig = arrays["sigT"][0]
    F = np.array([po.interpolate(vals, N, dx, mv, f32(s * f32(0.4))) for s in sig], np.float32)
    assert np.array_equal(psi, psi0 * (f32(1) - F))
    # scaling psi and every source row by 2 scales psi and the tally exactly
    dims2, arrays2 = _one_region_problem(G)
    psi_a, acc_a = _attenuate(dims2, arrays2, f32(3.7), f32(0.25), f32(0.8), psi0)
    arrays2b = dict(arrays2)
    arrays2b["fine_source"] = arrays2["fine_source"] * 2
    psi_b, acc_b = _attenuate(dims2, arrays2b, f32(3.7), f32(0.25), f32(0.8), psi0 * 2)
    assert np.array_equal(psi_b, psi_a * 2)
    assert np.array_equal(acc_b, acc_a * 2)


def test_superposition_in_psi_and_source():
    G = 4
    dims, arrays = _one_region_problem(G)
    psi0 = np.array([1.0, 0.2, 3.0, 0.7], np.float32)
    zero_src = dict(arrays, fine_source=np.zeros_like(arrays["fine_source"]))
    pa, aa = _attenuate(dims, arrays, f32(0.9), f32(0.6), f32(0.5), psi0)
    pb, ab = _attenuate(dims, zero_src, f32(0.9), f32(0.6), f32(0.5), psi0)
    pc, ac = _attenuate(dims, arrays, f32(0.9), f32(0.6), f32(0.5), np.zeros(G, np.float32))
    np.testing.assert_allclose(pa, pb + pc, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(aa, ab + ac, rtol=1e-5, atol=1e-6)


# ---------------------------------------------------------------- trace
def test_trace_pieces_equal_plane_crossings():
    """Interior segments of a single tall stack: pieces = crossings + 1 and
    the piece lengths add up to the 3D segment length."""
    rng = np.random.default_rng(3)
    vals, N, dx, mv = po.build_table(0.01, 10.0)
    nseg = 12
    lengths = rng.uniform(0.01, 0.5, nseg).astype(np.float32)
    npol, zs, height, cai, fai = 4, 1, 50.0, 10, 5
    dims = dict(ntracks_2D=1, n_polar=npol, z_stacked=zs, G=1, nsr=10, fai=fai, cai=cai, decomp=1,
                axial_exp=2, height=height, axial_z_sep=50.0, assembly_width=1.0,
                domain_height=height, ntracks=npol * zs, n_mat=1, table_N=N, table_dx=dx,
                table_maxVal=mv, leakage=0.0)
    polar = (np.pi * (np.arange(npol) + 0.5) / npol).astype(np.float32)
    arrays = dict(az_weight=np.ones(1, np.float32), seg_offset=np.array([0, nseg]),
                  seg_length=lengths, polar_angle=polar, p_weight=np.ones(npol, np.float32),
                  z_height=np.zeros(npol, np.float32), f_psi=np.zeros(npol, np.float32),
                  fine_source=np.ones(10 * fai, np.float32), fine_flux=np.zeros(10 * fai, np.float32),
                  sigT=np.ones(10, np.float32), vol=np.ones(10, np.float32),
                  material=np.zeros(10, np.int32), xs=np.ones(3, np.float32),
                  scatter=np.ones(1, np.float32), table=vals)
    o = po.Oracle(dims, arrays)
    dz = height / (cai * fai)
    for j in range(npol):
        z0 = 23.37 if j < npol // 2 else 26.61      # well inside the domain
        rec, _ = o.trace_stack(0, j, np.array([z0], np.float32))
        mu = math.cos(float(polar[j]))
        z = float(np.float32(z0))
        for n in range(nseg - 1):                    # the last segment resets z
            pieces = rec[rec["n"] == n]
            s3 = float(np.float32(lengths[n] / math.sin(float(polar[j]))))
            z1 = z + s3 * mu
            lo, hi = min(z, z1), max(z, z1)
            crossings = math.floor(hi / dz) - math.floor(lo / dz)
            assert len(pieces) == crossings + 1, (j, n)
            assert abs(float(pieces["ds"].astype(np.float64).sum()) - s3) < 1e-5
            z = z1


# ---------------------------------------------------------------- reductions
def test_reductions_on_constant_fields():
    G, fai, nsr = 8, 5, 16
    vals, N, dx, mv = po.build_table(0.01, 10.0)
    dims = dict(ntracks_2D=1, n_polar=2, z_stacked=1, G=G, nsr=nsr, fai=fai, cai=1, decomp=1,
                axial_exp=2, height=5.0, axial_z_sep=1.0, assembly_width=1.0, domain_height=5.0,
                ntracks=2, n_mat=2, table_N=N, table_dx=dx, table_maxVal=mv, leakage=0.0)
    arrays = dict(az_weight=np.ones(1, np.float32), seg_offset=np.array([0, 0]),
                  seg_length=np.ones(1, np.float32), polar_angle=np.array([0.7, 2.4], np.float32),
                  p_weight=np.ones(2, np.float32), z_height=np.zeros(2, np.float32),
                  f_psi=np.full(2 * G, 3.0, np.float32),
                  fine_source=np.ones(nsr * fai * G, np.float32),
                  fine_flux=np.ones(nsr * fai * G, np.float32), sigT=np.ones(nsr * G, np.float32),
                  vol=np.full(nsr, 0.5, np.float32), material=np.zeros(nsr, np.int32),
                  xs=np.ones(2 * G * 3, np.float32), scatter=np.ones(2 * G * G, np.float32),
                  table=vals)
    for red in (po.REDUCE_F32, po.REDUCE_F64):
        o = po.Oracle(dims, arrays, reduce=red)
        assert o.compute_keff() == 1.0                      # nuSigmaF == SigmaA
        o.renormalize_flux()
        total = nsr * fai * G * 0.5                          # sum of flux*vol*nuSigmaF
        nf = f32(1.0 / total)
        adjust = f32(float(nf * f32(4)) * math.pi * fai / 0.5)
        assert np.all(o["fine_flux"] == adjust)
        assert np.all(o["f_psi"] == f32(3.0) * nf)
        res = o.update_sources(1.0)
        # fission G*adjust (chi 1, k 1) plus scatter G*adjust, over 4 pi
        new = 2 * G * float(adjust) / (4 * math.pi)
        assert np.allclose(o["fine_source"], new, rtol=1e-6)
        assert res > 0


# ---------------------------------------------------------------- region hash
def test_region_hash_range_and_uniformity():
    nsr = 97
    counts = np.zeros(nsr)
    for t in range(200):
        for n in range(10):
            for p in range(3):
                q = po.qsr_hash(5, 2, t, n, p, nsr)
                assert 0 <= q < nsr
                counts[q] += 1
    assert po.qsr_hash(5, 2, 7, 3, 1, nsr) == po.qsr_hash(5, 2, 7, 3, 1, nsr)
    assert po.qsr_hash(5, 2, 7, 3, 1, nsr) != po.qsr_hash(5, 3, 7, 3, 1, nsr) or \
        po.qsr_hash(5, 2, 8, 3, 1, nsr) != po.qsr_hash(5, 2, 7, 3, 1, nsr)
    expected = counts.sum() / nsr
    chi2 = ((counts - expected) ** 2 / expected).sum()
    assert chi2 < 2 * nsr   # loose chi-square bound for 96 degrees of freedom
