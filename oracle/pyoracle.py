"""ctypes binding of the CPU oracle (moc_oracle.c) and the reference harness.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

The oracle restates the reference's sweep path (src/solver.c, src/utils.c;
see moc_oracle.h for the line map).  This module wraps it for numpy arrays
and also reads the binary dumps written by ref_driver.c.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmoc_oracle.so")
REF_SERIAL = os.path.join(HERE, "_ref", "ref_serial")
REF_OMP_LIB = os.path.join(HERE, "_ref", "libsimplemoc_ref_omp.so")

QSR_HASH, QSR_LIBC, QSR_STREAM = 0, 1, 2
TALLY_F32, TALLY_F64 = 0, 1
REDUCE_F32, REDUCE_F64 = 0, 1
EXP_TABLE, EXP_INTRINSIC = 0, 1

_fp = C.POINTER(C.c_float)
_lp = C.POINTER(C.c_long)
_ip = C.POINTER(C.c_int)


class OrcProblem(C.Structure):
    _fields_ = [
        ("ntracks_2D", C.c_long), ("n_polar", C.c_int), ("z_stacked", C.c_int),
        ("G", C.c_int), ("nsr", C.c_long), ("fai", C.c_int), ("cai", C.c_int),
        ("decomp", C.c_int), ("axial_exp", C.c_int), ("height", C.c_float),
        ("axial_z_sep", C.c_float), ("assembly_width", C.c_float),
        ("domain_height", C.c_float), ("ntracks", C.c_long), ("n_mat", C.c_long),
        ("az_weight", _fp), ("seg_offset", _lp), ("seg_length", _fp),
        ("polar_angle", _fp), ("p_weight", _fp), ("z_height", _fp),
        ("f_psi", _fp), ("b_psi", _fp), ("fine_source", _fp), ("fine_flux", _fp),
        ("sigT", _fp), ("vol", _fp), ("material", _ip), ("xs", _fp),
        ("scatter", _fp), ("table", _fp), ("table_N", C.c_int),
        ("table_dx", C.c_float), ("table_maxVal", C.c_float), ("leakage", _fp),
    ]


class OrcOpts(C.Structure):
    _fields_ = [
        ("qsr_mode", C.c_int), ("tally", C.c_int), ("reduce", C.c_int),
        ("exp_mode", C.c_int), ("seed", C.c_uint64), ("sweep_index", C.c_long),
        ("qsr_stream", C.POINTER(C.c_int32)), ("qsr_stream_len", C.c_long),
        ("qsr_stream_pos", C.c_long),
    ]


class OrcPiece(C.Structure):
    _fields_ = [("track", C.c_long), ("n", C.c_int), ("p", C.c_int),
                ("zh", C.c_float), ("ds", C.c_float), ("qsr", C.c_long)]


class OrcShape(C.Structure):
    _fields_ = [
        ("n_azimuthal", C.c_int), ("n_polar_angles", C.c_int),
        ("assembly_width", C.c_float), ("radial_ray_sep", C.c_float),
        ("height", C.c_float), ("axial_z_sep", C.c_float),
        ("decomp_assemblies_ax", C.c_int), ("n_2D_source_regions_per_assembly", C.c_long),
        ("cai", C.c_int), ("n_azimuthal_half", C.c_int), ("ntracks_2D", C.c_long),
        ("z_stacked", C.c_int), ("ntracks", C.c_long), ("domain_height", C.c_float),
        ("nsr", C.c_long),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.orc_pairwise_sum.argtypes = [_fp, C.c_long]
        L.orc_pairwise_sum.restype = C.c_float
        L.orc_table_N.argtypes = [C.c_float, C.c_float]
        L.orc_table_N.restype = C.c_int
        L.orc_build_table.argtypes = [C.c_float, C.c_float, _fp, _fp, _fp]
        L.orc_interpolate.argtypes = [_fp, C.c_int, C.c_float, C.c_float, C.c_float]
        L.orc_interpolate.restype = C.c_float
        L.orc_qsr_hash.argtypes = [C.c_uint64, C.c_long, C.c_long, C.c_long, C.c_long, C.c_long]
        L.orc_qsr_hash.restype = C.c_uint32
        L.orc_derive_shape.argtypes = [C.POINTER(OrcShape)]
        L.orc_trace_stack.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcOpts), C.c_long,
                                      C.c_int, _fp, C.POINTER(OrcPiece), C.c_long]
        L.orc_trace_stack.restype = C.c_long
        L.orc_attenuate.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcOpts), C.c_long,
                                    C.c_float, C.c_float, C.c_long, C.c_float, C.c_float,
                                    _fp, _fp, C.POINTER(C.c_double)]
        L.orc_transport_sweep.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcOpts)]
        L.orc_transport_sweep.restype = C.c_long
        L.orc_two_way_sweep.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcOpts)]
        L.orc_two_way_sweep.restype = C.c_long
        L.orc_two_way_stack.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcOpts), C.c_long,
                                        C.c_int, _fp, _fp, _fp]
        L.orc_two_way_stack.restype = C.c_long
        L.orc_boundary_vacuum.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcOpts)]
        L.orc_renormalize_flux.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcOpts)]
        L.orc_update_sources.argtypes = [C.POINTER(OrcProblem), C.c_float]
        L.orc_update_sources.restype = C.c_float
        L.orc_compute_keff.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcOpts)]
        L.orc_compute_keff.restype = C.c_float
        _dp = C.POINTER(C.c_double)
        L.orc_functional_f64.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcOpts), _dp, _dp, _dp, C.c_int]
        L.orc_functional_f64.restype = C.c_double
        L.orc_adjoint_f64.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcOpts), _dp, _dp, _dp,
                                      C.c_int, _dp, _dp]
        L.orc_adjoint_f64.restype = C.c_double
        L.orc_fit_coefficients.argtypes = [C.POINTER(OrcProblem), _fp]
        _lib = L
    return _lib


def _ptr(a, ct):
    if a is None:
        return C.cast(None, C.POINTER(ct))
    return a.ctypes.data_as(C.POINTER(ct))


# --------------------------------------------------------------------------
# numerics helpers
def pairwise_sum(v) -> float:
    v = np.ascontiguousarray(v, dtype=np.float32)
    return float(lib().orc_pairwise_sum(_ptr(v, C.c_float), v.size))


def build_table(precision=0.01, max_val=10.0):
    L = lib()
    N = L.orc_table_N(precision, max_val)
    vals = np.zeros(2 * N, np.float32)
    dx = C.c_float()
    mv = C.c_float()
    L.orc_build_table(precision, max_val, _ptr(vals, C.c_float), C.byref(dx), C.byref(mv))
    return vals, N, dx.value, mv.value


def interpolate(vals, N, dx, max_val, x) -> float:
    return float(lib().orc_interpolate(_ptr(vals, C.c_float), N, dx, max_val, x))


def qsr_hash(seed, sweep, track, n, p, nsr) -> int:
    return int(lib().orc_qsr_hash(seed, sweep, track, n, p, nsr))


def derive_shape(**kw) -> dict:
    s = OrcShape()
    for k, v in kw.items():
        setattr(s, k, v)
    lib().orc_derive_shape(C.byref(s))
    return {f: getattr(s, f) for f, _ in OrcShape._fields_}


# --------------------------------------------------------------------------
class Oracle:
    """Holds copies of a problem's arrays and runs the oracle on them.

    `arrays` keys: az_weight, seg_offset, seg_length, polar_angle, p_weight,
    z_height, f_psi, (b_psi), fine_source, fine_flux, sigT, vol, material, xs,
    scatter, table; `dims` keys: ntracks_2D, n_polar, z_stacked, G, nsr, fai,
    cai, decomp, axial_exp, height, axial_z_sep, assembly_width,
    domain_height, ntracks, n_mat, table_N, table_dx, table_maxVal, leakage.
    """

    def __init__(self, dims: dict, arrays: dict, qsr_mode=QSR_HASH, tally=TALLY_F64,
                 reduce=REDUCE_F64, exp_mode=EXP_TABLE, seed=0):
        self.d = dict(dims)
        a = {}
        f32 = ("az_weight", "seg_length", "polar_angle", "p_weight", "z_height", "f_psi",
               "b_psi", "fine_source", "fine_flux", "sigT", "vol", "xs", "scatter", "table")
        for k in f32:
            if arrays.get(k) is not None:
                a[k] = np.array(arrays[k], dtype=np.float32, copy=True).ravel()
        a["seg_offset"] = np.array(arrays["seg_offset"], dtype=np.int64, copy=True)
        a["material"] = np.array(arrays["material"], dtype=np.int32, copy=True)
        self.a = a
        self.leak = np.array([dims.get("leakage", 0.0)], np.float32)
        self.opts = OrcOpts(qsr_mode, tally, reduce, exp_mode, seed, 0, None, 0, 0)
        self._stream = None
        p = OrcProblem()
        for k in ("ntracks_2D", "n_polar", "z_stacked", "G", "nsr", "fai", "cai", "decomp",
                  "axial_exp", "height", "axial_z_sep", "assembly_width", "domain_height",
                  "ntracks", "n_mat", "table_N", "table_dx", "table_maxVal"):
            setattr(p, k, dims[k])
        for k in f32:
            setattr(p, k, _ptr(a.get(k), C.c_float))
        p.seg_offset = _ptr(a["seg_offset"], C.c_long)
        p.material = _ptr(a["material"], C.c_int)
        p.leakage = _ptr(self.leak, C.c_float)
        self.p = p

    # -- state access
    def __getitem__(self, k):
        return self.a[k]

    @property
    def leakage(self) -> float:
        return float(self.leak[0])

    def set_stream(self, stream):
        self._stream = np.ascontiguousarray(stream, dtype=np.int32)
        self.opts.qsr_stream = self._stream.ctypes.data_as(C.POINTER(C.c_int32))
        self.opts.qsr_stream_len = self._stream.size
        self.opts.qsr_stream_pos = 0

    # -- per-iteration path
    def transport_sweep(self) -> int:
        n = lib().orc_transport_sweep(C.byref(self.p), C.byref(self.opts))
        if n < 0:
            raise RuntimeError("oracle sweep failed")
        return int(n)

    def two_way_sweep(self) -> int:
        """src/solver.c:556-891 (needs a b_psi array)."""
        n = lib().orc_two_way_sweep(C.byref(self.p), C.byref(self.opts))
        if n < 0:
            raise RuntimeError("oracle two-way sweep failed")
        return int(n)

    def boundary_vacuum(self):
        lib().orc_boundary_vacuum(C.byref(self.p), C.byref(self.opts))

    def renormalize_flux(self):
        lib().orc_renormalize_flux(C.byref(self.p), C.byref(self.opts))

    def update_sources(self, keff: float) -> float:
        return float(lib().orc_update_sources(C.byref(self.p), keff))

    def compute_keff(self) -> float:
        return float(lib().orc_compute_keff(C.byref(self.p), C.byref(self.opts)))

    def iterate(self, n_iters: int, keff: float = 1.0, vacuum: bool = False,
                two_way: bool = False):
        """The reference's per-iteration call order (src/main.c:57-92)."""
        out = []
        for _ in range(n_iters):
            segs = self.two_way_sweep() if two_way else self.transport_sweep()
            if vacuum:
                self.boundary_vacuum()
            self.renormalize_flux()
            res = self.update_sources(keff)
            keff = self.compute_keff()
            out.append((segs, keff, res))
        return out

    # -- FP64 functional and adjoint
    def fit_coefficients(self):
        """The sweep's quadratic-fit coefficients, shape [nfsr*G, 3] (float32)."""
        n = self.d["nsr"] * self.d["fai"] * self.d["G"]
        out = np.zeros(n * 3, np.float32)
        lib().orc_fit_coefficients(C.byref(self.p), _ptr(out, C.c_float))
        return out.reshape(n, 3)

    def functional_f64(self, coef, sigma, weights, exp_mode=EXP_INTRINSIC) -> float:
        coef = np.ascontiguousarray(coef, np.float64)
        sigma = np.ascontiguousarray(sigma, np.float64)
        weights = np.ascontiguousarray(weights, np.float64)
        return float(lib().orc_functional_f64(C.byref(self.p), C.byref(self.opts),
                                              _ptr(coef, C.c_double), _ptr(sigma, C.c_double),
                                              _ptr(weights, C.c_double), exp_mode))

    def adjoint_f64(self, coef, sigma, weights, exp_mode=EXP_INTRINSIC):
        coef = np.ascontiguousarray(coef, np.float64)
        sigma = np.ascontiguousarray(sigma, np.float64)
        weights = np.ascontiguousarray(weights, np.float64)
        gs = np.zeros(sigma.size, np.float64)
        gc = np.zeros(coef.size, np.float64)
        J = lib().orc_adjoint_f64(C.byref(self.p), C.byref(self.opts), _ptr(coef, C.c_double),
                                  _ptr(sigma, C.c_double), _ptr(weights, C.c_double), exp_mode,
                                  _ptr(gs, C.c_double), _ptr(gc, C.c_double))
        return float(J), gs, gc.reshape(coef.shape)

    def trace_stack(self, i: int, j: int, z_state=None):
        """Pieces of stack (i, j) from the given (or current) z heights."""
        zs = self.d["z_stacked"]
        s = (i * self.d["n_polar"] + j) * zs
        z = np.array(self.a["z_height"][s:s + zs] if z_state is None else z_state,
                     dtype=np.float32, copy=True)
        L = lib()
        z0 = z.copy()
        n = L.orc_trace_stack(C.byref(self.p), C.byref(self.opts), i, j,
                              _ptr(z0, C.c_float), None, 0)
        buf = (OrcPiece * max(n, 1))()
        n2 = L.orc_trace_stack(C.byref(self.p), C.byref(self.opts), i, j,
                               _ptr(z, C.c_float), buf, n)
        assert n2 == n
        dt = np.dtype({"names": [f[0] for f in OrcPiece._fields_],
                       "formats": ["i8", "i4", "i4", "f4", "f4", "i8"],
                       "offsets": [getattr(OrcPiece, f[0]).offset for f in OrcPiece._fields_],
                       "itemsize": C.sizeof(OrcPiece)})
        rec = np.frombuffer(buf, dtype=dt, count=n).copy()
        return rec, z


    def two_way_stack(self, i: int, j: int, z_state, f_psi, b_psi, sweep_index: int):
        """One stack of the two-way sweep without tallies (src/solver.c:632-855),
        draws of sweep `sweep_index`: returns (pieces, z, f_psi, b_psi) from
        copies of the given heights [zs] and angular fluxes [zs*G]."""
        zs, G = self.d["z_stacked"], self.d["G"]
        z = np.array(z_state, np.float32, copy=True).ravel()
        fp = np.array(f_psi, np.float32, copy=True).ravel()
        bp = np.array(b_psi, np.float32, copy=True).ravel()
        assert z.size == zs and fp.size == zs * G and bp.size == zs * G
        saved = self.opts.sweep_index
        self.opts.sweep_index = sweep_index
        try:
            n = lib().orc_two_way_stack(C.byref(self.p), C.byref(self.opts), i, j,
                                        _ptr(z, C.c_float), _ptr(fp, C.c_float), _ptr(bp, C.c_float))
        finally:
            self.opts.sweep_index = saved
        if n < 0:
            raise RuntimeError("oracle two-way stack failed")
        return int(n), z, fp, bp


# --------------------------------------------------------------------------
# reference dump reader (format written by ref_driver.c)
_DT = {0: np.float32, 1: np.int64, 2: np.int32, 3: np.float64}


def read_dump(path) -> dict:
    out = {}
    with open(path, "rb") as f:
        buf = f.read()
    pos = 0
    while pos < len(buf):
        nl = int(np.frombuffer(buf, np.uint32, 1, pos)[0]); pos += 4
        name = buf[pos:pos + nl].decode(); pos += nl
        ty = int(np.frombuffer(buf, np.uint32, 1, pos)[0]); pos += 4
        cnt = int(np.frombuffer(buf, np.uint64, 1, pos)[0]); pos += 8
        dt = np.dtype(_DT[ty])
        out[name] = np.frombuffer(buf, dt, cnt, pos).copy(); pos += cnt * dt.itemsize
    return out


def dump_problem(dump: dict):
    """(dims, arrays) of the generated problem recorded in a dump."""
    d = dump["dims"]
    fd = dump["fdims"]
    G, T = int(d[3]), int(d[9])
    dims = dict(ntracks_2D=int(d[0]), n_polar=int(d[1]), z_stacked=int(d[2]), G=G,
                nsr=int(d[4]), fai=int(d[5]), cai=int(d[6]), decomp=int(d[7]),
                axial_exp=int(d[8]), ntracks=T, n_mat=int(d[10]), table_N=int(d[12]),
                height=float(fd[0]), axial_z_sep=float(fd[1]), assembly_width=float(fd[2]),
                domain_height=float(fd[3]), table_dx=float(fd[4]), table_maxVal=float(fd[5]),
                leakage=0.0)
    arrays = dict(az_weight=dump["az_weight"], seg_offset=dump["seg_offset"],
                  seg_length=dump["seg_length"], polar_angle=dump["polar_angle"],
                  p_weight=dump["p_weight"], z_height=dump["z_height0"],
                  f_psi=np.zeros(T * G, np.float32), b_psi=np.zeros(T * G, np.float32),
                  fine_source=dump["fine_source0"],
                  fine_flux=np.zeros(int(d[4]) * int(d[5]) * G, np.float32),
                  sigT=dump["sigT"], vol=dump["vol"], material=dump["material"],
                  xs=dump["xs"], scatter=dump["scatter"], table=dump["table"])
    return dims, arrays


def run_ref_serial(input_file, gen_seed, sweep_seed, iters, dump_path, two_way=False,
                   tracks=None) -> dict:
    """Run the reference harness (serial -O0 build) and read its dump; with
    two_way the sweep is the reference's two_way_transport_sweep, with
    tracks (a path) the 2D tracks come from that OpenMOC track file."""
    if not os.path.exists(REF_SERIAL):
        raise FileNotFoundError(REF_SERIAL)
    extra = (["two_way"] if two_way else []) + ([f"tracks={tracks}"] if tracks else [])
    subprocess.run([REF_SERIAL, str(input_file), str(gen_seed), str(sweep_seed), str(iters),
                    str(dump_path)] + extra, check=True, stdout=subprocess.DEVNULL)
    return read_dump(dump_path)
