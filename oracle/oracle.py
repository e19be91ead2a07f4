"""ctypes front end of the FP64 CPU oracle (oracle/liboracle.so) and of the
reference build (oracle/_ref/libsimcref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and the timed CPU baseline.  The
product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
MAP_DIR = os.path.join(REPO, "data", "maps")
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libsimcref.so")

H_NCOUNT = 22
S_NCOUNT = 52
SO_NCOUNT = 18
R_NCOUNT = 21
NCOUNT = {1: H_NCOUNT, 2: SO_NCOUNT, 3: R_NCOUNT, 4: R_NCOUNT, 5: S_NCOUNT}

HMS_COUNTERS = [
    "trials", "slit_hor", "slit_vert", "slit_oct", "Q1_in", "Q1_mid", "Q1_out", "Q2_in",
    "Q2_mid", "Q2_out", "Q3_in", "Q3_mid", "Q3_out", "D1_in", "D1_out", "hut", "dc1", "dc2",
    "scin", "cal", "successes", "coll"]
SHMS_COUNTERS = [
    "trials", "successes", "targ_hor", "targ_vert", "targ_oct", "slit_hor", "slit_vert",
    "slit_oct", "HB_in", "HB_men", "HB_mex", "HB_out", "Q1_in", "Q1_men", "Q1_mid", "Q1_mex",
    "Q1_out", "Q2_in", "Q2_men", "Q2_mid", "Q2_mex", "Q2_out", "Q3_in", "Q3_men", "Q3_mid",
    "Q3_mex", "Q3_out", "D1_in", "D1_flr", "D1_men", "D1_mid1", "D1_mid2", "D1_mid3",
    "D1_mid4", "D1_mid5", "D1_mid6", "D1_mid7", "D1_mex", "D1_out", "BP_in", "BP_out", "hut",
    "dc1", "dc2", "cer1", "cer2", "s1", "s2", "s3", "cal", "stop_id", "coll"]
SOS_COUNTERS = [
    "trials", "slit_hor", "slit_vert", "slit_oct", "quad_in", "quad_mid", "quad_out", "bm01_in",
    "bm01_out", "bm02_in", "bm02_out", "exit", "hut", "dc1", "dc2", "scin", "cal", "successes"]
HRS_COUNTERS = [
    "trials", "slit_hor", "slit_vert", "Q1_in", "Q1_mid", "Q1_out", "Q2_in", "Q2_mid", "Q2_out",
    "Q3_in", "Q3_mid", "Q3_out", "D1_in", "D1_out", "hut", "dc1", "dc2", "s1", "s2", "cal",
    "successes"]

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class Settings(C.Structure):
    _fields_ = [("p_spec", C.c_double), ("m2", C.c_double), ("ctau", C.c_double),
                ("ms_flag", C.c_int), ("wcs_flag", C.c_int), ("decay_flag", C.c_int),
                ("using_coll", C.c_int), ("fit_f32", C.c_int), ("wire_sigma", C.c_double),
                ("mom_offset", C.c_double), ("sieve", C.c_int)]


class EventsIn(C.Structure):
    _fields_ = [(n, _dp) for n in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")]


class EventsOut(C.Structure):
    _fields_ = [(n, _dp) for n in ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp", "y_fp",
                                   "dy_fp", "pathlen", "resmult", "m2")] + \
        [("ok", _u8p), ("stop", _u8p), ("used", _i32p)]


_lib = None
_ref = None
_maps = None


def default_threads():
    """OpenMP threads for the oracle: the CPUs this process may use, at most
    16 (a one-GPU job's share of the host on the GPU machines)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def build():
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.isdir(os.environ.get("SIMC_REFERENCE", "/root/reference")):
        subprocess.run(["make", "-s", "-C", HERE, "ref",
                        "REF=" + os.environ.get("SIMC_REFERENCE", "/root/reference")],
                       check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.or_maps_load.restype = C.c_void_p
        L.or_maps_load.argtypes = [C.c_char_p]
        L.or_run.restype = C.c_int
        L.or_run.argtypes = [C.c_void_p, C.c_int, C.POINTER(Settings), C.c_int64,
                             C.POINTER(EventsIn), C.POINTER(EventsOut), C.c_int, _dp,
                             C.c_int64, C.c_uint64, C.c_uint64, C.c_int, _i64p]
        L.or_powi.restype = C.c_double
        L.or_powi.argtypes = [C.c_double, C.c_int]
        L.or_philox_uniform.restype = C.c_double
        L.or_philox_uniform.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64]
        L.or_philox4x32_10.restype = None
        L.or_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint32)]
        L.or_recon_run.restype = C.c_int
        L.or_recon_run.argtypes = [C.c_void_p, C.c_int, C.c_int64] + [_dp] * 9
        L.or_pion_coll_absorb.restype = C.c_double
        L.or_pion_coll_absorb.argtypes = [C.c_double, C.c_double]
        L.or_bm_log.restype = C.c_double
        L.or_bm_log.argtypes = [C.c_double]
        L.or_bm_sincos2pi.restype = None
        L.or_bm_sincos2pi.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_philox_normals.restype = C.c_int64
        L.or_philox_normals.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_double, _dp]
        _lib = L
    return _lib


_tune_maps = {}


def maps(map_dir=None):
    """The oracle's maps: the default tree (data/maps), or a directory of the
    same layout (tune_map_dir)."""
    global _maps
    if map_dir is not None and map_dir != MAP_DIR:
        if map_dir not in _tune_maps:
            p = lib().or_maps_load(map_dir.encode())
            if not p:
                raise RuntimeError("oracle: cannot load COSY maps from " + map_dir)
            _tune_maps[map_dir] = p
        return _tune_maps[map_dir]
    if _maps is None:
        p = lib().or_maps_load(MAP_DIR.encode())
        if not p:
            raise RuntimeError("oracle: cannot load COSY maps from " + MAP_DIR)
        _maps = p
    return _maps


TUNE_DIR = os.path.join(REPO, "data", "maps_tunes")


def tune_files(tune):
    """(forward, recon) paths of one of the reference's other SHMS tunes."""
    return (os.path.join(TUNE_DIR, "shms", tune + "_forward.dat"),
            os.path.join(TUNE_DIR, "shms", tune + "_recon.dat"))


def tune_map_dir(root, tune):
    """A map directory under `root` holding the default HMS maps and the
    SHMS tune `tune` under the reference's file names (symbolic links)."""
    d = os.path.join(root, "maps_" + tune)
    os.makedirs(os.path.join(d, "shms"), exist_ok=True)
    if not os.path.exists(os.path.join(d, "hms")):
        os.symlink(os.path.join(MAP_DIR, "hms"), os.path.join(d, "hms"))
    fwd, rec = tune_files(tune)
    for src, name in ((fwd, "shms_forward.dat"), (rec, "shms_recon.dat")):
        dst = os.path.join(d, "shms", name)
        if not os.path.exists(dst):
            os.symlink(src, dst)
    return d


_ref_variants = {}


def ref_lib(variant=""):
    """The reference build (oracle/_ref); variant "sieve1" / "sieve2" /
    "sieve3": the builds with the mock sieve slit on (make -C oracle sieve)."""
    global _ref
    if variant:
        if variant not in _ref_variants:
            path = os.path.join(HERE, "_ref", f"libsimcref_{variant}.so")
            if not os.path.exists(path):
                return None
            L = C.CDLL(path)
            L.ref_run.restype = C.c_int
            L.ref_run.argtypes = ([C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                   C.c_double, C.c_int, C.c_int64] + [_dp] * 7 + [_dp] * 11 +
                                  [_u8p, _i32p, _dp, C.c_int64, C.c_char_p, _i64p,
                                   C.POINTER(C.c_uint64)])
            _ref_variants[variant] = L
        return _ref_variants[variant]
    if _ref is None:
        if not os.path.exists(REF_LIB):
            return None
        L = C.CDLL(REF_LIB)
        L.ref_run.restype = C.c_int
        L.ref_run.argtypes = ([C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                               C.c_double, C.c_int, C.c_int64] + [_dp] * 7 + [_dp] * 11 +
                              [_u8p, _i32p, _dp, C.c_int64, C.c_char_p, _i64p,
                               C.POINTER(C.c_uint64)])
        L.ref_lfit.restype = None
        L.ref_lfit.argtypes = [C.c_int64, C.c_int32] + [C.POINTER(C.c_float)] * 5
        _ref = L
    return _ref


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def _prep_outputs(n, ev):
    out = {k: np.array(ev[k], dtype=np.float64, copy=True)
           for k in ("dpp", "y", "dxdz", "dydz")}
    for k in ("x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen", "resmult", "m2"):
        out[k] = np.zeros(n)
    out["ok"] = np.zeros(n, np.uint8)
    out["stop"] = np.zeros(n, np.uint8)
    out["used"] = np.zeros(n, np.int32)
    return out


def run(arm, events, *, p_spec, m2, ms=True, wcs=True, decay=False, ctau=0.0, coll=False,
        fit_f32=True, wire_sigma=0.0, rng="philox", tape=None, stride=0, seed=0, event0=0,
        nthreads=None, map_dir=None, mom_offset=0.0, sieve=0):
    """Run the oracle over SoA events (dict of float64 arrays: dpp, x, y, z, dxdz,
    dydz, fry).  arm: 1 HMS, 2 SOS, 5 SHMS.  rng: 'tape' | 'philox' | 'ranlux'.
    map_dir: a map directory other than the default tree.  mom_offset: the
    adjoint form's momentum scale s_p (HMS, SHMS).  sieve: the mock sieve slit,
    a bit mask (1 the arm's, 2 the SHMS's in front of the bender, 3 both)."""
    L = lib()
    n = len(events["dpp"])
    ev = {k: np.ascontiguousarray(events[k], dtype=np.float64) for k in
          ("dpp", "x", "y", "z", "dxdz", "dydz", "fry") if k in events}
    ein = EventsIn(*[_p(ev.get(k)) for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")])
    out = _prep_outputs(n, ev)
    eout = EventsOut(*[_p(out[k]) for k in ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp",
                                             "y_fp", "dy_fp", "pathlen", "resmult", "m2")],
                     _p(out["ok"], _u8p), _p(out["stop"], _u8p), _p(out["used"], _i32p))
    S = Settings(p_spec, m2, ctau, int(ms), int(wcs), int(decay), int(coll), int(fit_f32),
                 wire_sigma, mom_offset, int(sieve))
    kind = {"tape": 0, "philox": 1, "ranlux": 2}[rng]
    cnt = np.zeros(64, np.int64)
    tp = None
    if kind == 0:
        tape = np.ascontiguousarray(tape, dtype=np.float64)
        tp = _p(tape)
    if nthreads is None:
        nthreads = default_threads()
    rc = L.or_run(maps(map_dir), arm, C.byref(S), n, C.byref(ein), C.byref(eout), kind, tp,
                  int(stride), int(seed), int(event0), int(nthreads), _p(cnt, _i64p))
    out["rc"] = rc
    out["counters"] = cnt[:NCOUNT[arm]]
    return out


def ref_run(arm, events, *, p_spec, m2, tape, stride, ms=True, wcs=True, decay=False,
            ctau=0.0, coll=False, sieve=0):
    """Run the reference Fortran (flang build) over the same events on a tape;
    sieve 1 / 2 / 3: its build with use_sieve / use_2sieve / both on."""
    L = ref_lib({0: "", 1: "sieve1", 2: "sieve2", 3: "sieve3"}[sieve])
    if L is None:
        raise FileNotFoundError(REF_LIB)
    n = len(events["dpp"])
    ev = {k: np.ascontiguousarray(events[k], dtype=np.float64) for k in
          ("dpp", "x", "y", "z", "dxdz", "dydz", "fry") if k in events}
    out = _prep_outputs(n, ev)
    tape = np.ascontiguousarray(tape, dtype=np.float64)
    cnt = np.zeros(64, np.int64)
    mask = np.zeros(n, np.uint64)
    rc = L.ref_run(arm, p_spec, m2, int(ms), int(wcs), int(decay), ctau, int(coll), n,
                   *[_p(ev.get(k)) for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")],
                   *[_p(out[k]) for k in ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp",
                                          "y_fp", "dy_fp", "pathlen", "resmult", "m2")],
                   _p(out["ok"], _u8p), _p(out["used"], _i32p), _p(tape), int(stride),
                   MAP_DIR.encode(), _p(cnt, _i64p), _p(mask, C.POINTER(C.c_uint64)))
    out["rc"] = rc
    out["counters"] = cnt[:NCOUNT[arm]]
    # bit c set: the event moved STOP counter c of the arm's common block
    out["stop_mask"] = mask
    return out


def ref_lfit(X, Y):
    """The reference's LFIT (cern/lfit.f, compiled into oracle/_ref) over the
    rows of X and Y (float32, n x L), KEY = 0.  Returns A, B, E (float32)."""
    L = ref_lib()
    if L is None:
        raise FileNotFoundError(REF_LIB)
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    n, npts = X.shape
    A, B, E = (np.zeros(n, np.float32) for _ in range(3))
    fp = C.POINTER(C.c_float)
    L.ref_lfit(n, npts, _p(X, fp), _p(Y, fp), _p(A, fp), _p(B, fp), _p(E, fp))
    return A, B, E


def lfit_f32(X, Y):
    """The oracle's LFIT restatement (REAL*4 points, REAL*8 sums) over the rows
    of X and Y, KEY = 0.  Returns A, B, E (float32)."""
    L = lib()
    fn = L.or_lfit_f32
    fp = C.POINTER(C.c_float)
    fn.restype = None
    fn.argtypes = [fp, fp, C.c_int, C.c_int, fp, fp, fp]
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    n, npts = X.shape
    A, B, E = (np.zeros(n, np.float32) for _ in range(3))
    a, b, e = C.c_float(), C.c_float(), C.c_float()
    for i in range(n):
        fn(X[i].ctypes.data_as(fp), Y[i].ctypes.data_as(fp), npts, 0, C.byref(a), C.byref(b), C.byref(e))
        A[i], B[i], E[i] = a.value, b.value, e.value
    return A, B, E


def philox_block(ctr, key):
    a = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox4x32_10(a, k, o)
    return list(o)


def recon(arm, x_fp, dx_fp, y_fp, dy_fp, fry, map_dir=None):
    """Reconstruction only: returns delta (%), x'_tar, y'_tar, y_tar (cm)."""
    n = len(x_fp)
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (x_fp, dx_fp, y_fp, dy_fp, fry)]
    outs = [np.zeros(n) for _ in range(4)]
    rc = lib().or_recon_run(maps(map_dir), arm, n, *[_p(v) for v in a], *[_p(v) for v in outs])
    if rc:
        raise RuntimeError(f"or_recon_run: {rc}")
    return tuple(outs)


def philox_normals(seed, event, n, nsigmax=99.0):
    """n normal deviates of one event's philox stream; returns (values, uniforms used)."""
    out = np.zeros(n)
    used = lib().or_philox_normals(seed, event, n, nsigmax, _p(out))
    return out, used


# ---- target step ---------------------------------------------------------------------
class ORTarget(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("A", "Z", "rho", "X0", "length", "angle", "zoffset")] + \
        [("can", C.c_int)]


class ORPrep(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("narm", "spectrometer", "using_eloss", "mc_smear")] + \
        [(k, C.c_double) for k in ("p_spec", "th_spec", "phi", "mass", "coulomb", "beam_energy")] + \
        [("offset", C.c_double * 5)]


VERTEX_FIELDS = ("x", "y", "z", "E", "p", "E_vertex", "theta", "delta", "xptar", "yptar", "dang_yp",
                 "dang_xp")
PREP_OUT_FIELDS = ("dpp", "x", "y", "z", "dxdz", "dydz", "fry", "eloss", "teff", "dang1", "dang2")


class ORVertexIn(C.Structure):
    _fields_ = [(k, _dp) for k in VERTEX_FIELDS]


class ORPrepOut(C.Structure):
    _fields_ = [(k, _dp) for k in PREP_OUT_FIELDS]


def _target_lib():
    L = lib()
    if not hasattr(L, "_target_ready"):
        L.or_target_prep.restype = C.c_int
        L.or_target_prep.argtypes = [C.POINTER(ORTarget), C.POINTER(ORPrep), C.c_int64, C.POINTER(ORVertexIn),
                                     C.POINTER(ORPrepOut), C.c_int, _dp, C.c_int64, C.c_uint64, C.c_uint64,
                                     C.c_int, _i32p]
        L.or_target_X0.restype = C.c_double
        L.or_target_X0.argtypes = [C.c_double, C.c_double]
        L._target_ready = True
    return L


def make_target(A=1.0, Z=1.0, rho=0.0723, X0=0.0, length=10.0, angle=0.0, zoffset=0.0, can=3):
    """Target description; X0 <= 0 takes init.f's value for (A, Z)."""
    if X0 <= 0.0:
        X0 = target_X0(A, Z)
    return ORTarget(A, Z, rho, X0, length, angle, zoffset, can)


def target_X0(A, Z):
    return _target_lib().or_target_X0(A, Z)


def make_prep(narm, spectrometer=1, using_eloss=True, mc_smear=True, p_spec=0.0, th_spec=0.0, phi=0.0,
              mass=0.51099906, coulomb=0.0, beam_energy=0.0, offset=(0.0,) * 5):
    return ORPrep(narm, spectrometer, int(using_eloss), int(mc_smear), p_spec, th_spec, phi, mass, coulomb,
                  beam_energy, (C.c_double * 5)(*offset))


def target_prep(target, prep, vertex, *, rng="philox", seed=0, event0=0, tape=None, stride=0, nthreads=0):
    """The oracle's target step over numpy vertex arrays (fields of VERTEX_FIELDS,
    absent = NULL).  Returns (dict of outputs incl. 'used', return code)."""
    n = len(vertex["z"])
    vin = {k: np.ascontiguousarray(vertex[k], dtype=np.float64) for k in VERTEX_FIELDS
           if vertex.get(k) is not None}
    out = {k: np.zeros(n) for k in PREP_OUT_FIELDS}
    used = np.zeros(n, np.int32)
    V = ORVertexIn(*[_p(vin.get(k)) for k in VERTEX_FIELDS])
    O = ORPrepOut(*[_p(out[k]) for k in PREP_OUT_FIELDS])
    if rng == "tape":
        tape = np.ascontiguousarray(tape, dtype=np.float64)
        kind, tp = 0, _p(tape)
    else:
        kind, tp = 1, None
    rc = _target_lib().or_target_prep(C.byref(target), C.byref(prep), n, C.byref(V), C.byref(O), kind, tp,
                                      int(stride), int(seed), int(event0), int(nthreads or default_threads()),
                                      _p(used, _i32p))
    out["used"] = used
    return out, rc


def ref_target_step(target, narm, spec, mass, zpos, energy, theta, tape, stride, p_musc=None,
                    beta_musc=None, typeflag=1):
    """The reference's trip_thru_target (+ target_musc when p_musc is given)
    from oracle/_ref over events on tape slots.  Returns dict eloss, teff,
    dang1, dang2, used."""
    L = ref_lib()
    if L is None:
        raise FileNotFoundError(REF_LIB)
    if not hasattr(L, "_target_ready"):
        L.ref_target_step.restype = C.c_int
        L.ref_target_step.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int64,
                                      _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i32p, _dp, C.c_int64]
        L._target_ready = True
    n = len(zpos)
    t = np.array([target.A, target.Z, target.rho, target.X0, target.length, target.angle, target.zoffset])
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (zpos, energy, theta)]
    pm = None if p_musc is None else np.ascontiguousarray(p_musc, dtype=np.float64)
    bm = None if beta_musc is None else np.ascontiguousarray(beta_musc, dtype=np.float64)
    out = {k: np.zeros(n) for k in ("eloss", "teff", "dang1", "dang2")}
    used = np.zeros(n, np.int32)
    tape = np.ascontiguousarray(tape, dtype=np.float64)
    L.ref_target_step(_p(t), int(target.can), narm, spec, mass, typeflag, n, *[_p(v) for v in a], _p(pm),
                      _p(bm), *[_p(out[k]) for k in ("eloss", "teff", "dang1", "dang2")], _p(used, _i32p),
                      _p(tape), int(stride))
    out["used"] = used
    return out


# ---------------------------------------------------------------------------
# target-field tracking (trg_track.f)
# ---------------------------------------------------------------------------
FIELD_MAP = os.path.join(REPO, "data", "fields", "trg_field_map.dat")
_FIELD_BYTES = 8 * (2 * 51 * 51 + 2 * 51 + 4)


def _field_lib():
    L = lib()
    if not hasattr(L, "_field_ready"):
        L.or_trg_init.restype = C.c_int
        L.or_trg_init.argtypes = [C.c_void_p, C.c_char_p] + [C.c_double] * 4
        L.or_trg_field.restype = None
        L.or_trg_field.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.or_track_from_tgt_run.restype = None
        L.or_track_from_tgt_run.argtypes = [C.c_void_p, C.c_int64] + [_dp] * 6 + [C.c_double, C.c_int, _u8p,
                                                                                C.c_int]
        L.or_track_to_tgt_run.restype = None
        L.or_track_to_tgt_run.argtypes = ([C.c_void_p, C.c_void_p, C.c_int, C.c_int64] + [_dp] * 11 +
                                          [C.c_double, C.c_double, C.c_double, C.c_int, _u8p, C.c_int])
        L._field_ready = True
    return L


def field(map_path=FIELD_MAP, theta_e=0.0, theta_p=0.0):
    """trgInit: map_path None = the uniform 5 T test field, "0" = zero field;
    theta_e / theta_p: field angle relative to the e / p arm (deg)."""
    buf = C.create_string_buffer(_FIELD_BYTES)
    rc = _field_lib().or_trg_init(buf, None if map_path is None else map_path.encode(), theta_e, 0.0,
                                  theta_p, 0.0)
    if rc:
        raise RuntimeError("oracle: cannot read field map " + str(map_path))
    return buf


def field_at(F, x, spect):
    B = np.zeros(3)
    _field_lib().or_trg_field(F, _p(np.ascontiguousarray(x, dtype=np.float64)), _p(B), spect)
    return B


def track_from_tgt(F, ev, mom, mass, spect, nthreads=0):
    """Per-event track_from_tgt over dict ev (x, y, z, dx, dy); returns new
    arrays and ok (uint8)."""
    n = len(ev["x"])
    out = {k: np.array(ev[k], dtype=np.float64, copy=True) for k in ("x", "y", "z", "dx", "dy")}
    ok = np.zeros(n, np.uint8)
    m = np.ascontiguousarray(mom, dtype=np.float64)
    _field_lib().or_track_from_tgt_run(F, n, *[_p(out[k]) for k in ("x", "y", "z", "dx", "dy")], _p(m), mass,
                                       spect, _p(ok, _u8p), int(nthreads or default_threads()))
    out["ok"] = ok
    return out


def track_to_tgt(F, arm, trk, first, frx, fry, mom, mass, ctheta, stheta, spect, nthreads=0):
    """Per-event track_to_tgt: trk = focal-plane track dict (x_fp, dx_fp,
    y_fp, dy_fp), first = first-guess recon dict (delta, y, dx, dy)."""
    n = len(trk["x_fp"])
    out = {k: np.array(first[k], dtype=np.float64, copy=True) for k in ("delta", "y", "dx", "dy")}
    ok = np.zeros(n, np.uint8)
    a = [np.ascontiguousarray(trk[k], dtype=np.float64) for k in ("x_fp", "dx_fp", "y_fp", "dy_fp")]
    f = [np.ascontiguousarray(v, dtype=np.float64) for v in (frx, fry, mom)]
    _field_lib().or_track_to_tgt_run(F, C.c_void_p(maps()), arm, n, *[_p(v) for v in a],
                                     *[_p(out[k]) for k in ("delta", "y", "dx", "dy")], *[_p(v) for v in f],
                                     mass, ctheta, stheta, spect, _p(ok, _u8p), int(nthreads or default_threads()))
    out["ok"] = ok
    return out


def ref_track_from_tgt(map_path, theta_e, theta_p, ev, mom, mass, spect):
    """The reference's trgInit + track_from_tgt (oracle/_ref)."""
    L = ref_lib()
    if L is None:
        raise FileNotFoundError(REF_LIB)
    L.ref_trg_init.restype = None
    L.ref_trg_init.argtypes = [C.c_char_p] + [C.c_double] * 4
    L.ref_track_from_tgt.restype = None
    L.ref_track_from_tgt.argtypes = [C.c_int64] + [_dp] * 6 + [C.c_double, C.c_int, _u8p]
    L.ref_trg_init(None if map_path is None else map_path.encode(), theta_e, 0.0, theta_p, 0.0)
    n = len(ev["x"])
    out = {k: np.array(ev[k], dtype=np.float64, copy=True) for k in ("x", "y", "z", "dx", "dy")}
    ok = np.zeros(n, np.uint8)
    L.ref_track_from_tgt(n, *[_p(out[k]) for k in ("x", "y", "z", "dx", "dy")],
                         _p(np.ascontiguousarray(mom, dtype=np.float64)), mass, spect, _p(ok, _u8p))
    out["ok"] = ok
    return out


def ref_track_to_tgt(map_path, theta_e, theta_p, arm, trk, first, frx, fry, mom, mass, ctheta, stheta, spect):
    L = ref_lib()
    if L is None:
        raise FileNotFoundError(REF_LIB)
    L.ref_trg_init.restype = None
    L.ref_trg_init.argtypes = [C.c_char_p] + [C.c_double] * 4
    L.ref_track_to_tgt.restype = C.c_int
    L.ref_track_to_tgt.argtypes = [C.c_int, C.c_int64] + [_dp] * 11 + [C.c_double] * 3 + [C.c_int, _u8p,
                                                                                         C.c_char_p]
    L.ref_trg_init(None if map_path is None else map_path.encode(), theta_e, 0.0, theta_p, 0.0)
    n = len(trk["x_fp"])
    out = {k: np.array(first[k], dtype=np.float64, copy=True) for k in ("delta", "y", "dx", "dy")}
    ok = np.zeros(n, np.uint8)
    a = [np.ascontiguousarray(trk[k], dtype=np.float64) for k in ("x_fp", "dx_fp", "y_fp", "dy_fp")]
    f = [np.ascontiguousarray(v, dtype=np.float64) for v in (frx, fry, mom)]
    rc = L.ref_track_to_tgt(arm, n, *[_p(v) for v in a], *[_p(out[k]) for k in ("delta", "y", "dx", "dy")],
                            *[_p(v) for v in f], mass, ctheta, stheta, spect, _p(ok, _u8p), MAP_DIR.encode())
    assert rc == 0, rc
    out["ok"] = ok
    return out


# ---------------------------------------------------------------------------
# the .bin ntuple written by the reference's NtupleInit / NtupleClose
# ---------------------------------------------------------------------------
NTUPLE_MODES = {"elastic": 1, "pion": 2, "semi": 3}


def ref_ntuple(path, mode, tgt_field, rows):
    """Writes `path` with the reference's NtupleInit (for `mode` and
    using_tgt_field) and one record per value of each row, as
    results_ntu_write does; closes it with NtupleClose."""
    L = ref_lib()
    if L is None:
        raise FileNotFoundError(REF_LIB)
    m = C.c_int32(NTUPLE_MODES[mode])
    t = C.c_int32(1 if tgt_field else 0)
    L.ref_ntuple_mode_(C.byref(m), C.byref(t))
    name = path.encode().ljust(80)
    if len(name) > 80:
        raise ValueError("path longer than CHARACTER*80")
    L.ntupleinit_(C.c_char_p(name), C.c_size_t(80))
    for r in np.ascontiguousarray(rows, dtype=np.float64):
        r = np.ascontiguousarray(r)
        L.ref_ntuple_event_(r.ctypes.data_as(C.c_void_p))
    L.ntupleclose_(C.c_char_p(name), C.c_size_t(80))


# ---- event generator: H(e,e'p) elastic, radiation off -------------------------
HEP_OUT_FIELDS = ("x", "y", "z", "rasterx", "rastery", "Ein", "eloss", "teff",
                  "e_xptar", "e_yptar", "e_theta", "e_phi", "e_E", "e_delta",
                  "p_xptar", "p_yptar", "p_theta", "p_phi", "p_P", "p_E", "p_delta",
                  "Q2", "nu", "q", "jacobian", "sigcc", "weight")


class ORHep(C.Structure):
    _fields_ = [("Ebeam", C.c_double), ("dEbeam", C.c_double), ("xwid", C.c_double), ("ywid", C.c_double),
                ("xoffset", C.c_double), ("yoffset", C.c_double), ("fr1", C.c_double), ("fr2", C.c_double),
                ("coulomb", C.c_double), ("e_yptar", C.c_double * 2), ("e_xptar", C.c_double * 2),
                ("e_theta", C.c_double), ("e_phi", C.c_double), ("e_P", C.c_double),
                ("p_theta", C.c_double), ("p_phi", C.c_double), ("p_P", C.c_double), ("Mh", C.c_double),
                ("fr_pattern", C.c_int), ("using_eloss", C.c_int), ("using_coulomb", C.c_int)]


class ORHepOut(C.Structure):
    _fields_ = [(k, _dp) for k in HEP_OUT_FIELDS] + [("ok", _u8p)]


def make_hep(**kw):
    """Generator settings (the or_hep fields; e_yptar, e_xptar as (min, max))."""
    h = ORHep()
    for k, v in kw.items():
        if k in ("e_yptar", "e_xptar"):
            getattr(h, k)[0], getattr(h, k)[1] = v
        elif k in ("fr_pattern", "using_eloss", "using_coulomb"):
            setattr(h, k, int(v))
        else:
            setattr(h, k, float(v))
    return h


def _gen_lib():
    L = _target_lib()
    if not hasattr(L, "_gen_ready"):
        L.or_gen_hep.restype = C.c_int
        L.or_gen_hep.argtypes = [C.POINTER(ORTarget), C.POINTER(ORHep), C.c_int64, C.POINTER(ORHepOut), C.c_int,
                                 _dp, C.c_int64, C.c_uint64, C.c_uint64, C.c_int, _i32p]
        L.or_physics_angles.restype = None
        L.or_physics_angles.argtypes = [C.c_double] * 4 + [C.POINTER(C.c_double)] * 2
        L.or_spectrometer_angles.restype = None
        L.or_spectrometer_angles.argtypes = [C.c_double] * 2 + [C.POINTER(C.c_double)] * 2 + [C.c_double] * 2
        L.or_fofa_best_fit.restype = None
        L.or_fofa_best_fit.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_sigep.restype = C.c_double
        L.or_sigep.argtypes = [C.c_double] * 4
        L.or_sigmott.restype = C.c_double
        L.or_sigmott.argtypes = [C.c_double] * 3
        L._gen_ready = True
    return L


def gen_hep(target, hep, n, *, rng="philox", seed=0, event0=0, tape=None, stride=0, nthreads=0):
    """The oracle's H(e,e'p) generator for n events.  Returns (dict of
    HEP_OUT_FIELDS + 'ok' + 'used', return code)."""
    out = {k: np.zeros(n) for k in HEP_OUT_FIELDS}
    out["ok"] = np.zeros(n, np.uint8)
    used = np.zeros(n, np.int32)
    O = ORHepOut(*([_p(out[k]) for k in HEP_OUT_FIELDS] + [_p(out["ok"], _u8p)]))
    if rng == "tape":
        tape = np.ascontiguousarray(tape, dtype=np.float64)
        kind, tp = 0, _p(tape)
    else:
        kind, tp = 1, None
    rc = _gen_lib().or_gen_hep(C.byref(target), C.byref(hep), n, C.byref(O), kind, tp, int(stride), int(seed),
                               int(event0), int(nthreads or default_threads()), _p(used, _i32p))
    out["used"] = used
    return out, rc


def physics_angles(theta0, phi0, dx, dy):
    th, ph = C.c_double(), C.c_double()
    _gen_lib().or_physics_angles(theta0, phi0, dx, dy, C.byref(th), C.byref(ph))
    return th.value, ph.value


def spectrometer_angles(theta0, phi0, theta, phi):
    dx, dy = C.c_double(), C.c_double()
    _gen_lib().or_spectrometer_angles(theta0, phi0, C.byref(dx), C.byref(dy), theta, phi)
    return dx.value, dy.value


def fofa_best_fit(qsquar):
    ge, gm = C.c_double(), C.c_double()
    _gen_lib().or_fofa_best_fit(qsquar, C.byref(ge), C.byref(gm))
    return ge.value, gm.value


def sigep(Q2, E, theta, Ein):
    return _gen_lib().or_sigep(Q2, E, theta, Ein)


def _ref_gen_lib():
    L = ref_lib()
    if L is None:
        raise FileNotFoundError(REF_LIB)
    if not hasattr(L, "_gen_ready"):
        _d = C.POINTER(C.c_double)
        L.physics_angles_.restype = None
        L.physics_angles_.argtypes = [_d] * 6
        L.spectrometer_angles_.restype = None
        L.spectrometer_angles_.argtypes = [_d] * 6
        L.ref_sigep_.restype = C.c_double
        L.ref_sigep_.argtypes = [_d] * 4
        L.ref_fofa_.restype = None
        L.ref_fofa_.argtypes = [_d] * 3
        L._gen_ready = True
    return L


def ref_physics_angles(theta0, phi0, dx, dy):
    """The reference's physics_angles (event.f) from oracle/_ref."""
    a = [C.c_double(v) for v in (theta0, phi0, dx, dy, 0.0, 0.0)]
    _ref_gen_lib().physics_angles_(*[C.byref(v) for v in a])
    return a[4].value, a[5].value


def ref_spectrometer_angles(theta0, phi0, theta, phi):
    """The reference's spectrometer_angles (event.f) from oracle/_ref."""
    a = [C.c_double(v) for v in (theta0, phi0, 0.0, 0.0, theta, phi)]
    _ref_gen_lib().spectrometer_angles_(*[C.byref(v) for v in a])
    return a[2].value, a[3].value


def ref_sigep(Q2, E, theta, Ein):
    """The reference's sigep (physics_proton.f) from oracle/_ref."""
    a = [C.c_double(v) for v in (Q2, E, theta, Ein)]
    return _ref_gen_lib().ref_sigep_(*[C.byref(v) for v in a])


def ref_fofa(qsquar):
    a = [C.c_double(v) for v in (qsquar, 0.0, 0.0)]
    _ref_gen_lib().ref_fofa_(*[C.byref(v) for v in a])
    return a[1].value, a[2].value


def gauss1_tape(tape, nsigmax):
    """gauss1 (gauss1.f) on the uniforms of `tape`: (value, uniforms used)."""
    L = lib()
    if not hasattr(L, "_g1_ready"):
        L.or_gauss1_tape.restype = C.c_double
        L.or_gauss1_tape.argtypes = [_dp, C.c_int64, C.c_double, _i64p]
        L._g1_ready = True
    t = np.ascontiguousarray(tape, dtype=np.float64)
    u = np.zeros(1, np.int64)
    v = L.or_gauss1_tape(_p(t), len(t), float(nsigmax), _p(u, _i64p))
    return v, int(u[0])


def ref_gen_hep(target, hep, n, tape, stride, earm=5, harm=1):
    """The reference's generate + complete_main (event.f, H(e,e'p) elastic,
    radiation off) from oracle/_ref on per-event tape slots.  Returns
    (dict of HEP_OUT_FIELDS + 'ok' + 'used', return code); the uniforms
    used include complete_ev's losses of the two scattered particles."""
    L = ref_lib()
    if L is None:
        raise FileNotFoundError(REF_LIB)
    if not hasattr(L, "_gen_hep_ready"):
        L.ref_gen_hep.restype = C.c_int
        L.ref_gen_hep.argtypes = [_dp, C.c_int32, C.c_int32, C.c_int32, _dp, _i32p, C.c_int64, _dp, C.c_int64,
                                  _dp, _i32p, _i32p]
        L._gen_hep_ready = True
    t = np.array([target.A, target.Z, target.rho, target.X0, target.length, target.angle, target.zoffset])
    par = np.array([hep.Ebeam, hep.dEbeam, hep.xwid, hep.ywid, hep.xoffset, hep.yoffset, hep.fr1, hep.fr2,
                    hep.coulomb, hep.e_yptar[0], hep.e_yptar[1], hep.e_xptar[0], hep.e_xptar[1], hep.e_theta,
                    hep.e_phi, hep.e_P, hep.p_theta, hep.p_phi, hep.p_P, hep.Mh])
    ipar = np.array([hep.fr_pattern, hep.using_eloss, hep.using_coulomb], np.int32)
    tape = np.ascontiguousarray(tape, dtype=np.float64)
    out = np.zeros((n, 27))
    ok = np.zeros(n, np.int32)
    used = np.zeros(n, np.int32)
    rc = L.ref_gen_hep(_p(t), int(target.can), earm, harm, _p(par), _p(ipar, _i32p), n, _p(tape), int(stride),
                       _p(out), _p(ok, _i32p), _p(used, _i32p))
    res = {k: out[:, j].copy() for j, k in enumerate(HEP_OUT_FIELDS)}
    res["ok"] = ok.astype(np.uint8)
    res["used"] = used
    return res, rc


# ---- radiative tails (radc.f, brem.f) --------------------------------------------------
class ORRad(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("rad_flag", "extrad_flag", "intcor_mode", "use_expon",
                                       "use_offshell_rad")] + \
        [("doing_tail", C.c_int * 3), ("e_arm", C.c_int), ("p_arm", C.c_int)] + \
        [(k, C.c_double) for k in ("dE_edge_test", "Egamma_gen_max")] + \
        [(k, C.c_double * 2) for k in ("edge_e_E", "edge_p_E", "edge_Em", "vertex_Trec")] + \
        [(k, C.c_double) for k in ("Egamma_tot_max", "etatzai", "targ_A")]


RAD_OUT_FIELDS = ("Egamma1", "Egamma2", "Egamma3", "gen_weight", "orig_Ein", "orig_e_E", "orig_e_delta",
                  "orig_p_E", "orig_p_P", "orig_p_delta", "eloss2", "teff2", "eloss3", "teff3", "hardcorfac")


class ORRadOut(C.Structure):
    _fields_ = [(k, _dp) for k in RAD_OUT_FIELDS] + [("ntail", _i32p)]


def make_rad(**kw):
    """Radiative settings (the or_rad fields; pairs as (min, max), doing_tail
    as three flags)."""
    r = ORRad()
    for k, v in kw.items():
        if k == "doing_tail":
            for i in range(3):
                r.doing_tail[i] = int(v[i])
        elif k in ("edge_e_E", "edge_p_E", "edge_Em", "vertex_Trec"):
            getattr(r, k)[0], getattr(r, k)[1] = v
        elif k in ("rad_flag", "extrad_flag", "intcor_mode", "use_expon", "use_offshell_rad", "e_arm", "p_arm"):
            setattr(r, k, int(v))
        else:
            setattr(r, k, float(v))
    return r


def _rad_lib():
    L = _gen_lib()
    if not hasattr(L, "_rad_ready"):
        L.or_gen_hep_rad.restype = C.c_int
        L.or_gen_hep_rad.argtypes = [C.POINTER(ORTarget), C.POINTER(ORHep), C.POINTER(ORRad), C.c_int64,
                                     C.POINTER(ORHepOut), C.POINTER(ORRadOut), C.c_int, _dp, C.c_int64,
                                     C.c_uint64, C.c_uint64, C.c_int, _i32p]
        L.or_rad_etatzai.restype = C.c_double
        L.or_rad_etatzai.argtypes = [C.c_double]
        L.or_rad_gamma.restype = C.c_double
        L.or_rad_gamma.argtypes = [C.c_double]
        L.or_lambda_dave.restype = C.c_double
        L.or_lambda_dave.argtypes = [C.c_int] * 3 + [C.c_double] * 5
        d = C.POINTER(C.c_double)
        L.or_brem.restype = C.c_double
        L.or_brem.argtypes = [C.POINTER(ORRad)] + [C.c_double] * 3 + [C.c_int, d, d, d]
        L.or_bremos.restype = C.c_double
        L.or_bremos.argtypes = [C.POINTER(ORRad)] + [C.c_double] * 14 + [C.c_int, d, d, d]
        L._rad_ready = True
    return L


def rad_etatzai(Z):
    return _rad_lib().or_rad_etatzai(float(Z))


def gen_hep_rad(target, hep, rad, n, *, rng="philox", seed=0, event0=0, tape=None, stride=0, nthreads=0):
    """The oracle's H(e,e'p) generator with radiative tails for n events.
    Returns (dict of HEP_OUT_FIELDS + RAD_OUT_FIELDS + 'ntail', 'ok', 'used',
    return code)."""
    out = {k: np.zeros(n) for k in HEP_OUT_FIELDS + RAD_OUT_FIELDS}
    out["ok"] = np.zeros(n, np.uint8)
    out["ntail"] = np.zeros(n, np.int32)
    used = np.zeros(n, np.int32)
    O = ORHepOut(*([_p(out[k]) for k in HEP_OUT_FIELDS] + [_p(out["ok"], _u8p)]))
    RO = ORRadOut(*([_p(out[k]) for k in RAD_OUT_FIELDS] + [_p(out["ntail"], _i32p)]))
    if rng == "tape":
        tape = np.ascontiguousarray(tape, dtype=np.float64)
        kind, tp = 0, _p(tape)
    else:
        kind, tp = 1, None
    rc = _rad_lib().or_gen_hep_rad(C.byref(target), C.byref(hep), C.byref(rad), n, C.byref(O), C.byref(RO), kind,
                                   tp, int(stride), int(seed), int(event0), int(nthreads or default_threads()),
                                   _p(used, _i32p))
    out["used"] = used
    return out, rc


def ref_gen_hep_rad(target, hep, rad, n, tape, stride):
    """The reference's generate with using_rad (generate_rad, radc.f) +
    complete_main from oracle/_ref on per-event tape slots.  Returns (dict of
    HEP_OUT_FIELDS + RAD_OUT_FIELDS + 'ntail', 'etatzai', 'ok', 'used',
    return code)."""
    L = ref_lib()
    if L is None:
        raise FileNotFoundError(REF_LIB)
    if not hasattr(L, "_gen_rad_ready"):
        L.ref_gen_hep_rad.restype = C.c_int
        L.ref_gen_hep_rad.argtypes = [_dp, C.c_int32, C.c_int32, C.c_int32, _dp, _i32p, _dp, _i32p, C.c_int64, _dp,
                                      C.c_int64, _dp, _i32p, _i32p]
        L._gen_rad_ready = True
    t = np.array([target.A, target.Z, target.rho, target.X0, target.length, target.angle, target.zoffset])
    par = np.array([hep.Ebeam, hep.dEbeam, hep.xwid, hep.ywid, hep.xoffset, hep.yoffset, hep.fr1, hep.fr2,
                    hep.coulomb, hep.e_yptar[0], hep.e_yptar[1], hep.e_xptar[0], hep.e_xptar[1], hep.e_theta,
                    hep.e_phi, hep.e_P, hep.p_theta, hep.p_phi, hep.p_P, hep.Mh])
    ipar = np.array([hep.fr_pattern, hep.using_eloss, hep.using_coulomb], np.int32)
    rpar = np.array([rad.dE_edge_test, rad.Egamma_gen_max, rad.edge_e_E[0], rad.edge_e_E[1], rad.edge_p_E[0],
                     rad.edge_p_E[1], rad.edge_Em[0], rad.edge_Em[1], rad.vertex_Trec[0], rad.vertex_Trec[1],
                     rad.Egamma_tot_max])
    irpar = np.array([rad.rad_flag, rad.extrad_flag, rad.intcor_mode, rad.use_expon, rad.use_offshell_rad,
                      rad.doing_tail[0], rad.doing_tail[1], rad.doing_tail[2]], np.int32)
    tape = np.ascontiguousarray(tape, dtype=np.float64)
    out = np.zeros((n, 44))
    ok = np.zeros(n, np.int32)
    used = np.zeros(n, np.int32)
    rc = L.ref_gen_hep_rad(_p(t), int(target.can), int(rad.e_arm), int(rad.p_arm), _p(par), _p(ipar, _i32p),
                           _p(rpar), _p(irpar, _i32p), n, _p(tape), int(stride), _p(out), _p(ok, _i32p),
                           _p(used, _i32p))
    res = {k: out[:, j].copy() for j, k in enumerate(HEP_OUT_FIELDS)}
    for j, k in enumerate(("Egamma1", "Egamma2", "Egamma3")):
        res[k] = out[:, 27 + j].copy()
    res["ntail"] = out[:, 30].astype(np.int32)
    for j, k in enumerate(("gen_weight", "orig_Ein", "orig_e_E", "orig_e_delta", "orig_p_E", "orig_p_P",
                           "orig_p_delta", "eloss2", "teff2", "eloss3", "teff3", "hardcorfac", "etatzai")):
        res[k] = out[:, 31 + j].copy()
    res["ok"] = ok.astype(np.uint8)
    res["used"] = used
    return res, rc
