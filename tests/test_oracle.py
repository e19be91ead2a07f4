"""Pins of the FP64 oracle (CPU only).

* golden vectors written by the reference's own Fortran (tests/golden/):
  the oracle must reproduce every output bit for bit on the same tapes;
* known answers and closed forms for the single routines (Philox KAT,
  RANLUX vs the C++ standard engine, LFIT on exact lines, the Marsaglia polar
  step, drifts as COSY classes, the absorption law, rotations);
* the map files: checksums and term counts against an independent count.
"""
import ctypes as C
import glob
import hashlib
import math
import os
import subprocess

import numpy as np
import pytest

from simc_gfortran_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
# the default builds' vectors, then those of the builds with the mock sieve slit
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz"))) + \
    sorted(glob.glob(os.path.join(HERE, "golden", "sieve", "*.npz")))


def _sieve(g):
    return int(g["sieve"]) if "sieve" in g.files else 0


# ---------------------------------------------------------------------------
# golden vectors of the reference
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_matches_reference_golden(oracle, path):
    g = np.load(path)
    stride = int(g["stride"])
    n = len(g["in_dpp"])
    tape = W.uniform_tape(n, stride, int(g["tape_seed"]))
    assert hashlib.sha256(tape.tobytes()).hexdigest() == str(g["tape_sha256"])
    p_spec, m2, ms, wcs, decay, coll, ctau = g["settings"]
    ev = {k: g["in_" + k] for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")}
    out = oracle.run(int(g["arm"]), ev, p_spec=p_spec, m2=m2, ms=bool(ms), wcs=bool(wcs),
                     decay=bool(decay), coll=bool(coll), ctau=ctau, rng="tape", tape=tape,
                     stride=stride, fit_f32=True, sieve=_sieve(g))
    assert out["rc"] == 0
    for k in ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen",
              "resmult", "m2", "ok", "used"):
        np.testing.assert_array_equal(out[k], g["out_" + k], err_msg=k)
    names = {1: oracle.HMS_COUNTERS, 2: oracle.SOS_COUNTERS, 3: oracle.HRS_COUNTERS, 4: oracle.HRS_COUNTERS,
             5: oracle.SHMS_COUNTERS}[int(g["arm"])]
    for i, name in enumerate(names):
        if name != "stop_id":  # a last-value register, not a counter
            assert out["counters"][i] == g["counters"][i], name


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_exit_points_match_reference_per_event(oracle, path):
    """Each event's exit point: the STOP counters the oracle moves for that
    event alone equal the counters the reference moved for it (recorded per
    event in the golden file), for every event."""
    g = np.load(path)
    stride = int(g["stride"])
    n = len(g["in_dpp"])
    tape = W.uniform_tape(n, stride, int(g["tape_seed"])).reshape(n, stride)
    p_spec, m2, ms, wcs, decay, coll, ctau = g["settings"]
    ev = {k: g["in_" + k] for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")}
    names = {1: oracle.HMS_COUNTERS, 2: oracle.SOS_COUNTERS, 3: oracle.HRS_COUNTERS, 4: oracle.HRS_COUNTERS,
             5: oracle.SHMS_COUNTERS}[int(g["arm"])]
    skip = names.index("stop_id") if "stop_id" in names else -1
    bad = []
    for i in range(n):
        out = oracle.run(int(g["arm"]), {k: v[i:i + 1] for k, v in ev.items()}, p_spec=p_spec, m2=m2,
                         ms=bool(ms), wcs=bool(wcs), decay=bool(decay), coll=bool(coll), ctau=ctau,
                         rng="tape", tape=tape[i], stride=stride, fit_f32=True, nthreads=1, sieve=_sieve(g))
        m = 0
        for c, v in enumerate(out["counters"]):
            if v and c != skip:
                m |= 1 << c
        ref = int(g["stop_mask"][i]) & ~((1 << skip) if skip >= 0 else 0)
        if m != ref:
            bad.append((i, int(out["stop"][0]), hex(m), hex(ref)))
    assert not bad, bad[:10]


# ---------------------------------------------------------------------------
# random numbers
# ---------------------------------------------------------------------------
def test_philox_known_answers(oracle):
    # Random123 kat_vectors, philox4x32_10
    kat = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
           ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
            (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, want in kat:
        assert tuple(oracle.philox_block(ctr, key)) == want


def test_philox_uniform_mapping(oracle):
    L = oracle.lib()
    blk = oracle.philox_block((0, 5, 0, 0x53494D43), (7, 0))
    for k in range(4):
        assert L.or_philox_uniform(7, 0, 5, k) == (blk[k] + 0.5) * 2.0 ** -32
    u = np.array([L.or_philox_uniform(1, 2, 3, k) for k in range(4000)])
    assert 0.0 < u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.02


@pytest.fixture(scope="module")
def ranlux_std(tmp_path_factory):
    d = tmp_path_factory.mktemp("ranlux")
    exe = str(d / "ranlux_std")
    subprocess.run(["g++", "-O1", "-std=c++17", os.path.join(HERE, "ranlux_std.cpp"), "-o", exe],
                   check=True)
    return exe


def test_std_ranlux_engine_known_answers(ranlux_std):
    # [rand.predef] fixes the 10000th value of ranlux24_base and ranlux24
    out = subprocess.run([ranlux_std, "1", "kat"], capture_output=True, text=True, check=True)
    assert out.stdout.split() == ["7937952", "9901578"]


@pytest.mark.parametrize("seed", [1, 314159265, 123456789])
def test_ranlux_matches_std_discard_block(oracle, ranlux_std, seed):
    """James's RANLUX (cern/ranlux.f), luxury 3, equals
    discard_block_engine<ranlux24_base, 223, 24> started from the same state."""
    L = oracle.lib()
    n = 3000
    isd = (C.c_int32 * 25)()
    rvec = (C.c_double * n)()
    L.or_ranlux_sequence(3, seed, isd, rvec, n)
    seeds = list(isd)[:24]
    i24 = abs(isd[24]) % 100
    assert i24 == 24 and isd[24] > 0  # fresh RLUXGO state: next write at 24, no carry
    # seeds(24) is the oldest value, seeds(1) the newest (transp of NEXT(), ranlux.f:64-69)
    state = " ".join(str(s) for s in seeds[::-1]) + " 0"
    out = subprocess.run([ranlux_std, str(n)], input=state + "\n", capture_output=True, text=True,
                         check=True)
    want = np.array([int(v) for v in out.stdout.split()], dtype=np.float64)
    got = np.array(list(rvec))
    # RANLUX pads values below 2^-12 with lower bits (ranlux.f:125-129)
    assert np.array_equal(np.floor(got * 2.0 ** 24), want)
    big = want >= 4096
    assert np.array_equal(got[big], want[big] * 2.0 ** -24)


# ---------------------------------------------------------------------------
# single routines
# ---------------------------------------------------------------------------
def gauss_on_tape(oracle, tape, nsig=99.0):
    L = oracle.lib()
    L.or_gauss1_tape.restype = C.c_double
    L.or_gauss1_tape.argtypes = [C.POINTER(C.c_double), C.c_int64, C.c_double, C.POINTER(C.c_int64)]
    t = (C.c_double * len(tape))(*tape)
    used = C.c_int64()
    v = L.or_gauss1_tape(t, len(tape), nsig, C.byref(used))
    return v, used.value


def test_gauss1_polar_step(oracle):
    # u = 0.75, 0.75 -> v = (0.5, 0.5), s = 1/2: g = 0.5 sqrt(4 ln 2) = sqrt(ln 2)
    v, used = gauss_on_tape(oracle, [0.75, 0.75])
    assert used == 2 and abs(v - math.sqrt(math.log(2.0))) < 1e-15
    # s > 1 is rejected and the next pair is taken; s == 0 is rejected too
    v, used = gauss_on_tape(oracle, [0.95, 0.95, 0.5, 0.5, 0.25, 0.5])
    assert used == 6 and abs(v - (-0.5) * math.sqrt(-2.0 * math.log(0.25) / 0.25)) < 1e-15
    # truncation at nsig rejects the pair
    v, used = gauss_on_tape(oracle, [0.5000001, 0.5, 0.75, 0.75], nsig=3.0)
    assert used == 4


# ---------------------------------------------------------------------------
# Box-Muller normals of the counter-based stream
# ---------------------------------------------------------------------------
def ulp(x):
    return math.ulp(x)


def test_bm_log_against_libm(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(11)
    us = list(rng.uniform(0.0, 1.0, 20000)) + [2.0 ** -33, 0.5, 0.25, 1.0 - 2.0 ** -33, 1.0 - 2.0 ** -20,
                                              math.sqrt(0.5), math.nextafter(math.sqrt(0.5), 1.0),
                                              0.7071067811865476, 0.9999999, 1e-9]
    worst = 0.0
    for u in us:
        if u <= 0.0:
            continue
        got, want = L.or_bm_log(u), math.log(u)
        worst = max(worst, abs(got - want) / ulp(want))
    assert worst <= 2.0, worst
    # powers of two are exact multiples of ln2 (two-part constant)
    for e in range(1, 34):
        assert abs(L.or_bm_log(2.0 ** -e) - (-e * math.log(2.0))) <= ulp(e * math.log(2.0))


def test_bm_sincos2pi_against_libm(oracle):
    L = oracle.lib()
    s, c = C.c_double(), C.c_double()
    exact = {0.0: (0.0, 1.0), 0.25: (1.0, 0.0), 0.5: (0.0, -1.0), 0.75: (-1.0, 0.0),
             0.125: (math.sqrt(0.5), math.sqrt(0.5)), 1.0 / 12: (0.5, math.sqrt(3) / 2)}
    for u, (ws, wc) in exact.items():
        L.or_bm_sincos2pi(u, C.byref(s), C.byref(c))
        assert abs(s.value - ws) <= 2e-16 and abs(c.value - wc) <= 2e-16, (u, s.value, c.value)
    rng = np.random.default_rng(12)
    worst = 0.0
    for u in rng.uniform(0.0, 1.0, 20000):
        L.or_bm_sincos2pi(u, C.byref(s), C.byref(c))
        # libm's argument 2 pi u carries up to ~1e-15 of rounding itself
        worst = max(worst, abs(s.value - math.sin(2 * math.pi * u)), abs(c.value - math.cos(2 * math.pi * u)))
        assert abs(s.value * s.value + c.value * c.value - 1.0) <= 5e-16
    assert worst <= 2e-15, worst


def test_bm_tables_match_their_generator():
    """simc_bm_tables.h is what gen_bm_tables.py writes (mpmath at 60 digits,
    one rounding; every log interval checked for an exact fma(m, invc, -1))."""
    import importlib.util
    csrc = os.path.join(os.path.dirname(HERE), "simc_gfortran_amd", "csrc")
    spec = importlib.util.spec_from_file_location("gen_bm_tables", os.path.join(csrc, "gen_bm_tables.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    with open(os.path.join(csrc, "simc_bm_tables.h")) as f:
        assert f.read() == gen.render()


def test_bm_log_stream_values_within_one_ulp(oracle):
    """On values of the stream, (w + 1/2) 2^-32, the logarithm is within one
    unit in the last place of the correctly rounded one, and sin/cos of 2 pi u
    within 1.2e-16 (50-digit references)."""
    import mpmath
    L = oracle.lib()
    mpmath.mp.dps = 40
    w = np.random.default_rng(21).integers(0, 2 ** 32, 4000, dtype=np.uint64)
    us = list((w.astype(np.float64) + 0.5) * 2.0 ** -32) + [0.5 * 2.0 ** -32, 1.0 - 0.5 * 2.0 ** -32]
    s, c = C.c_double(), C.c_double()
    for u in us:
        want = float(mpmath.log(mpmath.mpf(float(u))))
        assert abs(L.or_bm_log(float(u)) - want) <= ulp(want), u
        L.or_bm_sincos2pi(float(u), C.byref(s), C.byref(c))
        a = 2 * mpmath.pi * mpmath.mpf(float(u))
        assert abs(s.value - float(mpmath.sin(a))) <= 1.2e-16 and abs(c.value - float(mpmath.cos(a))) <= 1.2e-16


def test_philox_normals_pairing(oracle):
    """First pair of event 9, seed 5: u1, u2 = words 0, 1 of Philox block 0;
    the cosine value comes first, the sine value second."""
    blk = oracle.philox_block((0, 9, 0, 0x53494D43), (5, 0))
    u1, u2 = [(w + 0.5) * 2.0 ** -32 for w in blk[:2]]
    r = math.sqrt(-2.0 * math.log(u1))
    g, used = oracle.philox_normals(5, 9, 2)
    assert used == 2
    assert abs(g[0] - r * math.cos(2 * math.pi * u2)) <= 1e-14
    assert abs(g[1] - r * math.sin(2 * math.pi * u2)) <= 1e-14
    g3, used3 = oracle.philox_normals(5, 9, 3)
    assert used3 == 4 and (g3[:2] == g).all()


def test_philox_normals_are_standard_normal(oracle):
    from scipy import stats
    vals = np.concatenate([oracle.philox_normals(2024, e, 2000)[0] for e in range(100)])
    assert stats.kstest(vals, "norm").pvalue > 1e-3
    assert abs(vals.mean()) < 5 / math.sqrt(len(vals))
    assert abs(vals.var() - 1.0) < 5 * math.sqrt(2.0 / len(vals))
    # the two members of a pair are uncorrelated
    pairs = vals.reshape(-1, 2)
    assert abs(np.corrcoef(pairs[:, 0], pairs[:, 1])[0, 1]) < 5 / math.sqrt(len(pairs))


def lfit(oracle, z, y, f32):
    L = oracle.lib()
    n = len(z)
    if f32:
        T = C.c_float
        fn = L.or_lfit_f32
    else:
        T = C.c_double
        fn = L.or_lfit_f64
    fn.argtypes = [C.POINTER(T), C.POINTER(T), C.c_int, C.c_int, C.POINTER(T), C.POINTER(T), C.POINTER(T)]
    a, b, e = T(), T(), T()
    fn((T * n)(*z), (T * n)(*y), n, 0, C.byref(a), C.byref(b), C.byref(e))
    return a.value, b.value, e.value


@pytest.mark.parametrize("f32", [True, False])
def test_lfit_exact_line_and_zero_skip(oracle, f32):
    z = [-52.0, -50.0, -48.0, -46.0, 30.0, 32.0, 34.0, 36.0]
    y = [0.25 * zz + 3.0 for zz in z]   # exactly representable points on a line
    a, b, e = lfit(oracle, z, y, f32)
    assert a == 0.25 and b == 3.0 and e == 0.0
    # a zero hit is ignored (cern/lfit.f:30): same answer with an extra (z, 0) point
    a2, b2, _ = lfit(oracle, z + [10.0], y + [0.0], f32)
    assert (a2, b2) == (a, b)
    # fewer than two points: A = B = 0
    assert lfit(oracle, [1.0, 2.0], [0.0, 5.0], f32)[:2] == (0.0, 0.0)


def _lfit_golden():
    g = np.load(os.path.join(HERE, "golden", "lfit", "ref_lfit.npz"))
    X, Y = W.lfit_point_sets(int(g["n"]), int(g["seed"]))
    assert hashlib.sha256(X.tobytes() + Y.tobytes()).hexdigest() == str(g["points_sha256"])
    return X, Y, g


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def test_lfit_matches_reference_vectors(oracle):
    """or_lfit_f32 against the reference's own LFIT (cern/lfit.f:11-60, built
    as gfortran compiles it) on 12,000 REAL*4 point sets: A, B and E
    identical bit for bit (non-representable hits, skipped zero hits, the
    edge cases of the routine's branches)."""
    X, Y, g = _lfit_golden()
    A, B, E = oracle.lfit_f32(X, Y)
    for name, got, ref in (("A", A, g["A"]), ("B", B, g["B"]), ("E", E, g["E"])):
        np.testing.assert_array_equal(_bits(got), _bits(ref), err_msg=name)
    # the vectors reach every branch: no fit (A = B = E = 0), two points (E = 0),
    # a full fit
    assert ((g["A"] == 0) & (g["B"] == 0)).sum() > 100
    assert ((g["E"] == 0) & (g["A"] != 0)).sum() > 30
    assert (g["E"] > 0).sum() > 10000


def test_lfit_sensitivity_of_the_vectors(oracle):
    """The vectors tell the REAL*4-in / REAL*8-sums semantics from the
    all-float and the all-double variants of the same formulas."""
    X, Y, g = _lfit_golden()
    old_err = np.seterr(divide="ignore", invalid="ignore")  # sets with SUMXX = 0 give no fit
    sel = slice(None)
    x, y = X[sel].astype(np.float64), Y[sel].astype(np.float64)
    keep = y != 0.0
    cnt = keep.sum(axis=1)
    ok = cnt > 1
    x, y, keep, cnt = x[ok], y[ok], keep[ok], cnt[ok]
    refA, refB = g["A"][sel][ok], g["B"][sel][ok]

    def fit(t):
        xs, ys = x.astype(t), y.astype(t)
        c = cnt.astype(t)
        xm = (xs * keep).sum(axis=1, dtype=t) / c
        ym = (ys * keep).sum(axis=1, dtype=t) / c
        sx = (xs - xm[:, None]) * keep
        sy = (ys - ym[:, None]) * keep
        a = ((sx * sy).sum(axis=1, dtype=t) / (sx * sx).sum(axis=1, dtype=t)).astype(np.float32)
        b = (ym - a.astype(t) * xm).astype(np.float32)
        return a, b

    a32, b32 = fit(np.float32)
    assert (_bits(a32) != _bits(refA)).sum() + (_bits(b32) != _bits(refB)).sum() > 100
    # B from the double slope instead of the REAL*4 A
    xm = (x * keep).sum(axis=1) / cnt
    ym = (y * keep).sum(axis=1) / cnt
    sx = (x - xm[:, None]) * keep
    sy = (y - ym[:, None]) * keep
    a64 = (sx * sy).sum(axis=1) / (sx * sx).sum(axis=1)
    b_dbl = (ym - a64 * xm).astype(np.float32)
    np.seterr(**old_err)
    assert (_bits(b_dbl) != _bits(refB)).sum() > 50


def test_lfit_matches_live_reference(oracle):
    """The same comparison against the reference library itself on fresh
    point sets (only where oracle/_ref is built)."""
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (no reference tree)")
    X, Y = W.lfit_point_sets(20000, 77)
    ref = oracle.ref_lfit(X, Y)
    got = oracle.lfit_f32(X, Y)
    for name, a, b in zip("ABE", got, ref):
        np.testing.assert_array_equal(_bits(a), _bits(b), err_msg=name)


def test_pion_absorption_law(oracle):
    L = oracle.lib()
    p = 1500.0
    assert L.or_pion_coll_absorb(p, 0.0) == 1.0
    t1 = L.or_pion_coll_absorb(p, 0.3)
    t2 = L.or_pion_coll_absorb(p, 0.6)
    assert 0.0 < t1 < 1.0 and abs(t2 - t1 * t1) < 1e-15   # exp(-2 x) = exp(-x)^2
    assert L.or_pion_coll_absorb(20000.0, 0.3) < 1.0      # above the table: last entry


def test_powi_is_binary_exponentiation(oracle):
    L = oracle.lib()
    for x in (1.7, -0.3, 12.5, 1e-3):
        assert L.or_powi(x, 0) == 1.0
        assert L.or_powi(x, 1) == x
        assert L.or_powi(x, 2) == x * x
        assert L.or_powi(x, 4) == (x * x) * (x * x)
        for n in range(7):
            assert abs(L.or_powi(x, n) - x ** n) <= 4e-16 * abs(x ** n)
    assert L.or_powi(2.0, 6) == 64.0 and L.or_powi(2.0, -2) == 0.25


# ---------------------------------------------------------------------------
# the COSY maps
# ---------------------------------------------------------------------------
MAP_SHA = {
    "hms/forward_cosy.dat": "7e9a5bc770f46fda04f3fba1ff54b9b00f6e1e2e61aee027b6b5b700e3a729c7",
    "hms/recon_cosy.dat": "61cf89af0c3a5f7370ca69b84f2a2cd47d8b201c34eb4bbd761afd83e51d80ca",
    "shms/shms_forward.dat": "398595898491378e3d4ade794e1af6b324afd1fcccc20ec9bb5d4456ff75cffc",
    "shms/shms_recon.dat": "7477b37ed142de7743b50080f38d7e45074196e7387ee899f79d04c86bb655b0",
    # sos/forward_cosy.dat, sos/recon_cosy.dat of the reference tree, copied verbatim
    "sos/forward_cosy.dat": "ffb83fecf4edadeaed99a6c138a7a555a262922f690f674d2bba77dce922ed7b",
    "sos/recon_cosy.dat": "0093b61d39ec60aecc0e84d09dd2b03283b1c258dc2ccf56a2691669b82a17ea",
    # hrsr/ and hrsl/ hrs_forward_cosy.dat, hrs_recon_cosy.dat of the reference tree, copied
    # verbatim (the two arms ship identical files)
    "hrsr/hrs_forward_cosy.dat": "8bfd3cfafdd5aebaad2a60f9901e918193c64f929191e075a4c6711fd1d943f3",
    "hrsr/hrs_recon_cosy.dat": "e43c5816a93d9e3cbdd86681a3acb2c8e3a3c1a58d52f8eb3a0f5d8bd71b787d",
    "hrsl/hrs_forward_cosy.dat": "8bfd3cfafdd5aebaad2a60f9901e918193c64f929191e075a4c6711fd1d943f3",
    "hrsl/hrs_recon_cosy.dat": "e43c5816a93d9e3cbdd86681a3acb2c8e3a3c1a58d52f8eb3a0f5d8bd71b787d",
}


# the reference tree's other SHMS tunes (shms/), copied verbatim: their term
# structures differ from the default pair's
TUNE_SHA = {
    "shms/shms-2017-26cm-monte_quads_p18_forward.dat":
        "3ff63af667d2bf557b90a349348c93eb1b601d1c343eecdaec3a2dc271280221",
    "shms/shms-2017-26cm-monte_quads_p18_recon.dat":
        "e22c8049dc359745e4237414c145d92983d17a5bc4260621e10feead5ab3f0b3",
    "shms/shms-2011-26cm-monte_ideal_6ord_forward.dat":
        "c757b6f3446dc20108186118f3ca0f16565a3e63672ed596cf38a7398c87c71b",
    "shms/shms-2011-26cm-monte_ideal_6ord_recon.dat":
        "015f8e891b9f7a81bd008d189e8afb7a4787bfd4bb1f8f3ecbe48859247090d3",
}


def test_map_files_checksums():
    for rel, sha in MAP_SHA.items():
        with open(os.path.join(REPO, "data", "maps", rel), "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == sha, rel
    for rel, sha in TUNE_SHA.items():
        with open(os.path.join(REPO, "data", "maps_tunes", rel), "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == sha, rel


@pytest.mark.parametrize("tune", ["shms-2017-26cm-monte_quads_p18", "shms-2011-26cm-monte_ideal_6ord"])
def test_oracle_runs_other_tunes(oracle, tmp_path, tune):
    """The oracle reads any tune (transp_init and the recon reader are
    structure-free): a tune's map directory gives a run whose reconstruction
    differs from the default tune's and accepts a plausible fraction."""
    from simc_gfortran_amd import workloads as W
    d = oracle.tune_map_dir(str(tmp_path), tune)
    ev = W.generate(W.SHMS_C1, 4000)
    kw = dict(p_spec=W.SHMS_C1.p_spec, m2=W.SHMS_C1.m2, rng="philox", seed=3, fit_f32=False)
    a = oracle.run(5, ev, map_dir=d, **kw)
    b = oracle.run(5, ev, **kw)
    assert a["rc"] == 0 and 0.2 < a["ok"].mean() < 0.7
    both = (a["ok"] == 1) & (b["ok"] == 1)
    assert both.sum() > 500 and np.abs(a["dpp"][both] - b["dpp"][both]).max() > 1e-6


class FwdMap(C.Structure):
    _fields_ = [("n_classes", C.c_int), ("n_terms", C.c_int * 41), ("length", C.c_double * 41),
                ("adrift", C.c_int * 41), ("driftdist", C.c_double * 41),
                ("coeff", C.c_double * (41 * 1000 * 5)), ("expon", C.c_short * (41 * 1000 * 5))]


class ReconMap(C.Structure):
    _fields_ = [("n_terms", C.c_int), ("max_order", C.c_int), ("coeff", C.c_double * 4000),
                ("expon", C.c_short * 5000)]


class Maps(C.Structure):
    _fields_ = [("hms_fwd", C.POINTER(FwdMap)), ("shms_fwd", C.POINTER(FwdMap)),
                ("hms_rec", C.POINTER(ReconMap)), ("shms_rec", C.POINTER(ReconMap)),
                ("sos_fwd", C.POINTER(FwdMap)), ("sos_rec", C.POINTER(ReconMap)),
                ("hrsr_fwd", C.POINTER(FwdMap)), ("hrsl_fwd", C.POINTER(FwdMap)),
                ("hrsr_rec", C.POINTER(ReconMap)), ("hrsl_rec", C.POINTER(ReconMap))]


@pytest.fixture(scope="module")
def maps(oracle):
    return C.cast(C.c_void_p(oracle.maps()), C.POINTER(Maps)).contents


# Term counts per class after dropping the time-of-flight line, counted
# independently of both parsers (one TOF line per class).
HMS_TERMS = {1: 9, 2: 291, 3: 278, 4: 9, 5: 339, 6: 318, 7: 9, 8: 269, 9: 253, 10: 9, 11: 361, 12: 9}
SHMS_TERMS = {3: 212, 7: 279, 8: 284, 12: 287, 13: 288, 17: 271, 18: 280, 23: 220, 30: 218,
              **{k: 138 for k in range(24, 30)},
              **{k: 9 for k in (1, 2, 4, 5, 6, 9, 10, 11, 14, 15, 16, 19, 20, 21, 22, 31, 32)}}


def test_forward_map_structure(maps):
    h, s = maps.hms_fwd.contents, maps.shms_fwd.contents
    assert h.n_classes == 12 and s.n_classes == 32
    for cls, n in HMS_TERMS.items():
        assert h.n_terms[cls - 1] == n, cls
    for cls, n in SHMS_TERMS.items():
        assert s.n_terms[cls - 1] == n, cls
    # pure drifts are exactly the 9-term classes; their length matches the comment
    for m, terms in ((h, HMS_TERMS), (s, SHMS_TERMS)):
        for cls, n in terms.items():
            assert bool(m.adrift[cls - 1]) == (n == 9), cls
            if n == 9 and m.length[cls - 1] > 0:
                assert abs(m.driftdist[cls - 1] - m.length[cls - 1]) < 0.01
    assert abs(h.driftdist[0] - 216.075) < 1e-9


def test_recon_map_structure(maps):
    for r in (maps.hms_rec.contents, maps.shms_rec.contents):
        assert r.n_terms == 461 and r.max_order <= 6
        e = np.array(r.expon[:5 * r.n_terms]).reshape(-1, 5)
        assert not (e.sum(axis=1) == 0).any()          # no constant term
        assert len({tuple(x) for x in e}) == r.n_terms  # no duplicate monomial


def test_drift_class_is_a_field_free_drift(oracle, maps):
    """A pure-drift class evaluated as a COSY polynomial equals `project`:
    x += L x', y += L y' and the path-length term reproduces L(sqrt(1+x'^2+y'^2)-1)."""
    L = oracle.lib()
    L.or_transp_eval.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    fwd = C.cast(maps.hms_fwd, C.c_void_p)
    h = maps.hms_fwd.contents
    for cls in (1, 4, 7, 10, 12):
        Ld = h.driftdist[cls - 1]
        for xp, yp in ((0.01, -0.02), (-0.03, 0.005)):
            ray = (C.c_double * 5)(1.5, xp * 1000.0, -2.0, yp * 1000.0, 3.0)
            out = (C.c_double * 5)()
            L.or_transp_eval(fwd, cls, ray, out)
            assert abs(out[0] - (1.5 + Ld * xp)) < 1e-11
            assert abs(out[2] - (-2.0 + Ld * yp)) < 1e-11
            assert abs(out[1] - xp * 1000.0) < 1e-12 and abs(out[3] - yp * 1000.0) < 1e-12
            want_dl = Ld * (math.sqrt(1 + xp * xp + yp * yp) - 1.0)
            assert abs(-out[4] - want_dl) < 1e-6 * Ld * (xp * xp + yp * yp) + 1e-9


# Central path lengths mc_hrsr / mc_hrsl pass to transp (hrsl/mc_hrsl.f:244-259,
# :303-318, :388, :451-466).  transp compares each with the length the map's comments
# give and warns beyond 0.01 (shared/transp.f:125-130): the lengths agree.
HRS_LENGTHS = {2: 62.75333333, 3: 31.37666667, 5: 121.77333333, 6: 60.88666667, 8: 659.73445725,
               10: 121.7866667, 11: 60.89333333}


def test_hrs_maps_match_the_central_lengths(maps):
    for m in (maps.hrsr_fwd.contents, maps.hrsl_fwd.contents):
        assert m.n_classes == 12
        for cls, zd in HRS_LENGTHS.items():
            assert not m.adrift[cls - 1]
            assert m.length[cls - 1] > 0.01 and abs(m.length[cls - 1] - zd) <= 0.01, cls
        for cls in (1, 4, 7, 9, 12):  # the drifts mc_hrs? checks with adrift (:235, :294, :364, :442, :505)
            assert m.adrift[cls - 1], cls
    for r in (maps.hrsr_rec.contents, maps.hrsl_rec.contents):
        e = np.array(r.expon[:5 * r.n_terms]).reshape(-1, 5)
        # every monomial of degree <= 5 in 5 variables once (C(10, 5) = 252), the constant included
        assert r.n_terms == 252 and e.sum(axis=1).max() == 5
        assert len({tuple(x) for x in e}) == 252


def test_central_ray_stays_on_axis(oracle, maps):
    L = oracle.lib()
    L.or_transp_eval.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    for ptr, n in ((maps.hms_fwd, 12), (maps.shms_fwd, 32), (maps.sos_fwd, 10), (maps.hrsr_fwd, 12)):
        for cls in range(1, n + 1):
            out = (C.c_double * 5)()
            L.or_transp_eval(C.cast(ptr, C.c_void_p), cls, (C.c_double * 5)(0, 0, 0, 0, 0), out)
            assert max(abs(v) for v in out) < 1e-12, cls
