"""The oracle's target-field tracking (trg_track.f restated) on the CPU.

Pins: the reference's own trgInit / track_from_tgt / track_to_tgt outputs
(tests/golden/field/*.npz, written by tests/golden/make_golden.py field from
oracle/_ref), reproduced bit for bit; the field interpolation at grid nodes
and its rotation invariance; a straight line in zero field; and, in the
uniform 5 T test field along the beam, the conserved transverse slope and
the cyclotron turning angle 90 B L / (E v_z) that the equations of motion
(factor 90/E, trg_track.f:181-182) imply.
"""
import glob
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = sorted(glob.glob(os.path.join(HERE, "golden", "field", "*.npz")))


@pytest.fixture(scope="module")
def O():
    from oracle import oracle
    return oracle


def _field(O, g):
    mp = str(g["map"])
    path = O.FIELD_MAP if mp == "map" else (None if mp == "None" else mp)
    return O.field(path, float(g["theta_e"]), float(g["theta_p"]))


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p)[:-4] for p in GOLD])
def test_oracle_matches_reference_field_golden(O, path):
    g = np.load(path)
    F = _field(O, g)
    if str(g["kind"]) == "from":
        ev = {k: g["in_" + k] for k in ("x", "y", "z", "dx", "dy")}
        out = O.track_from_tgt(F, ev, g["mom"], float(g["mass"]), int(g["spect"]))
        keys = ("x", "y", "z", "dx", "dy", "ok")
    else:
        trk = {k: g["trk_" + k] for k in ("x_fp", "dx_fp", "y_fp", "dy_fp")}
        first = {k: g["in_" + k] for k in ("delta", "y", "dx", "dy")}
        out = O.track_to_tgt(F, int(g["arm"]), trk, first, g["frx"], g["fry"], g["mom"], float(g["mass"]),
                             float(g["ctheta"]), float(g["stheta"]), int(g["spect"]))
        keys = ("delta", "y", "dx", "dy", "ok")
        assert out["ok"].mean() > 0.9
    for k in keys:
        np.testing.assert_array_equal(out[k], g["out_" + k], err_msg=k)


def test_field_map_at_grid_nodes(O):
    """At a grid node (theta 0) the interpolation returns the map's B_z and
    B_r (trg_track.f:384-405); the map file's columns 3, 4 are B_z, B_r."""
    F = O.field(O.FIELD_MAP, 0.0, 0.0)
    rows = np.loadtxt(O.FIELD_MAP)
    for iz, ir in ((0, 0), (3, 2), (10, 5), (20, 30), (7, 44)):
        z, r = 2.0 * iz, 2.0 * ir
        row = rows[ir * 51 + iz]
        assert row[0] == z and row[1] == r
        B = O.field_at(F, [0.0, r, z], 1)
        assert B[2] == row[2]
        assert B[0] == 0.0
        assert abs(B[1] - row[3]) <= 1e-15 * max(1.0, abs(row[3]))
    # outside the grid (|z| or r beyond 100 cm) the field is zero
    assert not O.field_at(F, [0.0, 0.0, 101.0], 1).any()


def test_field_rotation_invariance(O):
    """The field angle rotates the map about the x axis: |B| at a rotated
    point equals |B| at the unrotated one."""
    F0 = O.field(O.FIELD_MAP, 0.0, 0.0)
    th = 17.0
    F1 = O.field(O.FIELD_MAP, 0.0, th)
    c, s = np.cos(th * 3.141592653 / 180.0), np.sin(th * 3.141592653 / 180.0)
    rng = np.random.default_rng(4)
    for _ in range(50):
        x = rng.uniform(-20, 20, 3)
        xr = np.array([x[0], s * x[2] + c * x[1], c * x[2] - s * x[1]])  # the field frame (:376-378)
        b0 = np.linalg.norm(O.field_at(F0, xr, 1))
        b1 = np.linalg.norm(O.field_at(F1, x, 1))
        assert abs(b0 - b1) <= 1e-13 * max(b0, 1e-3)


def _vertices(n, seed):
    rng = np.random.default_rng(seed)
    return dict(x=rng.uniform(-0.5, 0.5, n), y=rng.uniform(-0.5, 0.5, n), z=rng.uniform(-1.5, 1.5, n),
                dx=rng.uniform(-0.06, 0.06, n), dy=rng.uniform(-0.04, 0.04, n))


def test_zero_field_is_a_straight_line(O):
    F = O.field("0", 0.0, 0.0)
    ev = _vertices(500, 1)
    out = O.track_from_tgt(F, ev, np.full(500, 1500.0), 938.272, 1)
    assert out["ok"].all()
    np.testing.assert_allclose(out["z"], 100.0, rtol=0, atol=1e-9)
    np.testing.assert_allclose(out["x"], ev["x"] + (100.0 - ev["z"]) * ev["dx"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(out["y"], ev["y"] + (100.0 - ev["z"]) * ev["dy"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(out["dx"], ev["dx"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("sign", [+1, -1])
def test_uniform_field_turning_angle(O, sign):
    """Uniform 5 T along the beam over |z| <= 26 cm (ramping to 0 at 28 cm,
    r <= 16 cm; trg_track.f:311-323): v_z and |v_T| are conserved, and the
    transverse velocity turns by -(90/E) * integral(B dz) / v_z between the
    vertex and z = 100 cm."""
    F = O.field(None, 0.0, 0.0)
    n = 300
    ev = _vertices(n, 2)
    mass = 938.272
    p = np.random.default_rng(3).uniform(800.0, 2500.0, n)
    mom = sign * p
    out = O.track_from_tgt(F, ev, mom, mass, 1)
    assert out["ok"].all()
    st_in = np.hypot(ev["dx"], ev["dy"])
    st_out = np.hypot(out["dx"], out["dy"])
    np.testing.assert_allclose(st_out, st_in, rtol=2e-4)
    E = sign * np.sqrt(p * p + mass * mass)
    vel = p / np.sqrt(p * p + mass * mass) * 29.9792458
    vz = vel / np.sqrt(1 + ev["dx"] ** 2 + ev["dy"] ** 2)
    bdz = 5.0 * (26.0 - ev["z"]) + 5.0 * 1.0  # the linear ramp 26 -> 28 cm adds 1 cm
    want = -(90.0 / E) * bdz / vz
    got = np.arctan2(out["dy"], out["dx"]) - np.arctan2(ev["dy"], ev["dx"])
    got = (got + np.pi) % (2 * np.pi) - np.pi
    np.testing.assert_allclose(got, want, rtol=2e-3, atol=2e-4)


def test_field_map_checksum(O):
    """data/fields/trg_field_map.dat is the reference's trg_field_map.dat, copied verbatim."""
    import hashlib
    with open(O.FIELD_MAP, "rb") as f:
        assert hashlib.sha256(f.read()).hexdigest() == \
            "ee7952511d92012ff9cf74497b8cedce6fe3e3925b3ac54651af3bba741fa4fd"
