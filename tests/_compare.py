"""Comparison helpers shared by the parity tests."""
import numpy as np

# Scale of each coordinate, used as the absolute floor of the relative
# tolerance: |a - b| <= rtol * (max(|a|, |b|) + scale).
SCALES = {"x_fp": 1.0, "y_fp": 1.0, "dx_fp": 1e-2, "dy_fp": 1e-2, "dpp": 1.0, "y": 1.0,
          "dxdz": 1e-2, "dydz": 1e-2, "pathlen": 100.0}


def close(a, b, key, rtol):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    tol = rtol * (np.maximum(np.abs(a), np.abs(b)) + SCALES.get(key, 1.0))
    return np.abs(a - b) <= tol


def parity_report(gpu, ref, rtol, keys=None):
    """Flag agreement plus coordinate agreement on events both accept."""
    keys = keys or ["x_fp", "dx_fp", "y_fp", "dy_fp", "dpp", "y", "dxdz", "dydz", "pathlen"]
    stop_g = np.asarray(gpu["stop"])
    stop_r = np.asarray(ref["stop"])
    flag_mismatch = int((stop_g != stop_r).sum())
    both = (stop_g == 0) & (stop_r == 0)
    worst = {}
    bad = {}
    for k in keys:
        a = np.asarray(gpu[k])[both]
        b = np.asarray(ref[k])[both]
        ok = close(a, b, k, rtol)
        bad[k] = int((~ok).sum())
        d = np.abs(a - b) / (np.maximum(np.abs(a), np.abs(b)) + SCALES.get(k, 1.0))
        worst[k] = float(d.max()) if d.size else 0.0
    return {"flag_mismatch": flag_mismatch, "n_both_ok": int(both.sum()), "bad": bad,
            "worst": worst}


# Tolerances of the radiative generator's columns (relative, floor at the
# column's scale 1): 4 x the largest change that moving every elementary
# function of the radiative code (log, exp, pow, tgamma, sin, cos, asin,
# acos, atan2) by one ulp produces in the oracle over 5e4 events
# (tests/test_rad_oracle.py::test_rad_conditioning_envelope measures it).
# The photon energies are hi * y**(1/g) with g ~ 0.02-0.05 built from
# soft-photon deltas with cancellation, so last-bit differences of a libm
# reach them amplified ~1e7; the vertex kinematics after tail 1 and the
# deltas (E - P0)/P0 inherit them.  Columns not listed: 1e-12.
RAD_TOL = {"Egamma1": 2e-8, "Egamma2": 2e-8, "Egamma3": 2e-8, "gen_weight": 2e-9,
           "e_delta": 4e-9, "p_delta": 4e-9, "orig_e_delta": 4e-9, "orig_p_delta": 6e-9,
           "Ein": 4e-11, "e_E": 4e-11, "p_P": 6e-11, "p_E": 6e-11, "Q2": 8e-11, "nu": 8e-11, "q": 6e-11,
           "orig_e_E": 6e-11, "orig_p_E": 1.2e-10, "orig_p_P": 1.2e-10, "p_yptar": 2e-10, "p_theta": 1e-11,
           "p_xptar": 4e-12, "eloss3": 1e-11, "weight": 4e-12}
