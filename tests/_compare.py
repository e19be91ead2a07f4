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
