"""Two-sample histogram comparison of transport results (full random streams).

Histograms of the focal-plane and reconstructed target variables of the
accepted events of two runs with the same inputs and different random
streams are compared with the two-sample chi-square for equal sample sizes,
chi2 = sum (a_i - b_i)^2 / (a_i + b_i) over non-empty bins (ndf = bins - 1);
the accepted fractions are compared against their binomial error.
"""
import numpy as np

VARS = ("x_fp", "dx_fp", "y_fp", "dy_fp", "dpp", "dxdz", "dydz", "y")


def chi2_two_sample(a, b, bins=40):
    lo, hi = np.percentile(np.concatenate([a, b]), [0.5, 99.5])
    ha, _ = np.histogram(a, bins=bins, range=(lo, hi))
    hb, _ = np.histogram(b, bins=bins, range=(lo, hi))
    # equal-size samples are required by this form; rescale if they differ
    ka = np.sqrt(len(b) / len(a))
    kb = np.sqrt(len(a) / len(b))
    m = (ha + hb) > 0
    chi2 = float((((ka * ha - kb * hb) ** 2) / (ha + hb))[m].sum())
    return chi2, int(m.sum()) - 1


def compare(out_a, out_b, n, bins=40, nsig=5.0):
    """Returns (ok, report).  ok: every variable's chi2/ndf within nsig standard
    deviations of 1 and the accepted fractions within nsig binomial sigmas."""
    rep = {}
    ok = True
    ma = out_a["ok"] == 1
    mb = out_b["ok"] == 1
    fa, fb = ma.mean(), mb.mean()
    f = 0.5 * (fa + fb)
    sig = np.sqrt(max(f * (1 - f), 1e-12) * 2.0 / n)
    rep["accepted"] = (float(fa), float(fb), float((fa - fb) / sig))
    ok &= abs(fa - fb) <= nsig * sig
    for v in VARS:
        c2, ndf = chi2_two_sample(out_a[v][ma], out_b[v][mb], bins)
        z = (c2 - ndf) / np.sqrt(2.0 * ndf)
        rep[v] = (round(c2 / ndf, 3), round(float(z), 2))
        ok &= z <= nsig
    return bool(ok), rep
