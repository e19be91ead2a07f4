"""Write the golden vectors of tests/golden/*.npz from the reference itself.

The reference's Fortran hot path (mc_hms, mc_shms and everything they call)
is compiled from its own sources by `make -C oracle ref` into
oracle/_ref/libsimcref.so; this script drives it over small event samples on
per-event uniform tapes and stores inputs, tape recipe and every output.  The
tapes are regenerated from their seed by the tests (sha256 recorded here).

    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle import oracle as O  # noqa: E402
from simc_gfortran_amd import workloads as W  # noqa: E402

MPI = 139.57018

# name, arm, workload (event distribution), overrides
CASES = [
    ("hms_proton_ms_wcs", 1, W.HMS_C0, dict(p_spec=3000.0, m2=W.MP ** 2)),
    ("shms_electron_ms_wcs", 5, W.SHMS_C1, dict(p_spec=6000.0, m2=W.ME ** 2)),
    ("hms_proton_no_ms_no_wcs", 1, W.HMS_C0, dict(p_spec=3000.0, m2=W.MP ** 2, ms=False, wcs=False)),
    ("shms_electron_no_ms_no_wcs", 5, W.SHMS_C1, dict(p_spec=6000.0, m2=W.ME ** 2, ms=False, wcs=False)),
    ("hms_pion_coll_decay", 1, W.HMS_C0, dict(p_spec=2000.0, m2=MPI ** 2, coll=True, decay=True, ctau=780.45)),
    ("shms_pion_coll_decay", 5, W.SHMS_C1, dict(p_spec=2000.0, m2=MPI ** 2, coll=True, decay=True, ctau=780.45)),
]
N = 1500
STRIDE = 2048


def tape_for(i):
    return W.uniform_tape(N, STRIDE, 900 + i)


def main():
    for i, (name, arm, wl, kw) in enumerate(CASES):
        ev = W.generate(wl, N, seed=wl.seed + 17 * i)
        tape = tape_for(i)
        args = dict(ms=True, wcs=True)
        args.update(kw)
        out = O.ref_run(arm, ev, tape=tape, stride=STRIDE, **args)
        assert out["rc"] == 0, (name, out["rc"])
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"),
            arm=arm, tape_seed=900 + i, stride=STRIDE,
            tape_sha256=hashlib.sha256(tape.tobytes()).hexdigest(),
            settings=np.array([args.get("p_spec"), args.get("m2"), float(args.get("ms", True)),
                               float(args.get("wcs", True)), float(args.get("decay", False)),
                               float(args.get("coll", False)), args.get("ctau", 0.0)]),
            **{"in_" + k: ev[k] for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")},
            **{"out_" + k: out[k] for k in ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp", "y_fp",
                                             "dy_fp", "pathlen", "resmult", "m2", "ok", "used")},
            counters=out["counters"])
        print(name, "accepted", int(out["ok"].sum()), "of", N)


if __name__ == "__main__":
    main()
