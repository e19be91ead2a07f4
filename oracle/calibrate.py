"""Time the reference's own Fortran hot path (oracle/_ref, built by flang from
the reference's sources) against the oracle restatement on the same events
and tapes, single-threaded, in the container that has the reference tree.
The ratio converts the oracle's CPU rate (what bench.py can time on any
host) into an estimate of the reference's rate on that host.

    python oracle/calibrate.py   ->  profiles/cpu_calibration_reference_vs_oracle.json

TEST/MEASUREMENT INFRASTRUCTURE ONLY.
"""
import json
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path = [REPO] + [p for p in sys.path if os.path.abspath(p or ".") != HERE]

from oracle import oracle as O  # noqa: E402
from simc_gfortran_amd import workloads as W  # noqa: E402


def cpu_model():
    for line in open("/proc/cpuinfo"):
        if line.startswith("model name"):
            return line.split(":", 1)[1].strip()
    return platform.processor()


def best_of(fn, reps=3):
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        t.append(time.perf_counter() - t0)
    return min(t)


def main():
    if O.ref_lib() is None:
        sys.exit("oracle/_ref is not built (make -C oracle ref)")
    rows = []
    for wl, n in ((W.SHMS_C1, 20000), (W.HMS_C0, 20000)):
        stride = 2048
        ev = W.generate(wl, n, seed=31)
        tape = W.uniform_tape(n, stride, 32)
        kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True)
        ref = O.ref_run(wl.arm, ev, tape=tape, stride=stride, **kw)
        orc = O.run(wl.arm, ev, rng="tape", tape=tape, stride=stride, fit_f32=True, nthreads=1,
                    wire_sigma=0.0, **kw)
        same = all((ref[k] == orc[k]).all() for k in ("x_fp", "dx_fp", "y_fp", "dy_fp", "dpp", "y", "ok"))
        t_ref = best_of(lambda: O.ref_run(wl.arm, ev, tape=tape, stride=stride, **kw))
        t_orc = best_of(lambda: O.run(wl.arm, ev, rng="tape", tape=tape, stride=stride, fit_f32=True,
                                      nthreads=1, wire_sigma=0.0, **kw))
        rows.append({"workload": wl.name, "events": n, "outputs_identical": bool(same),
                     "reference_events_per_s": n / t_ref, "oracle_events_per_s": n / t_orc,
                     "reference_over_oracle": t_orc / t_ref})
        print(rows[-1])
    out = {"what": "single-thread events/s of the flang-built reference (mc_shms/mc_hms + huts + recon, "
                   "tape uniforms) and of the oracle on the same events and tapes",
           "cpu": cpu_model(), "nproc": os.cpu_count(),
           "compiler": "ROCm flang (reference, -O, the reference's dialect flags); gcc -O2 (oracle)",
           "rows": rows,
           "reference_over_oracle": sum(r["reference_over_oracle"] for r in rows) / len(rows)}
    path = os.path.join(REPO, "profiles", "cpu_calibration_reference_vs_oracle.json")
    json.dump(out, open(path, "w"), indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
