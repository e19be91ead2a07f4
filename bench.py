"""Throughput of the spectrometer transport + reconstruction path on one GPU.

Workload (one step): the SHMS single-arm configuration -- 1e7 electrons at a
6.0 GeV/c central momentum, delta ~ U(-15, 25) %, dx/dz ~ U(-0.06, 0.06),
dy/dz ~ U(-0.04, 0.04), z ~ U(-5, 5) cm, multiple scattering on, wire-chamber
smearing sigma = 0.030 cm, track fit on (FP64), through mc_shms + hut +
reconstruction.  Inputs are synthetic (numpy, fixed seed) and resident in HBM
before timing; each step draws fresh per-event Philox streams.

Prints one JSON line (the driver's contract) including the roofline of the
transport kernel and the CPU baseline (the FP64 oracle on the host cores over
a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import math
import json
import os
import platform
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


# Host cores a one-GPU job may use on the GPU hosts this runs on (their CPU
# share per GPU); the machine's nproc is reported beside it.
CPU_SHARE = 16


def host_threads():
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return min(avail, CPU_SHARE), avail


def reference_calibration():
    """Ratio reference rate / oracle rate measured where the reference tree is
    (oracle/calibrate.py), or None."""
    path = os.path.join(REPO, "profiles", "cpu_calibration_reference_vs_oracle.json")
    if not os.path.exists(path):
        return None
    return json.load(open(path))


def cpu_baseline(wl, seconds=12.0):
    """The FP64 oracle over a bounded sample of the workload on the host
    cores (all the job may use), and on one thread (the reference's execution
    model: simc.f:169-351 loops over events one at a time)."""
    from oracle import oracle as O
    from simc_gfortran_amd import workloads as W
    threads, avail = host_threads()
    chunk = 200_000
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True, wire_sigma=wl.wire_sigma,
              fit_f32=False)
    done, t_used, k = 0, 0.0, 0
    while t_used < seconds:
        ev = W.generate(wl, chunk, seed=wl.seed + 1000 + k)
        t0 = time.perf_counter()
        O.run(wl.arm, ev, rng="philox", seed=99, event0=done, nthreads=threads, **kw)
        t_used += time.perf_counter() - t0
        done += chunk
        k += 1
    ev = W.generate(wl, 40_000, seed=wl.seed + 999)
    t0 = time.perf_counter()
    O.run(wl.arm, ev, rng="philox", seed=99, nthreads=1, **kw)
    t1 = time.perf_counter() - t0
    out = {"value": done / t_used, "unit": "events/s", "cores": threads, "kind": "port",
           "sample": f"{done} events of the same workload ({k} x {chunk}), FP64 oracle "
                     f"(oracle/simc_oracle.c), OpenMP over {threads} threads",
           "cpu_model": cpu_model(), "nproc": os.cpu_count(), "affinity_cpus": avail,
           "cpu_share_per_gpu": CPU_SHARE,
           "single_thread_events_per_s": 40_000 / t1}
    cal = reference_calibration()
    if cal:
        r = cal["reference_over_oracle"]
        out["reference_over_oracle"] = r
        out["reference_estimate_events_per_s"] = out["value"] * r
        out["reference_estimate_single_thread_events_per_s"] = out["single_thread_events_per_s"] * r
        out["calibration"] = ("profiles/cpu_calibration_reference_vs_oracle.json: the flang-built reference "
                              "timed against the oracle on identical events, one thread, " + cal["cpu"])
    return out


def cpu_c0(seconds=8.0):
    """The HMS configuration of 1e5 events on the CPU: the oracle's target
    step (trip_thru_target, target_musc, transform) and HMS transport +
    reconstruction, on one thread and on the host cores."""
    from oracle import oracle as O
    from simc_gfortran_amd import workloads as W
    wl = W.HMS_C0
    n = 100_000
    v = W.generate_vertex(wl, n, seed=wl.seed)
    T = O.make_target(**W.LH2_CELL)
    P = O.make_prep(3, spectrometer=1, p_spec=wl.p_spec, th_spec=math.radians(wl.theta_deg),
                    phi=1.5 * math.pi, mass=wl.mass)
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True, fit_f32=False)
    threads, _ = host_threads()
    rates = {}
    for nt in (1, threads):
        m, t, k = 0, 0.0, 0
        while t < seconds / 2 or k == 0:
            sub = n if nt > 1 else 20_000
            vv = {key: a[:sub] for key, a in v.items()}
            t0 = time.perf_counter()
            prep, rc = O.target_prep(T, P, vv, rng="philox", seed=1001, event0=m, nthreads=nt)
            assert rc == 0
            O.run(wl.arm, prep, rng="philox", seed=1001, event0=m, nthreads=nt, **kw)
            t += time.perf_counter() - t0
            m += sub
            k += 1
        rates[nt] = m / t
    return {"workload": "hms_proton_3GeV_lh2_cell_target_step_plus_arm", "events": n,
            "cpu_events_per_s_1_thread": rates[1], "cpu_events_per_s": rates[threads], "cores": threads}


STAGE_NAMES = ["stage1_target_to_collimator", "stage2_magnets", "stage3_hut_fit", "recon_accepted"]


def pmc_summary(fetch_csv, write_csv, events):
    """HBM bytes per launch of each kernel stage from FETCH_SIZE / WRITE_SIZE
    (KiB per dispatch), each scaled by what the same counter reads for the
    calibration kernel (an elementwise multiply reading and writing M MiB)."""
    import csv

    def per_kernel(path, counter):
        acc = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            name = r["Kernel_Name"]
            key = None
            if "stage_kernel" in name:
                key = STAGE_NAMES[0] if "false, 1, false, true" in name else \
                    STAGE_NAMES[1] if "false, 1, true, true" in name else \
                    STAGE_NAMES[2] if "false, 2, true, true" in name else None
            elif "recon_list_kernel" in name:
                key = STAGE_NAMES[3]
            elif "at::native" in name and "Mul" in name:
                key = "copy"
            if key is None:
                continue
            acc.setdefault(key, {}).setdefault(r["Dispatch_Id"], 0.0)
            acc[key][r["Dispatch_Id"]] += float(r["Counter_Value"])
        return acc

    fetch = per_kernel(fetch_csv, "FETCH_SIZE")
    write = per_kernel(write_csv, "WRITE_SIZE")
    copy_f = max(fetch["copy"].values())  # KiB
    copy_w = max(write["copy"].values())
    # the probe copies M MiB: infer M from the write side (nearest MiB)
    mib = round(copy_w / 1024.0)
    f_fetch = copy_f / (mib * 1024.0)
    f_write = copy_w / (mib * 1024.0)
    res = {"events_per_step": events, "calibration_copy_MiB": mib,
           "fetch_factor": f_fetch, "write_factor": f_write, "kernels": {}}
    for k in STAGE_NAMES:
        fk = float(np.mean(list(fetch[k].values()))) * 1024.0 / f_fetch
        wk = float(np.mean(list(write[k].values()))) * 1024.0 / f_write
        res["kernels"][k] = {"fetch_bytes": fk, "write_bytes": wk, "hbm_bytes_per_launch": fk + wk,
                             "dispatches": len(fetch[k])}
    out = os.path.join(REPO, "profiles", "hbm_traffic_shms_1e7.json")
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res))


def plain_rate(ctx, wl, n, steps=3):
    """Events/s of one large single-setting run of workload wl (n events)."""
    import torch
    from simc_gfortran_amd import workloads as W
    dev = torch.device("cuda:0")
    ev = W.generate_device(wl, n, seed=wl.seed, device=dev)
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True, wire_sigma=wl.wire_sigma, fit="fp64")
    out = None
    ts = []
    for s in range(steps + 1):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        out = ctx.run(wl.arm, ev, rng="philox", seed=1234, event0=s * n, out=out, **kw)
        b.record()
        torch.cuda.synchronize()
        if s:
            ts.append(a.elapsed_time(b))
    return n / (min(ts) * 1e-3)


def bench_batched(ctx, line, args):
    """The batched form: 64 mixed HMS/SHMS settings, 1.5e8 events in one run,
    launch tile from the autotuner, the launch sequence captured in a HIP
    graph and replayed.  Compared with the same settings run one plain call
    each, and with the single-setting rates of large HMS and SHMS runs
    weighted by the batch's event mix."""
    import torch
    from simc_gfortran_amd import api
    from simc_gfortran_amd import workloads as W
    dev = torch.device("cuda:0")
    sets = W.mixed_settings(64)
    gap = 0
    total = sum(w.n for w in sets)
    f64 = dict(dtype=torch.float64, device=dev)
    ev = {k: torch.empty(total, **f64) for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")}
    settings, off = [], 0
    for w in sets:
        W.generate_device(w, w.n, seed=w.seed, device=dev, out=ev, at=off)
        # each setting is one SIMC run with its own seed (seed*1000 + s) and
        # its events numbered from 0
        settings.append(dict(arm=w.arm, p_spec=w.p_spec, m2=w.m2, wire_sigma=w.wire_sigma, fit="fp64",
                             seed=w.seed, event0=0, offset=off, count=w.n))
        off += w.n + gap
    b = api.Batch(ctx, settings)
    n_passes, n_lane = b.passes()
    out = b.new_outputs(ev)
    chunk, tune_ms = b.tune(ev, out)
    # capture the launch sequence and time graph replays
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        b.run(ev, out=out)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        b.run(ev, out=out)
    ts = []
    for r in range(4):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        if r:
            ts.append(e0.elapsed_time(e1))
    batch_ms = float(np.mean(ts))
    batch_rate = total / (batch_ms * 1e-3)
    accepted = int(out["ok"].sum().item())
    del g
    # the same settings, one plain call each
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    outs = b.new_outputs(ev)
    for rep in range(2):
        t0.record()
        for d in settings:
            sl = slice(d["offset"], d["offset"] + d["count"])
            ctx.run(d["arm"], {k: v[sl] for k, v in ev.items()}, rng="philox", seed=d["seed"],
                    event0=d["event0"], out={k: v[sl] for k, v in outs.items()},
                    **{k: d[k] for k in ("p_spec", "m2", "wire_sigma", "fit")})
        t1.record()
        torch.cuda.synchronize()
    per_setting_ms = t0.elapsed_time(t1)
    b.close()
    del ev, out, outs
    torch.cuda.empty_cache()
    # single-setting rates of large runs, weighted by the batch's event mix
    n_hms = sum(w.n for w in sets if w.arm == 1)
    n_shms = total - n_hms
    r_hms = plain_rate(ctx, W.HMS_C0, 10_000_000)
    r_shms = line["value"]
    single_mix = total / (n_hms / r_hms + n_shms / r_shms)
    return {"settings": len(sets), "events": total, "hms_events": n_hms, "shms_events": n_shms,
            "passes": n_passes, "passes_keyed_per_setting": n_lane,
            "events_per_s": batch_rate, "ms": batch_ms, "hip_graph": True, "tuned_chunk": chunk,
            "tune_best_ms": tune_ms, "accepted_fraction": accepted / total,
            "per_setting_plain_events_per_s": total / (per_setting_ms * 1e-3),
            "vs_per_setting_plain": batch_rate / (total / (per_setting_ms * 1e-3)),
            "single_setting_rates": {"hms_1e7": r_hms, "shms_1e7": r_shms},
            "single_setting_mix_events_per_s": single_mix,
            "vs_single_setting": batch_rate / single_mix,
            "data": "synthetic: 64 settings alternating HMS protons / SHMS electrons, p0 U(1,9) GeV/c, "
                    "counts log-uniform in [1e4,1e7] scaled to 1.5e8 total, Philox seed 2026*1000+s per setting"}


def bench_adjoint(ctx, n=10_000_000, reps=3):
    """The adjoint form on HMS, 1e7 events of the HMS workload: one pass
    (loss and every gradient, draws replayed, mask from a plain run) against
    the plain forward run of the same events."""
    import torch
    from simc_gfortran_amd import workloads as W
    dev = torch.device("cuda:0")
    wl = W.HMS_C0
    ev = W.generate_device(wl, n, seed=wl.seed, device=dev)
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True)
    out = None
    tf = []
    for r in range(reps + 1):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        out = ctx.run(wl.arm, ev, rng="philox", seed=77, fit="fp64", out=out, **kw)
        b.record()
        torch.cuda.synchronize()
        if r:
            tf.append(a.elapsed_time(b))
    mask = out["ok"].clone()
    coef = torch.from_numpy(ctx.recon_coefficients(wl.arm)).to(dev)
    g = torch.empty_like(coef)
    o9 = torch.empty(9, dtype=torch.float64, device=dev)
    ta = []
    for r in range(reps + 1):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        ctx.adjoint(wl.arm, ev, mask, coef, [0.0] * 6, seed=77, grad_coef=g, out9=o9, **kw)
        b.record()
        torch.cuda.synchronize()
        if r:
            ta.append(a.elapsed_time(b))
    fwd, adj = float(np.mean(tf)), float(np.mean(ta))
    res = o9.cpu().numpy()
    return {"workload": "hms_proton_3GeV_1e7", "events": n, "events_used": int(res[8]),
            "forward_ms": fwd, "adjoint_ms": adj, "adjoint_over_forward": adj / fwd,
            "adjoint_events_per_s": n / (adj * 1e-3), "loss": float(res[0]),
            "gradients": int(coef.numel()) + 6,
            "data": "synthetic HMS workload, Philox streams replayed, mask = ok of the forward run"}


def bench_c0(ctx, reps=20):
    """The HMS configuration as specified: 1e5 protons at 3 GeV/c from the
    10 cm LH2 cell through the target step and mc_hms + hut + reconstruction
    (one launch sequence per step; at this size launch latency matters)."""
    import torch
    from simc_gfortran_amd import api, workloads as W
    dev = torch.device("cuda:0")
    wl = W.HMS_C0
    n = 100_000
    v = W.generate_vertex(wl, n, seed=wl.seed)
    dv = {k: torch.from_numpy(a).to(dev) for k, a in v.items()}
    T = api.Target(**{k: W.LH2_CELL[k] for k in ("A", "Z", "rho", "X0", "length", "angle", "zoffset",
                                                   "can")}, reserved=0)
    P = api.PrepSettings(3, 1, 1, 1, wl.p_spec, math.radians(wl.theta_deg), 1.5 * math.pi, wl.mass, 0.0, 0.0,
                         (api.C.c_double * 5)())
    prep = {k: torch.empty(n, dtype=torch.float64, device=dev) for k in ("dpp", "x", "y", "z", "dxdz", "dydz",
                                                                          "fry")}
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True)
    out = None
    ts = []
    for r in range(reps + 2):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        ctx.target_prep(T, P, dv, rng="philox", seed=1001, event0=r * n, out=prep)
        out = ctx.run(wl.arm, prep, rng="philox", seed=1001, event0=r * n, fit="fp64", out=out, **kw)
        b.record()
        torch.cuda.synchronize()
        if r >= 2:
            ts.append(a.elapsed_time(b))
    ms = float(np.mean(ts))
    return {"workload": "hms_proton_3GeV_lh2_cell_1e5 (target step + arm)", "events": n, "ms": ms,
            "events_per_s": n / (ms * 1e-3), "accepted_fraction": float(out["ok"].float().mean())}


def bench_target(ctx, n=10_000_000, reps=5):
    """The target step before the HMS arm (1e7 protons from the 10 cm LH2
    cell, the HMS workload's vertex distributions): the preparation alone
    (vertex -> arm inputs, trip_thru_target + target_musc sampled per event)
    and the chain preparation + transport + reconstruction."""
    import torch
    from simc_gfortran_amd import api, workloads as W
    dev = torch.device("cuda:0")
    wl = W.HMS_C0
    v = W.generate_vertex(wl, n, seed=wl.seed)
    dv = {k: torch.from_numpy(a).to(dev) for k, a in v.items()}
    del v
    T = api.Target(**{k: W.LH2_CELL[k] for k in ("A", "Z", "rho", "X0", "length", "angle", "zoffset",
                                                   "can")}, reserved=0)
    P = api.PrepSettings(3, 1, 1, 1, wl.p_spec, math.radians(wl.theta_deg), 1.5 * math.pi, wl.mass, 0.0, 0.0,
                         (api.C.c_double * 5)())
    prep = {k: torch.empty(n, dtype=torch.float64, device=dev) for k in ("dpp", "x", "y", "z", "dxdz", "dydz",
                                                                          "fry")}
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True)
    out = None
    tp, tc = [], []
    for r in range(reps + 1):
        a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        a.record()
        ctx.target_prep(T, P, dv, rng="philox", seed=1001, out=prep)
        b.record()
        out = ctx.run(wl.arm, prep, rng="philox", seed=1001, fit="fp64", out=out, **kw)
        c.record()
        torch.cuda.synchronize()
        if r:
            tp.append(a.elapsed_time(b))
            tc.append(a.elapsed_time(c))
    p_ms, c_ms = float(np.mean(tp)), float(np.mean(tc))
    return {"workload": "hms_proton_3GeV_lh2_cell_1e7", "events": n, "prep_ms": p_ms,
            "prep_events_per_s": n / (p_ms * 1e-3), "chain_ms": c_ms, "chain_events_per_s": n / (c_ms * 1e-3),
            "accepted_fraction": float(out["ok"].float().mean()),
            "data": "synthetic vertices in a 10 cm LH2 cell (can 3), loss and scattering sampled per event"}


def bench_gen(ctx, n=10_000_000, chunks=10, reps=5):
    """H(e,e'p) elastic events (10.6 GeV beam, 10 cm LH2 cell, electron in the
    SHMS at 12.5 deg, proton in the HMS; simc_gen_hep, radiation off): the
    generator alone (1e7 events, all 27 output columns), and the coincidence
    chain generator -> target step of each particle -> SHMS and HMS transport
    + reconstruction over chunks x n = 1e8 generated events (each chunk's
    event0 continues the streams)."""
    import torch
    from simc_gfortran_amd import api, workloads as W
    dev = torch.device("cuda:0")
    Hs = W.HEP_ELASTIC
    H = api.hep_settings(**Hs)
    T = api.Target(**W.LH2_CELL)
    Pe = api.PrepSettings(2, 5, 1, 1, Hs["e_P"], Hs["e_theta"], Hs["e_phi"], W.ME, 0.0, Hs["Ebeam"],
                          (api.C.c_double * 5)())
    Pp = api.PrepSettings(3, 1, 1, 1, Hs["p_P"], Hs["p_theta"], Hs["p_phi"], W.MP, 0.0, Hs["Ebeam"],
                          (api.C.c_double * 5)())
    cols = ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")
    pe = {k: torch.empty(n, dtype=torch.float64, device=dev) for k in cols}
    pp = {k: torch.empty(n, dtype=torch.float64, device=dev) for k in cols}
    g = None
    tg = []
    for r in range(reps + 1):
        a, b = (torch.cuda.Event(enable_timing=True) for _ in range(2))
        a.record()
        g = ctx.gen_hep(T, H, n, seed=2001, out=g)
        b.record()
        torch.cuda.synchronize()
        if r:
            tg.append(a.elapsed_time(b))
    oe = op = None
    # the electron side (target step + SHMS) and the proton side (target step
    # + HMS) are independent once the generator has run: two streams (default
    # priority), joined before the next chunk reuses the buffers
    main = torch.cuda.current_stream(dev)
    s_e, s_p = torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)

    def chunk(k):
        nonlocal g, oe, op
        e0 = k * n
        g = ctx.gen_hep(T, H, n, seed=2001, event0=e0, out=g)
        ve = {"x": g["x"], "y": g["y"], "z": g["z"], "E": g["e_E"], "p": g["e_E"], "theta": g["e_theta"],
              "xptar": g["e_xptar"], "yptar": g["e_yptar"]}
        vp = {"x": g["x"], "y": g["y"], "z": g["z"], "E": g["p_E"], "p": g["p_P"], "theta": g["p_theta"],
              "delta": g["p_delta"], "xptar": g["p_xptar"], "yptar": g["p_yptar"]}
        s_e.wait_stream(main)
        s_p.wait_stream(main)
        with torch.cuda.stream(s_e):
            ctx.target_prep(T, Pe, ve, rng="philox", seed=2001, event0=e0, out=pe)
            oe = ctx.run(5, pe, rng="philox", seed=2003, event0=e0, fit="fp64", out=oe, p_spec=Hs["e_P"],
                         m2=W.ME ** 2, wire_sigma=0.03)
        with torch.cuda.stream(s_p):
            ctx.target_prep(T, Pp, vp, rng="philox", seed=2002, event0=e0, out=pp)
            op = ctx.run(1, pp, rng="philox", seed=2004, event0=e0, fit="fp64", out=op, p_spec=Hs["p_P"],
                         m2=W.MP ** 2)
        main.wait_stream(s_e)
        main.wait_stream(s_p)
        both = (oe["ok"] == 1) & (op["ok"] == 1)
        return both.sum(), torch.where(both, g["weight"], torch.zeros_like(g["weight"])).sum()

    chunk(0)  # warmup
    torch.cuda.synchronize()
    a, b = (torch.cuda.Event(enable_timing=True) for _ in range(2))
    a.record()
    sums = [chunk(k) for k in range(chunks)]
    b.record()
    torch.cuda.synchronize()
    c_ms = a.elapsed_time(b)
    g_ms = float(np.mean(tg))
    ncoin = int(sum(int(c) for c, _ in sums))
    wsum = float(sum(float(w) for _, w in sums))
    bytes_out = n * (27 * 8 + 1 + 4)
    total = n * chunks
    return {"workload": "hep_elastic_10p6GeV_shms12p5_hms_lh2", "gen_events": n, "gen_ms": g_ms,
            "gen_events_per_s": n / (g_ms * 1e-3), "gen_write_GBps": bytes_out / (g_ms * 1e-3) / 1e9,
            "chain_events": total, "chain_ms": c_ms, "chain_events_per_s": total / (c_ms * 1e-3),
            "coincidences": ncoin, "coincidence_fraction": ncoin / total, "weighted_coincidences": wsum,
            "data": "synthetic: generated per event (Philox), radiation off; beam spread uniform with the "
                    "RMS of a 5e-4 Gaussian"}


def _rad_vertices(g, particle, ok):
    """The target step's inputs of one scattered particle with radiation: ORIG
    energy and momentum (photons removed), the vertex energy for its losses,
    the vertex angles; events the generator rejected get their vertex values
    (they never count)."""
    import torch

    def pick(orig, vert):
        return torch.where(ok, g[orig], g[vert])
    if particle == "e":
        e = pick("orig_e_E", "e_E")
        return {"x": g["x"], "y": g["y"], "z": g["z"], "E": e, "p": e, "E_vertex": g["e_E"],
                "theta": g["e_theta"], "xptar": g["e_xptar"], "yptar": g["e_yptar"]}
    return {"x": g["x"], "y": g["y"], "z": g["z"], "E": pick("orig_p_E", "p_E"), "p": pick("orig_p_P", "p_P"),
            "E_vertex": g["p_E"], "theta": g["p_theta"], "delta": pick("orig_p_delta", "p_delta"),
            "xptar": g["p_xptar"], "yptar": g["p_yptar"]}


def bench_c2(ctx, n=10_000_000, chunks=10, reps=3):
    """The coincidence configuration as specified, radiative tails included:
    H(e,e'p) elastic at 10.6 GeV (10 cm LH2, electron in the SHMS at 12.5 deg,
    proton in the HMS), simc_gen_hep_rad with the reference's H(e,e'p)
    radiative options (rad_flag 0, extrad_flag 2, off-shell bremos) -> target
    step of the electron and the proton on their ORIG kinematics -> SHMS and
    HMS transport + reconstruction, over chunks x n = 1e8 generated events."""
    import torch
    from simc_gfortran_amd import api, workloads as W
    dev = torch.device("cuda:0")
    Hs = W.HEP_ELASTIC
    H = api.hep_settings(**Hs)
    Rs = api.rad_settings(**W.HEP_RAD)
    T = api.Target(**W.LH2_CELL)
    Pe = api.PrepSettings(2, 5, 1, 1, Hs["e_P"], Hs["e_theta"], Hs["e_phi"], W.ME, 0.0, Hs["Ebeam"],
                          (api.C.c_double * 5)())
    Pp = api.PrepSettings(3, 1, 1, 1, Hs["p_P"], Hs["p_theta"], Hs["p_phi"], W.MP, 0.0, Hs["Ebeam"],
                          (api.C.c_double * 5)())
    cols = ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")
    pe = {k: torch.empty(n, dtype=torch.float64, device=dev) for k in cols}
    pp = {k: torch.empty(n, dtype=torch.float64, device=dev) for k in cols}
    g = None
    tg = []
    for r in range(reps + 1):
        a, b = (torch.cuda.Event(enable_timing=True) for _ in range(2))
        a.record()
        g = ctx.gen_hep_rad(T, H, Rs, n, seed=2001, out=g)
        b.record()
        torch.cuda.synchronize()
        if r:
            tg.append(a.elapsed_time(b))
    gen_ok = float(g["ok"].float().mean())
    oe = op = None
    main = torch.cuda.current_stream(dev)
    s_e, s_p = torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)

    def chunk(k):
        nonlocal g, oe, op
        e0 = k * n
        g = ctx.gen_hep_rad(T, H, Rs, n, seed=2001, event0=e0, out=g)
        ok = g["ok"] == 1
        ve, vp = _rad_vertices(g, "e", ok), _rad_vertices(g, "p", ok)
        s_e.wait_stream(main)
        s_p.wait_stream(main)
        with torch.cuda.stream(s_e):
            ctx.target_prep(T, Pe, ve, rng="philox", seed=2001, event0=e0, out=pe)
            oe = ctx.run(5, pe, rng="philox", seed=2003, event0=e0, fit="fp64", out=oe, p_spec=Hs["e_P"],
                         m2=W.ME ** 2, wire_sigma=0.03)
        with torch.cuda.stream(s_p):
            ctx.target_prep(T, Pp, vp, rng="philox", seed=2002, event0=e0, out=pp)
            op = ctx.run(1, pp, rng="philox", seed=2004, event0=e0, fit="fp64", out=op, p_spec=Hs["p_P"],
                         m2=W.MP ** 2)
        main.wait_stream(s_e)
        main.wait_stream(s_p)
        both = ok & (oe["ok"] == 1) & (op["ok"] == 1)
        return both.sum(), torch.where(both, g["weight"], torch.zeros_like(g["weight"])).sum()

    chunk(0)  # warmup
    torch.cuda.synchronize()
    a, b = (torch.cuda.Event(enable_timing=True) for _ in range(2))
    a.record()
    sums = [chunk(k) for k in range(chunks)]
    b.record()
    torch.cuda.synchronize()
    c_ms = a.elapsed_time(b)
    g_ms = float(np.mean(tg))
    ncoin = int(sum(int(c) for c, _ in sums))
    wsum = float(sum(float(w) for _, w in sums))
    total = n * chunks
    return {"workload": "hep_elastic_rad_10p6GeV_shms12p5_hms_lh2", "gen_events": n, "gen_ms": g_ms,
            "gen_events_per_s": n / (g_ms * 1e-3), "gen_ok_fraction": gen_ok,
            "chain_events": total, "chain_ms": c_ms, "chain_events_per_s": total / (c_ms * 1e-3),
            "coincidences": ncoin, "coincidence_fraction": ncoin / total, "weighted_coincidences": wsum,
            "data": "synthetic: generated per event (Philox), radiative tails (rad_flag 0, extrad_flag 2, "
                    "bremos), beam spread uniform with the RMS of a 5e-4 Gaussian, Em edge from a 0-200 MeV cut"}


def cpu_c2(seconds=10.0):
    """The radiated coincidence chain on the CPU: the oracle's generator with
    tails, target step of both particles and both arms, one thread and the
    host cores."""
    from oracle import oracle as O
    from simc_gfortran_amd import workloads as W
    Hs = W.HEP_ELASTIC
    T, H = O.make_target(**W.LH2_CELL), O.make_hep(**Hs)
    R = O.make_rad(**dict(W.HEP_RAD, etatzai=O.rad_etatzai(W.LH2_CELL["Z"]), targ_A=W.LH2_CELL["A"]))
    Pe = O.make_prep(2, 5, True, True, p_spec=Hs["e_P"], th_spec=Hs["e_theta"], phi=Hs["e_phi"], mass=W.ME,
                     beam_energy=Hs["Ebeam"])
    Pp = O.make_prep(3, 1, True, True, p_spec=Hs["p_P"], th_spec=Hs["p_theta"], phi=Hs["p_phi"], mass=W.MP,
                     beam_energy=Hs["Ebeam"])
    threads, _ = host_threads()

    def vert(g, particle):
        ok = g["ok"] == 1
        if particle == "e":
            e = np.where(ok, g["orig_e_E"], g["e_E"])
            return {"x": g["x"], "y": g["y"], "z": g["z"], "E": e, "p": e, "E_vertex": g["e_E"], "theta": g["e_theta"],
                    "xptar": g["e_xptar"], "yptar": g["e_yptar"]}
        return {"x": g["x"], "y": g["y"], "z": g["z"], "E": np.where(ok, g["orig_p_E"], g["p_E"]),
                "p": np.where(ok, g["orig_p_P"], g["p_P"]), "E_vertex": g["p_E"], "theta": g["p_theta"],
                "delta": np.where(ok, g["orig_p_delta"], g["p_delta"]), "xptar": g["p_xptar"], "yptar": g["p_yptar"]}

    rates = {}
    for nt in (1, threads):
        m, t = 0, 0.0
        sub = 20_000 if nt == 1 else 200_000
        while t < seconds / 2 or m == 0:
            t0 = time.perf_counter()
            g, rc = O.gen_hep_rad(T, H, R, sub, rng="philox", seed=2001, event0=m, nthreads=nt)
            for part, P, arm, p0, m2, wire, sd in (("e", Pe, 5, Hs["e_P"], W.ME ** 2, 0.03, 2003),
                                                   ("p", Pp, 1, Hs["p_P"], W.MP ** 2, 0.0, 2004)):
                ev, rc2 = O.target_prep(T, P, vert(g, part), rng="philox", seed=sd - 2, event0=m, nthreads=nt)
                O.run(arm, ev, p_spec=p0, m2=m2, wire_sigma=wire, fit_f32=False, rng="philox", seed=sd, event0=m,
                      nthreads=nt)
            t += time.perf_counter() - t0
            m += sub
        rates[nt] = m / t
    return {"cpu_chain_events_per_s_1_thread": rates[1], "cpu_chain_events_per_s": rates[threads],
            "cpu_cores": threads}


def bench_host_entry(ctx, wl, ev, reps=3):
    """The headline workload through the host-array entry point
    (simc_arm_run_host): inputs copied from pageable host memory, outputs
    copied back, synchronous.  The PCIe-inclusive rate; never the headline
    value, which is measured with inputs resident in HBM."""
    ts = []
    out = None
    for r in range(reps + 1):
        t0 = time.perf_counter()
        out = ctx.run_host(wl.arm, ev, p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True, wire_sigma=wl.wire_sigma,
                           fit="fp64", seed=wl.seed, out=out)
        if r:
            ts.append(time.perf_counter() - t0)
    n = len(ev["dpp"])
    ms = float(np.mean(ts)) * 1e3
    return {"workload": wl.name, "events": n, "ms": ms, "events_per_s": n / (ms * 1e-3),
            "accepted_fraction": float(np.mean(out["ok"])),
            "data": "host numpy arrays in and out (pageable memory, allocated once), the library's chunked "
                    "pinned staging and transfers included"}


def bench_dropin(wl, calls=2000):
    """The per-event drop-in (libsimc_dropin.so's mc_shms_, one event per call
    as montecarlo() calls it, simc.f:1464-1488): microseconds per call, after
    a warm-up that creates its context."""
    import ctypes as C
    from simc_gfortran_amd import workloads as W
    os.environ.setdefault("SIMC_MAP_DIR", os.path.join(REPO, "data", "maps"))
    L = C.CDLL(os.path.join(REPO, "simc_gfortran_amd", "libsimc_dropin.so"))
    d = lambda v=0.0: C.c_double(v)
    i = lambda v=0: C.c_int32(v)
    ev = W.generate(wl, calls + 100, seed=5)
    args_f = [d() for _ in range(13)]
    p_spec, th, m2 = d(wl.p_spec), d(0.0), d(wl.m2)
    ms, wcs, dec, ok, spectr, coll = i(1), i(1), i(0), i(0), i(5), i(0)
    resmult, fry, pathlen = d(), d(), d()
    dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp = [d() for _ in range(10)]
    ts = []
    acc = 0
    for k in range(calls + 100):
        dpp.value, x.value, y.value = ev["dpp"][k], ev["x"][k], ev["y"][k]
        z.value, dxdz.value, dydz.value = ev["z"][k], ev["dxdz"][k], ev["dydz"][k]
        fry.value, pathlen.value, m2.value = ev["fry"][k], 0.0, wl.m2
        t0 = time.perf_counter()
        L.mc_shms_(C.byref(p_spec), C.byref(th), C.byref(dpp), C.byref(x), C.byref(y), C.byref(z), C.byref(dxdz),
                   C.byref(dydz), C.byref(x_fp), C.byref(dx_fp), C.byref(y_fp), C.byref(dy_fp), C.byref(m2),
                   C.byref(ms), C.byref(wcs), C.byref(dec), C.byref(resmult), C.byref(fry), C.byref(ok),
                   C.byref(pathlen), C.byref(spectr), C.byref(coll))
        if k >= 100:
            ts.append(time.perf_counter() - t0)
            acc += ok.value
    us = float(np.median(ts)) * 1e6
    return {"workload": wl.name, "calls": calls, "us_per_call_median": us,
            "us_per_call_mean": float(np.mean(ts)) * 1e6, "accepted_fraction": acc / calls,
            "data": "one event per mc_shms_ call through ctypes (the Fortran ABI), synchronous"}


def bench_arms(ctx, n=10_000_000, reps=5):
    """The other spectrometers at 1e7 events each (Philox streams, FP64 fit):
    HMS protons (3 GeV/c), SOS protons (1.5 GeV/c), HRS-R protons (2 GeV/c),
    HRS-L electrons (3 GeV/c); per-stage times from HIP events."""
    import torch
    from simc_gfortran_amd import workloads as W
    dev = torch.device("cuda:0")
    res = {}
    for wl in (W.HMS_C0, W.SOS_1p5, W.HRSR_2, W.HRSL_3):
        ev = W.generate_device(wl, n, seed=wl.seed, device=dev)
        out = None
        st = []
        for r in range(reps + 1):
            e = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
            for x in e:
                x.record()
            out = ctx.run(wl.arm, ev, p_spec=wl.p_spec, m2=wl.m2, rng="philox", seed=11 + r, fit="fp64", out=out,
                          stage_events=e)
            torch.cuda.synchronize()
            if r:
                st.append([e[k].elapsed_time(e[k + 1]) for k in range(4)])
        ms = np.mean(np.array(st), axis=0)
        res[wl.name] = {"arm": wl.arm, "events": n, "ms": float(ms.sum()), "events_per_s": n / (ms.sum() * 1e-3),
                        "stage_ms": [float(x) for x in ms], "accepted_fraction": float(out["ok"].float().mean())}
        del ev, out
    return res


# FP64 operations of one RK4 step of the target-field tracking (trgRK4,
# trg_track.f:461-498): four field evaluations (rotation 6, radius 4, grid
# indices and fractions 8, two bilinear interpolations 18, B_r / r and signs 4,
# back-rotation 6: 46, counting sqrt and divisions as one) each followed by the
# cross product and scaling (9), plus the stage combinations (72).
RK4_FLOP = 4 * (46 + 9) + 72


def bench_field(ctx, n=10_000_000, reps=5):
    """Target-field tracking (the polarized-target map, field at 10 degrees to
    the hadron arm): track_from_tgt for 1e7 HMS-workload protons from a 3 cm
    cell to z = 100 cm, and track_to_tgt (reconstruction through the field,
    up to 10 iterations of three plane searches and the HMS reconstruction)
    for the accepted events of an HMS run."""
    import torch
    from simc_gfortran_amd import api, workloads as W
    dev = torch.device("cuda:0")
    wl = W.HMS_C0
    g = torch.Generator(device=dev)
    g.manual_seed(77)
    u = lambda lo, hi: torch.rand(n, dtype=torch.float64, device=dev, generator=g) * (hi - lo) + lo  # noqa: E731
    ev0 = {"x": u(-0.5, 0.5), "y": u(-0.5, 0.5), "z": u(-1.5, 1.5), "dx": u(*wl.dxdz), "dy": u(*wl.dydz)}
    mom = wl.p_spec * (1 + u(*wl.dpp) / 100)
    F = api.TargetField(api.FIELD_MAP, 10.0, 10.0)
    ev = {k: torch.empty_like(v) for k, v in ev0.items()}
    ok = torch.empty(n, dtype=torch.uint8, device=dev)
    st = torch.cuda.current_stream(dev)
    tf = []
    for r in range(reps + 1):
        for k in ev:
            ev[k].copy_(ev0[k])
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        F.track_from_tgt(ev, mom, wl.mass, 1, ok=ok)
        b.record(st)
        torch.cuda.synchronize()
        if r:
            tf.append(a.elapsed_time(b))
    f_ms = float(np.mean(tf))
    # RK4 steps: the first search crosses |z0| (plus its direction step), the
    # second the 100 cm to the field-free plane, 1 cm of path per step
    slope = torch.sqrt(1 + ev0["dx"] ** 2 + ev0["dy"] ** 2)
    steps = float((((ev0["z"].abs() + 100.0) * slope).ceil() + 4).mean())
    flop = n * steps * RK4_FLOP
    # reconstruction through the field for the accepted events of an HMS run
    run = ctx.run(wl.arm, W.generate_device(wl, 2_000_000, seed=wl.seed, device=dev), p_spec=wl.p_spec,
                  m2=wl.m2, rng="philox", seed=3, fit="fp64")
    sel = run["ok"] == 1
    m = int(sel.sum())
    trk = {k: run[k][sel].contiguous() for k in ("x_fp", "dx_fp", "y_fp", "dy_fp")}
    first = {"delta": run["dpp"][sel], "y": run["y"][sel], "dx": run["dxdz"][sel], "dy": run["dydz"][sel]}
    fr = torch.zeros(m, dtype=torch.float64, device=dev)
    pm = wl.p_spec * (1 + first["delta"] / 100)
    th = math.radians(wl.theta_deg)
    tt = []
    for r in range(reps + 1):
        rec = {k: v.clone() for k, v in first.items()}
        ok2 = torch.ones(m, dtype=torch.uint8, device=dev)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        F.track_to_tgt(ctx, wl.arm, trk, rec, fr, fr, pm, wl.mass, math.cos(th), math.sin(th), 1, ok=ok2)
        b.record(st)
        torch.cuda.synchronize()
        if r:
            tt.append(a.elapsed_time(b))
    t_ms = float(np.mean(tt))
    F.close()
    return {"workload": "hms_proton_3GeV_polarized_target_field_1e7", "events": n, "from_tgt_ms": f_ms,
            "from_tgt_events_per_s": n / (f_ms * 1e-3), "rk4_steps_per_event": steps,
            "rk4_flop_per_step": RK4_FLOP, "from_tgt_TFLOPs": flop / (f_ms * 1e-3) / 1e12,
            "from_tgt_ok_fraction": float(ok.float().mean()),
            "to_tgt_events": m, "to_tgt_ms": t_ms, "to_tgt_events_per_s": m / (t_ms * 1e-3),
            "to_tgt_ok_fraction": float(ok2.float().mean()),
            "data": "synthetic: vertices uniform in a 3 cm cell, HMS angles and momenta; the reference's "
                    "trg_field_map.dat at 10 degrees; track_to_tgt on the accepted events of 2e6 HMS events"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--events", type=int, default=10_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the batched multi-setting and adjoint legs")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run one elementwise kernel reading and writing this many MiB "
                         "(HBM counter calibration)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of "
                         "`bench.py --copy-probe-mib M` into profiles/hbm_traffic_shms_1e7.json")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary, args.events)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this library has no multi-GPU path "
              "(events are independent and shard by event range)", file=sys.stderr)
        sys.exit(2)

    import torch
    from simc_gfortran_amd import api
    from simc_gfortran_amd import workloads as W

    wl = W.SHMS_C1
    n = args.events
    ev = W.generate(wl, n)
    dev = torch.device("cuda:0")
    d_ev = {k: torch.from_numpy(v).to(dev) for k, v in ev.items() if k != "z"}
    ctx = api.Spectrometers()
    out = {k: torch.empty(n, dtype=torch.float64, device=dev)
           for k in ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen")}
    out["ok"] = torch.empty(n, dtype=torch.uint8, device=dev)
    out["stop"] = torch.empty(n, dtype=torch.uint8, device=dev)
    out["resmult"] = out["m2"] = out["used"] = None
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True, wire_sigma=wl.wire_sigma,
              fit="fp64", out=out)
    stream = torch.cuda.current_stream(dev)
    if args.copy_probe_mib:
        # one elementwise kernel that reads M MiB and writes M MiB
        m = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(m, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    def step(s):
        out["pathlen"].zero_()
        ctx.run(wl.arm, d_ev, rng="philox", seed=1234, event0=s * n, **kw)

    for s in range(args.warmup):
        step(s)
    torch.cuda.synchronize()
    ev_start = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ev_stop = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    # stage boundaries of every timed step (recorded once here so that the
    # underlying HIP events exist before their handles are passed down)
    ev_stage = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(args.steps)]
    for evs in ev_stage:
        for e in evs:
            e.record(stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        out["pathlen"].zero_()
        ev_start[s].record(stream)
        ctx.run(wl.arm, d_ev, rng="philox", seed=1234, event0=(args.warmup + s) * n,
                stage_events=ev_stage[s], **kw)
        ev_stop[s].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = float(np.mean([a.elapsed_time(b) for a, b in zip(ev_start, ev_stop)]))
    stage_ms = [float(np.mean([evs[k].elapsed_time(evs[k + 1]) for evs in ev_stage])) for k in range(4)]

    # algorithmic work of the last launch per kernel stage, from its stop codes
    stop = out["stop"].to(torch.int64)
    hist = torch.bincount(stop, minlength=256).cpu().numpy()
    # (the library's stage 3 count is the reconstruction of the accepted
    # events, done by its own kernel; the hut holds no polynomial: its count
    # is the FP64 arithmetic of the reference's hut routines)
    fl = {k: float((hist * np.array(ctx.stop_flops(wl.arm, k), dtype=np.float64)).sum()) for k in (1, 2, 3)}
    hut = float((hist * np.array(ctx.hut_flops(wl.arm, ms=True, wcs=True), dtype=np.float64)).sum())
    stage_flops = [fl[1], fl[2], hut, fl[3]]
    names = STAGE_NAMES
    stages = {names[k]: {"ms": stage_ms[k], "algorithmic_flop": stage_flops[k],
                         "TFLOPs": stage_flops[k] / (stage_ms[k] * 1e-3) / 1e12} for k in range(4)}
    dom = int(np.argmax(stage_ms))
    accepted = int(hist[0])
    tflops = stage_flops[dom] / (stage_ms[dom] * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(REPO, "profiles", "hbm_traffic_shms_1e7.json")
    if os.path.exists(tpath):
        tr = json.load(open(tpath))
        if tr.get("events_per_step") == n:
            traffic = tr["kernels"].get(names[dom], {}).get("hbm_bytes_per_launch")
    value = args.steps * n / wall
    line = {
        "metric": "events/s transported+reconstructed (FP64, 1 GPU); batched & adjoint vs plain",
        "value": value,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "shms_single_arm_electron_6GeV_1e7", "events_per_step": n,
                   "p_spec_MeV": wl.p_spec, "theta_deg": wl.theta_deg,
                   "wire_sigma_cm": wl.wire_sigma, "ms": True, "wcs": True, "fit": "fp64",
                   "rng": "philox4x32-10 per event"},
        "roofline": {"bound": "fp64_valu", "achieved": tflops, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": tflops / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": names[dom], "kernel_ms": stage_ms[dom],
                     "algorithmic_flop_per_launch": stage_flops[dom],
                     "counted": "COSY polynomials (stages 1, 2 and the reconstruction); the hut's "
                                "steps: the FP64 arithmetic of the reference's hut routines",
                     "step_frac": sum(stage_flops) / (step_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS},
        "stages": stages,
        "step_ms_hip_events": step_ms,
        "accepted_fraction": accepted / n,
    }
    if not args.no_forms:
        del d_ev, out
        torch.cuda.empty_cache()
        line["c0"] = bench_c0(ctx)
        line["batched"] = bench_batched(ctx, line, args)
        line["adjoint"] = bench_adjoint(ctx)
        line["target_step"] = bench_target(ctx)
        line["target_field"] = bench_field(ctx)
        line["other_arms"] = bench_arms(ctx)
        line["generator"] = bench_gen(ctx)
        line["c2_radiative_chain"] = bench_c2(ctx)
        line["host_entry"] = bench_host_entry(ctx, wl, ev)
        line["dropin_per_event"] = bench_dropin(wl)
    if not args.no_cpu:
        cb = cpu_baseline(wl)
        line["cpu_baseline"] = cb
        line["gpu_over_cpu"] = value / cb["value"]
        if "reference_estimate_events_per_s" in cb:
            line["gpu_over_reference_estimate"] = value / cb["reference_estimate_events_per_s"]
        if "c2_radiative_chain" in line:
            c2 = cpu_c2()
            line["c2_radiative_chain"].update(c2)
            r = line["c2_radiative_chain"]
            r["gpu_over_cpu"] = r["chain_events_per_s"] / c2["cpu_chain_events_per_s"]
            r["gpu_over_cpu_1_thread"] = r["chain_events_per_s"] / c2["cpu_chain_events_per_s_1_thread"]
        if "c0" in line:
            c0 = cpu_c0()
            line["c0"].update(c0)
            line["c0"]["gpu_over_cpu"] = line["c0"]["events_per_s"] / c0["cpu_events_per_s"]
            line["c0"]["gpu_over_cpu_1_thread"] = line["c0"]["events_per_s"] / c0["cpu_events_per_s_1_thread"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
