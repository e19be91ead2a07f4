"""The adjoint form (optics-fit gradients) on the GPU.

* At the nominal parameters its loss is the loss of the plain run's own
  outputs over the same events (same draws, same mask), and in tape mode the
  loss of the FP64 oracle's outputs.
* dL/dC and the focal-plane-offset gradients equal the analytic expressions
  evaluated in numpy from the plain run's focal-plane tracks (the loss is
  quadratic in C: dL/dC[t,k] = 1/N sum_e w_k(e) m_t(h_e)).
* All six offset gradients and a sample of dL/dC agree with central finite
  differences (h = 1e-6) of the adjoint's own loss (same draws, fixed mask).
* The torch.autograd binding returns the same gradients, and descending them
  lowers the loss.
* The s_p and dtheta gradients agree with Richardson-extrapolated central
  differences of the ORACLE's loss (its own transport, with the maps seeing
  delta' and y' - dtheta and the scattering the true momentum), on events that
  all the perturbed oracle runs accept; the loss itself agrees to 1e-12.
* At the adjoint configuration's 1e7 events: the six offset gradients and the
  largest coefficient gradients against central differences of the adjoint's
  own loss.
"""
import os
import sys

import numpy as np
import pytest

from simc_gfortran_amd import workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N = 20000
SEED = 424242


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from simc_gfortran_amd import api
    c = api.Spectrometers()
    yield c
    c.close()


def recon_exponents(arm):
    sys.path.insert(0, os.path.join(REPO, "simc_gfortran_amd", "csrc"))
    import gen_cosy
    f = os.path.join(REPO, "data", "maps", "hms/recon_cosy.dat" if arm == 1 else "shms/shms_recon.dat")
    return np.array([e for c, e in gen_cosy.parse_recon(f)], dtype=np.int64)


def setup(ctx, wl, n=N):
    ev = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in W.generate(wl, n, seed=8).items()}
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True, wire_sigma=wl.wire_sigma)
    out = ctx.run(wl.arm, ev, rng="philox", seed=SEED, fit="fp64", **kw)
    torch.cuda.synchronize()
    coef = torch.from_numpy(ctx.recon_coefficients(wl.arm)).cuda()
    return ev, out, coef, dict(kw, seed=SEED)


def adj(ctx, wl, ev, mask, coef, off, kw):
    g, o9 = ctx.adjoint(wl.arm, ev, mask, coef, off, **kw)
    torch.cuda.synchronize()
    return g.cpu().numpy(), o9.cpu().numpy()


def plain_loss(ev, out):
    ok = out["ok"].cpu().numpy() == 1
    r = [(out[k].cpu().numpy() - ev[k].cpu().numpy())[ok] for k in ("dxdz", "y", "dydz", "dpp")]
    return float(np.mean(sum(x * x for x in r))), int(ok.sum())


@pytest.mark.parametrize("wl", [W.HMS_C0, W.SHMS_C1], ids=["hms", "shms"])
def test_loss_equals_plain_run(ctx, wl):
    ev, out, coef, kw = setup(ctx, wl)
    g, o9 = adj(ctx, wl, ev, out["ok"], coef, [0.0] * 6, kw)
    L, n_ok = plain_loss(ev, out)
    assert o9[7] == 0 and o9[8] == n_ok
    assert abs(o9[0] - L) <= 1e-12 * L, (o9[0], L)


@pytest.mark.parametrize("wl", [W.HMS_C0, W.SHMS_C1], ids=["hms", "shms"])
def test_gradients_equal_analytic(ctx, wl):
    ev, out, coef, kw = setup(ctx, wl)
    g, o9 = adj(ctx, wl, ev, out["ok"], coef, [0.0] * 6, kw)
    ok = out["ok"].cpu().numpy() == 1
    x = {k: out[k].cpu().numpy()[ok] for k in ("x_fp", "dx_fp", "y_fp", "dy_fp", "dxdz", "y", "dydz", "dpp")}
    t = {k: ev[k].cpu().numpy()[ok] for k in ("dxdz", "y", "dydz", "dpp", "fry")}
    h4 = t["fry"] / 100.0
    h4[np.abs(h4) <= 1e-30] = 1e-30
    h = np.stack([x["x_fp"] / 100.0, x["dx_fp"], x["y_fp"] / 100.0, x["dy_fp"], h4], axis=1)
    E = recon_exponents(wl.arm)
    M = np.prod(h[:, None, :] ** E[None, :, :], axis=2)              # [N, T]
    n = len(h)
    w = np.stack([2 * (x["dxdz"] - t["dxdz"]), 200 * (x["y"] - t["y"]),
                  2 * (x["dydz"] - t["dydz"]), 200 * (x["dpp"] - t["dpp"])], axis=1) / n
    G = M.T @ w                                                     # [T, 4]
    scale = np.abs(G).max()
    assert np.abs(g - G).max() <= 1e-9 * scale, np.abs(g - G).max() / scale
    # d s_k / d h_i from the exponents, then the four focal-plane offsets
    C = coef.cpu().numpy()
    dh = np.zeros((n, 4))
    for i in range(4):
        Ei = E.copy()
        Ei[:, i] = np.maximum(E[:, i] - 1, 0)
        Mi = np.prod(h[:, None, :] ** Ei[None, :, :], axis=2) * E[None, :, i]
        dh[:, i] = ((Mi @ C) * w).sum(axis=1)
    want = np.array([dh[:, 0].sum() / 100.0, dh[:, 2].sum() / 100.0, dh[:, 1].sum(), dh[:, 3].sum()])
    np.testing.assert_allclose(o9[1:5], want, rtol=1e-9, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("wl", [W.HMS_C0, W.SHMS_C1], ids=["hms", "shms"])
def test_finite_differences(ctx, wl):
    ev, out, coef, kw = setup(ctx, wl, 8000)
    mask = out["ok"]
    base = [0.01, -0.02, 2e-4, -1e-4, 1e-3, 5e-4]
    g, o9 = adj(ctx, wl, ev, mask, coef, base, kw)
    hstep = 1e-6
    for j in range(6):
        lo, hi = list(base), list(base)
        lo[j] -= hstep
        hi[j] += hstep
        fd = (adj(ctx, wl, ev, mask, coef, hi, kw)[1][0] - adj(ctx, wl, ev, mask, coef, lo, kw)[1][0]) / (2 * hstep)
        assert abs(fd - o9[1 + j]) <= 1e-5 * abs(o9[1 + j]) + 1e-8 * o9[0], (j, fd, o9[1 + j])
    rng = np.random.default_rng(0)
    flat = np.abs(g).ravel()
    picks = list(np.argsort(flat)[-6:]) + list(rng.choice(flat.size, 6, replace=False))
    for idx in picks:
        t_, k = divmod(int(idx), 4)
        c0 = float(coef[t_, k])
        hc = 1e-6 * max(1.0, abs(c0))
        cp, cm = coef.clone(), coef.clone()
        cp[t_, k] += hc
        cm[t_, k] -= hc
        fd = (adj(ctx, wl, ev, mask, cp, base, kw)[1][0] - adj(ctx, wl, ev, mask, cm, base, kw)[1][0]) / (2 * hc)
        assert abs(fd - g[t_, k]) <= 1e-5 * abs(g[t_, k]) + 1e-9 * o9[0] / hc, (t_, k, fd, g[t_, k])


def test_tape_mode_matches_oracle(ctx, oracle):
    wl = W.HMS_C0
    n, stride = 3000, 1024
    evn = W.generate(wl, n, seed=9)
    tape = W.uniform_tape(n, stride, 77)
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True)
    ref = oracle.run(wl.arm, evn, rng="tape", tape=tape, stride=stride, fit_f32=False, **kw)
    ok = ref["ok"] == 1
    L = float(np.mean(sum((ref[k] - evn[k])[ok] ** 2 for k in ("dxdz", "y", "dydz", "dpp"))))
    ev = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in evn.items()}
    coef = torch.from_numpy(ctx.recon_coefficients(wl.arm)).cuda()
    g, o9 = ctx.adjoint(wl.arm, ev, torch.from_numpy(ref["ok"]).cuda(), coef, [0.0] * 6, rng="tape",
                        tape=torch.from_numpy(tape).cuda(), stride=stride, **kw)
    o9 = o9.cpu().numpy()
    assert o9[8] == ok.sum() and o9[7] == 0
    assert abs(o9[0] - L) <= 1e-11 * L, (o9[0], L)


def test_autograd_binding_and_descent(ctx):
    from simc_gfortran_amd.autograd import OpticsProblem, optics_loss
    wl = W.HMS_C0
    ev, out, coef0, kw = setup(ctx, wl, 10000)
    problem = OpticsProblem(ctx, wl.arm, ev, out["ok"], **kw)
    coef = coef0.clone().requires_grad_(True)
    offsets = torch.tensor([0.05, -0.05, 1e-3, -1e-3, 2e-3, 1e-3], dtype=torch.float64, requires_grad=True)
    loss = optics_loss(problem, coef, offsets)
    loss.backward()
    g, o9 = adj(ctx, wl, ev, out["ok"], coef0, offsets.detach().tolist(), kw)
    # (atomic accumulation order differs between calls: equal to rounding)
    np.testing.assert_allclose(coef.grad.cpu().numpy(), g, rtol=1e-10, atol=1e-14 * np.abs(g).max())
    np.testing.assert_allclose(offsets.grad.numpy(), o9[1:7], rtol=1e-10, atol=1e-14 * np.abs(o9[1:7]).max())
    # steepest descent in the offsets with backtracking lowers the loss
    theta = offsets.detach().clone()
    L0 = float(loss.detach())
    L = L0
    for _ in range(3):
        th = theta.clone().requires_grad_(True)
        l = optics_loss(problem, coef0, th)
        l.backward()
        gr = th.grad
        l = l.detach()
        step = 1.0
        while step > 1e-12:
            cand = theta - step * float(l) / float((gr * gr).sum()) * gr
            lc = float(optics_loss(problem, coef0, cand))
            if lc < float(l):
                theta, L = cand, lc
                break
            step *= 0.5
    assert L < L0, (L, L0)


@pytest.mark.parametrize("wl", [W.HMS_C0, W.SHMS_C1], ids=["hms", "shms"])
def test_offset_gradients_against_oracle_differences(ctx, oracle, wl):
    """An independent tangent for s_p and dtheta: differences of the oracle's
    loss, the transport restated on the CPU (not the generated tangent
    polynomials of the adjoint)."""
    n = 20000
    evn = W.generate(wl, n, seed=18)
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True, wire_sigma=wl.wire_sigma)
    sp0, dth0 = 2e-3, 1e-3

    def oracle_run(sp, dth):
        e = dict(evn)
        e["dydz"] = evn["dydz"] - dth
        return oracle.run(wl.arm, e, rng="philox", seed=SEED, fit_f32=False, mom_offset=sp, **kw)

    def loss(r, sel):
        return float(np.mean(sum((r[k] - evn[k])[sel] ** 2 for k in ("dxdz", "y", "dydz", "dpp"))))

    h = {"sp": 2e-6, "dth": 2e-6}
    runs = {(0, 0): oracle_run(sp0, dth0)}
    for name in ("sp", "dth"):
        for m in (-2, -1, 1, 2):
            d = m * h[name] / 2
            runs[(name, m)] = oracle_run(sp0 + (d if name == "sp" else 0.0), dth0 + (d if name == "dth" else 0.0))
    sel = np.ones(n, bool)
    for r in runs.values():
        sel &= r["ok"] == 1
    assert sel.sum() > 0.25 * n
    ev = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in evn.items()}
    coef = torch.from_numpy(ctx.recon_coefficients(wl.arm)).cuda()
    mask = torch.from_numpy(sel.astype(np.uint8)).cuda()
    g, o9 = adj(ctx, wl, ev, mask, coef, [0.0, 0.0, 0.0, 0.0, sp0, dth0], dict(kw, seed=SEED))
    assert o9[7] == 0 and o9[8] == sel.sum()
    L0 = loss(runs[(0, 0)], sel)
    assert abs(o9[0] - L0) <= 1e-12 * L0, (o9[0], L0)
    for j, name in ((5, "sp"), (6, "dth")):
        d1 = (loss(runs[(name, 2)], sel) - loss(runs[(name, -2)], sel)) / (2 * h[name])
        d2 = (loss(runs[(name, 1)], sel) - loss(runs[(name, -1)], sel)) / h[name]
        rich = (4 * d2 - d1) / 3
        print(wl.name, name, "adjoint", o9[j], "oracle differences", rich, flush=True)
        assert abs(rich - o9[j]) <= 2e-6 * abs(o9[j]), (name, rich, o9[j])


def test_gradients_at_1e7_events(ctx):
    """The adjoint configuration's size: HMS, 1e7 events."""
    wl = W.HMS_C0
    n = 10_000_000
    print("adjoint 1e7", flush=True)
    ev, out, coef, kw = setup(ctx, wl, n)
    mask = out["ok"]
    base = [0.01, -0.02, 2e-4, -1e-4, 1e-3, 5e-4]
    g, o9 = adj(ctx, wl, ev, mask, coef, base, kw)
    assert o9[8] > 0.3 * n and o9[7] < 1e-4 * n
    hstep = 1e-6
    for j in range(6):
        lo, hi = list(base), list(base)
        lo[j] -= hstep
        hi[j] += hstep
        fd = (adj(ctx, wl, ev, mask, coef, hi, kw)[1][0] - adj(ctx, wl, ev, mask, coef, lo, kw)[1][0]) / (2 * hstep)
        assert abs(fd - o9[1 + j]) <= 1e-5 * abs(o9[1 + j]) + 1e-8 * o9[0], (j, fd, o9[1 + j])
    flat = np.abs(g).ravel()
    for idx in np.argsort(flat)[-4:]:
        t_, k = divmod(int(idx), 4)
        hc = 1e-6 * max(1.0, abs(float(coef[t_, k])))
        cp, cm = coef.clone(), coef.clone()
        cp[t_, k] += hc
        cm[t_, k] -= hc
        fd = (adj(ctx, wl, ev, mask, cp, base, kw)[1][0] - adj(ctx, wl, ev, mask, cm, base, kw)[1][0]) / (2 * hc)
        assert abs(fd - g[t_, k]) <= 1e-5 * abs(g[t_, k]) + 1e-9 * o9[0] / hc, (t_, k, fd, g[t_, k])
