"""The generated COSY polynomial code (simc_gfortran_amd/csrc/gen_cosy.py),
evaluated on the CPU: every forward class and reconstruction map of every
arm, translated from the generated device functions into Python (FMA as
multiply-add), against the direct sum out_k = sum_i c_ki prod_v r_v^e_vi of
the map file's terms (shared/transp.f:205-218, hms/mc_hms_recon.f:122-130);
the tangent functions of the adjoint form against central differences of the
value functions, and their values against the value functions bit for bit.
"""
import math
import os
import random
import re
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "simc_gfortran_amd", "csrc"))
import gen_cosy as G  # noqa: E402

MAPS = os.path.join(REPO, "data", "maps")
SPECS = {"hms": ("hms/forward_cosy.dat", "hms/recon_cosy.dat"),
         "shms": ("shms/shms_forward.dat", "shms/shms_recon.dat"),
         "sos": ("sos/forward_cosy.dat", "sos/recon_cosy.dat"),
         "hrs": ("hrsr/hrs_forward_cosy.dat", "hrsr/hrs_recon_cosy.dat")}
SCALE = (5.0, 0.05, 3.0, 0.05, 5.0)


@pytest.fixture(scope="module")
def functions(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("gen") / "cosy_gen")
    G.generate(MAPS, out)
    src = open(out + ".dev.inc").read()
    funcs = {}
    for m in re.finditer(r"__device__ __forceinline__ void (\w+)\((.*?)\)\n\{\n(.*?)\n\}\n", src, re.S):
        funcs[m.group(1)] = (m.group(2), m.group(3))
    return funcs


def _python(funcs, name):
    params, body = funcs[name]
    lines = []
    for ln in body.split("\n"):
        ln = ln.strip()
        if not ln or ln.startswith("asm(") or ln.startswith("int z") or "cosy_cptr" in ln:
            continue
        ln = re.sub(r"\bcb\d+\[", "c[", ln).replace("const double ", "").rstrip(";")
        lines.append("    " + ln.replace("__builtin_fma(", "fma("))
    names = [x.split()[-1] for x in params.split(",")[1:] if "&" not in x]
    outs = [x.split("&")[-1].strip() for x in params.split(",") if "&" in x]
    code = "def f(c, " + ", ".join(names) + "):\n" + "\n".join(lines) + "\n    return [" + ", ".join(outs) + "]\n"
    ns = {"fma": lambda a, b, c: a * b + c}
    exec(code, ns)
    return ns["f"]


def _pack(name, terms, nout):
    """The coefficient order the generated function reads."""
    return G.emit_poly(name, terms, nout, ["h0", "h1", "h2", "h3", "h4"], [])[0]


def _polys(arm):
    fw, rc = SPECS[arm]
    classes = G.parse_forward(os.path.join(MAPS, fw))
    out = [(f"cosy_{arm}_fwd_c{c}", classes[c - 1], 5) for c in G.TRANSP_CLASSES[arm]]
    return out + [(f"cosy_{arm}_recon", G.parse_recon(os.path.join(MAPS, rc)), 4)]


@pytest.mark.parametrize("arm", list(SPECS))
def test_generated_polynomials_equal_the_direct_sum(functions, arm):
    rng = random.Random(3)
    for name, terms, nout in _polys(arm):
        pack = _pack(name, terms, nout)
        c = [terms[ti][0][k] for ti, k in pack]
        assert sorted(pack) == sorted((ti, k) for ti, (cc, _) in enumerate(terms) for k in range(nout) if cc[k])
        f = _python(functions, name)
        for _ in range(10):
            r = [rng.uniform(-1, 1) * s for s in SCALE]
            got = f(c, *r)
            for k in range(nout):
                parts = [cc[k] * math.prod(r[v] ** e[v] for v in range(5)) for cc, e in terms]
                assert abs(got[k] - sum(parts)) <= 1e-13 * (sum(abs(p) for p in parts) + 1e-300), (name, k)


@pytest.mark.parametrize("arm", list(SPECS))
def test_tangent_polynomials(functions, arm):
    rng = random.Random(5)
    for name, terms, nout in _polys(arm)[:-1]:
        pack = _pack(name, terms, nout)
        c = [terms[ti][0][k] for ti, k in pack]
        f, ft = _python(functions, name), _python(functions, name + "_tan")
        for _ in range(3):
            r = [rng.uniform(-1, 1) * s for s in SCALE]
            da = [rng.uniform(-1, 1) * s for s in SCALE]
            db = [rng.uniform(-1, 1) * s for s in SCALE]
            res = ft(c, *r, *da, *db)
            assert res[:5] == f(c, *r), name
            for d, off in ((da, 5), (db, 10)):
                h = 1e-6
                for k in range(5):
                    fd = (f(c, *[r[v] + h * d[v] for v in range(5)])[k] -
                          f(c, *[r[v] - h * d[v] for v in range(5)])[k]) / (2 * h)
                    assert abs(fd - res[off + k]) <= 1e-6 * (abs(fd) + 1e-2), (name, k)


def test_planted_term_is_caught(functions):
    """A term dropped from the pack (the generator's plausible mistake) fails
    the comparison above."""
    name, terms, nout = _polys("hms")[0]
    pack = _pack(name, terms, nout)
    c = [terms[ti][0][k] for ti, k in pack]
    # a second-order term's coefficient (first-order ones dominate the sum)
    j = next(i for i, (ti, k) in enumerate(pack) if sum(terms[ti][1]) == 2)
    c[j] = 0.0
    f = _python(functions, name)
    r = [0.7 * s for s in SCALE]
    got = f(c, *r)
    bad = 0
    for k in range(nout):
        parts = [cc[k] * math.prod(r[v] ** e[v] for v in range(5)) for cc, e in terms]
        bad += abs(got[k] - sum(parts)) > 1e-13 * (sum(abs(p) for p in parts) + 1e-300)
    assert bad
