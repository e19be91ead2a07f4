"""Build-time generator of the COSY polynomial device code.

Reads the forward and reconstruction map files (the same files mc_hms /
mc_shms / mc_sos / mc_hrsr / mc_hrsl load at run time: hms/forward_cosy.dat,
hms/recon_cosy.dat, shms/shms_forward.dat, shms/shms_recon.dat,
sos/forward_cosy.dat, sos/recon_cosy.dat, hrsr/hrs_forward_cosy.dat,
hrsr/hrs_recon_cosy.dat -- the HRS-L files are the same) and writes cosy_gen.inc:

* one __device__ function per non-drift forward class and per recon map that
  evaluates  out_k = sum_i c_{k,i} prod_v r_v^{e_{v,i}}  (shared/transp.f:205-218,
  hms/mc_hms_recon.f:122-130) with
    - every monomial formed from a parent monomial by ONE multiply, walking the
      monomial trie depth first (the live set stays at the depth of the trie),
    - one FMA per NON-ZERO coefficient (about 2 of the 5 forward outputs of a
      term are non-zero: mid-plane symmetry),
    - coefficients read from a packed table with compile-time offsets, so the
      loads are uniform scalar loads through the constant address space and
      coefficient VALUES stay a run-time input; the reads are split into
      blocks of COEF_BLOCK coefficients whose loads overlap the arithmetic on
      the previous block (_block_fences);
* for the adjoint form: each forward class with two tangent directions
  carried along (same operations for the values), and the recon polynomial
  with all coefficients as run-time values, plus the gradient of its sums
  contracted with per-event weights with respect to the focal-plane inputs;
* the structure tables (exponents, non-zero masks, pack order) the host uses
  to check a map loaded at run time against the compiled structure and to
  pack its coefficients.

Only the exponent STRUCTURE is compiled in.  A map file with the same terms
and different coefficient values (a new tune, an optics fit) needs no rebuild.
"""
from __future__ import annotations

import os
import re
import sys

NVAR = 5
# coefficients per scalar-load block of the forward and reconstruction
# polynomials (see _block_fences); 0 = no blocking
COEF_BLOCK = 32
coef_block = COEF_BLOCK


def _read_records(path):
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            yield (line + " " * 132)[:132]


def _fval(field):
    s = field.replace(" ", "")
    if not s:
        return 0.0
    return float(s.replace("D", "E").replace("d", "e"))


def parse_forward(path):
    """Classes of a forward map (shared/transp.f:318-470): list of lists of
    (coeff[5], expon[5]) with the TOF line dropped; plus per-class drift flag."""
    recs = list(_read_records(path))
    i = 0
    while recs[i][0] == "!":
        i += 1
    classes = []
    while True:
        terms = []
        while recs[i][:4] != " ---":
            s = recs[i]
            c = [_fval(s[1 + 14 * k: 1 + 14 * (k + 1)]) for k in range(5)]
            d = [0 if ch == " " else int(ch) for ch in s[72:78]]
            if d[4] == 0:
                terms.append((c, [d[0], d[1], d[2], d[3], d[5]]))
            i += 1
        classes.append(terms)
        i += 1
        while i < len(recs) and (recs[i][0] == "!" or recs[i][:4] == " ---" or not recs[i].strip()):
            i += 1
        if i >= len(recs):
            break
    return classes


def parse_recon(path):
    recs = list(_read_records(path))
    i = 0
    while recs[i][0] == "!":
        i += 1
    terms = []
    while recs[i][:4] != " ---":
        s = recs[i]
        c = [_fval(s[1 + 16 * k: 1 + 16 * (k + 1)]) for k in range(4)]
        e = [0 if ch == " " else int(ch) for ch in s[66:71]]
        terms.append((c, e))
        i += 1
    return terms


# Forward classes evaluated through transp(): hms/mc_hms.f:268,276,294,302,320,328,351
# and shms/mc_shms.f:462,606,617,669,680,732,743,874,893,910,928,945,962,979,999.
# The other classes are pure drifts, which the arms step with project().
TRANSP_CLASSES = {
    "hms": [2, 3, 5, 6, 8, 9, 11],
    "shms": [3, 7, 8, 12, 13, 17, 18, 23, 24, 25, 26, 27, 28, 29, 30],
    "sos": [2, 3, 4, 5, 6, 7, 8, 9],          # sos/mc_sos.f:219-329
    "hrs": [2, 3, 5, 6, 8, 10, 11],           # hrsl/mc_hrsl.f:244-466 (HRS-R and HRS-L share the map)
}


def _last_var(e):
    for v in range(NVAR - 1, -1, -1):
        if e[v]:
            return v
    return 0


def _parent(e):
    v = _last_var(e)
    p = list(e)
    p[v] -= 1
    return tuple(p)


def emit_poly(fname, terms, nout, inputs, lines):
    """Emit one device function; returns the pack list [(term_index, out)]."""
    byexp = {}
    for ti, (c, e) in enumerate(terms):
        byexp.setdefault(tuple(e), []).append(ti)
    root = (0,) * NVAR
    needed = {root}
    for e in byexp:
        while e not in needed:
            needed.add(e)
            e = _parent(e)
    pack = []
    body = []
    counter = [0]

    def fmas(node, mname):
        for ti in byexp.get(node, []):
            c = terms[ti][0]
            for k in range(nout):
                if c[k] != 0.0:
                    idx = len(pack)
                    pack.append((ti, k))
                    if mname is None:
                        body.append(f"  o{k} += c[{idx}];")
                    else:
                        body.append(f"  o{k} = __builtin_fma({mname}, c[{idx}], o{k});")

    def walk(node, mname):
        fmas(node, mname)
        lv = _last_var(node) if node != root else 0
        for v in range(lv, NVAR):
            ch = list(node)
            ch[v] += 1
            ch = tuple(ch)
            if ch not in needed:
                continue
            if mname is None:
                cname = inputs[v]
            else:
                counter[0] += 1
                cname = f"m{counter[0]}"
                body.append(f"  const double {cname} = {mname} * {inputs[v]};")
            walk(ch, cname)

    walk(root, None)
    args = ", ".join(f"double {x}" for x in inputs)
    outs = ", ".join(f"double& o{k}" for k in range(nout))
    lines.append(f"__device__ __forceinline__ void {fname}(const double __attribute__((address_space(4)))* __restrict__ c, {args}, {outs})")
    lines.append("{")
    for k in range(nout):
        lines.append(f"  o{k} = 0.0;")
    lines.extend(_block_fences(body, coef_block))
    lines.append("}")
    lines.append("")
    return pack, counter[0]


_CREF = re.compile(r"\bc\[(\d+)\]")


def _block_fences(body, block):
    """Split the coefficient reads of a generated body into blocks of `block`
    consecutive coefficients, read through per-block pointers cb<j>.  Before
    the arithmetic of block j, one empty asm statement takes every
    coefficient of block j as an SGPR operand and yields the zero offset of
    pointer cb<j+1>:
      * all of block j has arrived there (one wait), and nothing later has
        been issued yet;
      * the loads of block j + 1 issue right after it and complete while
        block j is consumed.
    Scalar loads complete out of order, so a wait for one outstanding load
    waits for all of them; without this ordering the compiler issues a
    block's loads just before their use and waits for them at once.
    block = 0 leaves the body as is."""
    if not block:
        return body
    refs_of = {}
    order = []
    for line in body:
        for x in _CREF.findall(line):
            j = int(x) // block
            if j not in refs_of:
                order.append(j)
            refs_of.setdefault(j, []).append(int(x))
    if order != sorted(order):
        return body  # blocks not first read in order: leave unfenced
    out = ["  const cosy_cptr cb0 = c;"]
    done = set()
    for line in body:
        refs = [int(x) for x in _CREF.findall(line)]
        for idx in refs:
            j = idx // block
            if j in done:
                continue
            done.add(j)
            vals = sorted(set(refs_of[j]))
            ops = " ".join(f"%{k + 1}" for k in range(len(vals)))
            ins = ", ".join(f'"s"(cb{j}[{i}])' for i in vals)
            out.append(f"  int z{j + 1};")
            out.append(f'  asm("s_mov_b32 %0, 0 ; {ops}" : "=s"(z{j + 1}) : {ins});')
            out.append(f"  const cosy_cptr cb{j + 1} = c + z{j + 1};")
        if refs:
            line = _CREF.sub(lambda m: f"cb{int(m.group(1)) // block}[{m.group(1)}]", line)
        out.append(line)
    return out


def _trie(terms):
    byexp = {}
    for ti, (c, e) in enumerate(terms):
        byexp.setdefault(tuple(e), []).append(ti)
    root = (0,) * NVAR
    needed = {root}
    for e in byexp:
        while e not in needed:
            needed.add(e)
            e = _parent(e)
    return byexp, root, needed


def emit_poly_tangent(fname, terms, pack, inputs, lines):
    """The forward polynomial with two tangent directions (A, B) carried along:
    o_k and d o_k = J d r for dr = rA, rB.  The value part performs the same
    operations in the same order as emit_poly's function, and the packed
    coefficient table (pack order) is the same one."""
    byexp, root, needed = _trie(terms)
    pos = {tk: i for i, tk in enumerate(pack)}
    body = []
    counter = [0]

    def fmas(node, m):
        for ti in byexp.get(node, []):
            c = terms[ti][0]
            for k in range(5):
                if c[k] != 0.0:
                    idx = pos[(ti, k)]
                    if m is None:
                        body.append(f"  o{k} += c[{idx}];")
                    else:
                        body.append(f"  o{k} = __builtin_fma({m[0]}, c[{idx}], o{k});")
                        body.append(f"  a{k} = __builtin_fma({m[1]}, c[{idx}], a{k});")
                        body.append(f"  b{k} = __builtin_fma({m[2]}, c[{idx}], b{k});")

    def walk(node, m):
        fmas(node, m)
        lv = _last_var(node) if node != root else 0
        for v in range(lv, NVAR):
            ch = list(node)
            ch[v] += 1
            ch = tuple(ch)
            if ch not in needed:
                continue
            x, xa, xb = inputs[v], "a" + inputs[v], "b" + inputs[v]
            if m is None:
                cm = (x, xa, xb)
            else:
                counter[0] += 1
                n = counter[0]
                cm = (f"m{n}", f"ma{n}", f"mb{n}")
                body.append(f"  const double m{n} = {m[0]} * {x};")
                body.append(f"  const double ma{n} = __builtin_fma({m[1]}, {x}, {m[0]} * {xa});")
                body.append(f"  const double mb{n} = __builtin_fma({m[2]}, {x}, {m[0]} * {xb});")
            walk(ch, cm)

    walk(root, None)
    args = ", ".join(f"double {x}" for x in inputs)
    argsa = ", ".join(f"double a{x}" for x in inputs)
    argsb = ", ".join(f"double b{x}" for x in inputs)
    outs = ", ".join(f"double& {p}{k}" for p in "oab" for k in range(5))
    lines.append(f"__device__ __forceinline__ void {fname}(const double __attribute__((address_space(4)))* "
                 f"__restrict__ c, {args}, {argsa}, {argsb}, {outs})")
    lines.append("{")
    for p in "oab":
        for k in range(5):
            lines.append(f"  {p}{k} = 0.0;")
    lines.extend(body)
    lines.append("}")
    lines.append("")


def emit_recon_full(fname, terms, inputs, lines):
    """The reconstruction polynomial with EVERY coefficient a run-time value,
    c[4 t + k] for term t of the map file (zero coefficients included: they
    are parameters of an optics fit).  Same trie order and operations as
    emit_poly's function, so with the map's own coefficients the sums are
    bit-identical."""
    byexp, root, needed = _trie(terms)
    body = []
    counter = [0]

    def fmas(node, m):
        for ti in byexp.get(node, []):
            for k in range(4):
                ci = f"c[{4 * ti + k}]"
                body.append(f"  s{k} += {ci};" if m is None else f"  s{k} = __builtin_fma({m}, {ci}, s{k});")

    def walk(node, m):
        fmas(node, m)
        lv = _last_var(node) if node != root else 0
        for v in range(lv, NVAR):
            ch = list(node)
            ch[v] += 1
            ch = tuple(ch)
            if ch not in needed:
                continue
            if m is None:
                cm = inputs[v]
            else:
                counter[0] += 1
                cm = f"m{counter[0]}"
                body.append(f"  const double {cm} = {m} * {inputs[v]};")
            walk(ch, cm)

    walk(root, None)
    args = ", ".join(f"double {x}" for x in inputs)
    outs = ", ".join(f"double& s{k}" for k in range(4))
    lines.append(f"__device__ __forceinline__ void {fname}(const double __attribute__((address_space(4)))* "
                 f"__restrict__ c, {args}, {outs})")
    lines.append("{")
    for k in range(4):
        lines.append(f"  s{k} = 0.0;")
    lines.extend(body)
    lines.append("}")
    lines.append("")


def emit_recon_wgrad(fname, terms, inputs, lines):
    """Gradient of Q = sum_k w_k s_k (the reconstruction sums contracted with
    per-event weights) with respect to the first four inputs (the fifth,
    x_tar, is data): every monomial carries its value and four partial
    derivatives along the trie, every term its weight q_t = sum_k w_k c[4t+k]."""
    byexp, root, needed = _trie(terms)
    body = []
    counter = [0]
    ND = 4

    def fmas(node, m):
        if m is None:
            return  # the constant term has no gradient
        for ti in byexp.get(node, []):
            q = f"q{ti}"
            body.append(f"  const double {q} = __builtin_fma(w3, c[{4 * ti + 3}], __builtin_fma(w2, c[{4 * ti + 2}], "
                        f"__builtin_fma(w1, c[{4 * ti + 1}], w0 * c[{4 * ti}])));")
            for i in range(ND):
                if m[1 + i] is not None:
                    body.append(f"  d{i} = __builtin_fma({m[1 + i]}, {q}, d{i});")

    def walk(node, m):
        fmas(node, m)
        lv = _last_var(node) if node != root else 0
        for v in range(lv, NVAR):
            ch = list(node)
            ch[v] += 1
            ch = tuple(ch)
            if ch not in needed:
                continue
            x = inputs[v]
            counter[0] += 1
            n = counter[0]
            if m is None:
                cm = (x,) + tuple("1.0" if i == v else None for i in range(ND))
            else:
                names = [f"m{n}"]
                body.append(f"  const double m{n} = {m[0]} * {x};")
                for i in range(ND):
                    d = m[1 + i]
                    parts = []
                    if d is not None:
                        parts.append(f"{d} * {x}")
                    if i == v:
                        parts.append(m[0])
                    if not parts:
                        names.append(None)
                        continue
                    body.append(f"  const double m{n}d{i} = {' + '.join(parts)};")
                    names.append(f"m{n}d{i}")
                cm = tuple(names)
            walk(ch, cm)

    walk(root, None)
    args = ", ".join(f"double {x}" for x in inputs)
    lines.append(f"__device__ __forceinline__ void {fname}(const double __attribute__((address_space(4)))* "
                 f"__restrict__ c, {args}, double w0, double w1, double w2, double w3, "
                 f"double& d0, double& d1, double& d2, double& d3)")
    lines.append("{")
    for i in range(ND):
        lines.append(f"  d{i} = 0.0;")
    lines.extend(body)
    lines.append("}")
    lines.append("")


def generate(map_dir, out_path, block=COEF_BLOCK):
    global coef_block
    coef_block = block
    specs = [
        ("hms", os.path.join(map_dir, "hms", "forward_cosy.dat"),
         os.path.join(map_dir, "hms", "recon_cosy.dat")),
        ("shms", os.path.join(map_dir, "shms", "shms_forward.dat"),
         os.path.join(map_dir, "shms", "shms_recon.dat")),
        ("sos", os.path.join(map_dir, "sos", "forward_cosy.dat"),
         os.path.join(map_dir, "sos", "recon_cosy.dat")),
        ("hrs", os.path.join(map_dir, "hrsr", "hrs_forward_cosy.dat"),
         os.path.join(map_dir, "hrsr", "hrs_recon_cosy.dat")),
    ]
    dev, host = [], []
    dev.append("// Generated by gen_cosy.py from the COSY map files; do not edit.")
    dev.append("#pragma once")
    dev.append("typedef const double __attribute__((address_space(4)))* cosy_cptr;")
    host.append("// Generated by gen_cosy.py from the COSY map files; do not edit.")
    host.append("#pragma once")
    host.append("#include <cstdint>")
    host.append("namespace simc_gen {")
    host.append("struct PolyInfo { int cls; int n_terms; int nout; int coef_offset; int n_coef;"
                " const uint8_t (*expon)[5]; const uint16_t (*pack)[2]; int n_mul; };")
    summary = []
    for arm, fwd, rec in specs:
        classes = parse_forward(fwd)
        recon = parse_recon(rec)
        offset = 0
        infos = []
        for cls in TRANSP_CLASSES[arm]:
            terms = classes[cls - 1]
            fname = f"cosy_{arm}_fwd_c{cls}"
            pack, nmul = emit_poly(fname, terms, 5, ["r0", "r1", "r2", "r3", "r4"], dev)
            emit_poly_tangent(fname + "_tan", terms, pack, ["r0", "r1", "r2", "r3", "r4"], dev)
            infos.append((cls, terms, 5, offset, pack, nmul))
            offset += len(pack)
        fname = f"cosy_{arm}_recon"
        emit_recon_full(fname + "_full", recon, ["h0", "h1", "h2", "h3", "h4"], dev)
        emit_recon_wgrad(fname + "_wgrad", recon, ["h0", "h1", "h2", "h3", "h4"], dev)
        dev.append(f"#define COSY_{arm.upper()}_RECON_TERMS {len(recon)}")
        pack, nmul = emit_poly(fname, recon, 4, ["h0", "h1", "h2", "h3", "h4"], dev)
        infos.append((0, recon, 4, offset, pack, nmul))
        offset += len(pack)
        for cls, terms, nout, off, pack, nmul in infos:
            tag = f"{arm}_c{cls}" if cls else f"{arm}_recon"
            host.append(f"static const uint8_t expon_{tag}[{len(terms)}][5] = {{" +
                        ",".join("{" + ",".join(str(x) for x in e) + "}" for _, e in terms) + "};")
            host.append(f"static const uint16_t pack_{tag}[{len(pack)}][2] = {{" +
                        ",".join(f"{{{t},{k}}}" for t, k in pack) + "};")
            dev.append(f"#define COSY_{tag.upper()}_OFF {off}")
        host.append(f"static const PolyInfo polys_{arm}[] = {{")
        for cls, terms, nout, off, pack, nmul in infos:
            tag = f"{arm}_c{cls}" if cls else f"{arm}_recon"
            host.append(f"  {{{cls}, {len(terms)}, {nout}, {off}, {len(pack)}, expon_{tag}, pack_{tag}, {nmul}}},")
        host.append("};")
        host.append(f"static const int n_polys_{arm} = {len(infos)};")
        host.append(f"static const int n_coef_{arm} = {offset};")
        dev.append(f"#define COSY_{arm.upper()}_NCOEF {offset}")
        summary.append((arm, len(classes), len(infos), offset))
    host.append("}  // namespace simc_gen")
    with open(out_path + ".dev.inc", "w") as f:
        f.write("\n".join(dev) + "\n")
    with open(out_path + ".host.inc", "w") as f:
        f.write("\n".join(host) + "\n")
    return summary


if __name__ == "__main__":
    print(generate(sys.argv[1], sys.argv[2]))
