"""Emit generated process descriptions.

``python -m madflow_amd.procgen.emit`` regenerates, for g g > t t~ + 0..3 g and
u u~ > t t~ + 0..1 g:

* ``madflow_amd/processes/<name>.json`` — the process description (helicity table,
  external wavefunctions, HELAS program, jamp coefficients, colour matrix,
  denominator): the data the PyOut exporter bakes into ``matrix_<proc>.py``
  (``madgraph_plugin/PyOut_exporter.py:125-285``);
* ``madflow_amd/processes/matrix_1_<name>.py`` — the drop-in process module with
  ``class Matrix_1_<name>`` and ``smatrix(all_ps, *model_params)``
  (``madgraph_plugin/template_files/matrix_method_python.inc:60-138``);
* ``madflow_amd/csrc/procs/<name>.inc`` — straight-line HIP device code for the
  per-helicity amplitude program of the fused smatrix kernel.
"""

import json
import re as _re
import os
import sys

from fractions import Fraction

from .diagrams import gg_ttx_ng, qg_ttxq, qq_ttx_ng, schedule_ops

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROC_DIR = os.path.join(PKG, "processes")
INC_DIR = os.path.join(PKG, "csrc", "procs")

NGLUONS = (0, 1, 2, 3)        # g g > t t~ + ng g
QQ_NGLUONS = (0, 1, 2)        # u u~ > t t~ + ng g


# ---------------------------------------------------------------- HIP emission

def _exact(x):
    """A rational coefficient as an int when it is integral (the Gaussian-integer
    coefficients of the gluon processes stay ints), else as a Fraction."""
    x = Fraction(x)
    return int(x) if x.denominator == 1 else x


def _lit(x):
    """C++ literal of a rational coefficient in the kernel's real type R."""
    x = Fraction(x)
    if x.denominator == 1:
        return "R(%d)" % x.numerator
    return "R(%d.0 / %d.0)" % (x.numerator, x.denominator)


def _cname(name):
    return name.upper()


def _inv_index(i, j, n):
    if i > j:
        i, j = j, i
    return i * n - i * (i + 1) // 2 + (j - i - 1)


def den_expr(op, desc):
    """P^2 - M^2 of an off-shell current as a sum of the per-event invariants
    inv[k] = 2 k_i.k_j (outgoing momenta, computed once per event in FP64).

    P^2 = sum_S m_l^2 + sum_{i<j in S} 2 k_i.k_j; a top-quark line carries exactly one
    external top of mass M, so m_t^2 - M^2 cancels exactly and is left out.  For
    top-line currents the complementary leg set is used when it is shorter (momentum
    conservation gives the same P^2)."""
    n = desc["nexternal"]
    S = list(op["legs"])
    tops = {2, 3}
    if op["mass"] != "ZERO":
        assert len(tops & set(S)) == 1
        Sc = [l for l in range(n) if l not in S]
        if len(Sc) < len(S):
            S = Sc
    else:
        assert not (tops & set(S))
    terms = ["inv[%d]" % _inv_index(i, j, n) for a, i in enumerate(S) for j in S[a + 1:]]
    return "(" + " + ".join(terms) + ")"


def merged_ops(desc):
    """The amplitude program of the fused kernel.  Amplitudes that close on the same
    fermion pair with proportional jamp-coefficient vectors are evaluated once, on the
    sum of their vector currents (FFV1_0 is linear in V, and momentum conservation gives
    all of them the same momentum): amplitude ops with args [F1, F2, V_1 .. V_k],
    ``factors`` (unit Gaussian integers f_i, the sum is sum f_i V_i) and ``norm`` (the
    group's jamp coefficients).  The oracle keeps evaluating every Feynman amplitude."""
    coef = {}
    for c, terms in enumerate(desc["jamp"]):
        for (a, re, im) in terms:
            coef.setdefault(a, {})[c] = (Fraction(re), Fraction(im))
    groups = {}
    order = []
    for op in desc["ops"]:
        if op["kind"] != "amp":
            continue
        cs = sorted(coef.get(op["out"], {}).items())
        if not cs:
            continue
        # factor: the first coefficient c0; norm: every coefficient divided by c0
        r0, i0 = cs[0][1]
        n0 = r0 * r0 + i0 * i0
        norm = tuple((c, _exact((r * r0 + i * i0) / n0), _exact((i * r0 - r * i0) / n0)) for c, (r, i) in cs)
        key = (op["routine"], op["args"][0], op["args"][1], op["coup"], norm)
        if key not in groups:
            groups[key] = []
            order.append(key)
        groups[key].append((op, (_exact(r0), _exact(i0))))
    ops = [op for op in desc["ops"] if op["kind"] == "wf"]
    for key in order:
        routine, f1, f2, coup, norm = key
        members = groups[key]
        ops.append(dict(kind="amp", routine=routine, out=members[0][0]["out"],
                        args=[f1, f2] + [m[0]["args"][2] for m in members],
                        factors=[m[1] for m in members], coup=coup, norm=list(norm)))
    return ops


def device_program(desc):
    """The merged amplitude program of one helicity row, scheduled for few live currents."""
    return schedule_ops(merged_ops(desc), desc["nexternal"])


DUAL_OFF = 100000


def multi_program(desc, k):
    """One program for the 2^k helicity rows that differ only in the helicities of k
    external legs X_0 .. X_{k-1} (the k legs with the fewest dependent currents).  A
    current that depends on the legs in the set m is evaluated once per helicity
    assignment of m, 2^|m| times; a current that depends on none of them is computed once
    for all rows.  Every amplitude is evaluated once per row.  Copy ids: id + DUAL_OFF * r,
    r = the bits (bit b = leg X_b at helicity +1) of the assignment; amplitude copies
    carry ``jset`` = the row c, whose bit b is the helicity of X_b.  The single-row
    schedule is kept, the copies right after their original.  Returns (legs, ops)."""
    ops = device_program(desc)
    n = desc["nexternal"]
    size = {}
    for X in range(n):
        dep = {X}
        for op in ops:
            if op["kind"] == "wf" and any(a in dep for a in op["args"]):
                dep.add(op["out"])
        size[X] = len(dep)
    legs = sorted(range(n), key=lambda X: (size[X], X))[:k]
    mask = {X: 1 << b for b, X in enumerate(legs)}
    out = []

    def name(a, r):
        r &= mask.get(a, 0)
        return a + DUAL_OFF * r

    for op in ops:
        m = 0
        for a in op["args"]:
            m |= mask.get(a, 0)
        if op["kind"] == "wf":
            mask[op["out"]] = m
            subs = [r for r in range(1 << k) if r & ~m == 0]
        else:
            subs = list(range(1 << k))
        for r in subs:
            c = dict(op)
            c["args"] = [name(a, r) for a in op["args"]]
            if op["kind"] == "wf":
                c["out"] = op["out"] + DUAL_OFF * r
            elif r:
                c["jset"] = r
            out.append(c)
    return legs, out


def dual_program(desc):
    """multi_program with one paired leg: two helicity rows per pass.  Returns (X, ops)."""
    legs, ops = multi_program(desc, 1)
    return legs[0], ops


def _mom_expr(legs, overall, nin):
    """C expression of component c of overall * sum_{l in legs} k_l, k_l = (-1 if l is
    incoming else +1) * p[l] (the slot momentum of a current carrying those legs)."""
    terms = []
    for l in sorted(legs):
        sgn = overall * (-1 if l < nin else 1)
        terms.append(("+ " if sgn > 0 else "- ") + "p[%d][c]" % l)
    expr = " ".join(terms)
    return expr[2:] if expr.startswith("+ ") else "-" + expr[2:]


def _scaled(expr, re, im):
    """expr times a coefficient that is real or imaginary: a unit as a sign or an
    i-multiplication, any other rational as a product with its literal as well."""
    assert re == 0 or im == 0
    if im == 0:
        if re in (1, -1):
            return expr if re > 0 else "-" + expr
        return "(%s * %s)" % (_lit(re), expr)
    if im in (1, -1):
        return ("mul_i(%s)" if im > 0 else "mul_mi(%s)") % expr
    return "(%s * mul_i(%s))" % (_lit(im), expr)


def _jamp_update(w, indent, jn, c, re, im, a):
    """jamp[c] += (re + i im) * a for a real or imaginary rational coefficient."""
    if im == 0 and re in (1, -1):
        w("%s%s[%d] %s= %s;" % (indent, jn, c, "+" if re > 0 else "-", a))
    elif re == 0 and im in (1, -1):
        w("%s%s[%d] %s= mul_i(%s);" % (indent, jn, c, "+" if im > 0 else "-", a))
    else:
        w("%s%s[%d] += %s;" % (indent, jn, c, _scaled(a, re, im)))


def _pair_side(a, b):
    """For an amplitude and its next-row twin: 1, 2 or 3 if they differ only in F1, only in
    F2 or only in the vector current(s) (then one FFV1_0_pair call evaluates both), else 0."""
    if a["kind"] != "amp" or b["kind"] != "amp" or a.get("jset") or b.get("jset") != 1:
        return 0
    if a["routine"] != "FFV1_0" or b["routine"] != "FFV1_0":
        return 0
    if a["factors"] != b["factors"] or a["norm"] != b["norm"] or a["coup"] != b["coup"]:
        return 0
    d1, d2, dv = a["args"][0] != b["args"][0], a["args"][1] != b["args"][1], a["args"][2:] != b["args"][2:]
    if d1 + d2 + dv != 1:
        return 0
    return 1 if d1 else 2 if d2 else 3


def _emit_ops(w, prog, desc, legs_of):
    """Emit the statements of an amplitude program (wavefunction ops and amplitudes).  An
    amplitude followed by its twin for the next helicity row of a pass, differing only in
    one fermion, is emitted as one FFV1_0_pair call."""
    n = desc["nexternal"]
    skip = set()
    for i, op in enumerate(prog):
        if i in skip:
            continue
        side = _pair_side(op, prog[i + 1]) if i + 1 < len(prog) else 0
        if side:
            skip.add(i + 1)
            tw = prog[i + 1]
            coup = "prm.%s" % op["coup"]
            f1, f2, vs = op["args"][0], op["args"][1], op["args"][2:]
            w("  { cx<R> amp0, amp1;")
            if len(vs) == 1:
                (r0, i0), = op["factors"]
                va, vb = "w%d" % vs[0], "w%d" % tw["args"][2]
                post = "    const cx<R> amp = %s, ampt = %s;" % (_scaled("amp0", r0, i0), _scaled("amp1", r0, i0))
            else:
                for nm, vv in (("vs", vs), ("vst", tw["args"][2:])):
                    if nm == "vst" and side != 3:
                        continue
                    w("    W<R> %s;" % nm)
                    w("    #pragma unroll")
                    w("    for (int q = 2; q < 6; ++q) %s[q] = %s;" % (nm, " + ".join(
                        _scaled("w%d[q]" % v, r, i_) for v, (r, i_) in zip(vv, op["factors"]))))
                va, vb = "vs", "vst"
                post = "    const cx<R> amp = amp0, ampt = amp1;"
            if side == 2:
                w("    FFV1_0_pair2(w%d, w%d, w%d, %s, %s, amp0, amp1);" % (f1, f2, tw["args"][1], va, coup))
            elif side == 1:
                w("    FFV1_0_pair1(w%d, w%d, w%d, %s, %s, amp0, amp1);" % (f1, tw["args"][0], f2, va, coup))
            else:
                w("    FFV1_0_pairV(w%d, w%d, %s, %s, %s, amp0, amp1);" % (f1, f2, va, vb, coup))
            w(post)
            for jn, a in (("jamp", "amp"), ("jamp2", "ampt")):
                for (c, re, im) in op["norm"]:
                    _jamp_update(w, "    ", jn, c, re, im, a)
            w("  }")
            w("  __builtin_amdgcn_sched_barrier(0);")
            continue
        args = ", ".join("w%d" % a for a in op["args"])
        coup = "prm.%s" % op["coup"]
        if op["kind"] == "wf":
            mass = "ZERO" if op["mass"] == "ZERO" else "prm.%s" % op["mass"]
            width = "ZERO" if op["width"] == "ZERO" else "prm.%s" % op["width"]
            # FP32: P^2 - M^2 from the FP64 invariants; FP64: from the current's slot momentum,
            # the reference's own arithmetic (Lorentz square of the summed momenta)
            r = op["routine"]
            m = _re.match(r"VVVV(\d)P0_1$", r)
            slot_r = "VVVVP0_1s<%s>" % m.group(1) if m else r
            dexpr = den_expr(op, desc)
            nin = desc["ninitial"]
            w("  W<R> w%d;" % op["out"])
            # FP32: momenta from the external momenta of each current's legs, so that no
            # wavefunction has to keep its momentum slots in registers
            if r in ("FFV1_1", "FFV1_2"):
                w("  if constexpr (StableForms<R>::value) {")
                w("    R P[4];")
                w("    #pragma unroll")
                w("    for (int c = 0; c < 4; ++c) P[c] = %s;" % _mom_expr(op["legs"], -1, nin))
                w("    %s_p(%s, %s, %s, %s, %s, P, w%d);" % (r, args, coup, mass, width, dexpr, op["out"]))
                w("  } else %s(%s, %s, %s, %s, w%d);" % (r, args, coup, mass, width, op["out"]))
            elif r == "VVV1P0_1":
                a2, a3 = op["args"]
                w("  if constexpr (StableForms<R>::value) {")
                w("    R P2[4], P3[4];")
                w("    #pragma unroll")
                w("    for (int c = 0; c < 4; ++c) {")
                w("      P2[c] = %s;" % _mom_expr(legs_of[a2], 1, nin))
                w("      P3[c] = %s;" % _mom_expr(legs_of[a3], 1, nin))
                w("    }")
                w("    VVV1P0_1_p(%s, %s, %s, %s, %s, P2, P3, w%d);" % (args, coup, mass, width, dexpr, op["out"]))
                w("  } else VVV1P0_1(%s, %s, %s, %s, w%d);" % (args, coup, mass, width, op["out"]))
            else:
                w("  if constexpr (StableForms<R>::value) %s(%s, %s, %s, %s, %s, w%d);" % (
                    r, args, coup, mass, width, dexpr, op["out"]))
                w("  else %s(%s, %s, %s, %s, w%d);" % (slot_r, args, coup, mass, width, op["out"]))
        else:
            f1, f2, vs = op["args"][0], op["args"][1], op["args"][2:]
            jn = "jamp%d" % (op["jset"] + 1) if op.get("jset") else "jamp"
            if len(vs) == 1:
                (r0, i0), = op["factors"]
                w("  { const cx<R> amp0 = %s(w%d, w%d, w%d, %s);" % (op["routine"], f1, f2, vs[0], coup))
                w("    const cx<R> amp = %s;" % _scaled("amp0", r0, i0))
            else:
                w("  { W<R> vs;")
                w("    #pragma unroll")
                w("    for (int q = 2; q < 6; ++q) vs[q] = %s;" % " + ".join(
                    _scaled("w%d[q]" % v, r, i) for v, (r, i) in zip(vs, op["factors"])))
                w("    const cx<R> amp = %s(w%d, w%d, vs, %s);" % (op["routine"], f1, f2, coup))
            for (c, re, im) in op["norm"]:
                _jamp_update(w, "    ", jn, c, re, im, "amp")
            w("  }")
            # keep the machine scheduler from hoisting later currents above this point
            w("  __builtin_amdgcn_sched_barrier(0);")


def zero_helicity_pairs(desc):
    """Pairs of external legs (quark, antiquark) of the massless quark lines.  A vector
    coupling conserves chirality along a massless line, so every helicity row in which the
    incoming quark and antiquark carry the same helicity label vanishes identically (the
    helicity filtering MadGraph does at run time, decided here by the process)."""
    kinds = desc.get("kinds", [])
    if "quark" in kinds and "antiquark" in kinds:
        return [(kinds.index("quark"), kinds.index("antiquark"))]
    return []


def same_helicity_pairs(desc):
    """Pairs of external legs (incoming quark, outgoing quark) of a massless quark line
    that scatters: chirality conservation along the line makes every row in which the two
    carry different helicity labels vanish."""
    kinds = desc.get("kinds", [])
    for a, b in (("quark", "quark_out"), ("antiquark", "antiquark_out")):
        if a in kinds and b in kinds:
            return [(kinds.index(a), kinds.index(b))]
    return []


def emit_hip(desc, uname=None):
    """HIP device code: constants + a templated amplitude program for one helicity.

    Every wavefunction is an SSA value (``W<R> wN``) so the compiler can keep the live
    set in VGPRs; the jamp accumulators are an array indexed by constants.  ``uname`` is
    the name the kernels register under (the process name when they are its default).
    """
    name = uname or desc["name"]
    n = desc["nexternal"]
    ncol = desc["ncolor"]
    lines = []
    w = lines.append
    w("// Generated by madflow_amd/procgen/emit.py — do not edit.")
    w("// Process: %s  (%d diagrams, %d amplitudes, %d colour flows, %d helicities)" % (
        desc["process"], desc["ndiags"], desc["ngraphs"], ncol, desc["ncomb"]))
    w("#pragma once")
    w("namespace mfx {")
    w("namespace %s {" % name)
    w("constexpr int NEXTERNAL = %d;" % n)
    w("constexpr int NCOMB = %d;" % desc["ncomb"])
    w("constexpr int NCOLOR = %d;" % ncol)
    w("constexpr int NDIAGS = %d;" % desc["ndiags"])
    w("constexpr int NGRAPHS = %d;" % desc["ngraphs"])
    w("constexpr double DENOMINATOR = %d.0;" % desc["denominator"])
    w("constexpr int NINV = %d;  // invariants 2 k_i.k_j, i < j" % (n * (n - 1) // 2))
    hels = desc["helicities"]
    # helicities packed as bits: bit l set <=> leg l has helicity +1
    packed = []
    for row in hels:
        b = 0
        for l, h in enumerate(row):
            if h > 0:
                b |= 1 << l
        packed.append(b)
    w("__constant__ unsigned short HEL_BITS[NCOMB] = {%s};" % ", ".join(str(b) for b in packed))
    # colour matrix: numerators / row denominator (all rows share one denominator)
    dens = set(desc["color_denom"])
    assert len(dens) == 1
    den = dens.pop()
    w("constexpr double CF_DEN = %d.0;" % den)
    rows = []
    for r in desc["color_num"]:
        rows.append("{" + ", ".join("%d" % x for x in r) + "}")
    # compile-time constants: every colour-sum coefficient becomes an instruction operand
    # (inline constant or literal).  A __constant__ array would be loaded, hoisted out of
    # the helicity loop and kept live (spilled) across it.
    w("constexpr float CF_NUM_F[NCOLOR][NCOLOR] = {%s};" % ",\n    ".join(rows))
    w("constexpr double CF_NUM_D[NCOLOR][NCOLOR] = {%s};" % ",\n    ".join(rows))
    w("")
    # amplitude program
    w("// Fills jamp[NCOLOR] for one helicity row; p = momenta (E,px,py,pz) per leg.")
    w("template <class R>")
    w("__device__ __forceinline__ void amplitudes(const R (&p)[NEXTERNAL][4], const R (&inv)[NINV], unsigned hb,")
    w("                                           const int* helv, const Params<R>& prm, cx<R> (&jamp)[NCOLOR]) {")
    w("  const R ZERO = R(0);")
    for l, e in enumerate(desc["externals"]):
        mass = "ZERO" if e["mass"] == "ZERO" else "prm.mdl_MT"
        hel = "(helv ? R(helv[%d]) : (((hb >> %d) & 1u) ? R(1) : R(-1)))" % (l, l)
        if e["routine"] == "vxxxxx":
            w("  W<R> w%d; vxxxxx(p[%d], %s, %s, R(%d), prm.sqh, w%d);" % (l, l, mass, hel, e["sign"], l))
        else:
            w("  W<R> w%d; %s(p[%d], %s, %s, R(%d), w%d);" % (l, e["routine"], l, mass, hel, e["sign"], l))
    w("  #pragma unroll")
    w("  for (int c = 0; c < NCOLOR; ++c) jamp[c] = cx<R>(R(0), R(0));")
    legs_of = {l: [l] for l in range(n)}
    for op in desc["ops"]:
        if op["kind"] == "wf":
            legs_of[op["out"]] = op["legs"]
    _emit_ops(w, device_program(desc), desc, legs_of)
    w("}")
    w("")
    dual = desc["ncomb"] == 1 << n
    k = MULTI_K.get(desc["name"], 1) if dual else 0
    if dual:
        legs, progk = multi_program(desc, k)
        for key in list(legs_of):
            for r in range(1, 1 << k):
                legs_of[key + DUAL_OFF * r] = legs_of[key]
        nrow = 1 << k
        fname = "amplitudes%d" % nrow
        jnames = ["jamp"] + ["jamp%d" % (c + 1) for c in range(1, nrow)]
        w("// %d helicity rows at once: hb with the legs %s at helicity -1 / +1 (row c: bit b of c" % (
            nrow, ", ".join(str(X) for X in legs)))
        w("// is the helicity of the b-th of them).  Currents that do not depend on a leg are")
        w("// shared by the rows that differ only in that leg's helicity.")
        w("// J0 .. J%d: arrays cx<R>[NCOLOR] or accessors whose elements support =, += and -=." % (nrow - 1))
        w("template <class R, %s>" % ", ".join("class J%d" % c for c in range(nrow)))
        w("__device__ __forceinline__ void %s(const R (&p)[NEXTERNAL][4], const R (&inv)[NINV], unsigned hb," % fname)
        w("    const Params<R>& prm, %s) {" % ", ".join("J%d& %s" % (c, j) for c, j in enumerate(jnames)))
        w("  const R ZERO = R(0);")
        for l, e in enumerate(desc["externals"]):
            mass = "ZERO" if e["mass"] == "ZERO" else "prm.mdl_MT"
            if l in legs:
                b = legs.index(l)
                idents = [(l, "R(-1)"), (l + DUAL_OFF * (1 << b), "R(1)")]
            else:
                idents = [(l, "(((hb >> %d) & 1u) ? R(1) : R(-1))" % l)]
            for ident, hel in idents:
                if e["routine"] == "vxxxxx":
                    w("  W<R> w%d; vxxxxx(p[%d], %s, %s, R(%d), prm.sqh, w%d);" % (ident, l, mass, hel, e["sign"], ident))
                else:
                    w("  W<R> w%d; %s(p[%d], %s, %s, R(%d), w%d);" % (ident, e["routine"], l, mass, hel, e["sign"],
                                                                     ident))
        w("  #pragma unroll")
        w("  for (int c = 0; c < NCOLOR; ++c) {")
        for j in jnames:
            w("    %s[c] = cx<R>(R(0), R(0));" % j)
        w("  }")
        _emit_ops(w, progk, desc, legs_of)
        w("}")
        w("")
    w("struct Traits {")
    w("  static constexpr int NEXTERNAL = %d, NCOMB = %d, NCOLOR = %d, NDIAGS = %d, NGRAPHS = %d, NINV = %d;" % (
        n, desc["ncomb"], ncol, desc["ndiags"], desc["ngraphs"], n * (n - 1) // 2))
    signs = ", ".join("%d" % (-1 if l < desc["ninitial"] else 1) for l in range(n))
    w("  static __device__ __forceinline__ double leg_sign(int l) { const double s[%d] = {%s}; return s[l]; }" % (
        n, signs))
    w("  static constexpr double DENOMINATOR = %d.0, CF_DEN = %d.0;" % (desc["denominator"], den))
    w("  static constexpr const char* PROC_NAME = \"%s\";" % name)
    w("  template <class R> static __device__ __forceinline__ void amplitudes(")
    w("      const R (&p)[NEXTERNAL][4], const R (&inv)[NINV], unsigned hb, const int* helv, const Params<R>& prm,")
    w("      cx<R> (&jamp)[NCOLOR]) {")
    w("    %s::amplitudes(p, inv, hb, helv, prm, jamp);" % name)
    w("  }")
    w("  static constexpr int MULTI_K = %d;   // rows per pass: 2^MULTI_K (0: one row)" % k)
    if dual:
        w("  static constexpr int MULTI_LEG[%d] = {%s};" % (k, ", ".join(str(X) for X in legs)))
        w("  template <class R, %s> static __device__ __forceinline__ void %s(" % (
            ", ".join("class J%d" % c for c in range(nrow)), fname))
        w("      const R (&p)[NEXTERNAL][4], const R (&inv)[NINV], unsigned hb, const Params<R>& prm,")
        w("      %s) {" % ", ".join("J%d& %s" % (c, j) for c, j in enumerate(jnames)))
        w("    %s::%s(p, inv, hb, prm, %s);" % (name, fname, ", ".join(jnames)))
        w("  }")
    w("  static __device__ __forceinline__ unsigned hel_bits(int h) { return HEL_BITS[h]; }")
    pairs = zero_helicity_pairs(desc)
    same = same_helicity_pairs(desc)
    if dual:
        assert not any(X in pq for X in legs for pq in pairs + same), "a paired leg decides a vanishing row"
    if pairs or same:
        w("  // rows that vanish identically: a massless quark line conserves chirality, so its")
        w("  // two external legs need opposite helicity labels when it annihilates (quark and")
        w("  // antiquark) and equal labels when it scatters (bit l of hb: leg l at helicity +1)")
        w("  static constexpr bool FILTERED = true;")
        w("  static __device__ __forceinline__ bool live(unsigned hb) {")
        terms = ["(((hb >> %d) ^ (hb >> %d)) & 1u)" % pq for pq in pairs]
        terms += ["!(((hb >> %d) ^ (hb >> %d)) & 1u)" % pq for pq in same]
        w("    return %s;" % " && ".join(terms))
        w("  }")
    else:
        w("  static constexpr bool FILTERED = false;")
        w("  static __device__ __forceinline__ bool live(unsigned) { return true; }")
    w("  static __device__ __forceinline__ float cf(float, int i, int j) { return CF_NUM_F[i][j]; }")
    w("  static __device__ __forceinline__ double cf(double, int i, int j) { return CF_NUM_D[i][j]; }")
    w("};")
    w("}  // namespace %s" % name)
    w("}  // namespace mfx")
    return "\n".join(lines) + "\n"


def emit_hip_unit(uname):
    """Per-process translation unit: instantiates the fused kernels for one process."""
    return ("// Generated by madflow_amd/procgen/emit.py — do not edit.\n"
            "#include \"../mfx_smatrix.h\"\n"
            "#include \"%s.inc\"\n"
            "MFX_INSTANTIATE_PROCESS(%s)\n") % (uname, uname)


# ------------------------------------------------------ colour-ordered process emission

# Kernel families of each process; the first is the default, registered under the process
# name (what the drop-in module Matrix_1_<name> runs), the others under <name>_<family>.
# "fused": one straight-line Feynman-diagram program per helicity row (mfx_smatrix.h);
# "ordered": colour-ordered Berends-Giele currents per ordering and an MFMA colour sum
# (mfx_ordered.h).  For g g > t t~ g g the colour-ordered kernels are the faster form in
# both precisions; for g g > t t~ g g g the n! jamps of the fused form do not fit the
# register file next to a 1,890-amplitude program.
KERNELS = {"gg_ttx": ("fused",), "gg_ttxg": ("fused",), "gg_ttxgg": ("ordered", "fused"),
           "gg_ttxggg": ("ordered",), "uux_ttx": ("fused",), "uux_ttxg": ("fused",), "uux_ttxgg": ("fused",),
           "ug_ttxu": ("fused",), "uxg_ttxux": ("fused",)}
# helicity legs paired per pass of the fused kernel (2^k rows per pass), default 1
MULTI_K = {}


def unit_name(name, family):
    """Registry name of a process's kernels of one family."""
    return name if KERNELS[name][0] == family else "%s_%s" % (name, family)


def emit_ordered_inc(desc, proc_name):
    """Traits of the colour-ordered kernels: gluon legs, the permutation table of the
    colour basis, helicity rows, the integer colour matrix."""
    n = desc["nexternal"]
    gluons = [l for l, k in enumerate(desc["kinds"]) if k == "gluon"]
    ng = len(gluons)
    ns = _cname(proc_name)
    orderings = [tuple(o) for o in desc["orderings"]]
    slot = {leg: i for i, leg in enumerate(gluons)}
    perm = [[slot[leg] for leg in o] for o in orderings]
    packed = []
    for row in desc["helicities"]:
        b = 0
        for l, h in enumerate(row):
            if h > 0:
                b |= 1 << l
        packed.append(b)
    dens = set(desc["color_denom"])
    assert len(dens) == 1
    den = dens.pop()
    lines = []
    w = lines.append
    w("// Generated by madflow_amd/procgen/emit.py — do not edit.")
    w("// Process: %s through the colour-ordered kernels (%d colour orderings, %d helicities)" % (
        desc["process"], len(orderings), desc["ncomb"]))
    w("#pragma once")
    w("namespace mfx {")
    w("namespace %s {" % proc_name)
    w("__constant__ unsigned short HEL_BITS[%d] = {%s};" % (len(packed), ", ".join(str(b) for b in packed)))
    w("__constant__ unsigned char PERM[%d][%d] = {%s};" % (
        len(perm), ng, ",\n    ".join("{" + ", ".join(str(x) for x in r) + "}" for r in perm)))
    w("__constant__ float CF_NUM_F[%d][%d] = {%s};" % (
        len(orderings), len(orderings),
        ",\n    ".join("{" + ", ".join("%d" % x for x in r) + "}" for r in desc["color_num"])))
    w("struct Traits {")
    w("  static constexpr int NG = %d, NEXTERNAL = %d, NCOMB = %d, NCOLOR = %d, NDIAGS = %d, NGRAPHS = %d;" % (
        ng, n, desc["ncomb"], len(orderings), desc["ndiags"], desc["ngraphs"]))
    w("  static constexpr int TOP_LEG = 2, ATOP_LEG = 3;")
    w("  static constexpr double DENOMINATOR = %d.0, CF_DEN = %d.0;" % (desc["denominator"], den))
    w("  static constexpr const char* PROC_NAME = \"%s\";" % proc_name)
    signs = ", ".join("%d" % (-1 if l < desc["ninitial"] else 1) for l in range(n))
    w("  static __device__ __forceinline__ double leg_sign(int l) { const double s[%d] = {%s}; return s[l]; }" % (
        n, signs))
    sel = " : ".join("s == %d ? %d" % (i, leg) for i, leg in enumerate(gluons[:-1])) + " : %d" % gluons[-1]
    w("  static constexpr int gluon_leg(int s) { return %s; }" % sel)
    w("  // index of the invariant 2 k_x.k_y, x != y among gluon slots 0..%d and the top (slot %d)" % (ng - 1, ng))
    w("  static __device__ __forceinline__ int pair_index(int x, int y) {")
    w("    const int a = x < y ? x : y, b = x < y ? y : x;")
    w("    return a * %d - a * (a + 1) / 2 + (b - a - 1);" % (ng + 1))
    w("  }")
    w("  static __device__ __forceinline__ int perm(int c, int k) { return PERM[c][k]; }")
    w("  static __device__ __forceinline__ unsigned hel_bits(int h) { return HEL_BITS[h]; }")
    w("  static __device__ __forceinline__ float cf(float, int i, int j) { return CF_NUM_F[i][j]; }")
    w("  static __device__ __forceinline__ double cf(double, int i, int j) { return double(CF_NUM_F[i][j]); }")
    w("  template <class R>")
    w("  static __device__ __forceinline__ R hel(unsigned hb, const int* helv, int l) {")
    w("    return helv ? R(helv[l]) : (((hb >> l) & 1u) ? R(1) : R(-1));")
    w("  }")
    w("  template <class R>")
    w("  static __device__ __forceinline__ void fermions(const R (&p)[NEXTERNAL][4], unsigned hb, const int* helv,")
    w("                                                  const Params<R>& prm, W<R>& top, W<R>& atop) {")
    w("    oxxxxx(p[2], prm.mdl_MT, hel<R>(hb, helv, 2), R(1), top);")
    w("    ixxxxx(p[3], prm.mdl_MT, hel<R>(hb, helv, 3), R(-1), atop);")
    w("  }")
    w("  template <class R>")
    w("  static __device__ __forceinline__ void gluon(const R (&p)[NEXTERNAL][4], int s, unsigned hb, const int* helv,")
    w("                                               const Params<R>& prm, W<R>& g) {")
    w("    const int leg = gluon_leg(s);")
    w("    vxxxxx(p[leg], R(0), hel<R>(hb, helv, leg), leg < 2 ? R(-1) : R(1), prm.sqh, g);")
    w("  }")
    w("};")
    w("}  // namespace %s" % proc_name)
    w("}  // namespace mfx")
    return "\n".join(lines) + "\n"


def emit_ordered_unit(proc_name):
    return ("// Generated by madflow_amd/procgen/emit.py — do not edit.\n"
            "#include \"../mfx_ordered.h\"\n"
            "#include \"%s.ordered.inc\"\n"
            "MFX_INSTANTIATE_ORDERED(%s)\n") % (proc_name, proc_name)


# ---------------------------------------------------------- python module emission

MODULE_TEMPLATE = '''"""Drop-in process module for {process} (generated by madflow_amd/procgen/emit.py).

Keeps the interface of the PyOut exporter's ``matrix_<proc>.py``
(``madgraph_plugin/template_files/matrix_method_python.inc:36-138``): the class
``{class_name}`` with ``nexternal, ndiags, ncomb, helicities, denominator,
initial_states, mirror_initial_states`` and ``smatrix(all_ps, {params})``, plus
``get_model_param``.  The per-event work runs in HIP kernels behind the ``mfx_smatrix``
C entry point; this file only validates and forwards buffers.
{kernel_text}
"""

from madflow_amd.matrix import MatrixBase, model_for
from madflow_amd.parameters import Model

PROCESS_NAME = "{name}"


def get_model_param(model=None, param_card_path=None):
    """Return the Model (constants {consts}, alpha_s-dependent couplings {funcs})."""
    return model_for(PROCESS_NAME, param_card_path)


class {class_name}(MatrixBase):
    process_name = PROCESS_NAME
    nexternal = {nexternal}
    ndiags = {ndiags}
    ngraphs = {ngraphs}
    ncomb = {ncomb}
    ncolor = {ncolor}
    initial_states = [{initial_states}]
    mirror_initial_states = {mirror}
    helicities = {helicities}
    denominator = {denominator}
    params = {param_list}
    kernels = {kernels}

    def smatrix(self, all_ps, {params}):
        """Sum over all helicities of matrix(), divided by the denominator: f64[nev]."""
        return self._smatrix(all_ps, ({params},))

    def matrix(self, all_ps, hel, {params}):
        """Colour-summed |M|^2 of one helicity row ``hel`` (length nexternal): f64[nev]."""
        return self._matrix(all_ps, hel, ({params},))

    def __str__(self):
        return "{name}"
'''


def emit_module(desc):
    params = ", ".join(desc["params"])
    consts = [p for p in desc["params"] if not p.startswith("GC_")]
    funcs = [p for p in desc["params"] if p.startswith("GC_")]
    return MODULE_TEMPLATE.format(
        process=desc["process"], class_name=desc["class_name"], name=desc["name"], params=params,
        consts=consts, funcs=funcs, nexternal=desc["nexternal"], ndiags=desc["ndiags"],
        ngraphs=desc["ngraphs"], ncomb=desc["ncomb"], ncolor=desc["ncolor"],
        initial_states=",".join("(%d, %d)" % tuple(s) for s in desc["initial_states"]),
        mirror=desc["mirror_initial_states"], helicities=json.dumps(desc["helicities"]),
        denominator=desc["denominator"], param_list=json.dumps(desc["params"]),
        kernels=repr(tuple(KERNELS[desc["name"]])), kernel_text=_kernel_text(desc["name"]))


_FAMILY_TEXT = {"fused": "the fused Feynman-diagram kernel", "ordered": "the colour-ordered kernels"}


def _kernel_text(name):
    fam = KERNELS[name]
    t = "\nKernels: %s." % _FAMILY_TEXT[fam[0]]
    for f in fam[1:]:
        t += "\n``%s(kernels=%r)`` runs %s instead." % ("Matrix_1_" + name, f, _FAMILY_TEXT[f])
    return t


def emitted_files(desc):
    """{relative path: text} of every file generated for one process description."""
    name = desc["name"]
    out = {os.path.join("processes", name + ".json"): json.dumps(desc, separators=(",", ":")),
           os.path.join("processes", "matrix_1_%s.py" % name): emit_module(desc)}
    for fam in KERNELS[name]:
        u = unit_name(name, fam)
        if fam == "fused":
            out[os.path.join("csrc", "procs", u + ".inc")] = emit_hip(desc, u)
            out[os.path.join("csrc", "procs", u + ".hip")] = emit_hip_unit(u)
        else:
            out[os.path.join("csrc", "procs", u + ".ordered.inc")] = emit_ordered_inc(desc, u)
            out[os.path.join("csrc", "procs", u + ".hip")] = emit_ordered_unit(u)
    return out


def descriptions():
    """Process descriptions of every generated process, in emission order."""
    return [gg_ttx_ng(ng) for ng in NGLUONS] + [qq_ttx_ng(ng) for ng in QQ_NGLUONS] + [qg_ttxq(), qg_ttxq(anti=True)]


def main(argv=None):
    os.makedirs(PROC_DIR, exist_ok=True)
    os.makedirs(INC_DIR, exist_ok=True)
    for desc in descriptions():
        for rel, text in emitted_files(desc).items():
            with open(os.path.join(PKG, rel), "w") as f:
                f.write(text)
        print("emitted", desc["name"], desc["ndiags"], "diagrams", desc["ngraphs"], "amplitudes, kernels",
              ", ".join(unit_name(desc["name"], f) + " (" + f + ")" for f in KERNELS[desc["name"]]))


if __name__ == "__main__":
    main(sys.argv[1:])
