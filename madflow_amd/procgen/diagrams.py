"""Tree-level Feynman diagrams and HELAS call programs for g g > t t~ + k g.

MadGraph5 generates the diagrams of a process and the PyOut exporter writes them
as a sequence of HELAS calls (``madgraph_plugin/PyOut_helas_call_writer.py:37-130``)
followed by ``jamp`` lines (``madgraph_plugin/PyOut_exporter.py:334-375``).
MadGraph is not available here, so this module generates the same objects for the
one-fermion-line QCD processes ``g g > t t~ (g ...)``:

* a diagram is an ordered sequence of gluon *clusters* attached along the top line
  from the top (leg 2) to the anti-top (leg 3); every cluster is a rooted tree of
  three-gluon (``VVV1``) and four-gluon (``VVVV1/3/4``) vertices over its gluons;
* off-shell currents are memoised, so a current shared by several diagrams is
  computed once (MadGraph does the same with its wavefunction numbering);
* the amplitude vertex of a diagram is the top-line vertex where the cluster holding
  the highest-numbered gluon attaches.  For g g > t t~ this reproduces the mockup
  call sequence exactly (``python_package/madflow/tests/mockup_debug_me.py:516-528``);
* the colour factor of every amplitude is expanded into the fundamental-string
  basis, which gives the ``jamp`` coefficients (all of them +-1 or +-i), with the
  fermion-flow factor -1 that makes the g g > t t~ jamps identical to the mockup's
  ``jamp = [i*amp0 - amp1, -i*amp0 - amp2]`` (:530).

Diagram counts (3, 16, 123, 1240 for k = 0..3) match MadGraph's.
"""

from itertools import combinations
from .colour import TPoly, f_contract, all_orderings, colour_matrix, line_numerators

# four-gluon vertex: Lorentz structure k pairs with colour structure k
# (SM UFO vertex V_2: colours f(-1,1,2)f(3,4,-1), f(-1,1,3)f(2,4,-1), f(-1,1,4)f(2,3,-1)
#  with lorentz VVVV1, VVVV3, VVVV4 and coupling GC_12 = i gs^2)
V4_NAMES = ("VVVV1", "VVVV3", "VVVV4")


# ---------------------------------------------------------------------------
# trees
# ---------------------------------------------------------------------------

def _set_partitions(items, nblocks):
    """Unordered partitions of ``items`` into exactly ``nblocks`` non-empty blocks."""
    items = list(items)
    if nblocks == 1:
        if items:
            yield [tuple(items)]
        return
    if len(items) < nblocks:
        return
    first, rest = items[0], items[1:]
    # first goes into a block with some subset of rest
    for r in range(0, len(rest) + 1):
        for comb in combinations(rest, r):
            block = (first,) + comb
            remaining = [x for x in rest if x not in comb]
            for sub in _set_partitions(remaining, nblocks - 1):
                yield [block] + sub


def trees(leaves):
    """All rooted trees over ``leaves`` whose internal nodes have 2 (VVV) or 3 (VVVV)
    children.  A leaf is an int; a node is ``('V3', c1, c2)`` or ``('V4', c1, c2, c3)``
    with children ordered by their smallest leaf."""
    leaves = tuple(sorted(leaves))
    if len(leaves) == 1:
        return [leaves[0]]
    out = []
    for nb, tag in ((2, "V3"), (3, "V4")):
        for part in _set_partitions(leaves, nb):
            part = sorted(part, key=min)
            sub = [trees(b) for b in part]
            if nb == 2:
                for a in sub[0]:
                    for b in sub[1]:
                        out.append((tag, a, b))
            else:
                for a in sub[0]:
                    for b in sub[1]:
                        for c in sub[2]:
                            out.append((tag, a, b, c))
    return out


def tree_leaves(t):
    if isinstance(t, int):
        return (t,)
    return tuple(sorted(sum((tree_leaves(c) for c in t[1:]), ())))


def ordered_set_partitions(items):
    """All ordered sequences of non-empty disjoint blocks covering ``items``."""
    items = tuple(items)
    if not items:
        yield []
        return
    for r in range(1, len(items) + 1):
        for comb in combinations(items, r):
            rest = tuple(x for x in items if x not in comb)
            for tail in ordered_set_partitions(rest):
                yield [comb] + tail


def enumerate_diagrams(gluons):
    """Every tree diagram as a list of cluster trees ordered from top to anti-top."""
    diags = []
    for part in ordered_set_partitions(gluons):
        per_block = [trees(b) for b in part]

        def rec(i, acc):
            if i == len(part):
                diags.append(list(acc))
                return
            for t in per_block[i]:
                acc.append(t)
                rec(i + 1, acc)
                acc.pop()

        rec(0, [])
    return diags


# ---------------------------------------------------------------------------
# HELAS program
# ---------------------------------------------------------------------------

class Program:
    """Straight-line HELAS program: external wavefunctions, memoised off-shell
    currents, amplitudes and the jamp expansion."""

    def __init__(self, externals, nexternal):
        self.externals = externals          # list of dicts (kind, leg, mass, nsign)
        self.nexternal = nexternal
        self.ops = []                       # dicts: out, routine, args, coup, mass, width
        self.nwf = nexternal
        self.namps = 0
        self.memo = {}
        self.jamp_terms = []                # per amplitude: {ordering tuple: (re, im)}
        # external legs whose momenta flow through each wavefunction
        self.legs = {l: frozenset([l]) for l in range(nexternal)}

    def _new_wf(self):
        w = self.nwf
        self.nwf += 1
        return w

    def current(self, key, routine, args, coup, mass, width):
        if key in self.memo:
            return self.memo[key]
        w = self._new_wf()
        legs = frozenset().union(*[self.legs[a] for a in args])
        self.legs[w] = legs
        self.ops.append(dict(out=w, kind="wf", routine=routine, args=list(args),
                             coup=coup, mass=mass, width=width, legs=sorted(legs)))
        self.memo[key] = w
        return w

    def amplitude(self, routine, args, coup, tpoly):
        a = self.namps
        self.namps += 1
        self.ops.append(dict(out=a, kind="amp", routine=routine, args=list(args),
                             coup=coup, mass=None, width=None))
        self.jamp_terms.append(tpoly)
        return a


def _gluon_current(prog, tree):
    """Return (wavefunction index, colour polynomial) for an expanded tree.

    ``tree`` here is an *expanded* tree where V4 nodes carry their Lorentz index:
    ('V3', a, b) or ('V4', k, a, b, c)."""
    if isinstance(tree, int):
        return tree, TPoly.gen(tree)
    if tree[0] == "V3":
        wa, pa = _gluon_current(prog, tree[1])
        wb, pb = _gluon_current(prog, tree[2])
        w = prog.current(("G", tree), "VVV1P0_1", (wa, wb), "GC_10", "ZERO", "ZERO")
        return w, f_contract(pa, pb)
    k = tree[1]
    wa, pa = _gluon_current(prog, tree[2])
    wb, pb = _gluon_current(prog, tree[3])
    wc, pc = _gluon_current(prog, tree[4])
    w = prog.current(("G", tree), V4_NAMES[k] + "P0_1", (wa, wb, wc), "GC_12", "ZERO", "ZERO")
    # colour of structure k with leg 1 = parent, legs 2,3,4 = children a,b,c:
    #   k=0: f^{e 1 2} f^{3 4 e} -> -[A,[B,C]];  k=1 -> -[B,[A,C]];  k=2 -> -[C,[A,B]]
    if k == 0:
        poly = pa.commutator(pb.commutator(pc))
    elif k == 1:
        poly = pb.commutator(pa.commutator(pc))
    else:
        poly = pc.commutator(pa.commutator(pb))
    return w, poly.scale(-1, 0)


def _expand_v4(tree):
    """All Lorentz-expanded versions of a tree (each V4 node -> 3 structures)."""
    if isinstance(tree, int):
        return [tree]
    if tree[0] == "V3":
        return [("V3", a, b) for a in _expand_v4(tree[1]) for b in _expand_v4(tree[2])]
    out = []
    for k in range(3):
        for a in _expand_v4(tree[1]):
            for b in _expand_v4(tree[2]):
                for c in _expand_v4(tree[3]):
                    out.append(("V4", k, a, b, c))
    return out


FERMION_FLOW_FACTOR = (-1, 0)


def build_program(gluons, top_leg, atop_leg, externals, nexternal):
    """Generate the HELAS program for g g > t t~ + gluons.

    ``gluons`` are the external gluon leg numbers (0-based), ``top_leg`` is the
    outgoing top (wavefunction from oxxxxx), ``atop_leg`` the outgoing anti-top
    (ixxxxx)."""
    prog = Program(externals, nexternal)
    diags = enumerate_diagrams(gluons)
    # order diagrams: fewer clusters first (as MadGraph lists the s-channel-like
    # graphs first), then by cluster structure so shared currents stay adjacent
    def sort_key(d):
        return (len(d), [tree_leaves(t) for t in d], [repr(t) for t in d])
    diags.sort(key=sort_key)
    last_gluon = max(gluons)
    for d in diags:
        # locate the cluster holding the highest-numbered gluon
        close = next(i for i, t in enumerate(d) if last_gluon in tree_leaves(t))
        for expanded in _product_expand(d):
            cur = [_gluon_current(prog, t) for t in expanded]
            # top side chain: clusters 0 .. close-1 applied to the top wavefunction
            w_o = top_leg
            key = ()
            for i in range(close):
                key = key + (expanded[i],)
                w_o = prog.current(("O", key), "FFV1_1", (w_o, cur[i][0]), "GC_11",
                                   "mdl_MT", "mdl_WT")
            # anti-top side chain: clusters n-1 .. close+1 applied to the anti-top
            w_i = atop_leg
            key = ()
            for i in range(len(expanded) - 1, close, -1):
                key = key + (expanded[i],)
                w_i = prog.current(("I", key), "FFV1_2", (w_i, cur[i][0]), "GC_11",
                                   "mdl_MT", "mdl_WT")
            poly = TPoly.one()
            for _, p in cur:
                poly = poly * p
            poly = poly.scale(*FERMION_FLOW_FACTOR)
            prog.amplitude("FFV1_0", (w_i, w_o, cur[close][0]), "GC_11", poly)
    return prog, len(diags)


def schedule_ops(ops, nexternal, inputs=None):
    """Reorder a HELAS program to keep few wavefunctions live at once.

    Greedy list scheduling over amplitudes: next evaluate the amplitude whose missing
    currents are fewest (ties: the one whose evaluation retires the most live
    currents), computing its missing currents just before it.  Every wavefunction is
    still computed once and before its first use, so any evaluator sees the same
    values; only the order (and hence the register footprint of the fused kernel)
    changes.  For g g > t t~ g g the peak live set drops from 43 to 18 wavefunctions.
    ``inputs``: the ids available before the program (default: the external legs).
    """
    wf_def = {op["out"]: op for op in ops if op["kind"] == "wf"}
    amps = [op for op in ops if op["kind"] == "amp"]
    cons = {}
    for op in ops:
        for a in op["args"]:
            cons[a] = cons.get(a, 0) + 1
    computed = set(range(nexternal)) if inputs is None else set(inputs)
    remaining = dict(cons)

    def deps(w, acc):
        if w in computed or w in acc:
            return
        for a in wf_def[w]["args"]:
            deps(a, acc)
        acc.append(w)

    out = []
    pending = list(amps)
    while pending:
        best = None
        bestkey = None
        for i, amp in enumerate(pending):
            need = []
            for a in amp["args"]:
                deps(a, need)
            rem = {}
            for w in need:
                for a in wf_def[w]["args"]:
                    rem[a] = rem.get(a, remaining[a]) - 1
            for a in amp["args"]:
                rem[a] = rem.get(a, remaining[a]) - 1
            frees = sum(1 for v in rem.values() if v == 0)
            key = (len(need), -frees)
            if bestkey is None or key < bestkey:
                bestkey = key
                best = (i, need)
        i, need = best
        amp = pending.pop(i)
        for w in need:
            op = wf_def[w]
            out.append(op)
            computed.add(w)
            for a in op["args"]:
                remaining[a] -= 1
        out.append(amp)
        for a in amp["args"]:
            remaining[a] -= 1
    return out


def _product_expand(d):
    lists = [_expand_v4(t) for t in d]
    out = [[]]
    for l in lists:
        out = [o + [x] for o in out for x in l]
    return out


# ---------------------------------------------------------------------------
# process description
# ---------------------------------------------------------------------------

def helicity_table(kinds):
    """Helicity combinations in the emitted order (first leg slowest): vectors and
    fermions run (-1, +1), anti-fermions (incoming or outgoing) (+1, -1) — the order of
    the mockup's ``helicities`` (``mockup_debug_me.py:424-442``)."""
    states = []
    for k in kinds:
        if k in ("atop", "antiquark", "antiquark_out"):
            states.append((1, -1))
        else:
            states.append((-1, 1))
    rows = [[]]
    for st in states:
        rows = [r + [h] for r in rows for h in st]
    return rows


def gg_ttx_ng(ng):
    """Process description dict for g g > t t~ + ng gluons."""
    from math import factorial
    nexternal = 4 + ng
    kinds = ["gluon", "gluon", "top", "atop"] + ["gluon"] * ng
    externals = []
    for leg, k in enumerate(kinds):
        if k == "gluon":
            # initial-state bosons carry nsv = -1 (PyOut_helas_call_writer.py:57-61)
            externals.append(dict(leg=leg, routine="vxxxxx", mass="ZERO",
                                  sign=-1 if leg < 2 else +1))
        elif k == "top":
            externals.append(dict(leg=leg, routine="oxxxxx", mass="mdl_MT", sign=+1))
        else:
            externals.append(dict(leg=leg, routine="ixxxxx", mass="mdl_MT", sign=-1))
    gluons = [0, 1] + list(range(4, 4 + ng))
    prog, ndiags = build_program(gluons, 2, 3, externals, nexternal)
    prog.ops = schedule_ops(prog.ops, nexternal)
    orderings = all_orderings(gluons)
    index = {o: i for i, o in enumerate(orderings)}
    jamp = [[] for _ in orderings]
    for a, poly in enumerate(prog.jamp_terms):
        for mono, (re, im) in poly.terms.items():
            jamp[index[mono]].append((a, re, im))
    cf = colour_matrix(orderings)
    denoms, nums = line_numerators(cf)
    suffix = "g" * ng
    name = "gg_ttx" + suffix
    params = ["mdl_MT", "mdl_WT", "GC_10", "GC_11"] + (["GC_12"] if any(
        op["coup"] == "GC_12" for op in prog.ops) else [])
    return dict(
        name=name,
        process="g g > t t~" + " g" * ng,
        class_name="Matrix_1_" + name,
        nexternal=nexternal,
        ninitial=2,
        ndiags=ndiags,
        ngraphs=prog.namps,
        nwavefuncs=prog.nwf,
        ncolor=len(orderings),
        ncomb=2 ** nexternal,
        kinds=kinds,
        externals=externals,
        helicities=helicity_table(kinds),
        denominator=4 * 64 * factorial(ng),
        params=params,
        masses=["ZERO", "ZERO", "mdl_MT", "mdl_MT"] + ["ZERO"] * ng,
        ops=prog.ops,
        jamp=jamp,
        orderings=[list(o) for o in orderings],
        color_denom=denoms,
        color_num=nums,
        initial_states=[[21, 21]],
        mirror_initial_states=False,
    )


# ---------------------------------------------------------------------------
# q q~ > t t~ + gluons: two fermion lines
# ---------------------------------------------------------------------------

LIGHT = -1    # leaf of a bridge tree that stands for the light-quark line's current


def _cluster_sequences(glist):
    """Every ordered sequence of gluon clusters (one tree per block) over ``glist``."""
    import itertools
    out = []
    for part in ordered_set_partitions(glist):
        for combo in itertools.product(*[trees(b) for b in part]):
            out.append(list(combo))
    return out


def enumerate_qq_diagrams(gluons):
    """Tree diagrams of q q~ > t t~ + gluons as (light, pb, bridge, top, pt):

    * ``light``: the gluon clusters on the light-quark line, from the incoming antiquark
      (the end the fermion flow leaves by) to the incoming quark, and ``pb`` the position
      of the bridge vertex among them (clusters light[:pb] sit on the antiquark side);
    * ``bridge``: the gluon tree from the light line to the top line, over LIGHT and the
      gluons it collects through three- and four-gluon vertices;
    * ``top``: the clusters on the top line from the top to the anti-top, and ``pt`` the
      position of the bridge among them.

    Every tree-level QCD diagram has this form: the two lines are joined by exactly one
    gluon tree, and every other gluon hangs off one of the lines in a tree of its own.
    Diagram counts: 1, 5, 36 for 0, 1, 2 gluons."""
    import itertools
    diags = []
    for assign in itertools.product(range(3), repeat=len(gluons)):
        gl = [g for g, a in zip(gluons, assign) if a == 0]
        gb = [g for g, a in zip(gluons, assign) if a == 1]
        gt = [g for g, a in zip(gluons, assign) if a == 2]
        for lseq in _cluster_sequences(gl):
            for pb in range(len(lseq) + 1):
                for btree in trees([LIGHT] + gb):
                    for tseq in _cluster_sequences(gt):
                        for pt in range(len(tseq) + 1):
                            diags.append((lseq, pb, btree, tseq, pt))
    return diags


def _subst_leaf(tree, old, new):
    """An expanded tree with the leaf ``old`` replaced by ``new``."""
    if isinstance(tree, int):
        return new if tree == old else tree
    if tree[0] == "V3":
        return ("V3", _subst_leaf(tree[1], old, new), _subst_leaf(tree[2], old, new))
    return ("V4", tree[1]) + tuple(_subst_leaf(c, old, new) for c in tree[2:])


def _current_with_leaves(prog, tree, leaves):
    """_gluon_current for trees whose leaves may be off-shell currents: ``leaves`` maps
    such a leaf (a tuple tagged "L") to its (wavefunction index, colour polynomial)."""
    if isinstance(tree, tuple) and tree[0] == "L":
        return leaves[tree]
    if isinstance(tree, int):
        return _gluon_current(prog, tree)
    if tree[0] == "V3":
        wa, pa = _current_with_leaves(prog, tree[1], leaves)
        wb, pb = _current_with_leaves(prog, tree[2], leaves)
        w = prog.current(("G", tree), "VVV1P0_1", (wa, wb), "GC_10", "ZERO", "ZERO")
        return w, f_contract(pa, pb)
    k = tree[1]
    (wa, pa), (wb, pb), (wc, pc) = [_current_with_leaves(prog, c, leaves) for c in tree[2:]]
    w = prog.current(("G", tree), V4_NAMES[k] + "P0_1", (wa, wb, wc), "GC_12", "ZERO", "ZERO")
    if k == 0:
        poly = pa.commutator(pb.commutator(pc))
    elif k == 1:
        poly = pb.commutator(pa.commutator(pc))
    else:
        poly = pc.commutator(pa.commutator(pb))
    return w, poly.scale(-1, 0)


BRIDGE_LABEL = "b"


def build_program_qq(gluons, q_leg, qb_leg, top_leg, atop_leg, externals, nexternal):
    """HELAS program of q q~ > t t~ + gluons (incoming quark ``q_leg`` from ixxxxx,
    incoming antiquark ``qb_leg`` from oxxxxx).  Per diagram: the light line's clusters
    are applied to the antiquark (FFV1_1) and to the quark (FFV1_2) up to the bridge
    vertex, which FFV1P0_3 closes into the light current; the bridge tree carries it to the
    top line, whose amplitude closes at the bridge with FFV1_0.  The colour factor is the
    product of the two lines' generator strings with the bridge index contracted by the
    Fierz identity, expanded in the two-string basis of ``colour.fierz_contract``.
    Returns (program, number of diagrams, {amplitude: {basis element: coefficient}})."""
    from .colour import fierz_contract
    prog = Program(externals, nexternal)
    diags = enumerate_qq_diagrams(gluons)
    diags.sort(key=lambda d: (len(d[0]) + len(d[3]), repr(d)))
    colours = []
    for lseq, pb, btree, tseq, pt in diags:
        for lexp in _product_expand(lseq):
            curl = [_gluon_current(prog, t) for t in lexp]
            w_o, key = qb_leg, ()
            for i in range(pb):
                key = key + (lexp[i],)
                w_o = prog.current(("LO", key), "FFV1_1", (w_o, curl[i][0]), "GC_11", "ZERO", "ZERO")
            w_i, key = q_leg, ()
            for i in range(len(lexp) - 1, pb - 1, -1):
                key = key + (lexp[i],)
                w_i = prog.current(("LI", key), "FFV1_2", (w_i, curl[i][0]), "GC_11", "ZERO", "ZERO")
            lkey = ("L", tuple(lexp[:pb]), tuple(lexp[pb:]))
            w_l = prog.current(("LC", lkey), "FFV1P0_3", (w_i, w_o), "GC_11", "ZERO", "ZERO")
            lpoly = TPoly.one()
            for _, p in curl[:pb]:
                lpoly = lpoly * p
            lpoly = lpoly * TPoly.gen(BRIDGE_LABEL)
            for _, p in curl[pb:]:
                lpoly = lpoly * p
            leaves = {lkey: (w_l, TPoly.gen(BRIDGE_LABEL))}
            for bexp in _expand_v4(btree):
                bexp = _subst_leaf(bexp, LIGHT, lkey)
                for texp in _product_expand(tseq):
                    curt = [_gluon_current(prog, t) for t in texp]
                    w_b, bpoly = _current_with_leaves(prog, bexp, leaves)
                    w_top, key = top_leg, ()
                    for i in range(pt):
                        key = key + (texp[i],)
                        w_top = prog.current(("O", key), "FFV1_1", (w_top, curt[i][0]), "GC_11", "mdl_MT", "mdl_WT")
                    w_atop, key = atop_leg, ()
                    for i in range(len(texp) - 1, pt - 1, -1):
                        key = key + (texp[i],)
                        w_atop = prog.current(("I", key), "FFV1_2", (w_atop, curt[i][0]), "GC_11", "mdl_MT",
                                              "mdl_WT")
                    tpoly = TPoly.one()
                    for _, p in curt[:pt]:
                        tpoly = tpoly * p
                    tpoly = tpoly * bpoly
                    for _, p in curt[pt:]:
                        tpoly = tpoly * p
                    prog.amplitude("FFV1_0", (w_atop, w_top, w_b), "GC_11", None)
                    colours.append(fierz_contract(tpoly, lpoly, BRIDGE_LABEL))
    return prog, len(diags), colours


def qg_ttxq(anti=False):
    """Process description dict for u g > t t~ u, the crossing of u u~ > t t~ g (incoming
    gluon, outgoing quark): the same diagrams and colour algebra, with the outgoing quark's
    row spinor (oxxxxx, nsf = +1) in place of the incoming antiquark's and the incoming
    gluon's polarisation (vxxxxx, nsv = -1).  The same matrix element serves c, d and s
    quarks; the mirror process has the gluon in beam 1.

    anti=True gives u~ g > t t~ u~: the light line runs from the outgoing antiquark's
    column spinor (ixxxxx, nsf = -1) to the incoming antiquark's row spinor (oxxxxx,
    nsf = -1); diagrams, bridge and colour strings are those of the quark case with the
    line's two ends exchanged."""
    from .colour import colour_matrix_two_lines
    nexternal = 5
    if anti:
        kinds = ["antiquark", "gluon", "top", "atop", "antiquark_out"]
        light = [dict(leg=0, routine="oxxxxx", mass="ZERO", sign=-1),
                 dict(leg=4, routine="ixxxxx", mass="ZERO", sign=-1)]
        col_leg, row_leg = 4, 0
        name, process, pdg = "uxg_ttxux", "u~ g > t t~ u~", [-2, -4, -1, -3]
    else:
        kinds = ["quark", "gluon", "top", "atop", "quark_out"]
        light = [dict(leg=0, routine="ixxxxx", mass="ZERO", sign=+1),
                 dict(leg=4, routine="oxxxxx", mass="ZERO", sign=+1)]
        col_leg, row_leg = 0, 4
        name, process, pdg = "ug_ttxu", "u g > t t~ u", [2, 4, 1, 3]
    externals = [light[0],
                 dict(leg=1, routine="vxxxxx", mass="ZERO", sign=-1),
                 dict(leg=2, routine="oxxxxx", mass="mdl_MT", sign=+1),
                 dict(leg=3, routine="ixxxxx", mass="mdl_MT", sign=-1),
                 light[1]]
    prog, ndiags, colours = build_program_qq([1], col_leg, row_leg, 2, 3, externals, nexternal)
    prog.ops = schedule_ops(prog.ops, nexternal)
    basis = sorted({key for c in colours for key in c})
    index = {b: i for i, b in enumerate(basis)}
    jamp = [[] for _ in basis]
    for a, col in enumerate(colours):
        for key, (re, im) in sorted(col.items()):
            jamp[index[key]].append((a, _coef_json(re), _coef_json(im)))
    cf = colour_matrix_two_lines(basis)
    denoms, nums = line_numerators(cf)
    used = {op["coup"] for op in prog.ops}
    params = ["mdl_MT", "mdl_WT"] + [c for c in ("GC_10", "GC_11", "GC_12") if c in used]
    return dict(
        name=name,
        process=process,
        class_name="Matrix_1_" + name,
        nexternal=nexternal,
        ninitial=2,
        ndiags=ndiags,
        ngraphs=prog.namps,
        nwavefuncs=prog.nwf,
        ncolor=len(basis),
        ncomb=2 ** nexternal,
        kinds=kinds,
        externals=externals,
        helicities=helicity_table(kinds),
        denominator=4 * 3 * 8,
        params=params,
        masses=["ZERO", "ZERO", "mdl_MT", "mdl_MT", "ZERO"],
        ops=prog.ops,
        jamp=jamp,
        colour_basis=[[p, list(w1), list(w2)] for p, w1, w2 in basis],
        color_denom=denoms,
        color_num=nums,
        initial_states=[[q, 21] for q in pdg],
        mirror_initial_states=True,
    )


def _coef_json(x):
    """A rational coefficient for the process description: an int when integral, else the
    exact fraction as a string ("-1/6")."""
    from fractions import Fraction
    x = Fraction(x)
    return int(x) if x.denominator == 1 else "%d/%d" % (x.numerator, x.denominator)


def qq_ttx_ng(ng):
    """Process description dict for u u~ > t t~ + ng gluons.  With massless quarks and QCD
    vertices only, the same matrix element serves c c~, d d~ and s s~ (``initial_states``),
    and the mirror processes with the antiquark in beam 1 (``mirror_initial_states``)."""
    from math import factorial, lcm
    from fractions import Fraction
    from .colour import colour_matrix_two_lines
    nexternal = 4 + ng
    kinds = ["quark", "antiquark", "top", "atop"] + ["gluon"] * ng
    externals = [dict(leg=0, routine="ixxxxx", mass="ZERO", sign=+1),
                 dict(leg=1, routine="oxxxxx", mass="ZERO", sign=-1),
                 dict(leg=2, routine="oxxxxx", mass="mdl_MT", sign=+1),
                 dict(leg=3, routine="ixxxxx", mass="mdl_MT", sign=-1)]
    externals += [dict(leg=l, routine="vxxxxx", mass="ZERO", sign=+1) for l in range(4, nexternal)]
    gluons = list(range(4, nexternal))
    prog, ndiags, colours = build_program_qq(gluons, 0, 1, 2, 3, externals, nexternal)
    prog.ops = schedule_ops(prog.ops, nexternal)
    basis = sorted({key for c in colours for key in c})
    index = {b: i for i, b in enumerate(basis)}
    jamp = [[] for _ in basis]
    for a, col in enumerate(colours):
        for key, (re, im) in sorted(col.items()):
            jamp[index[key]].append((a, _coef_json(re), _coef_json(im)))
    cf = colour_matrix_two_lines(basis)
    denoms, nums = line_numerators(cf)
    name = "uux_ttx" + "g" * ng
    used = {op["coup"] for op in prog.ops}
    params = ["mdl_MT", "mdl_WT"] + [c for c in ("GC_10", "GC_11", "GC_12") if c in used]
    return dict(
        name=name,
        process="u u~ > t t~" + " g" * ng,
        class_name="Matrix_1_" + name,
        nexternal=nexternal,
        ninitial=2,
        ndiags=ndiags,
        ngraphs=prog.namps,
        nwavefuncs=prog.nwf,
        ncolor=len(basis),
        ncomb=2 ** nexternal,
        kinds=kinds,
        externals=externals,
        helicities=helicity_table(kinds),
        denominator=4 * 9 * factorial(ng),
        params=params,
        masses=["ZERO", "ZERO", "mdl_MT", "mdl_MT"] + ["ZERO"] * ng,
        ops=prog.ops,
        jamp=jamp,
        colour_basis=[[p, list(w1), list(w2)] for p, w1, w2 in basis],
        color_denom=denoms,
        color_num=nums,
        initial_states=[[2, -2], [4, -4], [1, -1], [3, -3]],
        mirror_initial_states=True,
    )
