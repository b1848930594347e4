"""Exact SU(3) colour algebra for one quark line plus gluons.

The generated process modules store, like the PyOut exporter does
(``madgraph_plugin/PyOut_exporter.py:301-324``), a colour matrix as integer
row numerators ``cf`` plus one integer denominator per row ``denom``, over the
colour basis of fundamental strings ``(T^{a_s1} ... T^{a_sk})_{i_t, j_tbar}``
(one basis element per ordering ``s`` of the gluons).  MadGraph computes those
numbers inside MG5; here they are computed from scratch:

* ``trace_sum(word)`` evaluates ``sum_{adjoint indices} Tr(T^{w1} ... T^{wn})``
  where every label appears exactly twice, with ``T = lambda/2`` and the Fierz
  identities

      sum_a Tr(X T^a Y T^a)   = 1/2 Tr(X) Tr(Y) - 1/(2N) Tr(XY)
      sum_a Tr(X T^a) Tr(Y T^a) = 1/2 Tr(XY)   - 1/(2N) Tr(X) Tr(Y)

  in exact rational arithmetic (N = 3).
* ``colour_matrix(orderings)`` returns ``CF[s][t] = Tr(T^s (T^t)^dagger)``,
  which for g g > t t~ reproduces the mockup's ``[[16,-2],[-2,16]]/3``
  (``python_package/madflow/tests/mockup_debug_me.py:506-507``).

The colour factor of a Feynman diagram is expanded into the same basis with
``TPoly`` (non-commutative polynomials in the external-gluon generators with
Gaussian-integer coefficients): every three-gluon vertex contributes
``sum_b T^b f^{b x y} = -i [T^x, T^y]``.
"""

from fractions import Fraction
from functools import lru_cache
from itertools import permutations

NC = 3


def _canon_trace(tr):
    """Rotate a single trace (tuple of labels) to its lexicographically smallest rotation."""
    if not tr:
        return tr
    n = len(tr)
    return min(tuple(tr[i:] + tr[:i]) for i in range(n))


def _canon(traces):
    """Canonical form of a product of traces, relabelling labels by first appearance."""
    traces = [_canon_trace(t) for t in traces]
    traces.sort()
    # relabel by first appearance, then re-canonicalise once (cheap, stable enough for memo)
    mapping = {}
    for t in traces:
        for x in t:
            if x not in mapping:
                mapping[x] = len(mapping)
    relab = sorted(_canon_trace(tuple(mapping[x] for x in t)) for t in traces)
    return tuple(relab)


@lru_cache(maxsize=None)
def _trace_prod(traces):
    """sum over adjoint labels of prod_k Tr(traces[k]); every label appears twice."""
    # find the first trace that still has a label
    ntr = len(traces)
    for k in range(ntr):
        if traces[k]:
            break
    else:
        return Fraction(NC) ** ntr
    tr = traces[k]
    lab = tr[0]
    # locate the partner of ``lab``
    pos_same = [i for i in range(1, len(tr)) if tr[i] == lab]
    half = Fraction(1, 2)
    inv2n = Fraction(1, 2 * NC)
    others = traces[:k] + traces[k + 1:]
    if pos_same:
        j = pos_same[0]
        # Tr(T^a Y T^a X) with Y = tr[1:j], X = tr[j+1:]
        y = tr[1:j]
        x = tr[j + 1:]
        t1 = _trace_prod(_canon(list(others) + [x, y]))
        t2 = _trace_prod(_canon(list(others) + [x + y]))
        return half * t1 - inv2n * t2
    # partner lives in another trace
    for m, tr2 in enumerate(others):
        if lab in tr2:
            break
    j = tr2.index(lab)
    # Tr(T^a X) Tr(T^a Y): X = tr[1:], Y = tr2 rotated so that label sits first
    x = tr[1:]
    y = tr2[j + 1:] + tr2[:j]
    rest = others[:m] + others[m + 1:]
    t1 = _trace_prod(_canon(list(rest) + [x + y]))
    t2 = _trace_prod(_canon(list(rest) + [x, y]))
    return half * t1 - inv2n * t2


def trace_sum(word):
    """sum over adjoint indices of Tr(T^{w_1} ... T^{w_n}); each label must appear twice."""
    return _trace_prod(_canon([tuple(word)]))


def colour_matrix(orderings):
    """Exact colour matrix CF[s][t] = sum_colours (T^s)_{ij} conj((T^t)_{ij}).

    (T^t)^dagger = T^{t_k} ... T^{t_1}, so CF[s][t] = sum Tr(T^{s_1}..T^{s_k} T^{t_k}..T^{t_1}).
    Relabelling the gluons maps rows onto each other, so only row 0 is traced and the
    others follow from CF[s][t] = CF[0][s^-1 t].
    """
    orderings = [tuple(o) for o in orderings]
    n = len(orderings)
    base = orderings[0]
    row0 = {}
    for t in orderings:
        row0[t] = trace_sum(list(base) + list(reversed(t)))
    cf = [[None] * n for _ in range(n)]
    for i, s in enumerate(orderings):
        # relabel s -> base: label s[k] becomes base[k]
        rel = {s[k]: base[k] for k in range(len(s))}
        for j, t in enumerate(orderings):
            tt = tuple(rel[x] for x in t)
            cf[i][j] = row0[tt]
    return cf


def line_numerators(cf):
    """Split a rational colour matrix into (row denominators, integer numerators)
    the way the emitted ``denom``/``cf`` constants are laid out
    (``madgraph_plugin/PyOut_exporter.py:308-323``): each row gets the least common
    denominator of its entries."""
    from math import lcm
    denoms = []
    nums = []
    for row in cf:
        d = 1
        for x in row:
            d = lcm(d, Fraction(x).denominator)
        denoms.append(d)
        nums.append([int(Fraction(x) * d) for x in row])
    return denoms, nums


class TPoly:
    """Non-commutative polynomial in generators T^{label}; coefficients are
    Gaussian integers stored as (re, im) int pairs."""

    __slots__ = ("terms",)

    def __init__(self, terms=None):
        self.terms = dict(terms or {})

    @staticmethod
    def gen(label):
        return TPoly({(label,): (1, 0)})

    @staticmethod
    def one():
        return TPoly({(): (1, 0)})

    def __mul__(self, other):
        out = {}
        for m1, (a, b) in self.terms.items():
            for m2, (c, d) in other.terms.items():
                m = m1 + m2
                re, im = a * c - b * d, a * d + b * c
                o = out.get(m, (0, 0))
                out[m] = (o[0] + re, o[1] + im)
        return TPoly({k: v for k, v in out.items() if v != (0, 0)})

    def scale(self, re, im):
        out = {}
        for m, (a, b) in self.terms.items():
            out[m] = (a * re - b * im, a * im + b * re)
        return TPoly(out)

    def __add__(self, other):
        out = dict(self.terms)
        for m, (a, b) in other.terms.items():
            o = out.get(m, (0, 0))
            out[m] = (o[0] + a, o[1] + b)
        return TPoly({k: v for k, v in out.items() if v != (0, 0)})

    def __sub__(self, other):
        return self + other.scale(-1, 0)

    def commutator(self, other):
        return self * other - other * self


def f_contract(x, y):
    """sum_b T^b f^{b x y} with x, y already contracted into polynomials: -i [X, Y]."""
    return x.commutator(y).scale(0, -1)


def all_orderings(gluons):
    """Colour basis: every ordering of the gluon labels, lexicographic in the given list.
    For g g > t t~ this gives [(0,1),(1,0)] as in the mockup's jamp order."""
    return [tuple(p) for p in permutations(gluons)]


# ------------------------------------------------------------ two fermion lines

# Colour basis of q q~ > t t~ + gluons: products of two fundamental strings.  A basis
# element is (pairing, w_top, w_light): the string w_top starts at the top's colour index,
# the string w_light at the incoming antiquark's (the fundamental end of the light line).
# pairing 0 ("straight"): w_top ends at the anti-top, w_light at the incoming quark;
# pairing 1 ("crossed"): w_top ends at the incoming quark, w_light at the anti-top.
_ENDS = ({"T": "aT", "L": "aL"}, {"T": "aL", "L": "aT"})


def fierz_contract(top, light, label):
    """sum over the adjoint index ``label`` of (top)_{top, anti-top} (light)_{qbar, q}, with
    ``top`` and ``light`` TPolys whose every monomial holds ``label`` once.  The Fierz
    identity T^b_{ab} T^b_{cd} = 1/2 d_{ad} d_{cb} - 1/(2N) d_{ab} d_{cd} gives
    (X b Y)(Z b W) = 1/2 (ZY)(XW) [crossed] - 1/(2N) (ZW)(XY) [straight].
    Returns {(pairing, w_top, w_light): (Fraction re, Fraction im)}."""
    out = {}
    half, sub = Fraction(1, 2), Fraction(-1, 2 * NC)
    for mt, (ar, ai) in top.terms.items():
        k = mt.index(label)
        Z, W = mt[:k], mt[k + 1:]
        for ml, (br, bi) in light.terms.items():
            j = ml.index(label)
            X, Y = ml[:j], ml[j + 1:]
            cr, ci = ar * br - ai * bi, ar * bi + ai * br
            for key, f in (((1, Z + Y, X + W), half), ((0, Z + W, X + Y), sub)):
                o = out.get(key, (Fraction(0), Fraction(0)))
                out[key] = (o[0] + f * cr, o[1] + f * ci)
    return {k: v for k, v in out.items() if v != (0, 0)}


def colour_matrix_two_lines(basis):
    """CF[s][t] = sum over colours of B_s conj(B_t) for two-string basis elements: the
    open indices close into cycles (string of s, then the reversed string of t that ends
    at the same anti-fundamental index, ...), each cycle one trace."""
    n = len(basis)
    cf = [[None] * n for _ in range(n)]
    for i, s in enumerate(basis):
        ws = {"T": tuple(s[1]), "L": tuple(s[2])}
        es = _ENDS[s[0]]
        for j, t in enumerate(basis):
            wt = {"T": tuple(t[1]), "L": tuple(t[2])}
            back = {v: k for k, v in _ENDS[t[0]].items()}
            traces, seen = [], set()
            for f0 in ("T", "L"):
                if f0 in seen:
                    continue
                word, f = (), f0
                while True:
                    seen.add(f)
                    word += ws[f]
                    f = back[es[f]]
                    word += tuple(reversed(wt[f]))
                    if f == f0:
                        break
                traces.append(word)
            cf[i][j] = _trace_prod(_canon(traces))
    return cf
