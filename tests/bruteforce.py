"""Independent brute-force |M|^2 for g g > t t~ + n g (test infrastructure, numpy).

Shares nothing with the generator or the oracle's ALOHA code except the external
wavefunctions (pinned separately by the Dirac-equation / polarisation tests):

* colour is explicit: Gell-Mann generators T = lambda/2, f^{abc} = -2i Tr([T^a,T^b]T^c),
  every external colour index is a tensor axis, and |M|^2 is summed over all of them;
* Lorentz algebra is explicit: 4x4 Dirac matrices in the chiral (HELAS) basis;
* the amplitude is built by colour-dressed Berends-Giele recursion over subsets of the
  gluons (all Feynman diagrams at once, no diagram list, no colour basis, no jamps);
* Feynman rules are the SM UFO ones (vertex = coupling x colour x Lorentz):
  q-q-g  GC_11 T^a gamma^mu;  g-g-g  GC_10 f^{abc} VVV1;  g-g-g-g  GC_12 sum_k c_k VVVV_k;
  propagators -i g_{mu nu}/p^2 and i (p-slash + m)/(p^2 - m^2 + i m Gamma).
  Momenta of the vertex legs are the outgoing momenta of the attached subsets (the
  slot convention of the HELAS wavefunctions).
"""

import itertools

import numpy as np

from oracle import oracle as O

# --------------------------------------------------------------------- colour
_L = np.zeros((8, 3, 3), dtype=complex)
_L[0][0, 1] = _L[0][1, 0] = 1
_L[1][0, 1] = -1j
_L[1][1, 0] = 1j
_L[2][0, 0] = 1
_L[2][1, 1] = -1
_L[3][0, 2] = _L[3][2, 0] = 1
_L[4][0, 2] = -1j
_L[4][2, 0] = 1j
_L[5][1, 2] = _L[5][2, 1] = 1
_L[6][1, 2] = -1j
_L[6][2, 1] = 1j
_L[7][0, 0] = _L[7][1, 1] = 1 / np.sqrt(3)
_L[7][2, 2] = -2 / np.sqrt(3)
T = _L / 2
F = np.zeros((8, 8, 8))
for a, b, c in itertools.product(range(8), repeat=3):
    comm = T[a] @ T[b] - T[b] @ T[a]
    F[a, b, c] = (-2j * np.trace(comm @ T[c])).real

# f(-1,1,2) f(3,4,-1), f(-1,1,3) f(2,4,-1), f(-1,1,4) f(2,3,-1) with legs (1,2,3,4)
C4 = [np.einsum("eab,cde->abcd", F, F),
      np.einsum("eac,bde->abcd", F, F),
      np.einsum("ead,bce->abcd", F, F)]

# ---------------------------------------------------------------------- Dirac
_s = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]], complex)]
_I2 = np.eye(2)
_Z2 = np.zeros((2, 2))
G = [np.block([[_Z2, _I2], [_I2, _Z2]])] + [np.block([[_Z2, s], [-s, _Z2]]) for s in _s]
METRIC = np.diag([1.0, -1.0, -1.0, -1.0])


def slash(v):
    """gamma^mu v_mu for an upper-index 4-vector v (complex allowed)."""
    vl = METRIC @ v
    return sum(G[m] * vl[m] for m in range(4))


def mdot(a, b):
    return a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3]


# ------------------------------------------------------------------ recursion
def amplitude_tensor(p, hel, mt, wt, gs, sqh):
    """Full colour tensor M[a_g1..a_gn, i_t, j_tbar] for one event and helicity row.

    p: (nexternal, 4) momenta (legs 0,1 incoming gluons, 2 top, 3 anti-top, 4.. gluons)."""
    n = p.shape[0]
    gl = [0, 1] + list(range(4, n))
    ng = len(gl)
    GC10, GC11, GC12 = -gs, 1j * gs, 1j * gs * gs
    ref = sqh != np.sqrt(0.5)
    eps, K = {}, {}
    for i, leg in enumerate(gl):
        w = O.wavefunction("v", p[leg:leg + 1], 0.0, hel[leg], -1 if leg < 2 else 1, ref_compat=ref)[0]
        eps[i] = w[2:6]
        K[i] = (-1 if leg < 2 else 1) * p[leg]
    ubar = O.wavefunction("o", p[2:3], mt, hel[2], 1)[0][2:6]    # row spinor of the outgoing top
    vcol = O.wavefunction("i", p[3:4], mt, hel[3], -1)[0][2:6]   # column spinor of the outgoing anti-top
    pt = p[2]

    # gluon currents J[S]: array (4, 8, 8**|S|); colour axes of S in increasing order
    J = {}
    for i in range(ng):
        arr = np.zeros((4, 8, 8), complex)
        for a in range(8):
            arr[:, a, a] = eps[i]
        J[(i,)] = arr

    def KS(S):
        return sum(K[i] for i in S)

    def merge(parts, arrs):
        """Combine the colour axes of several children into the sorted order of their union."""
        S = tuple(sorted(sum(parts, ())))
        return S

    def reorder(tensor, child_sets, S):
        # tensor has axes (..., sets concatenated in child order); reorder to sorted S
        order = sum(child_sets, ())
        perm = [order.index(x) for x in S]
        lead = tensor.ndim - len(order)
        return np.transpose(tensor, list(range(lead)) + [lead + q for q in perm])

    for size in range(2, ng + 1):
        for S in itertools.combinations(range(ng), size):
            KSv = KS(S)
            acc = np.zeros((4, 8) + (8,) * size, complex)
            first = S[0]
            rest = S[1:]
            # two children: S1 contains the first element
            for r in range(0, len(rest)):
                for comb in itertools.combinations(rest, r):
                    S1 = (first,) + comb
                    S2 = tuple(x for x in rest if x not in comb)
                    A = J[S1].reshape((4, 8) + (8,) * len(S1))
                    B = J[S2].reshape((4, 8) + (8,) * len(S2))
                    P1 = -KSv
                    P2 = KS(S1)
                    P3 = KS(S2)
                    # L^mu = e2^mu (P1.e3) - e2^mu (P2.e3) - e3^mu (P1.e2) + e3^mu (P3.e2) + (e2.e3)(P2-P3)^mu
                    Am = np.tensordot(METRIC, A, axes=(1, 0))
                    Bm = np.tensordot(METRIC, B, axes=(1, 0))
                    e3P = np.tensordot(P1 - P2, Bm, axes=(0, 0))       # (8, cols2)
                    e2P = np.tensordot(P3 - P1, Am, axes=(0, 0))       # (8, cols1)
                    d23 = np.tensordot(A, Bm, axes=(0, 0))             # (8,cols1,8,cols2)
                    n1, n2 = len(S1), len(S2)
                    ax1 = "".join("ABCDEFG"[q] for q in range(n1))
                    ax2 = "".join("HIJKLMN"[q] for q in range(n2))
                    L = (np.einsum("mb%s,c%s->mb%sc%s" % (ax1, ax2, ax1, ax2), A, e3P)
                         + np.einsum("mc%s,b%s->mb%sc%s" % (ax2, ax1, ax1, ax2), B, e2P)
                         + np.einsum("b%sc%s,m->mb%sc%s" % (ax1, ax2, ax1, ax2), d23, P2 - P3))
                    term = GC10 * np.einsum("abc,mb%sc%s->ma%s%s" % (ax1, ax2, ax1, ax2), F, L)
                    acc += reorder(term, [S1, S2], S)
            # three children (four-gluon vertex)
            if size >= 3:
                for part in _partitions3(S):
                    S1, S2, S3 = part
                    A = J[S1].reshape((4, 8) + (8,) * len(S1))
                    B = J[S2].reshape((4, 8) + (8,) * len(S2))
                    Cc = J[S3].reshape((4, 8) + (8,) * len(S3))
                    n1, n2, n3 = len(S1), len(S2), len(S3)
                    a1 = "".join("ABCDEFG"[q] for q in range(n1))
                    a2 = "".join("HIJKLMN"[q] for q in range(n2))
                    a3 = "OPQRSTU"[:n3]
                    Bm = np.tensordot(METRIC, B, axes=(1, 0))
                    Cm = np.tensordot(METRIC, Cc, axes=(1, 0))
                    d23 = np.tensordot(A, Bm, axes=(0, 0))    # (b,a1,c,a2)
                    d24 = np.tensordot(A, Cm, axes=(0, 0))    # (b,a1,d,a3)
                    d34 = np.tensordot(B, Cm, axes=(0, 0))    # (c,a2,d,a3)
                    out = "mb%sc%sd%s" % (a1, a2, a3)
                    t_c_d23 = np.einsum("md%s,b%sc%s->%s" % (a3, a1, a2, out), Cc, d23)   # e4 (e2.e3)
                    t_b_d24 = np.einsum("mc%s,b%sd%s->%s" % (a2, a1, a3, out), B, d24)    # e3 (e2.e4)
                    t_a_d34 = np.einsum("mb%s,c%sd%s->%s" % (a1, a2, a3, out), A, d34)    # e2 (e3.e4)
                    Ls = [t_c_d23 - t_b_d24,      # VVVV1: g14 g23 - g13 g24
                          t_c_d23 - t_a_d34,      # VVVV3: g14 g23 - g12 g34
                          t_b_d24 - t_a_d34]      # VVVV4: g13 g24 - g12 g34
                    for k in range(3):
                        term = GC12 * np.einsum("abcd,%s->ma%s%s%s" % (out, a1, a2, a3), C4[k], Ls[k])
                        acc += reorder(term, [S1, S2, S3], S)
            if size < ng or True:
                k2 = mdot(KSv, KSv)
                acc = acc * (-1j / k2)
            J[S] = acc.reshape(4, 8, -1)

    # quark line from the top: U[S] has axes (spinor 4, open colour 3, top colour 3, 8**|S|)
    U = {(): np.einsum("b,kl->bkl", ubar, np.eye(3)).reshape(4, 3, 3, 1)}

    def vertex_right(Urow, Jc, S0, Sc):
        """Urow (4,3,3,cols0) x GC11 T^c gamma.J -> (4,3,3,cols0,colsc)."""
        Jm = Jc.reshape(4, 8, -1)
        # gamma^mu J_mu: (4x4) per colour/col
        gJ = np.einsum("mxy,m...->xy...", np.array([G[m] * METRIC[m, m] for m in range(4)]), Jm)  # (4,4,8,colsc)
        tmp = np.einsum("bkl...,bcaZ->ckla...Z", Urow, gJ) if False else None
        out = np.einsum("bklA,byaZ->yklAaZ", Urow, gJ)            # (4,3,3,cols0,8,colsc)
        out = np.einsum("yklAaZ,akm->ymlAZ", out, T)              # colour flows k -> m
        return GC11 * out

    for size in range(1, ng + 1):
        for S in itertools.combinations(range(ng), size):
            acc = None
            for r in range(1, size + 1):
                for Sl in itertools.combinations(S, r):    # last attached cluster
                    S0 = tuple(x for x in S if x not in Sl)
                    v = vertex_right(U[S0], J[Sl], S0, Sl)      # (4,3,3,cols0,colsl)
                    v = v.reshape(4, 3, 3, *((8,) * (len(S0) + len(Sl))))
                    v = reorder(v, [S0, Sl], S).reshape(4, 3, 3, -1)
                    acc = v if acc is None else acc + v
            if size < ng:
                P = pt + KS(S)
                prop = 1j * (slash(P.astype(complex)) + mt * np.eye(4)) / (mdot(P, P) - mt * mt + 1j * mt * wt)
                acc = np.einsum("bklA,bc->cklA", acc, prop)
                U[S] = acc
            else:
                full = acc
    # the last vertex was already applied for size == ng; close with v(tbar)
    # recompute closing: M = sum_{Sl} U[S0] GC11 T gamma.J v
    M = None
    Sall = tuple(range(ng))
    for r in range(1, ng + 1):
        for Sl in itertools.combinations(Sall, r):
            S0 = tuple(x for x in Sall if x not in Sl)
            v = vertex_right(U[S0], J[Sl], S0, Sl)
            v = np.einsum("yklAZ,y->klAZ", v, vcol)
            v = v.reshape(3, 3, *((8,) * ng))
            v = reorder(v, [S0, Sl], Sall)
            M = v if M is None else M + v
    return M   # (i_open=tbar colour j, i_top, a_1..a_ng)


def _partitions3(S):
    S = list(S)
    first = S[0]
    rest = S[1:]
    out = []
    for r in range(0, len(rest)):
        for comb in itertools.combinations(rest, r):
            S1 = (first,) + comb
            R = [x for x in rest if x not in comb]
            if len(R) < 2:
                continue
            f2 = R[0]
            rr = R[1:]
            for r2 in range(0, len(rr)):
                for comb2 in itertools.combinations(rr, r2):
                    S2 = (f2,) + comb2
                    S3 = tuple(x for x in rr if x not in comb2)
                    if S3:
                        out.append((S1, S2, S3))
    return out


def matrix_hel(p, hel, mt=173.0, wt=1.4915000200271606, alpha_s=0.118, ref_compat=False):
    gs = 2 * np.sqrt(np.pi * alpha_s)
    sqh = (np.float32(np.sqrt(0.5)).item()) if ref_compat else np.sqrt(0.5)
    M = amplitude_tensor(p, hel, mt, wt, gs, sqh)
    return float(np.sum(np.abs(M) ** 2))


# ------------------------------------------------ explicit colour assignments
# The tensor recursion above keeps every external colour as an open axis, so its largest
# current grows as 8**n_gluons (32,768 complex per Lorentz component at five gluons, and
# its four-gluon vertex materialises 8**8).  The form below fixes every external gluon
# colour instead and batches the assignments: a current is (B, 4, 8) with B the number of
# assignments, the quark line (B, 4, 3, 3).  It shares the vertex algebra and the external
# wavefunctions with the recursion above, not its tensor bookkeeping, and the two are
# compared with each other at three and four gluons.
#
# Schur's lemma gives the colour sum with the first gluon's colour held at one value: the
# amplitude is an invariant tensor, so rho(a, a') = sum over every other colour of
# M(a, ...) conj(M(a', ...)) commutes with the (irreducible) adjoint representation and is
# c * delta(a, a'); the full sum is 8 * rho(a0, a0) for any a0.  fix_first=False sums all
# 8**n assignments instead (the tests use both and compare them).

# colour tensors of the three four-gluon Lorentz structures, grouped by structure:
# VVVV1 = t_c_d23 - t_b_d24, VVVV3 = t_c_d23 - t_a_d34, VVVV4 = t_b_d24 - t_a_d34
_K_CD23 = C4[0] + C4[1]
_K_BD24 = -C4[0] + C4[2]
_K_AD34 = -C4[1] - C4[2]
_GM = np.array([G[m] * METRIC[m, m] for m in range(4)])    # gamma_mu with a lowered index


def _colour_assignments(ng, fix_first):
    free = ng - 1 if fix_first else ng
    idx = np.indices((8,) * free).reshape(free, -1).T
    if fix_first:
        idx = np.concatenate([np.zeros((idx.shape[0], 1), dtype=idx.dtype), idx], axis=1)
    return idx                                             # (B, ng)


def amplitude_batched(p, hel, mt, wt, gs, sqh, fix_first=True):
    """Amplitudes M[B, j_tbar, i_top] for every assignment of external gluon colours.

    Returns (M, colours, weight): sum_B |M|^2 * weight is the colour-summed |M|^2."""
    n = p.shape[0]
    gl = [0, 1] + list(range(4, n))
    ng = len(gl)
    GC10, GC11, GC12 = -gs, 1j * gs, 1j * gs * gs
    ref = sqh != np.sqrt(0.5)
    cols = _colour_assignments(ng, fix_first)
    B = cols.shape[0]
    K = {}
    J = {}
    for i, leg in enumerate(gl):
        w = O.wavefunction("v", p[leg:leg + 1], 0.0, hel[leg], -1 if leg < 2 else 1, ref_compat=ref)[0]
        K[i] = (-1 if leg < 2 else 1) * p[leg]
        cur = np.zeros((B, 4, 8), complex)
        cur[np.arange(B), :, cols[:, i]] = w[2:6]
        J[(i,)] = cur
    ubar = O.wavefunction("o", p[2:3], mt, hel[2], 1)[0][2:6]
    vcol = O.wavefunction("i", p[3:4], mt, hel[3], -1)[0][2:6]

    def KS(S):
        return sum(K[i] for i in S)

    def lower(X):                                        # metric on the Lorentz axis 1
        return X * np.diag(METRIC)[None, :, None]

    for size in range(2, ng + 1):
        for S in itertools.combinations(range(ng), size):
            KSv = KS(S)
            acc = np.zeros((B, 4, 8), complex)
            first, rest = S[0], S[1:]
            for r in range(0, len(rest)):
                for comb in itertools.combinations(rest, r):
                    S1 = (first,) + comb
                    S2 = tuple(x for x in rest if x not in comb)
                    A, Bc = J[S1], J[S2]
                    P1, P2, P3 = -KSv, KS(S1), KS(S2)
                    Am, Bm = lower(A), lower(Bc)
                    e3P = np.einsum("m,Bmc->Bc", P1 - P2, Bm)
                    e2P = np.einsum("m,Bmb->Bb", P3 - P1, Am)
                    d23 = np.einsum("Bmb,Bmc->Bbc", A, Bm)
                    t1 = np.einsum("Bmb,Bab->Bma", A, np.einsum("abc,Bc->Bab", F, e3P))
                    t2 = np.einsum("Bmc,Bac->Bma", Bc, np.einsum("abc,Bb->Bac", F, e2P))
                    t3 = np.einsum("m,Ba->Bma", P2 - P3, np.einsum("abc,Bbc->Ba", F, d23))
                    acc += GC10 * (t1 + t2 + t3)
            if size >= 3:
                for S1, S2, S3 in _partitions3(S):
                    A, Bc, Cc = J[S1], J[S2], J[S3]
                    d12 = np.einsum("Bmb,Bmc->Bbc", A, lower(Bc)).reshape(B, 64)
                    d13 = np.einsum("Bmb,Bmd->Bbd", A, lower(Cc)).reshape(B, 64)
                    d23 = np.einsum("Bmc,Bmd->Bcd", Bc, lower(Cc)).reshape(B, 64)
                    z1 = (d12 @ _K_CD23.transpose(1, 2, 0, 3).reshape(64, 64)).reshape(B, 8, 8)   # (a, d)
                    z2 = (d13 @ _K_BD24.transpose(1, 3, 0, 2).reshape(64, 64)).reshape(B, 8, 8)   # (a, c)
                    z3 = (d23 @ _K_AD34.transpose(2, 3, 0, 1).reshape(64, 64)).reshape(B, 8, 8)   # (a, b)
                    acc += GC12 * (np.einsum("Bad,Bmd->Bma", z1, Cc) + np.einsum("Bac,Bmc->Bma", z2, Bc)
                                   + np.einsum("Bab,Bmb->Bma", z3, A))
            J[S] = acc * (-1j / mdot(KSv, KSv))

    GMt = _GM.reshape(4, 16)                             # (mu, b y)

    def vertex_right(Urow, Jc):
        """Urow (B, 4, 3, 3) x GC11 T^a gamma.J^a -> (B, 4, 3, 3):
        out[y, m, l] = sum_{b, k} U[b, k, l] X[b, y, k, m] with
        X[b, y, k, m] = sum_{mu, a} gamma_mu[b, y] J^{mu a} T^a[k, m]."""
        JT = (Jc.reshape(B * 4, 8) @ T.reshape(8, 9)).reshape(B, 4, 9)              # (B, mu, k m)
        X = (JT.transpose(0, 2, 1).reshape(B * 9, 4) @ GMt).reshape(B, 3, 3, 4, 4)  # (B, k, m, b, y)
        return GC11 * np.einsum("Bbkl,Bkmby->Byml", Urow, X)

    U = {(): np.broadcast_to(np.einsum("b,kl->bkl", ubar, np.eye(3)), (B, 4, 3, 3)).copy()}
    Sall = tuple(range(ng))
    M = np.zeros((B, 3, 3), complex)
    for size in range(1, ng + 1):
        for S in itertools.combinations(range(ng), size):
            acc = np.zeros((B, 4, 3, 3), complex)
            for r in range(1, size + 1):
                for Sl in itertools.combinations(S, r):
                    S0 = tuple(x for x in S if x not in Sl)
                    acc += vertex_right(U[S0], J[Sl])
            if S == Sall:
                M = np.einsum("Byml,y->Bml", acc, vcol)
            else:
                P = pt_plus = p[2] + KS(S)
                prop = 1j * (slash(P.astype(complex)) + mt * np.eye(4)) / (mdot(P, P) - mt * mt + 1j * mt * wt)
                U[S] = np.einsum("Bbkl,bc->Bckl", acc, prop)
    return M, cols, (8.0 if fix_first else 1.0)


def matrix_hel_batched(p, hel, mt=173.0, wt=1.4915000200271606, alpha_s=0.118, ref_compat=False,
                       fix_first=True):
    """Colour-summed |M|^2 of one helicity row from explicit colour assignments."""
    gs = 2 * np.sqrt(np.pi * alpha_s)
    sqh = (np.float32(np.sqrt(0.5)).item()) if ref_compat else np.sqrt(0.5)
    M, _, weight = amplitude_batched(p, hel, mt, wt, gs, sqh, fix_first=fix_first)
    return float(np.sum(np.abs(M) ** 2) * weight)


# ----------------------------------------------------- q q~ > t t~ (+ one gluon)
# Two fermion lines: every Feynman diagram written out with explicit Dirac matrices and
# explicit colour tensors (Gell-Mann T, the f^{abc} above), independent of the generator's
# diagram enumeration, its Fierz expansion and its two-string colour basis.
#   light line: row spinor v-bar of the incoming antiquark (oxxxxx), column spinor u of
#   the incoming quark (ixxxxx); colour string T[c_qbar, c_q].  Top line as above.

def _vvv(e2, e3, P2, P3):
    """Lorentz part of the three-gluon current of children e2 (momentum P2) and e3 (P3),
    the structure of the recursion above: parent momentum P1 = -(P2 + P3)."""
    P1 = -(P2 + P3)
    return e2 * mdot(P1 - P2, e3) + e3 * mdot(P3 - P1, e2) + mdot(e2, e3) * (P2 - P3)


def qq_amplitude_tensor(p, hel, mt, wt, gs, sqh):
    """M[c_qbar, c_q, c_t, c_tbar(, a)] of u u~ > t t~ (g): legs 0 u, 1 u~, 2 t, 3 t~, 4 g."""
    n = p.shape[0]
    ng = n - 4
    assert ng in (0, 1)
    GC10, GC11 = -gs, 1j * gs
    ref = sqh != np.sqrt(0.5)
    u = O.wavefunction("i", p[0:1], 0.0, hel[0], 1)[0][2:6]
    vb = O.wavefunction("o", p[1:2], 0.0, hel[1], -1)[0][2:6]
    ub = O.wavefunction("o", p[2:3], mt, hel[2], 1)[0][2:6]
    v = O.wavefunction("i", p[3:4], mt, hel[3], -1)[0][2:6]
    K = [-p[0], -p[1], p[2], p[3]]

    def fprop(P, m, w):
        return 1j * (slash(P.astype(complex)) + m * np.eye(4)) / (mdot(P, P) - m * m + 1j * m * w)

    KL, KT = K[0] + K[1], K[2] + K[3]
    # each line's current as a gluon with its propagator (colour T^b on the line, not included)
    Lprop = (-1j / mdot(KL, KL)) * GC11 * np.array([vb @ G[m] @ u for m in range(4)])
    Tprop = (-1j / mdot(KT, KT)) * GC11 * np.array([ub @ G[m] @ v for m in range(4)])
    TT = np.einsum("bij,bkl->ijkl", T, T)                   # sum_b T^b[qb, q] T^b[t, tb]
    if ng == 0:
        return GC11 * (ub @ slash(Lprop) @ v) * TT
    eps = O.wavefunction("v", p[4:5], 0.0, hel[4], 1, ref_compat=ref)[0][2:6]
    kg = p[4]
    se, sL, sT = slash(eps), slash(Lprop), slash(Tprop)
    A1 = GC11 ** 2 * ub @ se @ fprop(K[2] + kg, mt, wt) @ sL @ v           # gluon off the top
    A2 = GC11 ** 2 * ub @ sL @ fprop(K[2] + KL, mt, wt) @ se @ v           # off the anti-top
    A3 = GC11 ** 2 * vb @ se @ fprop(K[1] + kg, 0.0, 0.0) @ sT @ u        # off the antiquark
    A4 = GC11 ** 2 * vb @ sT @ fprop(K[1] + KT, 0.0, 0.0) @ se @ u        # off the quark
    J = GC10 * _vvv(Lprop, eps, KL, kg) * (-1j / mdot(KL + kg, KL + kg))  # the s-channel gluon
    A5 = GC11 * ub @ slash(J) @ v
    TaTb = np.einsum("aij,bjk->abik", T, T)                # (T^a T^b)[i, k]
    M = (A1 * np.einsum("abik,bqr->qrika", TaTb, T)        # (T^a T^b)[t, tb] T^b[qb, q]
         + A2 * np.einsum("baik,bqr->qrika", TaTb, T)      # (T^b T^a)[t, tb] T^b[qb, q]
         + A3 * np.einsum("abqr,bik->qrika", TaTb, T)      # (T^a T^b)[qb, q] T^b[t, tb]
         + A4 * np.einsum("baqr,bik->qrika", TaTb, T)      # (T^b T^a)[qb, q] T^b[t, tb]
         + A5 * np.einsum("cba,cik,bqr->qrika", F, T, T))  # f^{c b a} T^c[t, tb] T^b[qb, q]
    return M


def qq_matrix_hel(p, hel, mt=173.0, wt=1.4915000200271606, alpha_s=0.118, ref_compat=False):
    """Colour-summed |M|^2 of one helicity row of u u~ > t t~ (g)."""
    gs = 2 * np.sqrt(np.pi * alpha_s)
    sqh = (np.float32(np.sqrt(0.5)).item()) if ref_compat else np.sqrt(0.5)
    return float(np.sum(np.abs(qq_amplitude_tensor(p, hel, mt, wt, gs, sqh)) ** 2))


def qg_amplitude_tensor(p, hel, mt, wt, gs, sqh, anti=False):
    """M[c_row, c_col, c_t, c_tbar, a] of u g > t t~ u (legs 0 u in, 1 g in, 2 t, 3 t~, 4 u out):
    the five diagrams of u u~ > t t~ g written out again with the light line's row spinor
    that of the outgoing quark (oxxxxx, nsf = +1), its column spinor that of the incoming
    quark, and the incoming gluon's polarisation (vxxxxx, nsv = -1); momenta are taken
    outgoing (incoming legs negated).

    anti=True: u~ g > t t~ u~ (legs 0 u~ in, 4 u~ out).  The fermion arrow of the light line
    now runs from the outgoing antiquark (column spinor v, ixxxxx nsf = -1) to the incoming
    one (row spinor v-bar, oxxxxx nsf = -1), and the momentum along the arrow at the row
    end is -p0 instead of p4; the propagators below are written with that momentum."""
    GC10, GC11 = -gs, 1j * gs
    ref = sqh != np.sqrt(0.5)
    if anti:
        u = O.wavefunction("i", p[4:5], 0.0, hel[4], -1)[0][2:6]      # column: v(p4)
        ubo = O.wavefunction("o", p[0:1], 0.0, hel[0], -1)[0][2:6]    # row: v-bar(p0)
    else:
        u = O.wavefunction("i", p[0:1], 0.0, hel[0], 1)[0][2:6]
        ubo = O.wavefunction("o", p[4:5], 0.0, hel[4], 1)[0][2:6]
    ub = O.wavefunction("o", p[2:3], mt, hel[2], 1)[0][2:6]
    v = O.wavefunction("i", p[3:4], mt, hel[3], -1)[0][2:6]
    eps = O.wavefunction("v", p[1:2], 0.0, hel[1], -1, ref_compat=ref)[0][2:6]
    Kt, Ktb, kg = p[2], p[3], -p[1]
    Kqo = -p[0] if anti else p[4]         # momentum along the light line's arrow at its row end

    def fprop(P, m, w):
        return 1j * (slash(P.astype(complex)) + m * np.eye(4)) / (mdot(P, P) - m * m + 1j * m * w)

    KL, KT = -p[0] + p[4], Kt + Ktb
    Lprop = (-1j / mdot(KL, KL)) * GC11 * np.array([ubo @ G[m] @ u for m in range(4)])
    Tprop = (-1j / mdot(KT, KT)) * GC11 * np.array([ub @ G[m] @ v for m in range(4)])
    se, sL, sT = slash(eps), slash(Lprop), slash(Tprop)
    A1 = GC11 ** 2 * ub @ se @ fprop(Kt + kg, mt, wt) @ sL @ v           # gluon on the top
    A2 = GC11 ** 2 * ub @ sL @ fprop(Kt + KL, mt, wt) @ se @ v           # on the anti-top
    A3 = GC11 ** 2 * ubo @ se @ fprop(Kqo + kg, 0.0, 0.0) @ sT @ u       # on the outgoing quark
    A4 = GC11 ** 2 * ubo @ sT @ fprop(Kqo + KT, 0.0, 0.0) @ se @ u       # on the incoming quark
    J = GC10 * _vvv(Lprop, eps, KL, kg) * (-1j / mdot(KL + kg, KL + kg))  # t-channel three-gluon vertex
    A5 = GC11 * ub @ slash(J) @ v
    TaTb = np.einsum("aij,bjk->abik", T, T)
    return (A1 * np.einsum("abik,bqr->qrika", TaTb, T)
            + A2 * np.einsum("baik,bqr->qrika", TaTb, T)
            + A3 * np.einsum("abqr,bik->qrika", TaTb, T)
            + A4 * np.einsum("baqr,bik->qrika", TaTb, T)
            + A5 * np.einsum("cba,cik,bqr->qrika", F, T, T))


def qg_matrix_hel(p, hel, mt=173.0, wt=1.4915000200271606, alpha_s=0.118, ref_compat=False, anti=False):
    """Colour-summed |M|^2 of one helicity row of u g > t t~ u (anti: u~ g > t t~ u~)."""
    gs = 2 * np.sqrt(np.pi * alpha_s)
    sqh = (np.float32(np.sqrt(0.5)).item()) if ref_compat else np.sqrt(0.5)
    return float(np.sum(np.abs(qg_amplitude_tensor(p, hel, mt, wt, gs, sqh, anti)) ** 2))
