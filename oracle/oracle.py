"""ORACLE — test infrastructure only.

ctypes front end of the CPU oracle (``oracle/liboracle.so``, built by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module, and only as the checker.

Parity status of each oracle function (how it is pinned):

* wavefunctions (ixxxxx/oxxxxx/vxxxxx): Dirac equation, completeness, transversality,
  polarisation sums and the reference's branch structure — tests/test_oracle_helas.py
* ALOHA routines: explicit 4x4 gamma-matrix and metric contractions (numpy), and
  routine-equivalence identities — tests/test_oracle_aloha.py
* g g > t t~ smatrix: the committed mockup (restated by hand in
  ``matrix_gg_ttx_mockup``) and the textbook |M|^2 — tests/test_oracle_matrix.py
* g g > t t~ + 1..3 g: colour-dressed brute force over explicit colour indices,
  gauge (BRST) invariance, Lorentz invariance and Bose symmetry — tests/test_oracle_multijet.py
* RAMBO: the reference's closed-form massless volume and on-shell checks — tests/test_oracle_phasespace.py
* hadronic mapping (ramboflow x1/x2, _boost_to_lab): closed forms of the (x1, x2) weight's
  moments and the lab-frame parton momenta — tests/test_oracle_phasespace.py
* VEGAS: closed-form integrals and grid invariants — tests/test_oracle_vegas.py
* unweighting: the kept count against sum min(1, w/wmax), over-weight and zero-weight
  events — tests/test_oracle_vegas.py
"""

import ctypes
import json
from fractions import Fraction
import os
import subprocess
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
PROC_DIR = os.path.join(os.path.dirname(HERE), "madflow_amd", "processes")

ROUTINES = {
    "FFV1_0": 0, "FFV1_1": 1, "FFV1_2": 2, "VVV1P0_1": 3, "VVVV1P0_1": 4, "VVVV3P0_1": 5,
    "VVVV4P0_1": 6, "FFV1P0_3": 7, "VVV1_0": 8, "VVVV1_0": 9, "VVVV3_0": 10, "VVVV4_0": 11,
}
EXT_KIND = {"vxxxxx": 0, "oxxxxx": 1, "ixxxxx": 2}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def load_process(name):
    with open(os.path.join(PROC_DIR, name + ".json")) as f:
        return json.load(f)


class EncodedProcess:
    """Flat arrays of a process description in the layout oracle_process.hpp reads."""

    def __init__(self, desc):
        self.desc = desc
        real_params = [p for p in desc["params"] if not p.startswith("GC_")]
        self.real_params = real_params
        self.couplings = [p for p in desc["params"] if p.startswith("GC_")]
        ridx = {p: i for i, p in enumerate(real_params)}
        cidx = {p: i for i, p in enumerate(self.couplings)}

        def midx(name):
            return -1 if name in (None, "ZERO") else ridx[name]

        n = desc["nexternal"]
        self.nexternal = n
        self.hel = np.ascontiguousarray(np.array(desc["helicities"], dtype=np.float64))
        ext = []
        for e in desc["externals"]:
            ext += [EXT_KIND[e["routine"]], midx(e["mass"]), int(e["sign"])]
        self.ext = np.array(ext, dtype=np.int32)
        ops = []
        for op in desc["ops"]:
            args = list(op["args"]) + [-1] * (4 - len(op["args"]))
            ops += [ROUTINES[op["routine"]], op["out"]] + args + [
                cidx[op["coup"]], midx(op["mass"]), midx(op["width"])]
        self.ops = np.array(ops, dtype=np.int32)
        self.nops = len(desc["ops"])
        jt, jc = [], []
        for c, terms in enumerate(desc["jamp"]):
            for (a, re, im) in terms:
                # coefficients: ints, or exact fractions written as "p/q"
                jt += [c, a]
                jc += [float(Fraction(re)), float(Fraction(im))]
        self.jterms = np.array(jt, dtype=np.int32)
        self.jcoef = np.array(jc, dtype=np.float64)
        self.njterms = len(jt) // 2
        self.cf_num = np.ascontiguousarray(np.array(desc["color_num"], dtype=np.float64))
        self.cf_den = np.array(desc["color_denom"], dtype=np.float64)

    def args(self):
        d = self.desc
        return [
            ctypes.c_int(d["nexternal"]), ctypes.c_int(d["ncomb"]), _p(self.hel), _p(self.ext),
            ctypes.c_int(self.nops), _p(self.ops), ctypes.c_int(d["nwavefuncs"]), ctypes.c_int(d["ngraphs"]),
            ctypes.c_int(d["ncolor"]), ctypes.c_int(self.njterms), _p(self.jterms), _p(self.jcoef), _p(self.cf_num),
            _p(self.cf_den), ctypes.c_double(d["denominator"]),
        ]


_cache = {}


def encoded(name):
    if name not in _cache:
        _cache[name] = EncodedProcess(load_process(name))
    return _cache[name]


def default_params(alpha_s=0.118, mt=173.0, wt=1.4915000200271606, ref_gs=False):
    """Real parameters (mdl_MT, mdl_WT) and couplings GC_10=-gs, GC_11=i gs, GC_12=i gs^2.

    ref_gs=True uses the mockup's literal gs = 1.2177157847767195
    (mockup_debug_me.py:24-25) instead of 2 sqrt(pi alpha_s)."""
    gs = 1.2177157847767195 if ref_gs else 2.0 * np.sqrt(np.pi * alpha_s)
    rp = np.array([mt, wt], dtype=np.float64)
    coup = {"GC_10": complex(-gs, 0.0), "GC_11": complex(0.0, gs), "GC_12": complex(0.0, gs * gs)}
    return rp, coup


def _coup_array(enc, coup):
    return np.array(sum(([coup[c].real, coup[c].imag] for c in enc.couplings), []), dtype=np.float64)


def smatrix(name, all_ps, rp=None, coup=None, ref_compat=False, nthreads=0):
    enc = encoded(name)
    if rp is None:
        rp, coup = default_params()
    all_ps = np.ascontiguousarray(all_ps, dtype=np.float64)
    nev = all_ps.shape[0]
    out = np.zeros(nev)
    ca = _coup_array(enc, coup)
    rpa = np.ascontiguousarray(rp, dtype=np.float64)
    f = lib().mfo_smatrix
    f.restype = ctypes.c_int
    f(*enc.args(), _p(all_ps), ctypes.c_int64(nev), _p(rpa), _p(ca), ctypes.c_int(len(enc.couplings)),
      ctypes.c_int64(1), _p(out), ctypes.c_int(int(ref_compat)), ctypes.c_int(nthreads))
    return out


def matrix_hel(name, all_ps, helv, rp=None, coup=None, ref_compat=False, want_jamp=False):
    enc = encoded(name)
    if rp is None:
        rp, coup = default_params()
    all_ps = np.ascontiguousarray(all_ps, dtype=np.float64)
    nev = all_ps.shape[0]
    out = np.zeros(nev)
    jamp = np.zeros((nev, enc.desc["ncolor"], 2)) if want_jamp else None
    helv = np.ascontiguousarray(helv, dtype=np.float64)
    ca = _coup_array(enc, coup)
    rpa = np.ascontiguousarray(rp, dtype=np.float64)
    lib().mfo_matrix_hel(*enc.args(), _p(all_ps), ctypes.c_int64(nev), _p(helv), _p(rpa), _p(ca),
                         ctypes.c_int(len(enc.couplings)), _p(out),
                         _p(jamp) if want_jamp else None, ctypes.c_int(int(ref_compat)))
    if want_jamp:
        return out, jamp[..., 0] + 1j * jamp[..., 1]
    return out


def flops_per_me(name, all_ps_one=None):
    """Exact (add/sub, mul, div, sqrt) counts of one smatrix evaluation; FMA counts 2."""
    enc = encoded(name)
    rp, coup = default_params()
    if all_ps_one is None:
        all_ps_one = np.zeros((enc.nexternal, 4))
        all_ps_one[:, 0] = 1.0
    p = np.ascontiguousarray(all_ps_one, dtype=np.float64).reshape(-1)
    counts = np.zeros(4, dtype=np.uint64)
    ca = _coup_array(enc, coup)
    rpa = np.ascontiguousarray(np.concatenate([rp, [0, 0]]), dtype=np.float64)
    lib().mfo_flops(*enc.args(), _p(p), _p(rpa), _p(ca), ctypes.c_int(len(enc.couplings)), _p(counts))
    return dict(add=int(counts[0]), mul=int(counts[1]), div=int(counts[2]), sqrt=int(counts[3]),
                total=int(counts.sum()))


def smatrix_gg_ttx_mockup(all_ps, mt=173.0, wt=1.4915000200271606, gs=1.2177157847767195, ref_compat=True):
    all_ps = np.ascontiguousarray(all_ps, dtype=np.float64)
    nev = all_ps.shape[0]
    out = np.zeros(nev)
    jamp = np.zeros((nev, 2, 2))
    gc10 = np.array([-gs, 0.0])
    gc11 = np.array([0.0, gs])
    lib().mfo_smatrix_gg_ttx_mockup(_p(all_ps), ctypes.c_int64(nev), ctypes.c_double(mt), ctypes.c_double(wt),
                                    _p(gc10), _p(gc11), _p(out), _p(jamp), ctypes.c_int(int(ref_compat)))
    return out, jamp[..., 0] + 1j * jamp[..., 1]


def wavefunction(kind, p, mass, nhel, nsign, ref_compat=False):
    kinds = {"v": 0, "o": 1, "i": 2, "s": 3}
    k = kinds[kind]
    p = np.ascontiguousarray(p, dtype=np.float64)
    nev = p.shape[0]
    n = 3 if k == 3 else 6
    out = np.zeros((nev, n, 2))
    lib().mfo_wavefunction(ctypes.c_int(k), _p(p), ctypes.c_int64(nev), ctypes.c_double(mass),
                           ctypes.c_double(nhel), ctypes.c_double(nsign), _p(out), ctypes.c_int(int(ref_compat)))
    return out[..., 0] + 1j * out[..., 1]


AMP_ROUTINES = {"FFV1_0", "VVV1_0", "VVVV1_0", "VVVV3_0", "VVVV4_0"}


def aloha(routine, inputs, coup, M=0.0, W=0.0):
    """inputs: list of complex arrays (nev, 6)."""
    nev = inputs[0].shape[0]
    ins = np.zeros((len(inputs), nev, 6, 2))
    for i, a in enumerate(inputs):
        ins[i, :, :, 0] = a.real
        ins[i, :, :, 1] = a.imag
    c = np.array([coup.real, coup.imag])
    if routine in AMP_ROUTINES:
        out = np.zeros((nev, 2))
    else:
        out = np.zeros((nev, 6, 2))
    rc = lib().mfo_aloha(ctypes.c_int(ROUTINES[routine]), _p(ins), ctypes.c_int(len(inputs)), ctypes.c_int64(nev),
                         _p(c), ctypes.c_double(M), ctypes.c_double(W), _p(out))
    assert rc == 0
    return out[..., 0] + 1j * out[..., 1]


def rambo(xrand, n, sqrts, masses=None, ref_compat=False):
    xrand = np.ascontiguousarray(xrand, dtype=np.float64)
    nev = xrand.shape[0]
    s = np.ascontiguousarray(np.broadcast_to(np.asarray(sqrts, dtype=np.float64), (nev,)))
    m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
    p = np.zeros((nev, n, 4))
    wt = np.zeros(nev)
    lib().mfo_rambo(_p(xrand), ctypes.c_int64(nev), ctypes.c_int(xrand.shape[1]), ctypes.c_int(n), _p(s),
                    ctypes.c_int(1), _p(m) if m is not None else None, ctypes.c_int(int(ref_compat)), _p(p), _p(wt))
    return p, wt


def philox_uniforms(nev, ndim, seed, iteration, ev0=0):
    out = np.zeros((nev, ndim))
    lib().mfo_philox_uniforms(ctypes.c_uint64(ev0), ctypes.c_int64(nev), ctypes.c_int(ndim), ctypes.c_uint32(seed),
                              ctypes.c_uint32(iteration), _p(out))
    return out


def philox_raw(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().mfo_philox_raw(_p(c), _p(k), _p(out))
    return out


def unweight(w, wmax, seed, iteration, ev0=0):
    """Hit-or-miss keep mask (u8) of the device unweighting (same Philox stream)."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    keep = np.zeros(w.shape[0], dtype=np.uint8)
    lib().mfo_unweight(_p(w), ctypes.c_int64(w.shape[0]), ctypes.c_double(wmax), ctypes.c_uint32(seed),
                       ctypes.c_uint32(iteration), ctypes.c_uint64(ev0), _p(keep))
    return keep


def vegas_map(edges, y):
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    ndim, nb1 = edges.shape
    y = np.ascontiguousarray(y, dtype=np.float64)
    nev = y.shape[0]
    x = np.zeros_like(y)
    jac = np.zeros(nev)
    bins = np.zeros((nev, ndim), dtype=np.int32)
    lib().mfo_vegas_map(_p(edges), ctypes.c_int(ndim), ctypes.c_int(nb1 - 1), _p(y), ctypes.c_int64(nev), _p(x),
                        _p(jac), _p(bins))
    return x, jac, bins


def vegas_refine(edges, hist, alpha=1.5):
    edges = np.array(edges, dtype=np.float64, order="C")
    hist = np.ascontiguousarray(hist, dtype=np.float64)
    ndim, nb1 = edges.shape
    lib().mfo_vegas_refine(_p(edges), _p(hist), ctypes.c_int(ndim), ctypes.c_int(nb1 - 1), ctypes.c_double(alpha))
    return edges


def generate(cfg, edges, ev0, nev, iteration):
    """Oracle of mfx_ps_generate: returns (all_ps, wps, flag) for events ev0..ev0+nev-1."""
    n = cfg["n_out"]
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    masses = np.zeros(8)
    masses[:n] = cfg["masses"][:n]
    all_ps = np.zeros((nev, n + 2, 4))
    wps = np.zeros(nev)
    flag = np.zeros(nev, dtype=np.uint8)
    x12 = np.zeros((nev, 2))
    lib().mfo_generate(ctypes.c_int(n), ctypes.c_int(cfg["ndim"]), ctypes.c_int(cfg["nbins"]),
                       ctypes.c_int(cfg["mapping"]), ctypes.c_double(cfg["sqrts"]), _p(masses),
                       ctypes.c_double(cfg["gev2pb"]), ctypes.c_double(cfg["pt_min"]),
                       ctypes.c_double(cfg["abs_eta_max"]), ctypes.c_double(cfg["dr_min"]),
                       ctypes.c_int(cfg["cut_first_leg"]), ctypes.c_uint32(cfg["seed"]), _p(edges),
                       ctypes.c_uint64(ev0), ctypes.c_int64(nev), ctypes.c_uint32(iteration), _p(all_ps), _p(wps),
                       _p(flag), ctypes.c_double(cfg.get("shat_min", 0.0)), ctypes.c_int(cfg.get("lab_frame", 0)),
                       _p(x12))
    if cfg["mapping"] == 2:
        return all_ps, wps, flag, x12
    return all_ps, wps, flag


def eta(p):
    f = lib().mfo_eta
    f.restype = ctypes.c_double
    return f(_p(np.ascontiguousarray(p, dtype=np.float64)))


def delta_r(a, b):
    f = lib().mfo_delta_r
    f.restype = ctypes.c_double
    return f(_p(np.ascontiguousarray(a, dtype=np.float64)), _p(np.ascontiguousarray(b, dtype=np.float64)))
