"""CPU checks of the C ABI library (loads, exports every declared symbol, rejects bad
arguments without touching a GPU) and of the process generator."""

import ctypes
import json
import math
import os

import numpy as np
import pytest

from madflow_amd import _lib
from madflow_amd.procgen import colour
from madflow_amd.procgen.diagrams import gg_ttx_ng, enumerate_diagrams, trees

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    lib = _lib.load()
    missing = [s for s in _lib.exported_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_registry_and_argument_checks():
    lib = _lib.load()
    names = [lib.mfx_process_name(i).decode() for i in range(lib.mfx_process_count())]
    # sorted by (external legs, name): the process names select the default kernels,
    # gg_ttxgg_fused the fused kernel of g g > t t~ g g
    assert names == ["gg_ttx", "uux_ttx", "gg_ttxg", "ug_ttxu", "uux_ttxg", "uxg_ttxux", "gg_ttxgg",
                     "gg_ttxgg_fused", "uux_ttxgg", "gg_ttxggg"]
    info = np.zeros(5, dtype=np.int32)
    ttgg = lib.mfx_process_index(b"gg_ttxgg")
    assert lib.mfx_process_info(ttgg, info.ctypes.data_as(ctypes.c_void_p)) == 0
    assert list(info) == [6, 64, 24, 123, 159]
    assert lib.mfx_process_info(lib.mfx_process_index(b"gg_ttxggg"), info.ctypes.data_as(ctypes.c_void_p)) == 0
    assert list(info) == [7, 128, 120, 1240, 1890]
    assert lib.mfx_process_index(b"nope") < 0
    mw = np.array([173.0, 1.5])
    gc = np.zeros(6)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    # negative nev, bad precision, unknown process: rejected before any launch
    assert lib.mfx_smatrix(0, None, -1, vp(mw), vp(gc), None, None, 64, 1, None, 0, None) < 0
    assert lib.mfx_smatrix(0, None, 0, vp(mw), vp(gc), None, None, 16, 1, None, 0, None) < 0
    assert b"precision" in lib.mfx_last_error()
    assert lib.mfx_smatrix(99, None, 0, vp(mw), vp(gc), None, None, 64, 1, None, 0, None) < 0
    # empty batches are valid no-ops
    assert lib.mfx_smatrix(0, None, 0, vp(mw), vp(gc), None, None, 64, 1, None, 0, None) == 0
    assert lib.mfx_process_info(lib.mfx_process_index(b"gg_ttxgg_fused"), info.ctypes.data_as(ctypes.c_void_p)) == 0
    assert list(info) == [6, 64, 24, 123, 159]
    # the fused kernels need no workspace; the colour-ordered ones say how much
    assert lib.mfx_smatrix_workspace_size(lib.mfx_process_index(b"gg_ttxgg_fused"), 1000, 32, 0) == 0
    assert lib.mfx_smatrix_workspace_size(ttgg, 1000, 32, 0) == 1008 * 64 * 24 * 8
    t5 = lib.mfx_process_index(b"gg_ttxggg")
    assert lib.mfx_smatrix_workspace_size(t5, 1000, 32, 0) == 1008 * 128 * 120 * 8   # whole 16-event tiles
    assert lib.mfx_smatrix_workspace_size(t5, 1000, 64, 1) == 1008 * 120 * 16
    assert lib.mfx_smatrix_workspace_size(t5, 1024, 64, 1) == 1024 * 120 * 16
    # a missing workspace is reported before anything touches the (fake, aligned) pointers
    fake = ctypes.c_void_p(1 << 20)
    assert lib.mfx_smatrix(t5, fake, 10, vp(mw), vp(gc), None, fake, 32, 1, None, 0, None) == -8
    assert b"workspace" in lib.mfx_last_error()


def test_physical_input_guards():
    """Not enough energy for the outgoing masses is rejected on the host before any launch,
    as rambo(check_physical=True) raises (phasespace.py:152-159): mfx_rambo with one
    sqrt(s), mfx_ps_generate for every mapping, and a hadronic shat_min below the mass
    threshold.  The pointers are never dereferenced (fake, aligned addresses)."""
    from madflow_amd.vegas import PsParams, _sig
    lib = _lib.load()
    _sig(lib)
    fake = ctypes.c_void_p(1 << 20)
    m = np.array([173.0, 173.0, 10.0])
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert lib.mfx_rambo(fake, 10, 12, 3, None, ctypes.c_double(356.0), vp(m), fake, fake, None, None) == -9
    assert b"not enough energy" in lib.mfx_last_error()
    assert lib.mfx_rambo(fake, 10, 12, 3, None, ctypes.c_double(300.0), vp(m), fake, fake, None, None) == -9
    assert lib.mfx_rambo(fake, 10, 12, 3, None, ctypes.c_double(500.0), vp(-m), fake, fake, None, None) == -2
    # massless particles need only a positive energy
    assert lib.mfx_rambo(fake, 10, 12, 3, None, ctypes.c_double(0.0), None, fake, fake, None, None) == -9
    for mapping, ndim in ((0, 12), (1, 5), (2, 14)):
        p = PsParams()
        p.n_out, p.ndim, p.nbins, p.mapping = 3, ndim, 50, mapping
        p.sqrts = 300.0
        for k, v in enumerate(m):
            p.masses[k] = v
        p.gev2pb, p.cut_first_leg, p.shat_min = 1.0, 5, 356.0 ** 2
        npass = ctypes.c_int64(0)
        rc = lib.mfx_ps_generate(ctypes.byref(p), fake, 0, 10, 0, None, None, fake, fake, fake, fake, fake, None,
                                 fake, 1 << 30, ctypes.byref(npass), None)
        assert rc == -2 and b"energy" in lib.mfx_last_error(), (mapping, rc)
    p.sqrts = 13000.0
    p.shat_min = 300.0 ** 2          # below the (sum of masses)^2 threshold
    rc = lib.mfx_ps_generate(ctypes.byref(p), fake, 0, 10, 0, None, None, fake, fake, fake, fake, fake, None,
                             fake, 1 << 30, ctypes.byref(npass), None)
    assert rc == -2 and b"shat_min" in lib.mfx_last_error()


def test_register_cut_arguments_and_struct_layout(tmp_path):
    """register_cut's cuts in mfx_ps_params: the ctypes struct has the C header's size and
    field offsets (checked with the host compiler), and malformed cuts are rejected on the
    host before any launch: too many, an unknown variable, an incoming or missing
    particle, and neither bound (the pointers are never dereferenced)."""
    import shutil
    import subprocess
    from madflow_amd.vegas import MAX_CUTS, PsParams, _sig
    if shutil.which("gcc"):
        src = tmp_path / "layout.c"
        src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include <stdint.h>\n'
                       '#include "%s"\nint main(void){printf("%%zu %%zu %%zu", sizeof(mfx_ps_params), '
                       'offsetof(mfx_ps_params, ncuts), offsetof(mfx_ps_params, cuts)); return 0;}\n'
                       % os.path.join(ROOT, "include", "madflow_amd.h"))
        exe = tmp_path / "layout"
        subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
        size, off_n, off_c = map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split())
        assert (size, off_n, off_c) == (ctypes.sizeof(PsParams), PsParams.ncuts.offset, PsParams.cuts.offset)
    lib = _lib.load()
    _sig(lib)
    fake = ctypes.c_void_p(1 << 20)

    def run(**cut):
        p = PsParams()
        p.n_out, p.ndim, p.nbins, p.mapping, p.sqrts = 3, 5, 50, 1, 2000.0
        p.masses[0], p.masses[1] = 173.0, 173.0
        p.gev2pb, p.cut_first_leg = 1.0, 5
        p.ncuts = cut.pop("ncuts", 1)
        k = p.cuts[0]
        k.variable, k.particle, k.use_min, k.use_max = cut.get("variable", 0), cut.get("particle", 3), 1, 0
        k.min_val = 60.0
        if cut.get("no_bounds"):
            k.use_min = 0
        npass = ctypes.c_int64(0)
        return lib.mfx_ps_generate(ctypes.byref(p), fake, 0, 10, 0, None, None, fake, fake, fake, fake, fake, None,
                                   fake, 1 << 30, ctypes.byref(npass), None)
    assert run(ncuts=MAX_CUTS + 1) == -2 and b"ncuts" in lib.mfx_last_error()
    assert run(variable=3) == -2 and b"variable" in lib.mfx_last_error()
    assert run(particle=1) == -2 and b"outgoing" in lib.mfx_last_error()
    assert run(particle=5) == -2 and b"outgoing" in lib.mfx_last_error()
    assert run(no_bounds=True) == -2 and b"bound" in lib.mfx_last_error()


def test_rambo_wrapper_rejects_unphysical():
    """The Python wrapper raises ValueError for one sqrt(s) below the masses before it
    touches a device (the reference's message, phasespace.py:156-158)."""
    import torch
    from madflow_amd.phasespace_gpu import rambo
    x = torch.zeros((4, 12), dtype=torch.float64)
    with pytest.raises(ValueError, match="Not enough energy"):
        rambo(x, 3, 300.0, [173.0, 173.0, 10.0])


def test_product_path_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from madflow_amd.matrix import load_matrix
    m, _ = load_matrix("gg_ttx")
    with pytest.raises(_lib.MfxError):
        m.smatrix(torch.zeros((2, 4, 4), dtype=torch.float64), 173.0, 1.5, -1.2, 1.2j)


@pytest.mark.parametrize("ng,count", [(0, 3), (1, 16), (2, 123), (3, 1240)])
def test_diagram_counts(ng, count):
    """MadGraph's diagram counts for g g > t t~ + ng g."""
    assert len(enumerate_diagrams([0, 1] + list(range(4, 4 + ng)))) == count


def test_tree_counts():
    # rooted trees with 2- or 3-child nodes on k labelled leaves.  k = 4: 15 binary +
    # 6 (ternary root over a pair) + 4 (binary root over a ternary) = 25; k = 5: 105 binary
    # + 10 with two ternary nodes + 105 with one ternary and two binary nodes = 220
    assert [len(trees(range(k))) for k in range(1, 6)] == [1, 1, 4, 25, 220]


def test_colour_matrices():
    cf = colour.colour_matrix(colour.all_orderings([0, 1]))
    den, num = colour.line_numerators(cf)
    assert den == [3, 3] and num == [[16, -2], [-2, 16]]      # mockup_debug_me.py:506-507
    for gl, d0 in (([0, 1, 4], 9), ([0, 1, 4, 5], 54), ([0, 1, 4, 5, 6], 324)):
        cf = colour.colour_matrix(colour.all_orderings(gl))
        den, num = colour.line_numerators(cf)
        n = len(cf)
        assert set(den) == {d0}
        assert all(cf[i][j] == cf[j][i] for i in range(n) for j in range(n))
        # positive definite on a random vector (it is a Gram matrix)
        v = np.random.default_rng(1).normal(size=n)
        assert float(v @ np.array(cf, dtype=float) @ v) > 0
        # trace of CF = N_orderings * C_F^k N (sum over colours of |T^s|^2)
        assert sum(cf[i][i] for i in range(n)) == n * colour.trace_sum(list(gl) + list(reversed(gl)))
        # ... and every diagonal element is N C_F^k with C_F = 4/3 (Casimir of the fundamental)
        from fractions import Fraction
        assert all(cf[i][i] == 3 * Fraction(4, 3) ** len(gl) for i in range(n))


# Gell-Mann generators T^a = lambda^a / 2 (explicit 3x3 matrices, independent of the
# generator's trace algebra)
def _gell_mann():
    L = np.zeros((8, 3, 3), dtype=complex)
    L[0][0, 1] = L[0][1, 0] = 1
    L[1][0, 1], L[1][1, 0] = -1j, 1j
    L[2][0, 0], L[2][1, 1] = 1, -1
    L[3][0, 2] = L[3][2, 0] = 1
    L[4][0, 2], L[4][2, 0] = -1j, 1j
    L[5][1, 2] = L[5][2, 1] = 1
    L[6][1, 2], L[6][2, 1] = -1j, 1j
    L[7][0, 0] = L[7][1, 1] = 1 / np.sqrt(3)
    L[7][2, 2] = -2 / np.sqrt(3)
    return L / 2


@pytest.mark.parametrize("name", ["gg_ttx", "gg_ttxg", "gg_ttxgg", "gg_ttxggg"])
def test_colour_matrix_gell_mann_brute_force(name):
    """CF[s][t] = sum over every gluon colour and the quark colours i, j of
    (T^{a_s1} ... T^{a_sk})_ij conj((T^{a_t1} ... T^{a_tk})_ij), by explicit products of
    Gell-Mann matrices over all 8^k colour assignments (chunked over the first gluon's
    colour), against the integer numerators and row denominator of the committed process
    description, which the colour-ordered and fused kernels are emitted from."""
    with open(os.path.join(ROOT, "madflow_amd", "processes", name + ".json")) as f:
        d = json.load(f)
    T = _gell_mann()
    gl = sorted({leg for o in d["orderings"] for leg in o})
    k = len(gl)
    pos = {leg: i for i, leg in enumerate(gl)}
    nord = len(d["orderings"])
    cf = np.zeros((nord, nord), dtype=complex)
    for a0 in range(8):
        free = np.indices((8,) * (k - 1)).reshape(k - 1, -1).T
        cols = np.concatenate([np.full((len(free), 1), a0), free], axis=1)     # (8^(k-1), k)
        C = np.empty((nord, len(cols) * 9), dtype=complex)
        for s, word in enumerate(d["orderings"]):
            M = T[cols[:, pos[word[0]]]]
            for leg in word[1:]:
                M = M @ T[cols[:, pos[leg]]]
            C[s] = M.reshape(-1)
        cf += C @ C.conj().T
    want = np.array(d["color_num"], dtype=float) / np.array(d["color_denom"], dtype=float)[:, None]
    assert np.max(np.abs(cf.imag)) < 1e-12
    np.testing.assert_allclose(cf.real, want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("name,k_final", [("gg_ttx", 0), ("gg_ttxg", 1), ("gg_ttxgg", 2), ("gg_ttxggg", 3)])
def test_denominators(name, k_final):
    """smatrix averages over the initial state and divides by the symmetry factor of the
    final state: 2 x 2 gluon helicities, 8 x 8 gluon colours, k! for k identical final
    gluons (256, 256, 512, 1536)."""
    import math
    with open(os.path.join(ROOT, "madflow_amd", "processes", name + ".json")) as f:
        d = json.load(f)
    assert d["denominator"] == 4 * 64 * math.factorial(k_final)
    import importlib
    cls = getattr(importlib.import_module("madflow_amd.processes.matrix_1_%s" % name), "Matrix_1_%s" % name)
    assert cls.denominator == d["denominator"]


def test_trace_sum_closed_forms():
    from fractions import Fraction
    assert colour.trace_sum([0, 0]) == Fraction(4)              # C_F N
    assert colour.trace_sum([0, 1, 0, 1]) == Fraction(-2, 3)    # (C_F - N/2) C_F N
    assert colour.trace_sum([0, 1, 1, 0]) == Fraction(16, 3)


@pytest.mark.parametrize("kind,ng", [("gg", 0), ("gg", 1), ("gg", 2), ("gg", 3), ("qq", 0), ("qq", 1), ("qq", 2),
                                     ("qg", 0), ("qxg", 0)])
def test_emitted_files_are_current(kind, ng):
    """Every committed generated file (process description, drop-in module, HIP kernels of
    each kernel family) equals what the generator produces now."""
    from madflow_amd.procgen.emit import emitted_files
    from madflow_amd.procgen.diagrams import qg_ttxq, qq_ttx_ng
    d = gg_ttx_ng(ng) if kind == "gg" else (qq_ttx_ng(ng) if kind == "qq" else qg_ttxq(anti=kind == "qxg"))
    files = emitted_files(d)
    assert os.path.join("processes", d["name"] + ".json") in files
    for rel, text in files.items():
        with open(os.path.join(ROOT, "madflow_amd", rel)) as f:
            committed = f.read()
        if rel.endswith(".json"):
            assert json.loads(committed) == json.loads(text), rel
        else:
            assert committed == text, rel


def test_kernel_families_are_registered():
    """Each drop-in class's kernel families are compiled into the library under the names
    the class resolves them to."""
    import importlib
    lib = _lib.load()
    names = {lib.mfx_process_name(i).decode() for i in range(lib.mfx_process_count())}
    for name in ("gg_ttx", "gg_ttxg", "gg_ttxgg", "gg_ttxggg", "uux_ttx", "uux_ttxg", "uux_ttxgg", "ug_ttxu",
                 "uxg_ttxux"):
        mod = importlib.import_module("madflow_amd.processes.matrix_1_%s" % name)
        cls = getattr(mod, "Matrix_1_%s" % name)
        for fam in cls.kernels:
            assert cls(kernels=fam).registry_name in names, (name, fam)
        with pytest.raises(ValueError):
            cls(kernels="nope")
    from madflow_amd.processes.matrix_1_gg_ttxgg import Matrix_1_gg_ttxgg
    assert Matrix_1_gg_ttxgg().registry_name == "gg_ttxgg"
    assert Matrix_1_gg_ttxgg.kernels[0] == "ordered"


def test_jamp_coefficients_units():
    d = gg_ttx_ng(2)
    for terms in d["jamp"]:
        for _, re, im in terms:
            assert (abs(re), abs(im)) in ((1, 0), (0, 1))


@pytest.mark.parametrize("ng", [0, 1, 2])
def test_dual_row_program(ng):
    """The two-row program of the fused kernels: every single-row operation appears once,
    exactly the currents that depend on the paired leg X get a twin, every amplitude is
    evaluated for both rows, and each operation only reads values defined before it."""
    from madflow_amd.procgen.emit import device_program, dual_program, DUAL_OFF
    d = gg_ttx_ng(ng)
    n = d["nexternal"]
    single = device_program(d)
    X, dual = dual_program(d)
    dep = {X}
    for op in single:
        if op["kind"] == "wf" and any(a in dep for a in op["args"]):
            dep.add(op["out"])
    orig = [op for op in dual if not op.get("jset") and (op["kind"] == "amp" or op["out"] < DUAL_OFF)]
    key = lambda op: (op["kind"], op["out"], tuple(op["args"]))
    assert [key(op) for op in orig] == [key(op) for op in single]
    twins = [op for op in dual if op["kind"] == "wf" and op["out"] >= DUAL_OFF]
    assert sorted(op["out"] - DUAL_OFF for op in twins) == sorted(dep - {X})
    amps = [op for op in dual if op["kind"] == "amp"]
    assert len(amps) == 2 * sum(op["kind"] == "amp" for op in single)
    assert sum(1 for op in amps if op.get("jset")) == len(amps) // 2
    defined = set(range(n)) | {X + DUAL_OFF}
    for op in dual:
        assert all(a in defined for a in op["args"])
        if op.get("jset"):
            # the second row reads the twin of every X-dependent value and nothing else
            assert all((a - DUAL_OFF in dep) if a >= DUAL_OFF else (a not in dep) for a in op["args"])
        if op["kind"] == "wf":
            defined.add(op["out"])


@pytest.mark.parametrize("name", ["gg_ttx", "gg_ttxg", "gg_ttxgg", "gg_ttxggg", "uux_ttx", "uux_ttxg", "uux_ttxgg",
                                  "ug_ttxu", "uxg_ttxux"])
def test_process_module_model_params(name):
    """Every drop-in module's get_model_param() evaluates exactly the parameters its
    smatrix takes, in that order."""
    import importlib
    mod = importlib.import_module("madflow_amd.processes.matrix_1_%s" % name)
    cls = getattr(mod, "Matrix_1_%s" % name)
    model = mod.get_model_param()
    model.freeze_alpha_s(0.118)
    vals = list(model.evaluate())
    assert len(vals) == len(cls.params)


def test_qq_generator_counts_and_colour():
    """u u~ > t t~ + k g: 1, 5 and 36 diagrams and 2, 4 and 12 colour flows for k = 0, 1, 2
    (MadGraph's counts); the colour matrix of the two-string basis for k = 0 is
    [[N^2, N], [N, N^2]] = [[9, 3], [3, 9]], and the single amplitude's jamp coefficients
    are the Fierz factors -1/(2N) (straight) and 1/2 (crossed)."""
    from fractions import Fraction
    from madflow_amd.procgen.diagrams import qq_ttx_ng
    for k, (nd, nc) in enumerate([(1, 2), (5, 4), (36, 12)]):
        d = qq_ttx_ng(k)
        assert (d["ndiags"], d["ncolor"]) == (nd, nc)
        assert d["denominator"] == 36 * math.factorial(k)
    d = qq_ttx_ng(0)
    assert d["color_num"] == [[9, 3], [3, 9]] and d["color_denom"] == [1, 1]
    assert [[(a, Fraction(re), Fraction(im)) for a, re, im in t] for t in d["jamp"]] == [
        [(0, Fraction(-1, 6), 0)], [(0, Fraction(1, 2), 0)]]
    # gluon on each line, and through the three-gluon vertex: every flow is reached
    d = qq_ttx_ng(1)
    assert all(d["jamp"])
    assert sorted(op["routine"] for op in d["ops"] if op["kind"] == "wf") == sorted(
        ["FFV1P0_3", "FFV1P0_3", "FFV1P0_3", "FFV1_1", "FFV1_2", "FFV1_1", "FFV1_2", "VVV1P0_1"])


def test_two_line_colour_matrix_brute_force():
    """The two-string colour matrix of u u~ > t t~ g equals the explicit Gell-Mann sum
    sum over all colour indices of B_s conj(B_t)."""
    from madflow_amd.procgen.diagrams import qq_ttx_ng
    from madflow_amd.procgen.colour import colour_matrix_two_lines
    from tests.bruteforce import T
    d = qq_ttx_ng(1)
    basis = [(p, tuple(w1), tuple(w2)) for p, w1, w2 in d["colour_basis"]]
    eye = np.eye(3)

    def string(word, a):
        m = eye.astype(complex)
        for _ in word:
            m = m @ T[a]
        return m

    def tensor(b):
        # B[c_t, c_tbar, c_qbar, c_q, a]
        pairing, w1, w2 = b
        out = np.zeros((3, 3, 3, 3, 8), complex)
        for a in range(8):
            s1, s2 = string(w1, a), string(w2, a)
            if pairing == 0:
                out[..., a] = np.einsum("ij,kl->ijkl", s1, s2)
            else:   # w1 from the top to the quark, w2 from the antiquark to the anti-top
                out[..., a] = np.einsum("il,kj->ijkl", s1, s2)
        return out

    Bs = [tensor(b) for b in basis]
    brute = np.array([[np.sum(x * np.conj(y)).real for y in Bs] for x in Bs])
    exact = np.array(colour_matrix_two_lines(basis), dtype=float)
    np.testing.assert_allclose(brute, exact, atol=1e-12)


def test_initial_flavours_with_mirror():
    """The hadronic luminosity's flavour pairs follow madflow_exec.py:141-155: the initial
    states, then (mirror_initial_states) the same pairs with the beams swapped."""
    from madflow_amd.vegas import initial_flavours
    from madflow_amd.processes.matrix_1_uux_ttx import Matrix_1_uux_ttx
    from madflow_amd.processes.matrix_1_gg_ttx import Matrix_1_gg_ttx
    assert initial_flavours(Matrix_1_uux_ttx) == ([2, 4, 1, 3, -2, -4, -1, -3], [-2, -4, -1, -3, 2, 4, 1, 3])
    assert initial_flavours(Matrix_1_gg_ttx) is None
