"""CPU checks of the C ABI library (loads, exports every declared symbol, rejects bad
arguments without touching a GPU) and of the process generator."""

import ctypes
import json
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
    assert names[:2] == ["gg_ttx", "gg_ttxg"]
    assert set(names) >= {"gg_ttx", "gg_ttxg", "gg_ttxgg", "gg_ttxgg_ordered", "gg_ttxggg"}
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
    # the fused processes need no workspace; the colour-ordered ones say how much
    assert lib.mfx_smatrix_workspace_size(ttgg, 1000, 32, 0) == 0
    t5 = lib.mfx_process_index(b"gg_ttxggg")
    assert lib.mfx_smatrix_workspace_size(t5, 1000, 32, 0) == 1008 * 128 * 120 * 8   # whole 16-event tiles
    assert lib.mfx_smatrix_workspace_size(t5, 1000, 64, 1) == 1008 * 120 * 16
    assert lib.mfx_smatrix_workspace_size(t5, 1024, 64, 1) == 1024 * 120 * 16
    # a missing workspace is reported before anything touches the (fake, aligned) pointers
    fake = ctypes.c_void_p(1 << 20)
    assert lib.mfx_smatrix(t5, fake, 10, vp(mw), vp(gc), None, fake, 32, 1, None, 0, None) == -8
    assert b"workspace" in lib.mfx_last_error()


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
    for gl, d0 in (([0, 1, 4], 9), ([0, 1, 4, 5], 54)):
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


def test_trace_sum_closed_forms():
    from fractions import Fraction
    assert colour.trace_sum([0, 0]) == Fraction(4)              # C_F N
    assert colour.trace_sum([0, 1, 0, 1]) == Fraction(-2, 3)    # (C_F - N/2) C_F N
    assert colour.trace_sum([0, 1, 1, 0]) == Fraction(16, 3)


def test_emitted_descriptions_are_current():
    """The committed process descriptions equal what the generator produces now."""
    for ng in (0, 1, 2):
        d = gg_ttx_ng(ng)
        with open(os.path.join(ROOT, "madflow_amd", "processes", d["name"] + ".json")) as f:
            committed = json.load(f)
        assert json.loads(json.dumps(d)) == committed


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


@pytest.mark.parametrize("name", ["gg_ttx", "gg_ttxg", "gg_ttxgg", "gg_ttxgg_ordered", "gg_ttxggg"])
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
