"""GPU HELAS wavefunctions and ALOHA routines (the building-block C entry points
mfx_wavefunction / mfx_aloha) against the CPU oracle, on momenta that reach every branch
of the reference's wavefunctions (p at rest, along +z and -z, pT = 0, generic).

Bars: FP64 1e-13 relative to the largest component of the event (the device groups the
vertex algebra differently), FP32 2e-6 relative (5e-6 for the off-shell currents)."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O
from madflow_amd import _lib

MT = 173.0
KINDS = {"v": 0, "o": 1, "i": 2, "s": 3}


def _momenta(mass, n=64, seed=3):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3)) * 300
    special = np.array([[0, 0, 250.0], [0, 0, -250.0], [1e-3, 0, 400.0], [0, 0, 0], [0, 1e-4, -700.0]])
    v = np.concatenate([special, v])
    if mass == 0:
        v = v[np.linalg.norm(v, axis=1) > 0]
    e = np.sqrt((v * v).sum(1) + mass * mass)
    return np.ascontiguousarray(np.concatenate([e[:, None], v], axis=1))


def _device_wf(kind, p, mass, hel, nsf, precision, ref_compat=True):
    lib = _lib.load()
    n = 3 if kind == "s" else 6
    pt = torch.from_numpy(p).cuda()
    out = torch.zeros((p.shape[0], n, 2), dtype=torch.float64, device="cuda")
    _lib.check(lib.mfx_wavefunction(KINDS[kind], ctypes.c_void_p(pt.data_ptr()), p.shape[0], mass, hel, nsf,
                                    ctypes.c_void_p(out.data_ptr()), precision, int(ref_compat),
                                    _lib.stream_ptr(pt.device)))
    o = out.cpu().numpy()
    return o[..., 0] + 1j * o[..., 1]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,mass", [("v", 0.0), ("v", 80.0), ("o", MT), ("o", 0.0), ("i", MT), ("i", 0.0),
                                       ("s", 125.0)])
@pytest.mark.parametrize("precision", [64, 32])
def test_wavefunctions(dev, kind, mass, precision):
    print("case wavefunction", kind, mass, precision, flush=True)
    p = _momenta(mass)
    hels = [-1, 1] + ([0] if kind == "v" and mass > 0 else []) + ([4] if kind == "v" else [])
    for hel in hels:
        for nsf in (-1, 1):
            ref = O.wavefunction(kind, p, mass, hel, nsf, ref_compat=True)
            got = _device_wf(kind, p, mass, hel, nsf, precision)
            scale = np.abs(ref).max(axis=1, keepdims=True)
            rel = np.max(np.abs(got - ref) / scale, axis=1)
            if precision == 32:
                # within an ulp of the -z axis the reference's own FP64 branch choice (|p| + pz
                # rounding to 0 or not) is a phase convention that FP32 cannot reproduce; it
                # cancels in |M|^2.  Exactly on the axis both take the same branch.
                pt = np.hypot(p[:, 1], p[:, 2])
                rel = rel[(pt == 0) | (pt > 1e-6 * np.linalg.norm(p[:, 1:], axis=1))]
            err = rel.max()
            assert err < (1e-13 if precision == 64 else 2e-6), (kind, mass, hel, nsf, err)


ROUTINES = [("FFV1_0", ["i", "o", "v"], True), ("FFV1_1", ["o", "v"], False), ("FFV1_2", ["i", "v"], False),
            ("VVV1P0_1", ["v", "v"], False), ("VVVV1P0_1", ["v", "v", "v"], False),
            ("VVVV3P0_1", ["v", "v", "v"], False), ("VVVV4P0_1", ["v", "v", "v"], False),
            ("FFV1P0_3", ["i", "o"], False), ("VVV1_0", ["v", "v", "v"], True),
            ("VVVV1_0", ["v", "v", "v", "v"], True), ("VVVV3_0", ["v", "v", "v", "v"], True),
            ("VVVV4_0", ["v", "v", "v", "v"], True)]
IDS = {name: i for i, (name, _, _) in enumerate(ROUTINES)}


def _inputs(kinds, n=200, seed=9):
    rng = np.random.default_rng(seed)
    ins = []
    for j, k in enumerate(kinds):
        mass = MT if k in "io" else 0.0
        v = rng.normal(size=(n, 3)) * 400
        p = np.concatenate([np.sqrt((v * v).sum(1) + mass * mass)[:, None], v], axis=1)
        hel = rng.choice([-1, 1], size=n)
        w = np.stack([O.wavefunction(k, p[e:e + 1], mass, hel[e], 1 if j % 2 else -1)[0] for e in range(n)])
        ins.append(w)
    return ins


@pytest.mark.gpu
@pytest.mark.parametrize("name,kinds,is_amp", ROUTINES)
@pytest.mark.parametrize("precision", [64, 32])
def test_aloha_routines(dev, name, kinds, is_amp, precision):
    print("case aloha", name, precision, flush=True)
    ins = _inputs(kinds)
    n = ins[0].shape[0]
    coup = complex(0.3, 1.2)
    M, Wd = (MT, 1.4915) if name in ("FFV1_1", "FFV1_2") else (0.0, 0.0)
    ref = O.aloha(name, ins, coup, M, Wd)
    lib = _lib.load()
    arr = np.zeros((len(ins), n, 6, 2))
    for i, a in enumerate(ins):
        arr[i, ..., 0] = a.real
        arr[i, ..., 1] = a.imag
    tin = torch.from_numpy(arr).cuda()
    out = torch.zeros((n, 2) if is_amp else (n, 6, 2), dtype=torch.float64, device="cuda")
    c = np.array([coup.real, coup.imag])
    _lib.check(lib.mfx_aloha(IDS[name], ctypes.c_void_p(tin.data_ptr()), len(ins), n, c.ctypes.data_as(ctypes.c_void_p),
                             M, Wd, ctypes.c_void_p(out.data_ptr()), precision, _lib.stream_ptr(tin.device)))
    o = out.cpu().numpy()
    got = o[..., 0] + 1j * o[..., 1]
    if is_amp:
        err = np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-6 * np.abs(ref).max()))
    else:
        # momentum slots against the momentum scale, components against their own scale
        escale = sum(np.abs(a[:, 0].real) for a in ins)
        slot_err = np.max(np.abs(got[:, :2] - ref[:, :2]), axis=1) / escale
        assert slot_err.max() < (1e-15 if precision == 64 else 1e-6)
        scale = np.abs(ref[:, 2:]).max(axis=1, keepdims=True)
        rel = np.max(np.abs(got[:, 2:] - ref[:, 2:]) / scale, axis=1)
        if precision == 32:
            # the propagator amplifies the FP32 rounding of P^2 - M^2 by E^2 / |P^2 - M^2|:
            # compare where that factor is below 20 (FP64 covers every event)
            P = np.stack([ref[:, 0].real, ref[:, 1].real, ref[:, 1].imag, ref[:, 0].imag], axis=1)
            den = np.abs(P[:, 0] ** 2 - (P[:, 1:] ** 2).sum(1) - M * M)
            e2 = sum(np.abs(a[:, 0].real) for a in ins) ** 2
            rel = rel[den > 0.05 * e2]
            assert rel.size > n // 4
        err = rel.max()
    print(name, precision, "max rel err", err, flush=True)
    assert err < (1e-13 if precision == 64 else (5e-6 if not is_amp else 2e-5))
