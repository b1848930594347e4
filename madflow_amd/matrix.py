"""Host side of the drop-in process modules (``madflow_amd/processes/matrix_1_<proc>.py``).

``MatrixBase`` implements what the reference's generated class does in TensorFlow
(``madgraph_plugin/template_files/matrix_method_python.inc:80-138``) by handing
device buffers to the fused HIP kernel (``mfx_smatrix`` / ``mfx_matrix``):

* ``smatrix(all_ps, *params)``: ``all_ps`` is a ``(nev, nexternal, 4)`` float64 tensor
  on the GPU; ``params`` are ``mdl_MT, mdl_WT`` (floats) and the couplings (complex
  scalars, or complex tensors of length 1 or nev as returned by ``Model.evaluate``).
  Returns a float64 ``(nev,)`` tensor on the same device.
* ``matrix(all_ps, hel, *params)``: one helicity row, colour-summed, not averaged.

Each process has one or two kernel families in the library (``kernels`` of the class, the
first the default): ``"fused"``, one straight-line Feynman-diagram program per helicity
row, and ``"ordered"``, colour-ordered currents with an MFMA colour sum.  The default is
the faster form; ``Matrix_1_<proc>(kernels=...)`` selects another.

Precision is a runtime choice (``precision=64`` FP64 build, ``32`` FP32 build) of the
same kernels; ``ref_compat=True`` uses the reference's float32-rounded sqrt(1/2)
(``python_package/madflow/wavefunctions_flow.py:10``) so FP64 results match it to
round-off.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .parameters import make_model

_PROC_PARAMS = {
    "gg_ttx": ["mdl_MT", "mdl_WT", "GC_10", "GC_11"],
    "gg_ttxg": ["mdl_MT", "mdl_WT", "GC_10", "GC_11", "GC_12"],
    "gg_ttxgg": ["mdl_MT", "mdl_WT", "GC_10", "GC_11", "GC_12"],
    "gg_ttxggg": ["mdl_MT", "mdl_WT", "GC_10", "GC_11", "GC_12"],
    "uux_ttx": ["mdl_MT", "mdl_WT", "GC_11"],
    "uux_ttxg": ["mdl_MT", "mdl_WT", "GC_10", "GC_11"],
    "uux_ttxgg": ["mdl_MT", "mdl_WT", "GC_10", "GC_11", "GC_12"],
    "ug_ttxu": ["mdl_MT", "mdl_WT", "GC_10", "GC_11"],
    "uxg_ttxux": ["mdl_MT", "mdl_WT", "GC_10", "GC_11"],
}
_GC_SLOT = {"GC_10": 0, "GC_11": 1, "GC_12": 2}


def model_for(process_name, param_card_path=None):
    return make_model(_PROC_PARAMS[process_name], param_card_path)


def process_index(name):
    lib = _lib.load()
    idx = lib.mfx_process_index(name.encode())
    if idx < 0:
        raise _lib.MfxError("process %s is not compiled into %s" % (name, _lib.LIB_PATH))
    return idx


def _as_scalar(x, name):
    if isinstance(x, torch.Tensor):
        if x.numel() != 1:
            raise ValueError("%s must be a scalar" % name)
        return x.reshape(()).item()
    return x


class MatrixBase:
    process_name = None
    params = ()
    precision = 64
    ref_compat = True
    kernels = ("fused",)

    def __init__(self, precision=None, ref_compat=None, kernels=None):
        if precision is not None:
            self.precision = int(precision)
        if ref_compat is not None:
            self.ref_compat = bool(ref_compat)
        if kernels is not None and kernels not in self.kernels:
            raise ValueError("%s has the kernel families %s, not %r" % (type(self).__name__, self.kernels, kernels))
        self.kernel_family = kernels or self.kernels[0]
        self._idx = None
        self._ws = {}

    @property
    def registry_name(self):
        """Name of the kernels in the library's process registry."""
        if self.kernel_family == self.kernels[0]:
            return self.process_name
        return "%s_%s" % (self.process_name, self.kernel_family)

    @property
    def index(self):
        if self._idx is None:
            self._idx = process_index(self.registry_name)
        return self._idx

    # ---------------------------------------------------------------- helpers
    def _prepare(self, all_ps, params):
        all_ps = _lib.require_device_tensor(all_ps, torch.float64, "all_ps")
        if all_ps.dim() != 3 or all_ps.shape[1] != self.nexternal or all_ps.shape[2] != 4:
            raise ValueError("all_ps must have shape (nev, %d, 4), got %s" % (self.nexternal, tuple(all_ps.shape)))
        nev = all_ps.shape[0]
        if len(params) != len(self.params):
            raise TypeError("%s.smatrix expects %d model parameters %s" % (
                type(self).__name__, len(self.params), list(self.params)))
        vals = dict(zip(self.params, params))
        mw = np.array([float(_as_scalar(vals["mdl_MT"], "mdl_MT")), float(_as_scalar(vals["mdl_WT"], "mdl_WT"))])
        gc = np.zeros(6)
        per_event = None
        for name, slot in _GC_SLOT.items():
            if name not in vals:
                continue
            v = vals[name]
            if isinstance(v, torch.Tensor) and v.numel() > 1:
                if v.numel() != nev:
                    raise ValueError("%s must have length 1 or nev=%d" % (name, nev))
                if per_event is None:
                    per_event = torch.zeros((3, nev, 2), dtype=torch.float64, device=all_ps.device)
                vv = v.to(device=all_ps.device, dtype=torch.complex128).reshape(-1)
                per_event[slot, :, 0] = vv.real
                per_event[slot, :, 1] = vv.imag
            else:
                c = complex(_as_scalar(v, name))
                gc[2 * slot] = c.real
                gc[2 * slot + 1] = c.imag
        if per_event is not None:
            # broadcast frozen couplings into the per-event array too
            for name, slot in _GC_SLOT.items():
                if name in vals and not (isinstance(vals[name], torch.Tensor) and vals[name].numel() > 1):
                    per_event[slot, :, 0] = gc[2 * slot]
                    per_event[slot, :, 1] = gc[2 * slot + 1]
        return all_ps, nev, mw, gc, per_event

    # device workspace for processes evaluated by the colour-ordered kernels: at most this
    # many bytes are kept (the library splits events and helicity rows to fit)
    workspace_limit = 8 << 30

    def _workspace(self, nev, per_helicity, device):
        """(pointer, bytes) of a cached device workspace, grown on demand; (None, 0) when the
        process needs none.  One workspace is kept per (device, current stream) and is
        allocated while that stream is current, so the caching allocator orders its reuse
        after that stream's work, and calls on different streams never share one."""
        lib = _lib.load()
        need = int(lib.mfx_smatrix_workspace_size(self.index, nev, self.precision, int(per_helicity)))
        if need < 0:
            _lib.check(need)
        if need == 0:
            return None, 0
        size = min(need, self.workspace_limit)
        key = (device.index, torch.cuda.current_stream(device).cuda_stream)
        ws = self._ws.get(key)
        if ws is None or ws.numel() < size:
            self._ws.pop(key, None)
            ws = self._ws[key] = torch.empty(size, dtype=torch.uint8, device=device)
        return ctypes.c_void_p(ws.data_ptr()), ws.numel()

    def _smatrix(self, all_ps, params):
        all_ps, nev, mw, gc, per_event = self._prepare(all_ps, params)
        out = torch.empty(nev, dtype=torch.float64, device=all_ps.device)
        if nev == 0:
            return out
        lib = _lib.load()
        ws, ws_bytes = self._workspace(nev, False, all_ps.device)
        _lib.check(lib.mfx_smatrix(
            self.index, ctypes.c_void_p(all_ps.data_ptr()), nev, mw.ctypes.data_as(ctypes.c_void_p),
            gc.ctypes.data_as(ctypes.c_void_p),
            ctypes.c_void_p(per_event.data_ptr()) if per_event is not None else None,
            ctypes.c_void_p(out.data_ptr()), self.precision, int(self.ref_compat), ws, ws_bytes,
            _lib.stream_ptr(all_ps.device)))
        return out

    def _matrix(self, all_ps, hel, params, return_jamp=False):
        all_ps, nev, mw, gc, per_event = self._prepare(all_ps, params)
        if isinstance(hel, torch.Tensor):
            hel = hel.detach().cpu().numpy()
        helv = np.ascontiguousarray(np.asarray(hel, dtype=np.float64).round().astype(np.int32))
        if helv.shape != (self.nexternal,):
            raise ValueError("hel must have %d entries" % self.nexternal)
        out = torch.empty(nev, dtype=torch.float64, device=all_ps.device)
        jamp = torch.empty((nev, self.ncolor, 2), dtype=torch.float64, device=all_ps.device) if return_jamp else None
        if nev > 0:
            lib = _lib.load()
            ws, ws_bytes = self._workspace(nev, True, all_ps.device)
            _lib.check(lib.mfx_matrix(
                self.index, ctypes.c_void_p(all_ps.data_ptr()), nev, helv.ctypes.data_as(ctypes.c_void_p),
                mw.ctypes.data_as(ctypes.c_void_p), gc.ctypes.data_as(ctypes.c_void_p),
                ctypes.c_void_p(per_event.data_ptr()) if per_event is not None else None,
                ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(jamp.data_ptr()) if jamp is not None else None,
                self.precision, int(self.ref_compat), ws, ws_bytes, _lib.stream_ptr(all_ps.device)))
        if return_jamp:
            return out, torch.complex(jamp[..., 0], jamp[..., 1])
        return out


def load_matrix(process_name, **kw):
    """Instantiate the drop-in class of a generated process module by name (keywords:
    precision, ref_compat, kernels)."""
    import importlib
    mod = importlib.import_module("madflow_amd.processes.matrix_1_%s" % process_name)
    cls = getattr(mod, "Matrix_1_%s" % process_name)
    return cls(**kw), mod
