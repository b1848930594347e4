"""Model parameters and couplings (mirror of python_package/madflow/parameters.py).

``Model(constants, functions)`` keeps the reference's behaviour:

* ``constants`` is a namedtuple of alpha_s-independent parameters (``mdl_MT``, ``mdl_WT``),
* ``functions`` is a namedtuple of callables ``GC_x(G)`` of the strong coupling
  G = gs = 2 sqrt(pi alpha_s)  (``parameters.py:13-15``),
* ``evaluate(alpha_s)`` returns ``[*constants, *functions(gs)]`` in that order
  (``parameters.py:59-72``), ``freeze_alpha_s`` fixes alpha_s once (``:44-53``),
  freezing twice raises ``ValueError`` like the reference.

Couplings are returned as complex128 torch tensors of shape ``(n,)`` (one per event, or
length 1 when frozen) — the same shapes the reference passes to ``smatrix``.
"""

import collections
import math

import torch

GC_FUNCTIONS = {
    # SM UFO couplings used by the g g > t t~ + gluons processes
    "GC_10": lambda G: -G,
    "GC_11": lambda G: 1j * G,
    "GC_12": lambda G: 1j * G * G,
}


def alphas_to_gs(alpha_s):
    """gs = 2 sqrt(pi alpha_s) as complex128 (parameters.py:13-15)."""
    a = torch.as_tensor(alpha_s, dtype=torch.float64)
    return (2.0 * torch.sqrt(math.pi * a)).to(torch.complex128)


class Model:
    def __init__(self, constants, functions):
        self._tuple_constants = constants
        self._tuple_functions = functions
        self._constants = list(constants)
        self._to_evaluate = list(functions)
        self._frozen = []

    @property
    def frozen(self):
        return bool(self._frozen)

    def freeze_alpha_s(self, alpha_s):
        if self.frozen:
            raise ValueError("The model is already frozen")
        self._frozen = self._evaluate(torch.tensor([float(alpha_s)], dtype=torch.float64))

    def unfreeze(self):
        self._frozen = []

    def _evaluate(self, alpha_s):
        gs = alphas_to_gs(alpha_s).reshape(-1)
        results = [f(gs) for f in self._to_evaluate]
        if not results:
            return list(self._constants)
        return list(self._constants) + results

    def get_masses(self):
        return [v for k, v in self._tuple_constants._asdict().items() if k.startswith("mdl_M")]

    def parse_parameter(self, parameter_name):
        if parameter_name == "ZERO":
            return 0.0
        if hasattr(self._tuple_constants, parameter_name):
            return getattr(self._tuple_constants, parameter_name)
        if hasattr(self._tuple_functions, parameter_name):
            return getattr(self._tuple_functions, parameter_name)
        raise AttributeError(f"The model class does not contain parameter {parameter_name}")

    def evaluate(self, alpha_s=None):
        if self.frozen:
            return self._frozen
        if alpha_s is None:
            raise ValueError("alpha_s is required unless the model is frozen")
        return self._evaluate(alpha_s)


# Default SM values of the parameters these processes use (the reference's
# mockup: mockup_debug_me.py:22-23).
DEFAULT_CARD = {"mdl_MT": 173.0, "mdl_WT": 1.4915000200271606}


def read_param_card(path):
    """Minimal SLHA reader: MASS 6 -> mdl_MT, DECAY 6 -> mdl_WT."""
    vals = dict(DEFAULT_CARD)
    block = None
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            tok = line.split()
            if tok[0].upper() == "BLOCK":
                block = tok[1].upper()
                continue
            if tok[0].upper() == "DECAY" and len(tok) >= 3 and tok[1] == "6":
                vals["mdl_WT"] = float(tok[2])
                block = None
                continue
            if block == "MASS" and len(tok) >= 2 and tok[0] == "6":
                vals["mdl_MT"] = float(tok[1])
    return vals


def make_model(param_names, param_card_path=None):
    """Model for a process whose smatrix takes ``param_names`` (masses/widths then couplings)."""
    card = read_param_card(param_card_path) if param_card_path else dict(DEFAULT_CARD)
    consts = [p for p in param_names if not p.startswith("GC_")]
    funcs = [p for p in param_names if p.startswith("GC_")]
    C = collections.namedtuple("constants", consts)
    F = collections.namedtuple("functions", funcs)
    return Model(C(*[card[c] for c in consts]), F(*[GC_FUNCTIONS[f] for f in funcs]))
