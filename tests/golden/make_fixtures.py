"""Write the golden vectors under tests/golden/ from the CPU oracle.

For each process: 16 phase-space points (oracle RAMBO on fixed uniforms, incoming
gluons along +-z), |M|^2 summed over helicities in FP64 with the reference's constants
(``ref_compat``: float32-rounded sqrt(1/2), gs literal 1.2177157847767195,
Gamma_t = 1.4915000200271606) and with exact constants, and the per-helicity jamps of
the first non-vanishing helicity row.  gg -> tt~ additionally stores the hand-restated
reference mockup (tests/mockup_debug_me.py:453-541) on the same points.

The fixtures pin the oracle against later edits and give the GPU tests a stored
expectation that does not depend on rebuilding the oracle.  Run from the repo root:

    python -m tests.golden.make_fixtures
"""

import json
import os

import numpy as np

from oracle import oracle as O
from tests.helpers import phase_space_points

HERE = os.path.dirname(os.path.abspath(__file__))
PROCS = ["gg_ttx", "gg_ttxg", "gg_ttxgg", "gg_ttxggg"]
NEV = 16


def _coup_values(proc, coup):
    return {k: [v.real, v.imag] for k, v in coup.items()}


def make(proc):
    ps, wt = phase_space_points(proc, NEV, seed=2024)
    rp_ref, coup_ref = O.default_params(ref_gs=True)
    rp_ex, coup_ex = O.default_params(ref_gs=False)
    me_ref = O.smatrix(proc, ps, rp_ref, coup_ref, ref_compat=True)
    me_ex = O.smatrix(proc, ps, rp_ex, coup_ex, ref_compat=False)
    desc = O.load_process(proc)
    hel = None
    for row in desc["helicities"]:
        m, j = O.matrix_hel(proc, ps, np.array(row, dtype=np.float64), rp_ref, coup_ref, ref_compat=True,
                            want_jamp=True)
        if np.all(m > 0):
            hel = row
            break
    out = dict(process=proc, nev=NEV, seed=2024, sqrts=float(ps[0, 0, 0] * 2),
               params_ref_compat=dict(mdl_MT=rp_ref[0], mdl_WT=rp_ref[1], couplings=_coup_values(proc, coup_ref)),
               params_exact=dict(mdl_MT=rp_ex[0], mdl_WT=rp_ex[1], couplings=_coup_values(proc, coup_ex)),
               all_ps=ps.tolist(), rambo_weight=wt.tolist(), smatrix_ref_compat=me_ref.tolist(),
               smatrix_exact=me_ex.tolist(), helicity=list(hel), matrix_hel_ref_compat=m.tolist(),
               jamp_re=np.real(j).tolist(), jamp_im=np.imag(j).tolist())
    if proc == "gg_ttx":
        out["mockup_ref_compat"] = O.smatrix_gg_ttx_mockup(ps, ref_compat=True)[0].tolist()
    with open(os.path.join(HERE, proc + ".json"), "w") as f:
        json.dump(out, f, indent=None, separators=(",", ":"))
        f.write("\n")
    print(proc, "written")


if __name__ == "__main__":
    for p in PROCS:
        make(p)
