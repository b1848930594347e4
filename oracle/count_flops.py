"""ORACLE — test infrastructure only.

Writes madflow_amd/processes/flops_per_me.json: the exact number of real floating-point
operations (add/sub, mul, div, sqrt; an FMA counts 2) that the oracle executes for one
smatrix evaluation of each process (all helicities, all diagrams, jamps, colour sum).
This is the algorithmic work per matrix element used for the roofline figures.
Run: python -m oracle.count_flops
"""

import json
import os

from oracle import oracle as O

OUT = os.path.join(os.path.dirname(O.HERE), "madflow_amd", "processes", "flops_per_me.json")


def main():
    res = {}
    for name in ("gg_ttx", "gg_ttxg", "gg_ttxgg", "gg_ttxggg", "uux_ttx", "uux_ttxg", "uux_ttxgg", "ug_ttxu",
                 "uxg_ttxux"):
        try:
            O.load_process(name)
        except FileNotFoundError:
            continue
        res[name] = O.flops_per_me(name)
        # rows that do not vanish identically (the kernels skip the others; the count above
        # is over every row, as the reference evaluates them)
        d = O.load_process(name)
        kinds = d.get("kinds", [])
        q = [i for i, k in enumerate(kinds) if k in ("quark", "antiquark", "quark_out", "antiquark_out")]
        # a scattering line keeps the label, an annihilating one flips it
        same = "quark_out" in kinds or "antiquark_out" in kinds
        res[name]["live_rows"] = sum(1 for h in d["helicities"]
                                     if len(q) < 2 or (h[q[0]] == h[q[1]]) == same)
        res[name]["rows"] = len(d["helicities"])
        print(name, res[name])
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
