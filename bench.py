"""Benchmark: g g > t t~ g g on one MI355X.

Metric: unweighted events/sec at sigma/sigma = 1e-3 and ME evals/sec.
``value`` is the ME evaluations per second of the fused FP32 smatrix kernel over a
device-resident batch of post-cut momenta (one step = one kernel pass over the batch).
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f32_valu": 157.3, "f64_valu": 78.6}  # TFLOP/s, AMD published MI355X vector peaks


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nev", type=int, default=1 << 20)
    ap.add_argument("--precision", type=int, default=32)
    ap.add_argument("--process", default="gg_ttxgg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print("bench.py measures one GPU only: every box here has a single MI355X; "
              "multi-GPU sharding is described in DESIGN.md but not run", file=sys.stderr)
        return 2

    import torch
    from madflow_amd.matrix import load_matrix
    from madflow_amd import phasespace_gpu

    dev = torch.device("cuda:0")
    name = a.process
    m, mod = load_matrix(name, precision=a.precision, ref_compat=True)
    model = mod.get_model_param()
    model.freeze_alpha_s(0.118)
    params = model.evaluate()
    log("process", name, "nev", a.nev, "precision", a.precision)
    ps = phasespace_gpu.sample_momenta(name, a.nev, seed=3, device=dev)
    torch.cuda.synchronize()
    for _ in range(a.warmup):
        m.smatrix(ps, *params)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(a.steps):
        out = m.smatrix(ps, *params)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev0.elapsed_time(ev1)
    ms_step = ms_total / a.steps
    me_per_s = a.nev / (ms_step * 1e-3)
    log("ms/step %.3f  ME/s %.4g  mean|M|^2 %.4g" % (ms_step, me_per_s, out.mean().item()))
    with open(os.path.join(ROOT, "madflow_amd", "processes", "flops_per_me.json")) as f:
        flops = json.load(f)[name]["total"]
    achieved = flops * me_per_s / 1e12
    peak = PEAK["f32_valu"] if a.precision == 32 else PEAK["f64_valu"]
    res = {
        "metric": "unweighted events/sec at sigma/sigma=1e-3 and ME evals/sec, gg->ttgg on 1 MI355X",
        "value": me_per_s,
        "unit": "ME evals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if a.precision == 32 else "f64",
        "data": "synthetic",
        "config": {"workload": "%s smatrix, sqrt(s)=3 TeV RAMBO points" % name, "batch": a.nev},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None},
        "wall_s": wall,
    }
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
