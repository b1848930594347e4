"""Benchmark: g g > t t~ g g on one MI355X.

Workload (default): LO g g > t t~ g g at a fixed partonic sqrt(s) = 3 TeV. SM couplings:
m_t = 173, Gamma_t = 1.4915, alpha_s = 0.118. Cuts: pT(g) > 30 GeV and Delta R(g,g) > 0.4.
Phase space is RAMBO on diet on the 8-dimensional VEGAS cube (50 bins per dimension),
with 1.6e7 points per VEGAS iteration. Wavefunctions, amplitudes and the colour sum are
FP32; phase space and VEGAS accumulation are FP64. |M|^2 runs through the drop-in class
Matrix_1_gg_ttxgg, whose default kernels are the colour-ordered ones (k_ordered_jamps +
MFMA k_colour_mfma); --kernels fused runs the same physics through the fused
Feynman-diagram kernel.

One step is one VEGAS iteration on the frozen grid, after the warm-up iterations have
adapted it: Philox uniforms -> VEGAS map -> phase space -> cuts -> compaction -> fused
smatrix kernels -> compensated accumulation. Everything is device resident.

Reported (one JSON line):
  value                      ME evaluations per second over the K timed steps (whole pipeline)
  me_kernel_evals_per_s      the smatrix kernel alone (HIP events around its launches)
  roofline                   smatrix kernel: oracle flop count per ME x MEs / kernel time
  unweighted                 N_unw = sum w / max w over the frozen-grid sample, per second:
                             "steady" over the timed steps, and "to_target" with the warm-up
                             counted and the run continued (extrapolated) to sigma/sigma = 1e-3
  flow_comparison            plain VEGAS with FP64 matrix elements vs VEGAS + learned
                             RQ-spline proposal with FP32 matrix elements, both to
                             sigma/sigma = 1e-3 (set-up counted), and their steady rates
  cpu_baseline               the CPU oracle (C++ FP64, OpenMP over events) on a bounded
                             sample of the same post-cut events

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-flow] [--no-cpu-baseline]
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {32: 157.3, 64: 78.6}   # MI355X published vector peaks (FP32 packed, FP64)
# HBM bytes per matrix element, (process, precision): rocprofv3 FETCH_SIZE and WRITE_SIZE
# passes scaled by a 1 GiB copy (profiles/pmc_k_smatrix_gg_ttxgg_f32.md for the fused
# Feynman-diagram kernel, profiles/pmc_k_ordered_gg_ttxgg_f32.md for the colour-ordered
# kernels, jamps staged through HBM included; 1 GiB copy calibration)
HBM_BYTES_PER_ME = {("gg_ttxgg_fused", 32): 28180.3, ("gg_ttxgg", 32): 24776.7}
# VALU wave instructions issued per ME (SQ_INSTS_VALU summed over the call's kernels / MEs,
# same profiles): the executed-work side of the roofline.  A wave64 VALU instruction holds
# a 16-lane SIMD for 4 cycles; 4 SIMDs per CU at the 2.4 GHz peak engine clock.
VALU_WAVE_INSTS_PER_ME = {("gg_ttxgg", 32): 18802.7}
VALU_CYCLES_PER_INST = 4
ENGINE_HZ = 2.4e9
METRIC = "unweighted events/sec at sigma/sigma=1e-3 and ME evals/sec, gg->ttgg on 1 MI355X"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_threads():
    """Threads for the CPU baseline: OMP_NUM_THREADS when set (16 on a one-GPU box, the
    share of the host's cores that box gives one GPU), else the cores this process may run
    on; and the host's core count beside it."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else avail
    return min(threads, avail), {"host_cores": os.cpu_count(), "cores_available": avail,
                                 "threads_from": "OMP_NUM_THREADS" if env.isdigit() and int(env) > 0 else "affinity"}


def cpu_baseline(proc, dense, npass, params, names, budget_s=12.0):
    """Time the CPU oracle (FP64 C++ built -O3 -march=x86-64-v3, OpenMP over events) on
    the first passing events of the last chunk: doubles the sample until one timed call
    takes budget_s / 3."""
    import numpy as np
    from oracle import oracle as O
    cores, host = cpu_threads()
    nmax = min(npass, 1 << 20)
    ps = dense[:nmax].cpu().numpy()
    rp = np.array([float(params[0]), float(params[1])], dtype=np.float64)
    coup = {nm: complex(x.item()) if hasattr(x, "item") else complex(x) for nm, x in zip(names[2:], params[2:])}
    n = 256
    while True:
        t0 = time.perf_counter()
        O.smatrix(proc, ps[:n], rp, coup, ref_compat=True, nthreads=cores)
        t = time.perf_counter() - t0
        if t > budget_s / 3 or n >= nmax:
            break
        n = min(nmax, n * 2)
    out = {"value": n / t, "unit": "ME evals/s", "cores": cores, "kind": "port",
           "sample": "%d post-cut %s events of the benchmark stream, FP64 oracle smatrix "
                     "(C++ -O3 -march=x86-64-v3, OpenMP over events, %d threads), %.1f s" % (n, LABEL[proc], cores, t)}
    out.update(host)
    return out


def cpu_c1_end_to_end(cores, n=1_000_000, warmup=2, iters=5, target=1e-3):
    """The survey's C1 run on the CPU oracle: gg -> tt~ at sqrt(s) = 1 TeV, RAMBO on the
    8-dim VEGAS cube (50 bins), FP64, n points x (warmup + iters) iterations: oracle
    Philox -> map -> RAMBO (one thread), oracle smatrix (OpenMP over events), numpy sums
    and f^2 histograms, oracle grid refinement.  Returns sigma, ME evals/s and unweighted
    events/s to the target with the warm-up counted (N_unw = sum w / max w)."""
    import numpy as np
    from oracle import oracle as O
    from madflow_amd.vegas import combine
    cfg = dict(n_out=2, ndim=8, nbins=50, mapping=0, sqrts=1000.0, masses=[173.0, 173.0], gev2pb=389379365.6,
               pt_min=0.0, abs_eta_max=0.0, dr_min=0.0, cut_first_leg=4, seed=1)
    edges = np.tile(np.linspace(0.0, 1.0, 51), (8, 1))
    rp, coup = O.default_params(ref_gs=True)
    res, n_me, t_me = [], 0, 0.0
    t0 = time.perf_counter()
    t_setup = None
    for it in range(warmup + iters):
        ps, wps, flag = O.generate(cfg, edges, 0, n, it)
        me = np.zeros(n)
        ok = flag == 1
        t1 = time.perf_counter()
        me[ok] = O.smatrix("gg_ttx", ps[ok], rp, coup, ref_compat=True, nthreads=cores)
        t_me += time.perf_counter() - t1
        n_me += int(ok.sum())
        f = me * wps
        mean = f.mean()
        var = max((np.mean(f * f) - mean * mean) / (n - 1), 0.0)
        if it < warmup:
            y = O.philox_uniforms(n, 8, cfg["seed"], it)
            bins = np.minimum((y * 50).astype(int), 49)
            hist = np.stack([np.bincount(bins[:, d], weights=f * f, minlength=50) for d in range(8)])
            edges = O.vegas_refine(edges, hist, alpha=1.5)
            if it == warmup - 1:
                t_setup = time.perf_counter() - t0
        else:
            res.append(dict(I=mean, var=var, sum_f=float(f.sum()), wmax=float(f.max()), n=n))
            sig, err = combine(res)
            if err / sig <= target:
                break
    wall = time.perf_counter() - t0
    sig, err = combine(res)
    n_unw = sum(r["sum_f"] for r in res) / max(r["wmax"] for r in res)
    return {"workload": "gg->tt, sqrt(s)=1 TeV, RAMBO 8-dim VEGAS cube, 50 bins, FP64, %d points/iteration, "
                        "%d warm-up + %d iterations" % (n, warmup, len(res)),
            "sigma_pb": sig, "err_pb": err, "rel_err": err / sig, "reached_target": err / sig <= target,
            "seconds": wall, "setup_s": t_setup, "me_evals_per_s": n_me / t_me,
            "unweighted_events": n_unw, "unweighted_per_s_to_target": n_unw / wall}


def _unweighted(res, setup_s, sample_s, target):
    """sigma, error, N_unw = sum w / max w, and unweighted events per second: steady (the
    sample alone) and to the target (set-up counted; if the sample stopped short of the
    target, its length is extrapolated at the measured variance)."""
    from madflow_amd.vegas import combine
    s, e = combine(res)
    rel = e / s if s else float("inf")
    sum_w = math.fsum(r["sum_f"] for r in res)
    wmax = max(r["wmax"] for r in res)
    n_unw = sum_w / wmax if wmax > 0 else 0.0
    npts = sum(r["n"] for r in res)
    t_sample = sample_s * max(1.0, (rel / target) ** 2)
    n_unw_target = n_unw * t_sample / sample_s
    return dict(sigma_pb=s, err_pb=e, rel_err=rel, iterations=len(res), points=npts, setup_s=setup_s,
                sample_s=sample_s, reached_target=rel <= target, seconds_to_target=setup_s + t_sample,
                n_unweighted=n_unw, unweighting_efficiency=sum_w / (npts * wmax) if wmax > 0 else 0.0,
                unweighted_per_s_steady=n_unw / sample_s,
                unweighted_per_s_to_target=n_unw_target / (setup_s + t_sample))


def _run_to_target(iterate, target, max_iter, tag):
    from madflow_amd.vegas import combine
    res = []
    for k in range(max_iter):
        res.append(iterate(k))
        s, e = combine(res)
        log("%s iteration %d: %.6g +- %.3g (combined rel %.2e)" % (tag, k + 1, res[-1]["I"],
                                                                    math.sqrt(res[-1]["var"]), e / s if s else 0))
        if s and e / s <= target:
            break
    return res


def _extend_to(iterate, res, sampled_s, seconds, torch):
    """Keep iterating (frozen proposal) until the sampling time reaches ``seconds``; returns
    the sampled iterations with the sampling time at which each ended (the first
    len(res) iterations end together at sampled_s, when the target was reached)."""
    out = [(sampled_s, r) for r in res]
    t0 = time.perf_counter()
    extra = 0.0
    while sampled_s + extra < seconds:
        r = iterate(len(out))
        torch.cuda.synchronize()
        extra = time.perf_counter() - t0
        out.append((sampled_s + extra, r))
    return out


def _rate_at(timed, t_cut, wall):
    """Unweighted-event figures of the iterations finished by sampling time t_cut, per
    second of ``wall`` (the sampling time, or set-up plus sampling time)."""
    res = [r for t, r in timed if t <= t_cut + 1e-9]
    if not res:
        return dict(iterations=0, points=0, sample_s=t_cut, wall_s=wall, n_unweighted=0.0,
                    unweighting_efficiency=0.0, unweighted_per_s=0.0)
    sum_w = math.fsum(r["sum_f"] for r in res)
    wmax = max(r["wmax"] for r in res)
    npts = sum(r["n"] for r in res)
    n_unw = sum_w / wmax if wmax > 0 else 0.0
    return dict(iterations=len(res), points=npts, sample_s=t_cut, wall_s=wall, n_unweighted=n_unw,
                unweighting_efficiency=sum_w / (npts * wmax) if wmax > 0 else 0.0, unweighted_per_s=n_unw / wall)


def run_main(a, torch):
    """Plain VEGAS at the main precision: W warm-up iterations, then exactly K timed ones."""
    from madflow_amd.vegas import MEIntegrand, VegasFlow
    if a.mapping == "multichannel":
        from madflow_amd.multichannel import MultiChannelIntegrand
        ig = MultiChannelIntegrand(a.process, precision=a.precision, seed=3, events_limit=a.events_limit,
                                   kernels=a.kernels, **ME_ONLY_KIN[a.process])
    else:
        ig = MEIntegrand(a.process, precision=a.precision, mapping=a.mapping, seed=3, events_limit=a.events_limit,
                         kernels=a.kernels, **ME_ONLY_KIN[a.process])
    vf = VegasFlow(ig.ndim, a.nev)
    vf.compile(ig)
    log("warm-up: %d VEGAS iterations of %d points (grid adaptation)" % (a.warmup, a.nev))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vf.run_integration(a.warmup, log=log)
    torch.cuda.synchronize()
    t_warm = time.perf_counter() - t0
    vf.freeze_grid()
    ig.me_timer = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        vf.run_integration(1, log=log)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    res = vf.history[-a.steps:]
    me_ms = sum(e0.elapsed_time(e1) for e0, e1, _ in ig.me_timer)
    me_n = sum(n for _, _, n in ig.me_timer)
    launches = len(ig.me_timer)
    ig.me_timer = None
    return dict(ig=ig, res=res, wall=wall, t_warm=t_warm, me_ms=me_ms, me_n=me_n, launches=launches,
                npass=sum(r["npass"] for r in res), unw=_unweighted(res, t_warm, wall, a.target))


def train_flow(a, torch, ig):
    """Training pairs (uniform points of the frozen VEGAS cube and their integrand values,
    plus a held-out tenth for checkpoint selection) and the flow trained on them."""
    from madflow_amd.flow import RQSFlow
    from madflow_amd.phasespace_gpu import philox_uniforms
    D = ig.ndim
    # on the multi-channel phase space the flow learns channel 0's (the flat channel's) cube
    chunk = ig.cube0_chunk if hasattr(ig, "cube0_chunk") else ig.chunk
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ntot = a.flow_pairs + a.flow_pairs // 10
    y = torch.empty((ntot, D), dtype=torch.float32, device="cuda")
    f = torch.empty(ntot, dtype=torch.float64, device="cuda")
    done = 0
    while done < ntot:
        n = min(a.events_limit, ntot - done)
        yy = philox_uniforms(n, D, 99, 0, ev0=done, dtype=torch.float32)
        _, _, fb = chunk(done, n, 0, keep_f=True, y_in=yy)
        y[done:done + n] = yy
        f[done:done + n] = fb
        done += n
    torch.cuda.synchronize()
    t_pairs = time.perf_counter() - t0
    t0 = time.perf_counter()
    flow = RQSFlow(D, layers=8, seed=a.flow_seed)
    stage_s = []
    nt = a.flow_pairs
    s0 = 1000 * (a.flow_seed - 1)   # minibatch-draw seeds of the stages
    flow.train(y[:nt], f[:nt], steps=a.flow_steps, batch=a.flow_batch, lr=a.flow_lr, seed=s0, log=log,
               val=(y[nt:], f[nt:]), eval_every=a.flow_eval_every, mix=a.flow_mix, beta=a.flow_beta,
               patience=a.flow_patience)
    flow.stage_steps = [flow.steps_run]
    torch.cuda.synchronize()
    stage_s.append(time.perf_counter() - t0)
    for rnd in range(a.flow_refine_rounds if a.flow_refine_steps > 0 else 0):
        t1 = time.perf_counter()
        # further stages: points drawn from the current proposal, weighted by f/q', carry
        # far more of the information where the integrand is large than uniform points
        done = 0
        yq = torch.empty((nt, D), dtype=torch.float32, device="cuda")
        wq = torch.empty(nt, dtype=torch.float64, device="cuda")
        fq = torch.empty(nt, dtype=torch.float64, device="cuda") if a.flow_refine_beta else None
        while done < nt:
            n = min(a.events_limit, nt - done)
            u = philox_uniforms(n, D, 98 - rnd, 0, ev0=done, dtype=torch.float32)
            ys, winv = flow.sample_mixture(u, u, a.flow_mix)
            ys = ys.contiguous()
            _, _, fb = chunk(done, n, 0, keep_f=True, y_in=ys, w_in=winv.contiguous())
            yq[done:done + n] = ys
            wq[done:done + n] = fb
            if fq is not None:
                fq[done:done + n] = fb / winv
            done += n
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        flow.train(yq, wq, steps=a.flow_refine_steps, batch=a.flow_batch, lr=a.flow_lr * 0.5, seed=s0 + 1 + rnd,
                   log=log,
                   val=(y[nt:], f[nt:]), eval_every=a.flow_eval_every // 2, mix=a.flow_mix,
                   f=fq, beta=a.flow_refine_beta)
        flow.stage_steps.append(flow.steps_run)
        torch.cuda.synchronize()
        stage_s += [t2 - t1, time.perf_counter() - t2]
        del yq, wq, fq
    del y, f
    torch.cuda.synchronize()
    flow.stage_seconds = stage_s      # stage 1, then per later stage: its pairs, its training
    return flow, t_pairs, time.perf_counter() - t0


def flow_grid(a, torch, main=None):
    """The flow leg's frozen FP32 VEGAS grid: flow_grid_warmup adaptation iterations of
    flow_grid_nev points (nev when 0), counted in its set-up.  The flow corrects what a
    short adaptation leaves, so its leg adapts on smaller iterations than the plain leg of
    the reference's schedule; the plain leg is also run with a short warm-up of its own
    (plain_leg), and the gains are reported against both.  Returns the integrand and the
    adaptation time."""
    from madflow_amd.vegas import MEIntegrand, VegasFlow
    ig = MEIntegrand(a.process, precision=32, mapping=a.mapping, seed=7, events_limit=a.events_limit,
                     kernels=a.kernels, **ME_ONLY_KIN[a.process])
    vf = VegasFlow(ig.ndim, a.flow_grid_nev or a.nev)
    vf.compile(ig)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vf.run_integration(a.flow_grid_warmup, log=log)
    vf.freeze_grid()
    torch.cuda.synchronize()
    return ig, time.perf_counter() - t0


def flow_train_only(a, torch):
    """The flow leg's set-up alone (grid adaptation, training pairs, training): one JSON
    line of timings, for profiling the training kernels."""
    ig, t_grid = flow_grid(a, torch)
    flow, t_pairs, t_train = train_flow(a, torch, ig)
    print(json.dumps({"flow_train_only": True, "grid_warmup_s": t_grid, "train_pairs_s": t_pairs, "train_s": t_train,
                      "stage_steps_run": flow.stage_steps, "batch": a.flow_batch,
                      "ms_per_step": t_train * 1e3 / max(1, sum(flow.stage_steps))}))
    return 0


def flow_iteration(a, torch, ig, flow, k):
    """One iteration of a.nev points from the defensive mixture of the flow and the
    frozen VEGAS cube, each weighted by f(y)/q'(y)."""
    return ig.proposal_iteration(flow, a.nev, 100 + k, a.flow_mix)


def plain_leg(a, torch, warm_nev, tag):
    """Plain VEGAS with FP64 matrix elements: a.warmup adaptation iterations of warm_nev points
    (the reference's schedule runs them with the iteration size itself,
    madflow_exec.py:489-497), then frozen iterations of a.nev points to the target, then on
    to equal sampling time and equal wall time."""
    from madflow_amd.vegas import MEIntegrand, VegasFlow
    sync = torch.cuda.synchronize
    ig64 = MEIntegrand(a.process, precision=64, mapping=a.mapping, seed=5, events_limit=a.events_limit,
                       kernels=a.kernels, **ME_ONLY_KIN[a.process])
    vf = VegasFlow(ig64.ndim, warm_nev)
    vf.compile(ig64)
    sync()
    t0 = time.perf_counter()
    vf.run_integration(a.warmup, log=log)
    vf.freeze_grid()
    sync()
    setup64 = time.perf_counter() - t0
    vf.n_events = a.nev

    def it64(k):
        vf.run_integration(1)
        return vf.history[-1]
    t0 = time.perf_counter()
    res64 = _run_to_target(it64, a.target, a.max_iter, tag)
    sync()
    out = _unweighted(res64, setup64, time.perf_counter() - t0, a.target)
    timed64 = _extend_to(it64, list(res64), out["sample_s"], max(a.equal_time_s, a.equal_wall_s - setup64), torch)
    out["equal_time"] = _rate_at(timed64, a.equal_time_s, a.equal_time_s)
    out["equal_wall"] = _rate_at(timed64, a.equal_wall_s - setup64, a.equal_wall_s)
    out["warmup"] = "%d x %d points" % (a.warmup, warm_nev)
    del ig64, vf
    return out


def run_flow_compare(a, torch, main):
    """Plain VEGAS with FP64 matrix elements (the reference's schedule, and a short warm-up)
    against VEGAS + learned proposal with FP32 matrix elements on the same cube, each run to
    sigma/sigma = target.  The flow path adapts an FP32 grid of its own (counted in its
    set-up), generates the training pairs, trains, and samples y from the defensive
    mixture with weights f(y)/q'(y).  After the target every leg keeps sampling to equal
    sampling time (equal_time_s) and to equal wall time with its set-up counted
    (equal_wall_s)."""
    sync = torch.cuda.synchronize
    # ---- plain VEGAS, FP64 matrix elements: the reference's schedule, then a short warm-up
    plain = plain_leg(a, torch, a.nev, "plain fp64")
    plain_short = plain_leg(a, torch, a.short_warmup_nev, "plain fp64, short warm-up") if a.short_warmup_nev else None
    # ---- VEGAS + flow, FP32 matrix elements, on a frozen grid of its own
    ig, t_grid = flow_grid(a, torch, main)
    flow, t_pairs, t_train = train_flow(a, torch, ig)
    setup_flow = t_grid + t_pairs + t_train
    t0 = time.perf_counter()
    itf = lambda k: flow_iteration(a, torch, ig, flow, k)
    resf = _run_to_target(itf, a.target, a.max_iter, "flow fp32")
    sync()
    learned = _unweighted(resf, setup_flow, time.perf_counter() - t0, a.target)
    timedf = _extend_to(itf, list(resf), learned["sample_s"], max(a.equal_time_s, a.equal_wall_s - setup_flow), torch)
    learned["equal_time"] = _rate_at(timedf, a.equal_time_s, a.equal_time_s)
    learned["equal_wall"] = _rate_at(timedf, a.equal_wall_s - setup_flow, a.equal_wall_s)
    learned.update(train_pairs_s=t_pairs, train_s=t_train, grid_warmup_s=t_grid,
                   grid_warmup="%d x %d points" % (a.flow_grid_warmup, a.flow_grid_nev or a.nev),
                   best_step=getattr(flow, "best_step", None), mixture=a.flow_mix)

    def gain(k, base=None):
        base = plain if base is None else base
        return learned[k] / base[k] if base[k] > 0 else None
    mc = run_multichannel_leg(a, torch) if (a.multichannel_leg and a.mapping == "diet") else None
    mcf = run_flow_mc_leg(a, torch) if (a.flow_mc_leg and a.mapping == "diet") else None
    mcc = run_channel_flow_leg(a, torch) if (a.channel_flow_leg and a.mapping == "diet") else None
    out = {
        "workload": "%s, %s, %d-dim %s VEGAS cube, %d points/iteration, target sigma/sigma=%g" % (
            LABEL[a.process], _kin_text(a.process), ig.ndim, a.mapping, a.nev, a.target),
        "plain_vegas_fp64": plain,
        "plain_vegas_fp64_short_warmup": plain_short,
        "vegas_flow_fp32": learned,
        "flow": {"layers": 8, "bins": 16, "hidden": "2x128", "weights": "bf16 (fp32 master)",
                 "train_pairs": a.flow_pairs, "validation_pairs": a.flow_pairs // 10,
                 "stage_steps_max": [a.flow_steps] + [a.flow_refine_steps] * (a.flow_refine_rounds
                                                                              if a.flow_refine_steps else 0),
                 "stage_steps_run": list(getattr(flow, "stage_steps", [])),
                 "stage_seconds": list(getattr(flow, "stage_seconds", [])),
                 "batch": a.flow_batch, "defensive_mixture": a.flow_mix, "patience": a.flow_patience, "seed": a.flow_seed,
                 "loss_beta": [a.flow_beta] + [a.flow_refine_beta] * (a.flow_refine_rounds if a.flow_refine_steps else 0)},
        "gain_unweighted_per_s_to_target": gain("unweighted_per_s_to_target"),
        "gain_unweighted_per_s_to_target_vs_short_warmup": (gain("unweighted_per_s_to_target", plain_short)
                                                            if plain_short else None),
        "gain_unweighted_per_s_equal_time": (learned["equal_time"]["unweighted_per_s"] /
                                             plain["equal_time"]["unweighted_per_s"]
                                             if plain["equal_time"]["unweighted_per_s"] > 0 else None),
        "gain_unweighted_per_s_equal_wall": (learned["equal_wall"]["unweighted_per_s"] /
                                             plain["equal_wall"]["unweighted_per_s"]
                                             if plain["equal_wall"]["unweighted_per_s"] > 0 else None),
        "sigma_pull": (learned["sigma_pb"] - plain["sigma_pb"]) / math.hypot(learned["err_pb"], plain["err_pb"]),
    }
    if mc is not None:
        # the multi-channel phase space with plain VEGAS per channel (no learned part)
        out["multichannel_fp32"] = mc
        out["gain_multichannel_per_s_to_target"] = (mc["unweighted_per_s_to_target"] / plain["unweighted_per_s_to_target"]
                                                    if plain["unweighted_per_s_to_target"] > 0 else None)
        if plain_short:
            out["gain_multichannel_per_s_to_target_vs_short_warmup"] = (
                mc["unweighted_per_s_to_target"] / plain_short["unweighted_per_s_to_target"]
                if plain_short["unweighted_per_s_to_target"] > 0 else None)
        for k in ("equal_time", "equal_wall"):
            out["gain_multichannel_per_s_" + k] = (mc[k]["unweighted_per_s"] / plain[k]["unweighted_per_s"]
                                                   if plain[k]["unweighted_per_s"] > 0 else None)
        out["sigma_pull_multichannel"] = ((mc["sigma_pb"] - plain["sigma_pb"]) /
                                          math.hypot(mc["err_pb"], plain["err_pb"]))
    if mcc is not None:
        # a flow per dominant channel of the multi-channel phase space
        out["multichannel_channel_flows_fp32"] = mcc
        for k, key in (("to_target", "unweighted_per_s_to_target"),):
            out["gain_multichannel_channel_flows_per_s_" + k] = (mcc[key] / plain[key] if plain[key] > 0 else None)
        for k in ("equal_time", "equal_wall"):
            out["gain_multichannel_channel_flows_per_s_" + k] = (
                mcc[k]["unweighted_per_s"] / plain[k]["unweighted_per_s"] if plain[k]["unweighted_per_s"] > 0 else None)
        out["sigma_pull_multichannel_channel_flows"] = ((mcc["sigma_pb"] - plain["sigma_pb"]) /
                                                        math.hypot(mcc["err_pb"], plain["err_pb"]))
    if mcf is not None:
        # the learned proposal as a channel of the multi-channel phase space
        out["multichannel_flow_fp32"] = mcf
        for k, key in (("to_target", "unweighted_per_s_to_target"),):
            out["gain_multichannel_flow_per_s_" + k] = (mcf[key] / plain[key] if plain[key] > 0 else None)
        out["gain_multichannel_flow_per_s_equal_time"] = (mcf["equal_time"]["unweighted_per_s"] /
                                                          plain["equal_time"]["unweighted_per_s"]
                                                          if plain["equal_time"]["unweighted_per_s"] > 0 else None)
        out["gain_multichannel_flow_per_s_equal_wall"] = (mcf["equal_wall"]["unweighted_per_s"] /
                                                          plain["equal_wall"]["unweighted_per_s"]
                                                          if plain["equal_wall"]["unweighted_per_s"] > 0 else None)
        out["sigma_pull_multichannel_flow"] = ((mcf["sigma_pb"] - plain["sigma_pb"]) /
                                               math.hypot(mcf["err_pb"], plain["err_pb"]))
    return out


def run_flow_mc_leg(a, torch):
    """FP32 matrix elements on the multi-channel phase space with the learned proposal as
    one more channel: the channels' warm-up (grids and weights), then the flow trained on
    channel 0's cube as in the flow leg, attached with weight flow_mc_alpha; all of it
    counted in the set-up."""
    from madflow_amd.multichannel import MultiChannelIntegrand
    from madflow_amd.vegas import VegasFlow
    sync = torch.cuda.synchronize
    ig = MultiChannelIntegrand(a.process, precision=32, seed=9, events_limit=a.events_limit, kernels=a.kernels,
                               **ME_ONLY_KIN[a.process])
    vf = VegasFlow(ig.ndim, a.nev)
    vf.compile(ig)
    sync()
    t0 = time.perf_counter()
    vf.run_integration(a.warmup, log=log)
    vf.freeze_grid()
    sync()
    t_warm = time.perf_counter() - t0
    flow, t_pairs, t_train = train_flow(a, torch, ig)
    ig.attach_flow(flow, a.flow_mc_alpha)
    setup = t_warm + t_pairs + t_train

    def it(k):
        return ig.iteration(a.nev, 200 + k, train=False)
    t0 = time.perf_counter()
    res = _run_to_target(it, a.target, a.max_iter, "multichannel + flow fp32")
    sync()
    out = _unweighted(res, setup, time.perf_counter() - t0, a.target)
    timed = _extend_to(it, list(res), out["sample_s"], max(a.equal_time_s, a.equal_wall_s - setup), torch)
    out["equal_time"] = _rate_at(timed, a.equal_time_s, a.equal_time_s)
    out["equal_wall"] = _rate_at(timed, a.equal_wall_s - setup, a.equal_wall_s)
    out.update(warmup_s=t_warm, train_pairs_s=t_pairs, train_s=t_train, alpha_learned=ig.alpha_learned,
               alpha=[float(x) for x in ig.alpha], stage_steps_run=list(getattr(flow, "stage_steps", [])))
    return out


def run_channel_flow_leg(a, torch):
    """FP32 matrix elements on the multi-channel phase space with a flow per channel for the
    a.channel_flows channels of largest variance share (measured on one frozen iteration):
    each flow is trained on its channel's cube as in the flow leg (uniform points, their
    integrand values alpha_c f / sum_j alpha_j g_j) and draws that channel's points in a
    defensive mixture with the uniform; the channel weights stay those of the warm-up.
    Warm-up, the probe iteration and the training are counted in the set-up."""
    import numpy as np
    from madflow_amd.multichannel import MultiChannelIntegrand
    from madflow_amd.vegas import VegasFlow
    sync = torch.cuda.synchronize
    ig = MultiChannelIntegrand(a.process, precision=32, seed=7, events_limit=a.events_limit, kernels=a.kernels,
                               **ME_ONLY_KIN[a.process])
    vf = VegasFlow(ig.ndim, a.mc_warmup_nev or a.nev)
    vf.compile(ig)
    sync()
    t0 = time.perf_counter()
    vf.run_integration(a.warmup, log=log)
    vf.freeze_grid()
    probe = ig.iteration(a.mc_warmup_nev or a.nev, 50, train=False)
    sync()
    t_warm = time.perf_counter() - t0
    vp = np.array(probe["var_parts"])
    chosen = [int(c) for c in np.argsort(-vp)[:a.channel_flows]]

    class _Cube:
        """Channel c's cube as the flow leg's training sees an integrand."""
        def __init__(self, c):
            self.c, self.ndim, self.events_limit = c, ig.ndim, ig.events_limit

        def chunk(self, ev0, n, iteration, keep_f=False, y_in=None, w_in=None):
            r, k, _, f = ig.chunk_channel(self.c, ev0, n, iteration, 1.0, keep_f=True, y_in=y_in, w_in=w_in)
            return r, k, f
    t_train, steps = 0.0, []
    for c in chosen:
        sync()
        t1 = time.perf_counter()
        flow, _, _ = train_flow(a, torch, _Cube(c))
        sync()
        t_train += time.perf_counter() - t1
        ig.attach_channel_flow(c, flow, a.channel_flow_mix)
        steps.append(list(getattr(flow, "stage_steps", [])))
    setup = t_warm + t_train

    def it(k):
        return ig.iteration(a.nev, 200 + k, train=False)
    t0 = time.perf_counter()
    res = _run_to_target(it, a.target, a.max_iter, "multichannel + channel flows fp32")
    sync()
    out = _unweighted(res, setup, time.perf_counter() - t0, a.target)
    timed = _extend_to(it, list(res), out["sample_s"], max(a.equal_time_s, a.equal_wall_s - setup), torch)
    out["equal_time"] = _rate_at(timed, a.equal_time_s, a.equal_time_s)
    out["equal_wall"] = _rate_at(timed, a.equal_wall_s - setup, a.equal_wall_s)
    vps = np.mean([r["var_parts"] for _, r in timed], axis=0)
    out.update(warmup_s=t_warm, train_s=t_train, channels_with_flow=chosen, stage_steps_run=steps,
               variance_share_before=[float(v) for v in vp / vp.sum()],
               # each channel's variance after, at the probe's number of points, over the total before
               variance_after_over_total_before=[float(v) for v in vps / vp.sum() * (a.nev / probe["n"])],
               mixture=a.channel_flow_mix)
    return out


def run_multichannel_leg(a, torch):
    """FP32 matrix elements on the multi-channel phase space (flat diet, a t-channel per
    gluon, an s-channel per gluon pair, one VEGAS grid each, adapted channel weights): the
    same warm-up, run to the target and equal-time extension as the plain leg."""
    from madflow_amd.multichannel import MultiChannelIntegrand
    from madflow_amd.vegas import VegasFlow
    sync = torch.cuda.synchronize
    ig = MultiChannelIntegrand(a.process, precision=32, seed=7, events_limit=a.events_limit, kernels=a.kernels,
                               **ME_ONLY_KIN[a.process])
    vf = VegasFlow(ig.ndim, a.mc_warmup_nev or a.nev)
    vf.compile(ig)
    sync()
    t0 = time.perf_counter()
    vf.run_integration(a.warmup, log=log)
    vf.freeze_grid()
    sync()
    setup = time.perf_counter() - t0
    vf.n_events = a.nev

    def it(k):
        vf.run_integration(1)
        return vf.history[-1]
    t0 = time.perf_counter()
    res = _run_to_target(it, a.target, a.max_iter, "multichannel fp32")
    sync()
    out = _unweighted(res, setup, time.perf_counter() - t0, a.target)
    timed = _extend_to(it, list(res), out["sample_s"], max(a.equal_time_s, a.equal_wall_s - setup), torch)
    out["equal_time"] = _rate_at(timed, a.equal_time_s, a.equal_time_s)
    out["equal_wall"] = _rate_at(timed, a.equal_wall_s - setup, a.equal_wall_s)
    out["warmup"] = "%d x %d points" % (a.warmup, a.mc_warmup_nev or a.nev)
    out["channels"] = [list(c) for c in ig.channels]
    out["alpha"] = [float(x) for x in ig.alpha]
    return out


ME_ONLY_KIN = {"gg_ttx": dict(sqrts=1000.0), "gg_ttxg": dict(sqrts=2000.0, pt_min=30.0, abs_eta_max=2.5),
               "gg_ttxgg": dict(sqrts=3000.0, pt_min=30.0, dr_min=0.4),
               "gg_ttxggg": dict(sqrts=5000.0, pt_min=30.0, dr_min=0.4),
               "uux_ttx": dict(sqrts=1000.0), "uux_ttxg": dict(sqrts=2000.0, pt_min=30.0, abs_eta_max=2.5),
               "uux_ttxgg": dict(sqrts=3000.0, pt_min=30.0, dr_min=0.4),
               "ug_ttxu": dict(sqrts=2000.0, pt_min=30.0, abs_eta_max=2.5),
               "uxg_ttxux": dict(sqrts=2000.0, pt_min=30.0, abs_eta_max=2.5)}
LABEL = {"gg_ttx": "gg->tt", "gg_ttxg": "gg->ttg", "gg_ttxgg": "gg->ttgg", "gg_ttxggg": "gg->ttggg",
         "uux_ttx": "uu~->tt", "uux_ttxg": "uu~->ttg", "uux_ttxgg": "uu~->ttgg", "ug_ttxu": "ug->ttu",
         "uxg_ttxux": "u~g->ttu~"}


def kernel_info(proc, kernels, precision):
    """(registry name, kernel label, family) of a process's smatrix kernels."""
    from madflow_amd.matrix import load_matrix
    m, _ = load_matrix(proc, kernels=kernels)
    real = "float" if precision == 32 else "double"
    if m.kernel_family == "ordered":
        return m.registry_name, "k_ordered_jamps+k_colour_mfma<%s,%s>" % (m.registry_name, real), "ordered"
    return m.registry_name, "k_smatrix<%s,%s>" % (m.registry_name, real), "fused"


def _kin_text(proc):
    k = ME_ONLY_KIN[proc]
    t = ["sqrt(s)=%g TeV" % (k["sqrts"] / 1000.0)]
    if k.get("pt_min"):
        t.append("pT(g)>%g GeV" % k["pt_min"])
    if k.get("abs_eta_max"):
        t.append("|eta(g)|<%g" % k["abs_eta_max"])
    if k.get("dr_min"):
        t.append("dR(g,g)>%g" % k["dr_min"])
    return ", ".join(t)


def me_only(a, torch):
    """ME evaluations per second of one process's smatrix on a device-resident batch of
    post-cut events (generated by the integrand pipeline on a flat grid): median over
    repeats of the HIP-event time around one smatrix call."""
    from madflow_amd.vegas import MEIntegrand
    proc = a.me_only
    base = proc
    ig = MEIntegrand(base, precision=a.precision, mapping="rambo", seed=3, events_limit=a.batch,
                     **ME_ONLY_KIN[proc])
    ig.chunk(0, a.batch, 0)
    n = ig.last_chunk_npass
    ps = ig._bufs["dense"][:n]
    from madflow_amd.matrix import load_matrix
    m, _ = load_matrix(proc, precision=a.precision, ref_compat=True, kernels=a.kernels)
    params = ig.params if len(m.params) == len(ig.params) else ig.params[:len(m.params)]
    m.smatrix(ps, *params)
    times = []
    for _ in range(a.repeats):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        m.smatrix(ps, *params)
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1) * 1e-3)
    times.sort()
    t = times[len(times) // 2]
    with open(os.path.join(ROOT, "madflow_amd", "processes", "flops_per_me.json")) as fh:
        fl = json.load(fh)[base]
    flops = fl["total"]
    # rows that vanish identically (massless quark lines) are skipped by the kernels: the
    # oracle's count over every row overstates the work done by rows / live rows
    live = fl.get("live_rows", 1) / fl.get("rows", 1)
    peak = PEAK_TFLOPS[a.precision]
    print(json.dumps({"process": proc, "kernels": m.registry_name, "precision": a.precision, "events": n,
                      "repeats": a.repeats,
                      "median_s": t, "me_evals_per_s": n / t, "flops_per_me": flops,
                      "achieved_tflops": flops * n / t / 1e12, "frac_of_peak": flops * n / t / 1e12 / peak,
                      "live_row_fraction": live,
                      "frac_of_peak_live_rows": live * flops * n / t / 1e12 / peak}))
    return 0


def parse_args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nev", type=int, default=16_000_000)
    ap.add_argument("--events-limit", type=int, default=1 << 22)
    ap.add_argument("--process", default="gg_ttxgg", choices=["gg_ttx", "gg_ttxg", "gg_ttxgg", "gg_ttxggg"],
                    help="process of the VEGAS / flow workload (the benchmark metric is quoted on gg_ttxgg)")
    ap.add_argument("--kernels", default=None, choices=["fused", "ordered"],
                    help="kernel family of the matrix element (default: the drop-in class's, the faster form; "
                         "colour-ordered for gg_ttxgg)")
    ap.add_argument("--precision", type=int, default=32, choices=[32, 64])
    ap.add_argument("--mapping", default="diet", choices=["diet", "rambo", "multichannel"],
                    help="phase space of the plain VEGAS leg; multichannel skips the flow leg")
    ap.add_argument("--target", type=float, default=1e-3)
    ap.add_argument("--max-iter", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-flow", action="store_true")
    ap.add_argument("--no-multichannel-leg", dest="multichannel_leg", action="store_false",
                    help="skip the multi-channel FP32 leg of the flow comparison")
    ap.add_argument("--flow-mc-leg", action="store_true",
                    help="add a leg with the learned proposal as a channel of the multi-channel phase space")
    ap.add_argument("--flow-mc-alpha", type=float, default=0.5,
                    help="weight of the learned channel in the multi-channel mixture")
    ap.add_argument("--channel-flow-leg", action="store_true",
                    help="add a multi-channel leg with a flow for each of the --channel-flows channels of largest "
                         "variance share")
    ap.add_argument("--channel-flows", type=int, default=2)
    ap.add_argument("--channel-flow-mix", type=float, default=0.1,
                    help="weight of the uniform component in each channel flow's defensive mixture")
    ap.add_argument("--flow-pairs", type=int, default=5_000_000)
    ap.add_argument("--flow-grid-warmup", type=int, default=3,
                    help="grid adaptation iterations of the flow leg's FP32 grid, counted in its set-up (the plain "
                         "FP64 leg adapts for --warmup iterations)")
    ap.add_argument("--flow-grid-nev", type=int, default=4_000_000,
                    help="points per adaptation iteration of the flow leg's grid (0: --nev, like the plain leg)")
    ap.add_argument("--short-warmup-nev", type=int, default=4_000_000,
                    help="points per warm-up iteration of the second plain FP64 leg (0: no second leg); the first "
                         "follows the reference's schedule, warm-up iterations of --nev points")
    ap.add_argument("--mc-warmup-nev", type=int, default=4_000_000,
                    help="points per warm-up iteration of the multi-channel leg (0: --nev)")
    ap.add_argument("--flow-train-only", action="store_true",
                    help="run only the flow leg's grid adaptation, training pairs and training, print timings, exit")
    ap.add_argument("--flow-steps", type=int, default=20000)
    ap.add_argument("--flow-batch", type=int, default=8192)
    ap.add_argument("--flow-lr", type=float, default=1e-3)
    ap.add_argument("--flow-mix", type=float, default=0.3,
                    help="weight of the uniform (plain VEGAS) component of the defensive mixture q' = (1-m) q + m")
    ap.add_argument("--flow-eval-every", type=int, default=250)
    ap.add_argument("--flow-seed", type=int, default=1, help="initial weights and minibatch draws of the flow")
    ap.add_argument("--flow-patience", type=int, default=1,
                    help="stage 1 stops after this many held-out evaluations without a new best")
    ap.add_argument("--flow-refine-steps", type=int, default=1000)
    ap.add_argument("--flow-refine-rounds", type=int, default=1)
    ap.add_argument("--flow-beta", type=float, default=0.0,
                    help="stage-1 loss: 0 = KL, beta > 0 = alpha-divergence with alpha = 1 + beta")
    ap.add_argument("--flow-refine-beta", type=float, default=3.0, help="the same for the later stages")
    ap.add_argument("--equal-time-s", type=float, default=6.0,
                    help="after reaching the target, keep sampling both methods to this much sampling time")
    ap.add_argument("--equal-wall-s", type=float, default=14.0,
                    help="... and to this much wall time with each method's set-up counted")
    ap.add_argument("--me-only", default=None, metavar="PROC",
                    help="time only the smatrix kernel of PROC on pre-generated post-cut events and exit")
    ap.add_argument("--batch", type=int, default=1 << 22)
    ap.add_argument("--repeats", type=int, default=20)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer 4 times (calibrates HBM byte counters) and exit")
    return ap.parse_args(argv)


def main(argv=None):
    a = parse_args(argv)
    if a.gpus != 1:
        print("bench.py measures one GPU: the event loop shards by Philox counter range, but this build "
              "runs and reports a single MI355X only (--gpus 1)", file=sys.stderr)
        return 2
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    if a.me_only:
        return me_only(a, torch)
    if a.flow_train_only:
        return flow_train_only(a, torch)
    if a.copy_probe:
        src = torch.ones(a.copy_probe << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_per_copy": src.numel() * 4, "copies": 4}))
        return 0
    with open(os.path.join(ROOT, "madflow_amd", "processes", "flops_per_me.json")) as fh:
        flops = json.load(fh)[a.process]["total"]
    reg_name, kernel_label, family = kernel_info(a.process, a.kernels, a.precision)
    m = run_main(a, torch)
    kernel_s = m["me_ms"] * 1e-3
    achieved = flops * m["me_n"] / kernel_s / 1e12
    peak = PEAK_TFLOPS[a.precision]
    me_rate = m["npass"] / m["wall"]
    u = m["unw"]
    log("sigma = %.6g +- %.3g pb (rel %.2e)  ME/s %.4g  kernel ME/s %.4g  unweighted/s steady %.4g  eff %.3g" % (
        u["sigma_pb"], u["err_pb"], u["rel_err"], me_rate, m["me_n"] / kernel_s, u["unweighted_per_s_steady"],
        u["unweighting_efficiency"]))
    res = {
        "metric": METRIC,
        "value": me_rate,
        "unit": "ME evals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": m["wall"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if a.precision == 32 else "f64",
        "data": "synthetic (Philox4x32-10 points, seed 3)",
        "config": {"workload": "%s LO, %s, %s phase space on a "
                               "%d-dim 50-bin VEGAS grid, %d points/iteration, %s amplitudes and colour sum, FP64 "
                               "phase space and VEGAS sums" % (LABEL[a.process], _kin_text(a.process), a.mapping,
                                                               m["ig"].ndim, a.nev,
                                                               "FP32" if a.precision == 32 else "FP64"),
                   "points_per_step": a.nev, "passing_per_step": m["npass"] // a.steps,
                   "matrix_element": "Matrix_1_%s (library kernels %s)" % (a.process, reg_name),
                   "me_kernels": ("colour-ordered (Berends-Giele currents per ordering, MFMA colour sum)"
                                  if family == "ordered" else "fused Feynman-diagram program")},
        "me_kernel_evals_per_s": m["me_n"] / kernel_s,
        "unweighted": u,
        "unweighted_events_per_s": u["unweighted_per_s_to_target"],
        "unweighted_events_per_s_method": "plain VEGAS, %s matrix elements" % ("FP32" if a.precision == 32 else "FP64"),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     "traffic": (HBM_BYTES_PER_ME[(reg_name, a.precision)] * m["me_n"] / max(1, m["launches"])
                                 if (reg_name, a.precision) in HBM_BYTES_PER_ME
                                 and a.mapping == "diet" else None),
                     "traffic_unit": "bytes per launch (PMC)",
                     "kernel": kernel_label,
                     "flops_per_me": flops, "mes_per_launch": m["me_n"] / max(1, m["launches"]),
                     "kernel_ms_per_launch": m["me_ms"] / max(1, m["launches"])},
    }
    if (reg_name, a.precision) in VALU_WAVE_INSTS_PER_ME and kernel_s > 0:
        # executed VALU issue against the issue rate of every SIMD: the oracle-flop fraction
        # above counts the reference algorithm's work, this one the instructions run
        nsimd = 4 * torch.cuda.get_device_properties(0).multi_processor_count
        insts = VALU_WAVE_INSTS_PER_ME[(reg_name, a.precision)] * m["me_n"]
        res["roofline"]["valu_issue_frac"] = insts * VALU_CYCLES_PER_INST / (nsimd * ENGINE_HZ) / kernel_s
        res["roofline"]["valu_wave_insts_per_me"] = VALU_WAVE_INSTS_PER_ME[(reg_name, a.precision)]
    if not a.no_cpu_baseline:
        b = m["ig"]._bufs
        res["cpu_baseline"] = cpu_baseline(a.process, b["dense"], m["ig"].last_chunk_npass,
                                           m["ig"].params,
                                           list(m["ig"].matrix.params))
        res["cpu_baseline"]["c1_end_to_end"] = cpu_c1_end_to_end(res["cpu_baseline"]["cores"])
    if not a.no_flow and a.mapping != "multichannel":
        fc = run_flow_compare(a, torch, m)
        res["flow_comparison"] = fc
        # the product path: FP32 matrix elements with the learned proposal, to the target
        res["unweighted_events_per_s"] = fc["vegas_flow_fp32"]["unweighted_per_s_to_target"]
        res["unweighted_events_per_s_method"] = "VEGAS + learned proposal, FP32 matrix elements (flow_comparison)"
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
