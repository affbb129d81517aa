"""Benchmark: H_eff evaluations/s (+ simulated ns/s) on one MI355X.

Workload: the bulk configuration, 256x256x64 cells of 4 nm (internal order
64x256x256, padded 128x512x512 demag convolution), Ms = 8e5 A/m,
A = 1.3e-11 J/m, Ku = 5e5 J/m^3 along z, alpha = 0.1, seeded random unit m,
FP32.  One "step" = one effective-field evaluation m -> H_eff (demag FFT
convolution + exchange + anisotropy + Zeeman), inputs resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--rk-steps R] [--no-cpu]

Prints one JSON line.  The roofline object is for the dominant pass, timed
with HIP events on the stream the passes run on; the cpu_baseline is the
FP64 CPU oracle (same method, same pass structure) on one evaluation.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "2_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0


def pass_bytes(n, p):
    """Algorithmic HBM bytes per launch of each pass (zero padding not
    counted): inputs and outputs read/written once, 4 B per real, 8 B per
    complex, 24 B per stored octant kernel point (6 components)."""
    n0, n1, n2 = n
    p0, p1, p2 = p
    h2 = p2 // 2 + 1
    N = n0 * n1 * n2
    xrow = 24 * n0 * n1 * h2           # 3 components of the x-transformed rows
    ycol = 24 * n0 * p1 * h2           # 3 components after the y transform
    kern = 24 * (p0 // 2 + 1) * (p1 // 2 + 1) * h2
    if p0 == 1:
        return {"xf": 12 * N + xrow, "y2d": 2 * xrow + kern, "xi": xrow + 24 * N}
    return {"xf": 12 * N + xrow, "yf": xrow + ycol, "z": 2 * ycol + kern,
            "yi": ycol + xrow, "xi": xrow + 24 * N}


def cpu_baseline(cfg):
    """One FP64 H_eff evaluation of the same configuration on the host cores
    (the oracle: same kernel, same zero-padded FFT convolution, same local
    terms), kernel construction excluded from the clock."""
    import numpy as np
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    prob = O.Problem(cfg["n"], cfg["cell"], msat=cfg["msat"], aex=cfg["aex"],
                     alpha=cfg["alpha"], bext=cfg["bext"], ku=cfg["ku"], anis_u=cfg["anis_u"])
    m = prob.normalized(cfg["m"].astype(np.float64).copy())
    prob.heff(m)                       # warm-up (first touch of the FFT buffers)
    # repeat whole evaluations until about 10 s of CPU work have been timed
    n_eval, t0 = 0, time.perf_counter()
    while True:
        prob.heff(m)
        n_eval += 1
        dt = time.perf_counter() - t0
        if dt >= 10.0 or n_eval >= 50:
            break
    return {"value": n_eval / dt, "unit": "H_eff evals/s", "cores": threads, "kind": "port",
            "sample": "%d FP64 H_eff evaluations of the full 256x256x64 grid after 1 warm-up "
                      "(%.1f s), kernel construction excluded" % (n_eval, dt)}


def summarize_profiles(kt_db, fetch_db=None, write_db=None):
    """Markdown summary of rocprofv3 results (rocpd SQLite): per-kernel average
    duration from a --kernel-trace run and, when given, FETCH_SIZE/WRITE_SIZE
    per dispatch from --pmc runs, scaled by what the same counters read for the
    1 GiB copy probe that bench.py runs by default."""
    import collections
    import sqlite3

    def short(name):
        return name.split("(")[0].replace("void ", "")

    c = sqlite3.connect(kt_db)
    rows = list(c.execute("select name, total_calls, total_duration, average, percentage "
                          "from top_kernels order by total_duration desc"))
    lines = ["| kernel | calls | avg duration (us) | share |", "|---|---|---|---|"]
    for name, calls, tot, avg, pct in rows[:16]:
        lines.append("| `%s` | %d | %.1f | %.1f%% |" % (short(name), calls, avg, pct))
    out = "\n".join(lines)

    def counter(db):
        agg = collections.defaultdict(list)
        cc = sqlite3.connect(db)
        for name, val in cc.execute("select k.name, p.counter_value from pmc_events p "
                                    "join kernels k on k.dispatch_id = p.dispatch_id"):
            agg[short(name)].append(val * 1024.0)          # counters are in KiB
        return {k: sum(v) / len(v) for k, v in agg.items()}, agg

    def probe_scale(per_dispatch):
        # the copy probe's 1 GiB copies (the runtime's copy kernel, one
        # dispatch each) are the largest copy dispatches of the run; the
        # benchmark's own small copies are left out
        vals = [v for k, vs in per_dispatch.items() if "copyBuffer" in k for v in vs]
        if not vals or max(vals) <= 0:
            return 1.0
        big = [v for v in vals if v >= max(vals) / 8]
        return float(1 << 30) / (sum(big) / len(big))

    if fetch_db and write_db:
        fmean, fall = counter(fetch_db)
        wmean, wall = counter(write_db)
        fscale = probe_scale(fall)
        wscale = probe_scale(wall)
        lines = ["", "Counter scale factors from the 1 GiB copy probe: FETCH_SIZE x %.3f, "
                 "WRITE_SIZE x %.3f" % (fscale, wscale), "",
                 "| kernel | FETCH_SIZE (MB/dispatch) | WRITE_SIZE (MB/dispatch) | "
                 "scaled read+write (MB) |", "|---|---|---|---|"]
        for k in sorted(fmean, key=lambda k: -fmean[k] - wmean.get(k, 0)):
            if k.startswith("mmx::"):
                f, w = fmean[k], wmean.get(k, 0.0)
                lines.append("| `%s` | %.1f | %.1f | %.1f |" % (k, f / 1e6, w / 1e6,
                                                                (f * fscale + w * wscale) / 1e6))
        out += "\n".join(lines)
    return out


def measured_traffic(pass_name, summary="profiles/demag_passes_bulk_256x256x64.md"):
    """HBM bytes per launch of a pass kernel from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs, scaled by the factors
    that make the 1 GiB copy probe read and write 1 GiB), or None."""
    path = os.path.join(ROOT, summary)
    if not os.path.exists(path):
        return None
    key = "`mmx::%s_fast<" % pass_name
    in_table = False
    for line in open(path):
        if "scaled read+write" in line:
            in_table = True
        elif in_table and line.startswith("| " + key) and line.count("|") == 5:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            try:
                return float(cells[3]) * 1e6
            except ValueError:
                return None
    return None


def _time_evals(sim, steps, warmup):
    import torch
    for _ in range(warmup):
        sim.update("H_eff")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.update("H_eff")
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def run_precision(args):
    """The precision configuration: 1024x1024x32 film of 2 nm cubes, FP64
    Newell kernel built and transformed on the GPU, FP32 convolution and
    stepping, smooth seeded texture.  Reports evals/s and, against the FP64
    oracle (same Newell method, FP64 FFT convolution), the relative error of
    B_demag and H_eff (max and rms over cells, relative to max |field|), of
    the energies (compensated FP64 sums on the GPU) and of <m>."""
    import numpy as np
    import torch
    import mmx
    from mmx import synth

    def say(*a):
        print("[precision]", *a, file=sys.stderr, flush=True)

    cfg = synth.config(3)
    n, cell = cfg["n"], cfg["cell"]
    say("sim", n)
    sim = mmx.Sim(n, cell)
    sim.set_kernel(mmx.MMX_KERNEL_NEWELL)
    synth.setup_sim(sim, cfg)
    t_build = time.perf_counter()
    sim.update("H_eff")
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t_build
    say("kernel built in %.2f s" % t_build)
    steps = max(1, min(args.steps, 100))
    elapsed = _time_evals(sim, steps, max(1, min(args.warmup, 5)))
    ms_per = 1e3 * elapsed / steps
    passes = sim.time_passes(reps=20)
    p = tuple(2 * x if x > 1 else x for x in n)
    nbytes = pass_bytes(n, p)
    dom = max(passes, key=passes.get)
    achieved = nbytes[dom] / (passes[dom] * 1e-3) / 1e9

    h_gpu = sim.getarray("H_eff").astype(np.float64)
    b_gpu = sim.getarray("B").astype(np.float64)
    names = ["E_zeeman", "E_ex", "E_demag", "E"]
    e_gpu = {k: sim.getv(k) for k in names}
    m_avg_gpu = sim.getv("<m>")

    rk = None
    if args.rk_steps > 0:
        nrk = min(args.rk_steps, 200)
        sim.step()
        torch.cuda.synchronize()
        t_start, e0 = sim.getv("t"), sim.getv("evals")
        t0 = time.perf_counter()
        sim.steps(nrk)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        rk = {"steps": nrk, "steps_per_s": nrk / wall,
              "simulated_ns_per_s": (sim.getv("t") - t_start) * 1e9 / wall,
              "evals_per_step": (sim.getv("evals") - e0) / nrk, "m_maxerror": cfg["maxerr"]}

    report, cpu = None, None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        say("GPU done (%.3f ms/eval); oracle Newell kernel on %d threads" % (ms_per, threads))
        t0 = time.perf_counter()
        prob = O.Problem(n, cell, msat=cfg["msat"], aex=cfg["aex"], alpha=cfg["alpha"],
                         bext=cfg["bext"], kernel="newell")
        t_kernel = time.perf_counter() - t0
        say("oracle kernel %.1f s; oracle H_eff" % t_kernel)
        m = prob.normalized(cfg["m"].astype(np.float64).copy())
        t0 = time.perf_counter()
        h, hd, he, _ = prob.heff(m, terms=True)
        t_eval = time.perf_counter() - t0
        say("oracle H_eff %.1f s" % t_eval)
        b = O.MU0 * hd

        def err(a, ref):
            d = np.abs(a - ref)
            s = np.abs(ref).max()
            return {"max_rel": float(d.max() / s), "rms_rel": float(np.sqrt((d * d).mean()) / s)}

        V = float(np.prod(cell))
        ms = cfg["msat"]
        e_ref = {"E_zeeman": -V * ms * sum(m[c].sum() * cfg["bext"][c] for c in range(3)),
                 "E_ex": -0.5 * V * O.MU0 * ms * float((m * he).sum()),
                 "E_demag": -0.5 * V * ms * float((m * b).sum())}
        e_ref["E"] = sum(e_ref.values())
        m_avg = [float(x) for x in m.reshape(3, -1).mean(axis=1)]
        report = {
            "B_demag": err(b_gpu, b), "H_eff": err(h_gpu, h),
            "energies": {k: {"gpu": e_gpu[k], "oracle": e_ref[k],
                             "rel": abs(e_gpu[k] - e_ref[k]) / max(abs(e_ref[k]), 1e-300)}
                         for k in names},
            "<m>_abs": max(abs(a - b_) for a, b_ in zip(m_avg_gpu, m_avg)),
            "oracle_kernel_s": t_kernel,
        }
        cpu = {"value": 1.0 / t_eval, "unit": "H_eff evals/s", "cores": threads, "kind": "port",
               "sample": "1 FP64 H_eff evaluation of the full 1024x1024x32 grid (%.1f s), Newell "
                         "kernel construction (%.1f s) excluded" % (t_eval, t_kernel)}

    return {
        "metric": "H_eff evals/s + simulated ns/s at 256^2x64 FP32; surrogate vs exact, same GPU",
        "value": steps / elapsed, "unit": "H_eff evals/s", "n_gpus": 1, "steps": steps,
        "warmup": max(1, min(args.warmup, 5)), "ms_per_step": ms_per, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "precision film 1024x1024x32 cells of 2 nm, Ms 1e6, A 1.5e-11, "
                               "FP64 Newell kernel built and transformed on the GPU, smooth "
                               "16-mode texture (internal 32x1024x1024, padded 64x2048x2048)",
                   "steps_are": "H_eff evaluations"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": cpu,
        "passes_ms": passes, "pass_bytes": nbytes,
        "eval_algorithmic_GBps": sum(nbytes.values()) / (ms_per * 1e-3) / 1e9,
        "kernel_build_s": t_build,
        "rk12": rk,
        "precision": report,
    }


def run_batch(args):
    """The batching configuration: 64 independent films of seeded uneven sizes
    (32x32x1 .. 256x128x4, cell 4x4x3 nm, seeded Ms, alpha, random m, Zhang-Li
    at 1e12 A/m^2 along x) advanced by the same simulated time, once as one
    batch (one launch per pass for all, device step control, hipGraph
    replays, autotuned tiles) and once one simulation after the other with
    their own engines, on the same GPU."""
    import numpy as np
    import torch
    import mmx
    from mmx import synth

    def say(*a):
        print("[batch]", *a, file=sys.stderr, flush=True)

    cfgs = synth.batch_configs(64)

    def make():
        out = []
        for c in cfgs:
            x = mmx.Sim(c["n"], c["cell"])
            synth.setup_sim(x, c)
            out.append(x)
        return out

    warm, T = 2e-13, 1e-12
    sa = make()
    b = mmx.Batch(sa, tries_per_graph=8)
    t0 = time.perf_counter()
    lds = b.tune(reps=3)
    t_tune = time.perf_counter() - t0
    b.run(warm)
    steps0 = sum(x.getv("step") for x in sa)
    evals0 = sum(x.getv("evals") for x in sa)
    say("tuned %s in %.2f s; timing the batch" % (lds, t_tune))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run(T)
    torch.cuda.synchronize()
    wall_b = time.perf_counter() - t0
    steps_b = sum(x.getv("step") for x in sa) - steps0
    evals_b = sum(x.getv("evals") for x in sa) - evals0

    if args.no_sequential:
        return {"batch_only": True, "wall_s": wall_b, "evals_per_s": evals_b / wall_b,
                "rk_steps": int(steps_b), "tuned_lds_kb": lds}
    sb = make()
    for x in sb:
        x.run(warm)
    steps0 = sum(x.getv("step") for x in sb)
    evals0 = sum(x.getv("evals") for x in sb)
    say("batch %.3f s; timing the simulations one by one" % wall_b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for x in sb:
        x.run(T)
    torch.cuda.synchronize()
    wall_s = time.perf_counter() - t0
    steps_s = sum(x.getv("step") for x in sb) - steps0
    evals_s = sum(x.getv("evals") for x in sb) - evals0
    dm = max(float(np.abs(np.array(x.getv("<m>")) - np.array(y.getv("<m>"))).max())
             for x, y in zip(sa, sb))
    cells = sum(int(np.prod(c["n"])) for c in cfgs)
    info = b.info()
    return {
        "metric": "H_eff evals/s + simulated ns/s at 256^2x64 FP32; surrogate vs exact, same GPU",
        "value": evals_b / wall_b, "unit": "H_eff evals/s summed over 64 simulations",
        "n_gpus": 1, "steps": int(steps_b), "warmup": 0,
        "ms_per_step": 1e3 * wall_b / max(1, steps_b / 64), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batch of 64 films 32x32x1..256x128x4 (seeded sizes, Ms, alpha, "
                               "random m), cell 4x4x3 nm, Zhang-Li 1e12 A/m^2 along x, P 0.56, "
                               "xi 0.05; %d cells in total; each advanced %.0e s" % (cells, T),
                   "steps_are": "accepted RK12 steps, summed over the simulations"},
        "roofline": {"bound": "latency", "achieved": None, "peak": None, "unit": None,
                     "frac": None, "traffic": None,
                     "note": "working set per simulation < 50 MB (Infinity Cache); judged by "
                             "launches per try and the speed-up over one-by-one stepping"},
        "cpu_baseline": None,
        "batch": {"wall_s": wall_b, "rk_steps": int(steps_b), "evals": int(evals_b),
                  "simulated_ns_per_s_summed": 64 * T * 1e9 / wall_b,
                  "launches_per_try": 11, "tries_per_graph": 8, "tuned_lds_kb": lds,
                  "tune_s": t_tune, "blocks_per_pass": {k: v["blocks"] for k, v in info.items()}},
        "one_by_one": {"wall_s": wall_s, "rk_steps": int(steps_s), "evals": int(evals_s),
                       "evals_per_s": evals_s / wall_s},
        "speedup_evals_per_s": (evals_b / wall_b) / (evals_s / wall_s),
        "max_abs_d<m>_batch_vs_one_by_one": dm,
    }


def surrogate_costs(reps=200):
    """The surrogate leg of the default line, bounded to a few seconds: on the
    learned-correction film (256x256x4 cells of 4x4x3 nm) the cost of one
    exact H_eff evaluation against the surrogate's corrector field (2x-coarsened
    exact demag + the 4-layer conv net, seeded weights: the cost does not
    depend on the training), wall-clock timed between device
    synchronisations.  The surrogate replaces one of RK12's two field
    evaluations per step, so the surrogate/exact cost ratio r bounds the
    simulated-time speed-up at 2 / (1 + r) even with a perfect net.  Training and the trajectory comparison against the exact
    solver run as `--workload surrogate`."""
    import torch
    import mmx
    from mmx import synth

    cfg = synth.config(5)
    x = mmx.Sim(cfg["n"], cfg["cell"])
    synth.setup_sim(x, cfg)
    sur = mmx.Surrogate(x, seed=5)
    dm = torch.from_numpy(x.getarray("m")).cuda()
    bb = torch.empty_like(dm)

    def timed(f):
        # the calls run on the simulation's own stream: a device-wide
        # synchronisation on both sides, as for the headline figure
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e6 / reps

    exact = timed(lambda: x.update("H_eff"))
    coarse = timed(lambda: sur.predict_b(dm.data_ptr(), bb.data_ptr(), coarse_only=True))
    full = timed(lambda: sur.predict_b(dm.data_ptr(), bb.data_ptr()))
    x.close()
    return {"workload": "film 256x256x4 cells of 4x4x3 nm (learned-correction configuration)",
            "exact_eval_us": exact, "coarse_field_us": coarse, "coarse_plus_net_us": full,
            "cost_ratio_surrogate_over_exact": full / exact,
            "speedup_bound_simulated_ns_per_s": 2.0 / (1.0 + full / exact),
            "timed": "%d calls each, wall clock between device synchronisations, after "
                     "one warm-up" % reps,
            "trajectory_comparison": "python bench.py --workload surrogate (training on the "
                                     "GPU, then 1 ns exact and surrogate-assisted runs)"}


def run_film(args):
    """The plain method on the small film (the reference's own CPU-runnable
    case): 128x32x1 cells of 3.90625x3.90625x3 nm, Ms 8e5, A 1.3e-11, alpha
    0.02, seeded random m, 10 mT along a seeded in-plane direction, demag +
    exchange, RK12 at m_maxerror 1e-5, FP32, run 1 ns with <m> every 10 ps;
    reports simulated ns/s, steps and the final H_eff against the FP64 oracle
    at the same m.  cpu_baseline: the oracle's RK12 (FP64, same controller)
    from the same state for about 10 s."""
    import numpy as np
    import torch
    import mmx
    from mmx import synth
    from oracle import oracle as O

    cfg = synth.config(1)
    n, cell = cfg["n"], cfg["cell"]
    s = mmx.Sim(n, cell)
    synth.setup_sim(s, cfg)
    s.setarray("m", cfg["m"])
    s.update("H_eff")                  # kernel construction
    torch.cuda.synchronize()
    t0, step0 = s.getv("t"), s.getv("step")
    table = []
    ts = time.perf_counter()
    for q in range(1, 101):            # <m> every 10 ps
        s.run(max(0.0, t0 + q * 1e-11 - s.getv("t")))
        table.append(s.getv("<m>"))
    torch.cuda.synchronize()
    wall = time.perf_counter() - ts
    steps = s.getv("step") - step0
    sim_ns = (s.getv("t") - t0) * 1e9
    m_end = s.getarray("m").astype(np.float64)
    h_gpu = s.getarray("H_eff").astype(np.float64)
    prob = O.Problem(n, cell, msat=cfg["msat"], aex=cfg["aex"], alpha=cfg["alpha"],
                     bext=cfg["bext"])
    h_orc = prob.heff(m_end)
    h_err = float(np.abs(h_gpu - h_orc).max() / np.abs(h_orc).max())
    s.close()

    cpu = None
    if not args.no_cpu:
        mo = cfg["m"].astype(np.float64).copy()
        prob.set("maxerr", cfg["maxerr"])
        prob.step(mo)                  # warm-up
        tc0, sc0 = prob.get("t"), prob.get("step")
        tw0 = time.perf_counter()
        while time.perf_counter() - tw0 < 10.0:
            prob.step(mo)
        wc = time.perf_counter() - tw0
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": (prob.get("t") - tc0) * 1e9 / wc, "unit": "simulated ns/s",
               "cores": threads, "kind": "port",
               "sample": "%d accepted FP64 RK12 steps of the oracle from the same state (%.1f s)"
                         % (prob.get("step") - sc0, wc)}
    return {
        "metric": "H_eff evals/s + simulated ns/s at 256^2x64 FP32; surrogate vs exact, same GPU",
        "value": sim_ns / wall, "unit": "simulated ns/s", "n_gpus": 1, "steps": int(steps),
        "warmup": 1, "ms_per_step": 1e3 * wall / max(1, steps), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "film 128x32x1 cells of 3.90625x3.90625x3 nm, Ms 8e5, A 1.3e-11, "
                               "alpha 0.02, seeded random m, 10 mT in-plane, demag + exchange, "
                               "RK12 m_maxerror 1e-5, 1 ns with <m> every 10 ps",
                   "steps_are": "accepted RK12 steps"},
        "cpu_baseline": cpu,
        "simulated_ns": sim_ns, "wall_s": wall, "us_per_step": 1e6 * wall / max(1, steps),
        "final_<m>": table[-1], "<m>_every_100ps": table[9::10],
        "final_H_eff_rel_err_vs_fp64_oracle": h_err,
    }


def run_surrogate_kernels(args):
    """The surrogate's kernels alone at the learned-correction configuration's
    size (film 256x256x4), for profiling them: `--steps` training steps (the
    exact target field, net forward, backward, weight gradients, Adam) and as
    many corrector-field predictions (coarse field + net) on a seeded texture,
    each timed over a device synchronisation."""
    import torch
    import mmx
    from mmx import synth

    cfg = synth.config(5)
    n, cell = cfg["n"], cfg["cell"]
    x = mmx.Sim(n, cell)
    synth.setup_sim(x, cfg)
    sur = mmx.Surrogate(x, seed=5)
    x.setarray("m", synth.training_texture(1, n, cell))
    m_ptr = x.device_array("m")
    dm = torch.from_numpy(x.getarray("m")).cuda()
    bb = torch.empty_like(dm)
    sur.train_step(m_ptr, 1e-4)
    sur.predict_b(dm.data_ptr(), bb.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sur.train_step(m_ptr, 1e-4)
    torch.cuda.synchronize()
    t_train = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sur.predict_b(dm.data_ptr(), bb.data_ptr())
    torch.cuda.synchronize()
    t_pred = (time.perf_counter() - t0) / args.steps
    return {"workload": "surrogate kernels, film 256x256x4 cells of 4x4x3 nm", "steps": args.steps,
            "train_step_us": 1e6 * t_train, "predict_us": 1e6 * t_pred,
            "train_step_includes": "exact target field, forward, backward-data, weight "
                                   "gradients, Adam"}


def run_surrogate(args):
    """The learned-correction configuration: 256x256x4 film of 4x4x3 nm cells,
    Ms 8e5, A 1.3e-11, alpha 0.02, seeded random m.  The net is trained on the
    GPU on pairs (m, exact field) from exact trajectories of seeded vortex,
    domain-wall and random textures; then the same state is run for the same
    simulated time with the exact solver and with the surrogate in the RK12
    corrector (exact first stage, judge every 8 steps, fallback above a 1e-3
    relative residual).  Reports simulated ns/s of both and |d<m>| at the end,
    and the fine-field error of the coarse field with and without the net on
    unseen textures."""
    import numpy as np
    import torch
    import mmx
    from mmx import synth
    import mmx.train as T

    def say(*a):
        print("[surrogate]", *a, file=sys.stderr, flush=True)

    cfg = synth.config(5)
    n, cell = cfg["n"], cfg["cell"]

    def make():
        x = mmx.Sim(n, cell)
        synth.setup_sim(x, cfg)
        return x

    trainer = make()
    sur = mmx.Surrogate(trainer, seed=5)
    t0 = time.perf_counter()
    hist = T.train(sur, trainer, nsamples=args.train_samples, steps_per_sample=2,
                   samples_per_trajectory=100, lr=5e-4, lr_final=5e-5, alpha=0.1, log=say)
    t_train = time.perf_counter() - t0
    say("trained on %d pairs in %.1f s, loss %.3e -> %.3e" % (
        len(hist), t_train, np.mean([l for _, l in hist[:50]]), np.mean([l for _, l in hist[-50:]])))

    # gain of the net on unseen textures
    plan = mmx.DemagPlan(n, cell)
    gains = []
    for k in range(5):
        trainer.setarray("m", synth.training_texture(100003 + 17 * k, n, cell))
        trainer.steps(20)
        dm = torch.from_numpy(trainer.getarray("m")).cuda()
        ex, bn, bco = torch.empty_like(dm), torch.empty_like(dm), torch.empty_like(dm)
        plan.update_b(ex.data_ptr(), dm.data_ptr(), cfg["msat"])
        sur.predict_b(dm.data_ptr(), bn.data_ptr())
        sur.predict_b(dm.data_ptr(), bco.data_ptr(), coarse_only=True)
        torch.cuda.synchronize()
        e = ex.cpu().numpy().astype(np.float64)
        en = float(np.linalg.norm(bn.cpu().numpy() - e) / np.linalg.norm(e))
        ec = float(np.linalg.norm(bco.cpu().numpy() - e) / np.linalg.norm(e))
        gains.append({"coarse_only": ec, "with_net": en})

    def run(use_sur, tol=1e-3, m0=None, bext=None, alpha=None, pert=0.0, T=None):
        x = make()
        if m0 is not None:
            x.setarray("m", m0)
        if bext is not None:
            x.setv("B_ext", bext)
        if alpha is not None:
            x.setv("alpha", alpha)
        if pert:
            m = x.getarray("m").astype(np.float64)
            m += pert * np.random.default_rng(1).normal(size=m.shape)
            x.setarray("m", (m / np.sqrt((m * m).sum(axis=0))).astype(np.float32))
        s2 = None
        if use_sur:
            s2 = mmx.Surrogate(x, seed=5)
            s2.set_params(sur.get_params())
            x.set_surrogate(s2, tol=tol, check_every=8)
        x.step()
        torch.cuda.synchronize()
        ts = time.perf_counter()
        T_ = T if T is not None else args.sim_time
        t_start = x.getv("t")
        for q in range(1, 11):        # in tenths, with a progress line each
            x.run(max(0.0, t_start + q * T_ / 10 - x.getv("t")))
            say("  t %.3e s, %d steps, %.1f s wall" % (x.getv("t"), x.getv("step"),
                                                       time.perf_counter() - ts))
        torch.cuda.synchronize()
        wall = time.perf_counter() - ts
        out = {"wall_s": wall, "simulated_ns_per_s": x.getv("t") * 1e9 / wall,
               "steps": x.getv("step"), "<m>": x.getv("<m>")}
        if use_sur:
            out.update({k: x.getv("surrogate_" + k) for k in ("on", "residual", "checks",
                                                              "fallbacks", "evals")})
            out["tolerance"] = tol
        return out

    def dmax(a, b):
        return float(np.abs(np.array(a["<m>"]) - np.array(b["<m>"])).max())

    # (a) the stated case: seeded random m, no field
    ex_run = run(False)
    say("exact: %s" % ex_run)
    chaos = run(False, pert=1e-7)
    say("exact, m perturbed by 1e-7: |d<m>| %.3e" % dmax(ex_run, chaos))
    su_run = run(True)
    say("surrogate (tol 1e-3): %s" % su_run)
    dm = dmax(ex_run, su_run)
    # (b) deterministic dynamics: a relaxed vortex gyrating under 5 mT in-plane
    relax = make()
    relax.setarray("m", synth.vortex_m(77, n, cell))
    relax.setv("alpha", 0.5)
    for q in range(5):
        relax.run(4e-11)
        say("relaxing the vortex: t %.2e, %d steps" % (relax.getv("t"), relax.getv("step")))
    mv = relax.getarray("m")
    bx = tuple(synth.to_internal((5e-3, 0.0, 0.0)))
    gy = {"exact": run(False, m0=mv, bext=bx)}
    gy["surrogate_tol_1e-3"] = run(True, m0=mv, bext=bx)
    gy["surrogate_tol_0.5"] = run(True, tol=0.5, m0=mv, bext=bx)
    for k in ("surrogate_tol_1e-3", "surrogate_tol_0.5"):
        gy[k]["max_abs_d<m>_vs_exact"] = dmax(gy["exact"], gy[k])
        gy[k]["speedup_simulated_ns_per_s"] = (gy[k]["simulated_ns_per_s"] /
                                               gy["exact"]["simulated_ns_per_s"])
    say("vortex gyration: %s" % gy)
    # field evaluations alone
    evals = {}
    for name, f in (("exact", lambda x: x.update("H_eff")),):
        x = make()
        f(x)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(200):
            f(x)
        torch.cuda.synchronize()
        evals[name + "_eval_us"] = (time.perf_counter() - t1) / 200 * 1e6
    x = make()
    dmm = torch.from_numpy(x.getarray("m")).cuda()
    bb = torch.empty_like(dmm)
    for mode, key in ((True, "coarse_field_us"), (False, "coarse_plus_net_us")):
        sur.predict_b(dmm.data_ptr(), bb.data_ptr(), coarse_only=mode)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(200):
            sur.predict_b(dmm.data_ptr(), bb.data_ptr(), coarse_only=mode)
        torch.cuda.synchronize()
        evals[key] = (time.perf_counter() - t1) / 200 * 1e6
    say("evaluation costs %s" % evals)

    return {
        "metric": "H_eff evals/s + simulated ns/s at 256^2x64 FP32; surrogate vs exact, same GPU",
        "value": su_run["simulated_ns_per_s"], "unit": "simulated ns/s (surrogate-assisted RK12)",
        "n_gpus": 1, "steps": int(su_run["steps"]), "warmup": 1,
        "ms_per_step": 1e3 * su_run["wall_s"] / max(1, su_run["steps"]),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "learned demag correction: film 256x256x4 cells of 4x4x3 nm, Ms 8e5, "
                               "A 1.3e-11, alpha 0.02, seeded random m; 2x-coarsened exact demag + "
                               "4-layer 3x3 conv net (width 32, BF16 MFMA) trained on %d pairs from "
                               "exact trajectories; corrector stage on the surrogate, exact first "
                               "stage, judge every 8 steps, fallback above 1e-3" % len(hist),
                   "steps_are": "accepted RK12 steps", "simulated_time_s": args.sim_time},
        "roofline": {"bound": "latency", "achieved": None, "peak": None, "unit": None,
                     "frac": None, "traffic": None,
                     "note": "262k cells: each pass moves < 40 MB; judged by the speed-up over "
                             "the exact path on the same GPU"},
        "cpu_baseline": None,
        "exact": ex_run, "surrogate": su_run,
        "speedup_simulated_ns_per_s": su_run["simulated_ns_per_s"] / ex_run["simulated_ns_per_s"],
        "max_abs_d<m>_after_run": dm,
        "chaos_baseline_exact_vs_exact_perturbed_1e-7": dmax(ex_run, chaos),
        "vortex_gyration_5mT": gy,
        "evaluation_costs": evals,
        "training": {"pairs": len(hist), "seconds": t_train,
                     "loss_first50": float(np.mean([l for _, l in hist[:50]])),
                     "loss_last50": float(np.mean([l for _, l in hist[-50:]]))},
        "field_error_unseen_textures": gains,
    }


def run_demagexch(args):
    """tests/fftbench.py of the reference: 1024x1024x1 cells of 5 nm,
    micromagnetism + demagexch, Msat 800e3, Aex 1.3e-11, m = (1, 1, 0); H_dex
    forced to update 100 times (debug_invalidate + debug_update).  Also H_eff
    with and without demagexch loaded (the module is folded into the demag
    convolution, so both should take the same time)."""
    import numpy as np
    import torch
    import mmx

    def say(*a):
        print("[demagexch]", *a, file=sys.stderr, flush=True)

    n, cell, reps = (1, 1024, 1024), (5e-9, 5e-9, 5e-9), 100
    m = np.zeros((3,) + n, np.float32)
    m[1] = m[2] = 1.0 / np.sqrt(2.0)     # user (1, 1, 0) in internal (z, y, x) order

    def make(dex):
        s = mmx.Sim(n, cell)
        s.setv("Msat", 800e3)
        s.setv("Aex", 1.3e-11)
        s.setv("alpha", 1.0)
        s.setv("dt", 1e-12)
        if dex:
            s.load("demagexch")
        s.setarray("m", m)
        return s

    def timed(s, name):
        for _ in range(3):
            s.update(name)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            s.update(name)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    t0 = time.perf_counter()
    sd = make(True)
    sd.update("H_dex")
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    t_hdex = timed(sd, "H_dex")
    t_heff_dex = timed(sd, "H_eff")
    sd.close()
    sp = make(False)
    t_heff = timed(sp, "H_eff")
    sp.close()
    say("H_dex %.1f us, H_eff with demagexch %.1f us, without %.1f us" %
        (1e6 * t_hdex, 1e6 * t_heff_dex, 1e6 * t_heff))
    return {
        "metric": "H_dex updates/s (tests/fftbench.py: 1024x1024x1 demagexch, 100 updates)",
        "value": 1.0 / t_hdex, "unit": "H_dex updates/s", "n_gpus": 1, "steps": reps,
        "warmup": 3, "ms_per_step": 1e3 * t_hdex, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "demagexch film 1024x1024x1 cells of 5 nm (padded 1x2048x2048), "
                               "Msat 800e3, Aex 1.3e-11, m = (1, 1, 0)",
                   "steps_are": "H_dex updates (one convolution with the combined kernel)"},
        "kernels_build_s": t_build,
        "heff_ms_with_demagexch": 1e3 * t_heff_dex, "heff_ms_without": 1e3 * t_heff,
    }


def run_masked(args):
    """The 256x256x64 bulk workload with per-cell terms, one configuration at
    a time: uniform parameters (the headline), a cylinder geometry as an Msat
    mask (256-cell disk through all 64 layers, zero outside), the same with an
    Aex map on top, the same with B_ext masked to an 8-cell stripe across x
    (an antenna), uniform parameters with the Zhang-Li torque, and with the
    Brown thermal field at 300 K (redrawn every RK12 step).  H_eff
    evaluations/s, per-pass times and adaptive RK12 steps/s of each; the xi
    pass is where the configurations differ (its three epilogue variants)."""
    import numpy as np
    import torch
    import mmx
    from mmx import synth

    def say(*a):
        print("[masked]", *a, file=sys.stderr, flush=True)

    cfg = synth.config(2)
    n = cfg["n"]
    _, y, x = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    r2 = (x - (n[2] - 1) / 2.0) ** 2 + (y - (n[1] - 1) / 2.0) ** 2
    disk = (r2 <= (n[2] / 2.0) ** 2).astype(np.float32)
    aex = np.random.default_rng(5).uniform(0.5, 1.5, size=n).astype(np.float32)
    rows = {}
    stripe = np.zeros((3,) + n, np.float32)
    stripe[:, :, :, n[2] // 2 - 4:n[2] // 2 + 4] = 1.0
    for case in ("uniform", "msat_disk", "msat_disk_aex_map", "msat_disk_bext_stripe",
                 "zhang_li", "thermal_300K"):
        sim = mmx.Sim(n, cfg["cell"])
        synth.setup_sim(sim, cfg)
        if case.startswith("msat_disk"):
            sim.setmask("Msat", disk, 1)
        if case == "msat_disk_aex_map":
            sim.setmask("Aex", aex, 1)
        if case == "msat_disk_bext_stripe":
            sim.setmask("B_ext", stripe, 3)
        if case == "zhang_li":
            sim.load("zhang-li")
            sim.setv("j", (1e12, 0.0, 0.0))
        if case == "thermal_300K":
            sim.load("temperature/brown")
            sim.setv("Temp", 300.0)
        sim.update("H_eff")
        elapsed = _time_evals(sim, args.steps, args.warmup)
        passes = sim.time_passes(reps=max(10, min(args.steps, 50)))
        rk = per_step = None
        if args.rk_steps > 0:
            sim.setv("t", 0.0)
            sim.setv("dt", 1e-15)
            sim.step()
            torch.cuda.synchronize()
            e0 = sim.getv("evals")
            t0 = time.perf_counter()
            sim.steps(args.rk_steps)
            torch.cuda.synchronize()
            rk = args.rk_steps / (time.perf_counter() - t0)
            per_step = (sim.getv("evals") - e0) / args.rk_steps
        sim.close()
        rows[case] = {"evals_per_s": args.steps / elapsed,
                      "passes_us": {k: round(v * 1e3, 1) for k, v in passes.items()},
                      "rk12_steps_per_s": rk, "rk12_evals_per_step": per_step}
        say(case, "%.1f evals/s" % rows[case]["evals_per_s"], rows[case]["passes_us"],
            "rk12 %s steps/s, %s evals/step" % (None if rk is None else round(rk, 1),
                                                None if per_step is None else round(per_step, 2)))
    head = rows["msat_disk"]
    return {
        "metric": "H_eff evals/s at 256^2x64 FP32 with an Msat-masked cylinder geometry",
        "value": head["evals_per_s"], "unit": "H_eff evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 / head["evals_per_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "bulk 256x256x64 cells, 4 nm, Ku along z, seeded random m; "
                               "Msat mask: a 256-cell disk through all 64 layers (zero outside)",
                   "steps_are": "H_eff evaluations"},
        "configurations": rows,
    }


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize":
        print(summarize_profiles(*sys.argv[2:5]))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--rk-steps", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("bulk", "film", "precision", "batch", "surrogate",
                                           "surrogate-kernels", "demagexch", "masked"),
                    default="bulk",
                    help="bulk: 256x256x64 (the headline line); film: 128x32x1 RK12 for 1 ns "
                         "beside the oracle's RK12; precision: 1024x1024x32 film with the FP64 "
                         "Newell kernel and per-field errors against the FP64 oracle; batch: 64 "
                         "uneven films; surrogate: training and surrogate-assisted runs; "
                         "surrogate-kernels: the surrogate's kernels alone; demagexch; masked: "
                         "the bulk grid with an Msat-masked cylinder, an Aex map, Zhang-Li, "
                         "the thermal field")
    ap.add_argument("--train-samples", type=int, default=20000,
                    help="surrogate workload: training pairs from exact trajectories")
    ap.add_argument("--sim-time", type=float, default=1e-9,
                    help="surrogate workload: simulated time of the exact and surrogate runs")
    ap.add_argument("--no-sequential", action="store_true",
                    help="batch workload: skip the one-by-one comparison run")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="leave out the surrogate evaluation-cost leg")
    ap.add_argument("--no-copy-probe", action="store_true",
                    help="skip the 1 GiB device-to-device copy whose bandwidth is reported "
                         "next to the HBM peak (and calibrates profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the reference's FFT is single-device, "
              "fftplanX.go:36-39); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import mmx
    from mmx import synth

    mmx.check(mmx.lib().mmx_init(0))
    if args.workload == "precision":
        print(json.dumps(run_precision(args)))
        return
    if args.workload == "batch":
        print(json.dumps(run_batch(args)))
        return
    if args.workload == "surrogate":
        print(json.dumps(run_surrogate(args)))
        return
    if args.workload == "film":
        print(json.dumps(run_film(args)))
        return
    if args.workload == "surrogate-kernels":
        print(json.dumps(run_surrogate_kernels(args)))
        return
    if args.workload == "demagexch":
        print(json.dumps(run_demagexch(args)))
        return
    if args.workload == "masked":
        print(json.dumps(run_masked(args)))
        return
    cfg = synth.config(2)
    sim = mmx.Sim(cfg["n"], cfg["cell"])
    synth.setup_sim(sim, cfg)

    # warm-up (also builds the demag kernel)
    t_build = time.perf_counter()
    sim.update("H_eff")
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t_build
    for _ in range(args.warmup):
        sim.update("H_eff")
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.update("H_eff")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per = 1e3 * elapsed / args.steps

    # per-pass clock (HIP events on the sim's stream)
    passes = sim.time_passes(reps=max(10, min(args.steps, 50)))
    p = tuple(2 * x if x > 1 else x for x in cfg["n"])
    nbytes = pass_bytes(cfg["n"], p)
    dom = max(passes, key=passes.get)
    achieved = nbytes[dom] / (passes[dom] * 1e-3) / 1e9
    total_bytes = sum(nbytes.values())
    eval_gbs = total_bytes / (ms_per * 1e-3) / 1e9

    # adaptive RK12: simulated time per wall time
    sim_ns = None
    rk_rate = None
    if args.rk_steps > 0:
        sim.setarray("m", cfg["m"])
        sim.setv("t", 0.0)
        sim.setv("dt", 1e-15)
        sim.step()
        torch.cuda.synchronize()
        t_start = sim.getv("t")
        e0 = sim.getv("evals")
        t0 = time.perf_counter()
        sim.steps(args.rk_steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        sim_ns = (sim.getv("t") - t_start) * 1e9 / wall
        rk_rate = args.rk_steps / wall
        evals_per_step = (sim.getv("evals") - e0) / args.rk_steps

    copy_probe = None
    if not args.no_copy_probe:
        # measured stream-copy bandwidth on this box: 1 GiB read + 1 GiB written
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        reps = 5
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(reps):
            dst.copy_(src)
        ev1.record()
        torch.cuda.synchronize()
        sec = ev0.elapsed_time(ev1) * 1e-3 / reps
        copy_probe = {"bytes_read": src.numel() * 4, "bytes_written": dst.numel() * 4,
                      "GBps": 2 * src.numel() * 4 / sec / 1e9,
                      "frac_of_peak": 2 * src.numel() * 4 / sec / 1e9 / HBM_PEAK_GBS,
                      "timed": "%d copies by HIP events after one warm-up" % reps}
        del src, dst

    sur = None if args.no_surrogate else surrogate_costs()

    cpu = None if args.no_cpu else cpu_baseline(cfg)

    out = {
        "metric": "H_eff evals/s + simulated ns/s at 256^2x64 FP32; surrogate vs exact, same GPU",
        "value": args.steps / elapsed,
        "unit": "H_eff evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "bulk 256x256x64 cells, 4 nm, Ku along z, seeded random m "
                               "(internal 64x256x256, padded 128x512x512)",
                   "steps_are": "H_eff evaluations",
                   "exact_path_only": True},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom),
                     "traffic_source": "profiles/demag_passes_bulk_256x256x64.md (FETCH_SIZE x2 + "
                                       "WRITE_SIZE x1 per dispatch)"},
        "cpu_baseline": cpu,
        "passes_ms": passes,
        "pass_bytes": nbytes,
        "eval_algorithmic_GBps": eval_gbs,
        "eval_hbm_frac": eval_gbs / HBM_PEAK_GBS,
        "rk12": {"steps": args.rk_steps, "steps_per_s": rk_rate, "simulated_ns_per_s": sim_ns,
                 "evals_per_step": evals_per_step if args.rk_steps > 0 else None,
                 "m_maxerror": cfg["maxerr"]},
        "kernel_build_s": t_build,
        "copy_probe": copy_probe,
        "surrogate": sur,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
