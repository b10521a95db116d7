"""Benchmark: Lanczos iterations/s and H*v bandwidth on one MI355X.

Workload (default): nstates=3, nmodes=6, basis [12,10,10,8,8,6] (N=1,382,400),
ordr=2, niter=2000, partial re-orthogonalisation (reorthog=1, chkorthog=100),
seeded synthetic couplings.  One step = one whole Lanczos job: start vector,
2000 iterations (H*v, recurrence, omega recurrence, re-orthogonalisation) and
the tridiagonal eigensolve with intensities.  The mixed-precision job (H*v
FP32, recurrence FP64) is the headline; the FP64 job is timed beside it and
the converged eigenvalues of the two are compared (cm^-1).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Lanczos iter/s + H*v HBM GB/s (%peak); mixed vs FP64 at <=1e-3 cm^-1 error"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 peak
AU2WAVE = 219474.6313705


def run_job(eng, niter, prec, reorthog, chkorthog, seed):
    eng.lanczos_init(niter=niter, reorthog=reorthog, prec=prec, chkorthog=chkorthog, seed=seed)
    eng.run(1, niter)
    a, b, _ = eng.coefficients()
    ev, it, bj, _ = eng.ritz(a, b)
    return a, b, ev, it, bj


def timed(eng, args, prec, sync):
    """W untimed jobs, K timed jobs (captured iteration graphs), then one
    more job with every H*v launch bracketed by HIP events on the library's
    stream for the kernel's average duration."""
    for k in range(args.warmup):
        t0 = time.perf_counter()
        run_job(eng, args.niter, prec, args.reorthog, args.chkorthog, args.seed)
        print(f"# {prec} warmup {k + 1}/{args.warmup}: {time.perf_counter() - t0:.2f} s", file=sys.stderr,
              flush=True)
    sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        out = run_job(eng, args.niter, prec, args.reorthog, args.chkorthog, args.seed)
        print(f"# {prec} step {k + 1}/{args.steps} done at {time.perf_counter() - t0:.2f} s", file=sys.stderr,
              flush=True)
    sync()
    dt = time.perf_counter() - t0
    pro = eng.pro_stats()
    eng.profile(True)
    eng.profile_read()
    run_job(eng, args.niter, prec, args.reorthog, args.chkorthog, args.seed)
    hv_ms, hv_n = eng.profile_read()
    eng.profile(False)
    print(f"# {prec} profiled job: H*v {hv_ms / max(hv_n, 1) * 1e3:.1f} us x {hv_n}", file=sys.stderr, flush=True)
    return dt, hv_ms, hv_n, out, pro


def compare(ev64, it64, bj64, ev, it, bj, tol_conv=1e-9):
    """Converged FP64 roots matched to the nearest root of the other run."""
    conv = np.nonzero(bj64 < tol_conv)[0]
    dE, dI = [], []
    for k in conv:
        j = int(np.argmin(np.abs(ev - ev64[k])))
        dE.append(abs(ev[j] - ev64[k]) * AU2WAVE)
        dI.append(abs(it[j] - it64[k]))
    return (max(dE) if dE else float("nan")), (max(dI) if dI else float("nan")), len(conv)


def cpu_baseline(prob, seconds, seed=12345):
    """The oracle (scalar FP64 restatement) on this host's CPU, 1 thread, on
    the first iterations of the same job; also returns its coefficients."""
    from oracle.oracle import Oracle
    orc = Oracle(prob)
    t0 = time.perf_counter()
    orc.hv(np.ones(orc.n))
    t_hv = time.perf_counter() - t0
    k = max(2, int(seconds / max(t_hv * 1.1, 1e-6)))
    t0 = time.perf_counter()
    r = orc.lanczos(k, reorthog=prob.reorthog, chkorthog=prob.chkorthog, seed=seed)
    dt = time.perf_counter() - t0
    return dict(value=k / dt, unit="Lanczos iter/s", cores=1, kind="port",
                sample=f"first {k} Lanczos iterations of the same workload (FP64 oracle, 1 thread), {dt:.1f} s"), r




def batched_cfg5(sync, cpu_seconds):
    """The uneven 64-problem batch (400 iterations each, FP32 H*v, FP64
    recurrence): one batched run against the same problems run one after the
    other through single contexts with the same generic kernels, and the
    FP64 oracle on the host over a bounded sample."""
    from nadvibs_amd.batch import Batch
    from nadvibs_amd.engine import Engine
    from nadvibs_amd.synth import batch_workloads
    probs = batch_workloads()
    niter = probs[0].niter
    total = sum(p.n for p in probs)
    with Batch(probs) as bt:
        bt.init(niter, "f32")
        rep = bt.tune()
        bt.run(1, 2)
        bt.coefficients(0)
        bt.init(niter, "f32")
        sync()
        t0 = time.perf_counter()
        bt.run()
        sync()
        a0, _ = bt.coefficients(0)
        t_batch = time.perf_counter() - t0
    print(f"# batched cfg5: {t_batch * 1e3:.1f} ms ({rep})", file=sys.stderr, flush=True)
    engs = []
    for p in probs:
        e = Engine(p)
        e.select_hv_kernel(True)
        e.lanczos_init(niter, prec="f32")
        e.run(1, 2)   # capture the iteration graph
        e.coefficients()
        engs.append(e)

    def one_by_one(wait_each):
        for e in engs:
            e.lanczos_init(niter, prec="f32")
        sync()
        t0 = time.perf_counter()
        for e in engs:
            e.run(1, niter)
            if wait_each:
                e.coefficients()
        for e in engs:
            e.coefficients()
        return time.perf_counter() - t0

    t_seq = one_by_one(True)     # one problem after the other
    t_streams = one_by_one(False)  # every problem on its own stream, all enqueued at once
    a1, _, _ = engs[0].coefficients()
    for e in engs:
        e.close()
    print(f"# cfg5 one by one: {t_seq * 1e3:.1f} ms, concurrent streams: {t_streams * 1e3:.1f} ms",
          file=sys.stderr, flush=True)
    # host: the oracle over all problems for k iterations, k doubled until the sample lasts cpu_seconds / 3
    from oracle.oracle import Oracle
    k, dt = 1, 0.0
    while True:
        t0 = time.perf_counter()
        for p in probs:
            Oracle(p).lanczos(k)
        dt = time.perf_counter() - t0
        if dt > cpu_seconds / 3 or k >= niter:
            break
        k *= 2
    return {"workload": "cfg5: 64 problems, nstates U{2,3}, nmodes U{4..6}, basis U{4..12}, ordr 2, "
                        f"total N={total}", "niter": niter, "hv_precision": "f32",
            "batched_s": round(t_batch, 4), "sequential_s": round(t_seq, 4), "streams_s": round(t_streams, 4),
            "speedup_vs_sequential": round(t_seq / t_batch, 2), "speedup_vs_streams": round(t_streams / t_batch, 2),
            "problem_iter_per_s": round(len(probs) * niter / t_batch, 1),
            "alpha1_agree": bool(abs(a0[0] - a1[0]) <= 1e-12 * abs(a1[0])), "tune": rep,
            "cpu_baseline": {"value": round(len(probs) * k / dt, 2), "unit": "problem-iterations/s", "cores": 1,
                             "kind": "port", "sample": f"all 64 problems, {k} iterations each (FP64 oracle), "
                                                       f"{dt:.1f} s"}}


def adjoint_cfg4(sync):
    """cfg4: the cfg2 operator, 500 iterations; loss |G*sticks - target|_2
    with a Gaussian of sigma 15 cm^-1 on 4096 points above the lowest root;
    target = the spectrum of the couplings scaled by 1 + N(0, 0.05).
    Gradient with respect to every nadvibs.in coefficient and frequency
    through the Krylov-projected adjoint, in FP64 and with FP32 H*v, for two
    runs: without re-orthogonalisation, the basis recomputed from checkpoints
    every 50 iterations, and with partial re-orthogonalisation (kept basis).
    Each over the 0-6000 cm^-1 window and over the low window whose roots have
    converged by 500 iterations (there the two precisions compute the same
    loss, so their gradients can be compared)."""
    import torch
    from nadvibs_amd.adjoint import spectrum_of_parameters
    from nadvibs_amd.synth import workload
    p = workload("cfg2")
    niter = 500
    rng = np.random.default_rng(1004)
    w0 = torch.tensor(p.aomega, dtype=torch.float64)
    c0 = torch.tensor(p.concoef, dtype=torch.float64)
    pt0 = torch.tensor(p.potterms, dtype=torch.float64)
    scale = 1 + 0.05 * torch.from_numpy(rng.standard_normal(pt0.shape))
    res = {"workload": "cfg4: cfg2 operator, 500 iterations, sigma 15 cm^-1, 4096 points",
           "gradient": "Krylov-projected (nv_spectrum_adjoint + nv_krylov_gradient), generic H*v"}
    runs = (("checkpoints_50", dict(reorthog=0, checkpoint=50)), ("pro_kept_basis", dict(reorthog=1)))
    for label, hi in (("window_0_6000", 6000.0), ("window_0_1500", 1500.0)):
        grid = torch.linspace(0.0, hi, 4096, dtype=torch.float64)
        with torch.no_grad():
            target = spectrum_of_parameters(p, w0, c0, pt0 * scale, niter, grid, 15.0, "lowest", reorthog=1)
        for run, kw in runs:
            out, grads, losses = {}, {}, {}
            for prec in ("f64", "f32"):
                w, c, pt = (t.clone().requires_grad_(True) for t in (w0, c0, pt0))
                sync()
                t0 = time.perf_counter()
                env = spectrum_of_parameters(p, w, c, pt, niter, grid, 15.0, "lowest", prec=prec, **kw)
                loss = torch.linalg.vector_norm(env - target)
                t1 = time.perf_counter()
                loss.backward()
                t2 = time.perf_counter()
                grads[prec] = torch.cat([w.grad.reshape(-1), c.grad.reshape(-1), pt.grad.reshape(-1)]).numpy()
                losses[prec] = float(loss)
                out[prec] = {"forward_s": round(t1 - t0, 3), "backward_s": round(t2 - t1, 3), "loss": float(loss)}
                print(f"# adjoint cfg4 {label} {run} {prec}: forward {t1 - t0:.2f} s, backward {t2 - t1:.2f} s, "
                      f"loss {float(loss):.6g}", file=sys.stderr, flush=True)
            d = grads["f32"] - grads["f64"]
            out["parameters"] = int(grads["f64"].size)
            out["nonzero_gradients"] = int(np.count_nonzero(grads["f64"]))
            out["loss_rel_diff_f32_vs_f64"] = float(f"{abs(losses['f32'] - losses['f64']) / losses['f64']:.3e}")
            out["grad_max_rel_diff_f32_vs_f64"] = float(f"{np.max(np.abs(d)) / np.max(np.abs(grads['f64'])):.3e}")
            res[f"{label}_{run}"] = out
    return res


def cfg3_probe(sync, iters=8):
    """cfg3 (nstates 4, nmodes 10, basis [8,8,6,6,6,5,5,4,4,4], N = 88,473,600,
    ordr 3, reorthog 0): the first iterations in FP64 and with FP32 H*v, and
    the H*v kernel alone (back-to-back launches).  The 1760-step operator runs
    through the generic kernel (shaped in FP32, step table in FP64).  The full
    3000-iteration job is not run here; its time is the per-iteration figure
    times 3000."""
    from nadvibs_amd.engine import Engine
    from nadvibs_amd.synth import workload
    p = workload("cfg3")
    out = {"workload": f"cfg3: nstates={p.nstates} nmodes={p.nmodes} basis={p.nfunc} ordr={p.ordr} N={p.n}",
           "iterations_timed": iters}
    coef = {}
    with Engine(p, device=0) as eng:
        out["hv_kernel"] = eng.hv_kernel("f32")
        flops = eng.hv_flops()
        out["hv_flops"] = flops
        for prec, el in (("f64", 8), ("f32", 4)):
            eng.lanczos_init(niter=iters, reorthog=0, prec=prec)
            eng.run(1, 1)
            sync()
            t0 = time.perf_counter()
            eng.run(2, iters)
            a, b, _ = eng.coefficients()
            dt = (time.perf_counter() - t0) / (iters - 1)
            coef[prec] = (a, b)
            eng.lanczos_init(niter=4, reorthog=0, prec=prec)
            eng.update_lanczos(1)
            us = eng.profile_hv(5)
            out[prec] = {"ms_per_iteration": round(dt * 1e3, 2), "hv_avg_launch_us": round(us, 1),
                         "hv_gbps": round(p.n * 2 * el / us / 1e3, 1),
                         "hv_tflops": round(flops / us / 1e6, 2)}
            print(f"# cfg3 {prec}: {dt * 1e3:.1f} ms/iteration, H*v {us:.0f} us", file=sys.stderr, flush=True)
    a64, b64 = coef["f64"]
    a32, b32 = coef["f32"]
    out["alpha_max_rel_diff_f32_vs_f64"] = float(f"{np.max(np.abs(a32 - a64) / np.abs(a64)):.3e}")
    out["speedup_f32_vs_f64"] = round(out["f64"]["ms_per_iteration"] / out["f32"]["ms_per_iteration"], 3)
    return out


def cfg3_full(sync, niter=3000):
    """The whole cfg3 job (3000 iterations, reorthog 0) in FP32 and FP64 H*v,
    with the converged roots of the lowest 3000 cm^-1 compared (about 5
    minutes; not part of the default run)."""
    from nadvibs_amd.engine import Engine
    from nadvibs_amd.synth import workload
    p = workload("cfg3")
    runs = {}
    with Engine(p, device=0) as eng:
        for prec in ("f32", "f64"):
            eng.lanczos_init(niter=niter, reorthog=0, prec=prec)
            sync()
            t0 = time.perf_counter()
            for i0 in range(1, niter + 1, 500):
                i1 = min(niter, i0 + 499)
                eng.run(i0, i1)
                sync()
                print(f"# cfg3 {prec}: {i1} iterations at {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                      flush=True)
            a, b, _ = eng.coefficients()
            dt = time.perf_counter() - t0
            runs[prec] = (dt,) + tuple(eng.ritz(a, b)[:3])
    dt32, ev32, it32, bj32 = runs["f32"]
    dt64, ev64, it64, bj64 = runs["f64"]
    low = (ev64 - ev64[0]) * AU2WAVE < 3000.0
    keep = (bj64 < 1e-9) & low
    dE, _, nconv = compare(np.where(keep, ev64, np.inf), it64, np.where(keep, bj64, 1.0), ev32, it32, bj32)
    return {"workload": f"cfg3: nstates={p.nstates} nmodes={p.nmodes} basis={p.nfunc} ordr={p.ordr} N={p.n}",
            "niter": niter, "reorthog": 0, "f32_s": round(dt32, 2), "f64_s": round(dt64, 2),
            "f32_iter_per_s": round(niter / dt32, 2), "f64_iter_per_s": round(niter / dt64, 2),
            "speedup_f32_vs_f64": round(dt64 / dt32, 3), "n_converged_below_3000_cm1": nconv,
            "max_abs_dE_cm1": float(f"{dE:.3e}"),
            "note": "without re-orthogonalisation converged roots have ghost copies, so intensities are not compared"}


def load_traffic(path):
    """HBM bytes per H*v launch measured with rocprofv3 counters, if a
    summary with a 'traffic_bytes_per_launch:' line exists."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def pmc_probe(args):
    """Kernels for the HBM-traffic counters (run under rocprofv3 --pmc):
    a device copy of known size (the calibration) and repeated H*v launches
    of the mixed job's first iteration."""
    import torch
    from nadvibs_amd.engine import Engine
    from nadvibs_amd.synth import workload
    prob = workload(args.workload)
    src = torch.empty(64 * 1024 * 1024 // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    with Engine(prob, device=0) as eng:
        kernel = eng.hv_kernel("f32")
        eng.lanczos_init(niter=4, reorthog=0, prec="f32")
        eng.update_lanczos(1)
        for _ in range(args.steps):
            eng.Hv(1)
        eng.lanczos_step(1)
        eng.coefficients(n=1)
    print(json.dumps({"copy_bytes": 64 * 1024 * 1024, "hv_launches": args.steps, "n": prob.n,
                      "hv_kernel": kernel}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--niter", type=int, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-problem batched workload")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the spectrum-gradient workload")
    ap.add_argument("--no-cfg3", action="store_true", help="skip the 88.5M-row ordr-3 probe")
    ap.add_argument("--cfg3-full", action="store_true",
                    help="only the whole 3000-iteration cfg3 job in both precisions (about 5 minutes)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="only the calibration copy and H*v launches, for rocprofv3 --pmc")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the engine runs on a single "
              f"MI355X; multi-GPU sharding is a design only)", file=sys.stderr)
        sys.exit(2)

    if args.pmc_probe:
        pmc_probe(args)
        return
    if args.cfg3_full:
        import torch
        print(json.dumps(cfg3_full(torch.cuda.synchronize)), flush=True)
        return

    import torch
    from nadvibs_amd.engine import Engine
    from nadvibs_amd.synth import workload

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    prob = workload(args.workload)
    args.niter = args.niter or prob.niter
    args.reorthog, args.chkorthog, args.seed = prob.reorthog, prob.chkorthog, 2024
    sync = torch.cuda.synchronize
    N = prob.n
    print(f"# workload {args.workload}: N={N} niter={args.niter} reorthog={args.reorthog}", file=sys.stderr,
          flush=True)
    with Engine(prob, device=0) as eng:
        kinfo = {p: eng.hv_kernel(p) for p in ("f32", "f64")}
        print(f"# H*v kernels: {kinfo}", file=sys.stderr, flush=True)
        dt32, hv32_ms, hv32_n, out32, pro32 = timed(eng, args, "f32", sync)
        print(f"# f32 job {dt32 / args.steps * 1e3:.1f} ms", file=sys.stderr, flush=True)
        res64 = None
        if not args.no_fp64:
            dt64, hv64_ms, hv64_n, out64, pro64 = timed(eng, args, "f64", sync)
            res64 = (dt64, hv64_ms, hv64_n, out64, pro64)
            print(f"# f64 job {dt64 / args.steps * 1e3:.1f} ms", file=sys.stderr, flush=True)
        # H*v kernel time alone: 200 back-to-back launches between one pair of HIP events
        # on the library's stream (the per-launch events of the profiled job add their own gaps)
        burst = {}
        for p in ("f32", "f64"):
            eng.lanczos_init(niter=4, reorthog=0, prec=p)
            eng.update_lanczos(1)
            burst[p] = eng.profile_hv(200)
        print(f"# H*v back-to-back: {burst}", file=sys.stderr, flush=True)
        hv_flops = eng.hv_flops()
        # the same job without re-orthogonalisation: H*v and the recurrence only
        plain = {}
        for p in ("f32", "f64"):
            run_job(eng, args.niter, p, 0, args.chkorthog, args.seed)
            sync()
            t0 = time.perf_counter()
            run_job(eng, args.niter, p, 0, args.chkorthog, args.seed)
            sync()
            plain[p] = args.niter / (time.perf_counter() - t0)
        print(f"# reorthog 0: {plain}", file=sys.stderr, flush=True)
        # the same job with the H*v input rounded to BF16 (8-bit mantissa, FP32 arithmetic):
        # its eigenvalue error against FP64, beside the FP32 input's
        try:
            run_job(eng, args.niter, "bf16", args.reorthog, args.chkorthog, args.seed)
            sync()
            t0 = time.perf_counter()
            outbf = run_job(eng, args.niter, "bf16", args.reorthog, args.chkorthog, args.seed)
            sync()
            bf16 = (args.niter / (time.perf_counter() - t0), outbf)
        except RuntimeError as e:
            bf16 = str(e)
        print(f"# bf16 input: {bf16 if isinstance(bf16, str) else round(bf16[0], 1)}", file=sys.stderr, flush=True)
        # the first coefficients of BF16 and FP32 input against FP64 (no re-orthogonalisation;
        # past a few tens of iterations rounding differences of any kind grow)
        first = {}
        for p in ("f64", "f32", "bf16"):
            eng.lanczos_init(niter=10, reorthog=0, prec=p, seed=args.seed)
            eng.run(1, 10)
            first[p] = eng.coefficients()[0]

    iters = args.steps * args.niter
    value = iters / dt32
    hv_avg_s = burst["f32"] / 1e6
    hv_bytes = N * (4 + 4)  # x in FP32 read, off-diagonal y in FP32 written
    achieved = hv_bytes / hv_avg_s / 1e9
    rec = {
        "metric": METRIC, "value": round(value, 2), "unit": "Lanczos iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt32 / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: nstates={prob.nstates} nmodes={prob.nmodes} "
                               f"basis={prob.nfunc} ordr={prob.ordr} N={N}",
                   "niter": args.niter, "reorthog": args.reorthog, "chkorthog": args.chkorthog,
                   "hv_precision": "f32", "recurrence_precision": "f64", "step": "one whole Lanczos job"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(os.path.join(ROOT, "profiles", f"hv_f32_{args.workload}_traffic.txt")),
                     "kernel": "nv_hv_iter_f32 (" + kinfo["f32"] + ")", "avg_launch_us": round(hv_avg_s * 1e6, 2),
                     "bytes_per_launch": hv_bytes, "launches": 200,
                     "in_job_avg_launch_us": round(hv32_ms / max(hv32_n, 1) * 1e3, 2)},
        "hv_gbps": round(achieved, 1),
        "hv_compute": {"flops_per_launch": hv_flops, "achieved_tflops": round(hv_flops / hv_avg_s / 1e12, 2),
                       "peak_tflops_fp32": FP32_PEAK_TFLOPS,
                       "frac": round(hv_flops / hv_avg_s / 1e12 / FP32_PEAK_TFLOPS, 4),
                       "frac_of_unpacked_fp32": round(hv_flops / hv_avg_s / 1e12 / (FP32_PEAK_TFLOPS / 2), 4)},
        "pro": {"events_per_job": pro32[0], "vectors_touched_per_job": pro32[1]},
    }
    if res64 is not None:
        dt64, hv64_ms, hv64_n, out64, _ = res64
        v64 = iters / dt64
        hv64_s = burst["f64"] / 1e6
        dE, dI, nconv = compare(out64[2], out64[3], out64[4], out32[2], out32[3], out32[4])
        rec["fp64"] = {"value": round(v64, 2), "ms_per_step": round(dt64 / args.steps * 1e3, 3),
                       "hv_avg_launch_us": round(hv64_s * 1e6, 2),
                       "hv_gbps": round(N * 16 / hv64_s / 1e9, 1)}
        rec["speedup_vs_fp64"] = round(value / v64, 3)
        rec["max_abs_dE_cm1"] = float(f"{dE:.3e}")
        rec["max_abs_dI"] = float(f"{dI:.3e}")
        rec["n_converged_fp64"] = nconv
        rec["error_target_met"] = bool(dE <= 1e-3)
    if res64 is not None:
        a_rel = {p: float(f"{np.max(np.abs(first[p] - first['f64']) / np.abs(first['f64'])):.3e}")
                 for p in ("f32", "bf16")}
        if isinstance(bf16, str):
            rec["bf16_input"] = {"failed": bf16, "alpha_max_rel_diff_first_10": a_rel,
                                 "note": "H*v input rounded to BF16, FP32 arithmetic, FP64 diagonal and recurrence"}
        else:
            out64 = res64[3]
            dEb, dIb, _ = compare(out64[2], out64[3], out64[4], bf16[1][2], bf16[1][3], bf16[1][4])
            rec["bf16_input"] = {"iter_per_s": round(bf16[0], 1), "max_abs_dE_cm1": float(f"{dEb:.3e}"),
                                 "max_abs_dI": float(f"{dIb:.3e}"), "error_target_met": bool(dEb <= 1e-3),
                                 "alpha_max_rel_diff_first_10": a_rel,
                                 "note": "H*v input rounded to BF16, FP32 arithmetic, FP64 diagonal and recurrence"}
    rec["reorthog0"] = {"f32_iter_per_s": round(plain["f32"], 1), "f64_iter_per_s": round(plain["f64"], 1),
                        "speedup_f32_vs_f64": round(plain["f32"] / plain["f64"], 3),
                        "note": "same job without partial re-orthogonalisation (ghost copies are not compared)"}
    if not args.no_batch:
        rec["batched_cfg5"] = batched_cfg5(sync, 0.0 if args.no_cpu else args.cpu_seconds)
    if not args.no_adjoint:
        rec["adjoint_cfg4"] = adjoint_cfg4(sync)
    if not args.no_cfg3:
        rec["cfg3"] = cfg3_probe(sync)
    if not args.no_cpu:
        rec["cpu_baseline"], ref = cpu_baseline(prob, args.cpu_seconds, args.seed)
        # the GPU jobs' first coefficients against the FP64 CPU run of the same iterations
        k = ref["alpha"].size
        check = {"iterations": k,
                 "alpha_max_rel_diff_f32": float(f"{np.max(np.abs(out32[0][:k] - ref['alpha']) / np.abs(ref['alpha'])):.3e}")}
        if res64 is not None:
            check["alpha_max_rel_diff_f64"] = float(
                f"{np.max(np.abs(res64[3][0][:k] - ref['alpha']) / np.abs(ref['alpha'])):.3e}")
        rec["vs_cpu_fp64"] = check
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
