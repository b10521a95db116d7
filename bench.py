"""Benchmark: Lanczos iterations/s and H*v bandwidth on one MI355X.

Headline workload: nstates=4, nmodes=10, basis [8,8,6,6,6,5,5,4,4,4]
(N = 88,473,600), ordr=3, niter=3000, no re-orthogonalisation, seeded
synthetic couplings (cubic N(0,10) cm^-1, the rest as the 1.38M-row
workload).  One step = one slice of SLICE (default 100) Lanczos iterations of
the 3000-iteration job: H*v, recurrence and the device-side coefficient sums,
captured in graphs.  The mixed-precision job (H*v FP32, recurrence FP64) is
the headline; the FP64 job runs beside it, both to the end, and their
converged eigenvalues are compared (cm^-1).  W warmup slices, K timed slices,
then the rest of the job (untimed; its first SLICE iterations with every H*v
bracketed by HIP events).

Also reported: the 1.38M-row workload's 2000-iteration jobs with partial
re-orthogonalisation (mixed vs FP64), the 64-problem batch, the spectrum
gradient, and the FP64 CPU restatement with OpenMP on this host's cores.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--slice 100]
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


def compare_windows(ev64, bj64, ev, bj, tol_conv=1e-9):
    """Eigenvalue error of the converged roots (bji < tol_conv hartree in the
    FP64 run), each matched to the nearest root of the other run: over the
    0-6000 cm^-1 window above the lowest root and over every converged root,
    each once for every FP64-converged root and once for the roots converged
    in both runs.  Without re-orthogonalisation a root converged in one run
    can still be converging in the other (its Ritz value then differs by up
    to its residual), which the second form leaves out."""
    rel = (ev64 - ev64[0]) * AU2WAVE
    out = {}
    for name, win in (("0_6000", rel <= 6000.0), ("all", np.ones(ev64.size, bool))):
        d, db, nb = [], [], 0
        for k in np.nonzero((bj64 < tol_conv) & win)[0]:
            j = int(np.argmin(np.abs(ev - ev64[k])))
            e = abs(ev[j] - ev64[k]) * AU2WAVE
            d.append(e)
            if bj[j] < tol_conv:
                db.append(e)
                nb += 1
        out[name] = (max(d) if d else float("nan"), len(d))
        out[name + "_both" if name != "all" else "both"] = (max(db) if db else float("nan"), nb)
    return out


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
    # the chosen form's batched H*v time (x read in FP32, y written in FP32 per element)
    import re
    times = {m.group(1): float(m.group(2)) for m in re.finditer(r" (\d+/\d/\w+) ([\d.]+) us;", rep)}
    chosen = re.search(r"chosen (\S+)", rep).group(1)
    hv_us = times.get(chosen, float("nan"))
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
            "batched_hv_us": round(hv_us, 1), "batched_hv_gbps": round(total * 8 / hv_us / 1e3, 1),
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
            refused = None
            for prec in ("f64", "f32"):
                w, c, pt = (t.clone().requires_grad_(True) for t in (w0, c0, pt0))
                sync()
                t0 = time.perf_counter()
                env = spectrum_of_parameters(p, w, c, pt, niter, grid, 15.0, "lowest", prec=prec, **kw)
                loss = torch.linalg.vector_norm(env - target)
                t1 = time.perf_counter()
                import warnings
                with warnings.catch_warnings(record=True) as caught:
                    warnings.simplefilter("always")
                    loss.backward()
                t2 = time.perf_counter()
                flags = [str(x.message) for x in caught if "orthogonality" in str(x.message)]
                if flags:  # the library flags a basis that lost orthogonality
                    refused = flags[0]
                grads[prec] = torch.cat([w.grad.reshape(-1), c.grad.reshape(-1), pt.grad.reshape(-1)]).numpy()
                losses[prec] = float(loss)
                out[prec] = {"forward_s": round(t1 - t0, 3), "backward_s": round(t2 - t1, 3), "loss": float(loss)}
                print(f"# adjoint cfg4 {label} {run} {prec}: forward {t1 - t0:.2f} s, backward {t2 - t1:.2f} s, "
                      f"loss {float(loss):.6g}", file=sys.stderr, flush=True)
            if refused is not None:
                out["orthogonality_warning"] = refused
                print(f"# adjoint cfg4 {label} {run}: {refused}", file=sys.stderr, flush=True)
            d = grads["f32"] - grads["f64"]
            out["parameters"] = int(grads["f64"].size)
            out["nonzero_gradients"] = int(np.count_nonzero(grads["f64"]))
            out["loss_rel_diff_f32_vs_f64"] = float(f"{abs(losses['f32'] - losses['f64']) / losses['f64']:.3e}")
            out["grad_max_rel_diff_f32_vs_f64"] = float(f"{np.max(np.abs(d)) / np.max(np.abs(grads['f64'])):.3e}")
            res[f"{label}_{run}"] = out
    return res


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
    of the first iteration, FP32 then FP64."""
    import torch
    from nadvibs_amd.engine import Engine
    from nadvibs_amd.synth import workload
    prob = workload(args.workload)
    src = torch.empty(64 * 1024 * 1024 // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    kernels = {}
    with Engine(prob, device=0) as eng:
        for prec in ("f32", "f64"):
            kernels[prec] = eng.hv_kernel(prec)
            eng.lanczos_init(niter=4, reorthog=0, prec=prec)
            eng.update_lanczos(1)
            for _ in range(args.steps):
                eng.Hv(1)
            eng.lanczos_step(1)
            eng.coefficients(n=1)
    print(json.dumps({"copy_bytes": 64 * 1024 * 1024, "hv_launches": args.steps, "n": prob.n,
                      "hv_kernels": kernels}), flush=True)


def lscpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """CPU threads this process may use (its affinity set), at most 16 (the
    share of host CPUs a one-GPU job gets)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline_rows(prob, seconds):
    """The FP64 CPU restatement (oracle, OpenMP over contiguous row blocks as
    the reference's Global-Arrays row partition) on this host: one H*v of the
    workload over 8 equally spaced blocks of contiguous rows (all states),
    with every usable core and with one core, the blocks sized to about
    seconds/4 of work each way, scaled to the whole vector.  One Lanczos
    iteration is one H*v plus four vector passes, which the sample leaves out
    (they cost well under 1% of the H*v here)."""
    from oracle.oracle import Oracle
    orc = Oracle(prob)
    D = orc.dimen
    x = np.random.default_rng(1).standard_normal(orc.n)
    out = {}

    def sample(threads, rows):
        b = max(1, rows // 8)
        t0 = time.perf_counter()
        for k in range(8):
            r0 = min(D - b, (2 * k + 1) * D // 16)
            orc.hv_par(x, threads, rows=(r0, r0 + b))
        return time.perf_counter() - t0, 8 * b

    for threads in (host_threads(), 1):
        rows = 2000 * threads
        while True:
            dt, rows = sample(threads, rows)
            if dt >= seconds / 4 or rows >= D // 2:
                break
            rows = int(rows * min(8.0, max(2.0, seconds / 3 / max(dt, 1e-3))))
        t_full = dt * D / rows
        out[threads] = (1.0 / t_full, rows, dt)
        print(f"# cpu baseline: {threads} thread(s), {rows} rows in {dt:.2f} s -> {t_full:.1f} s per H*v",
              file=sys.stderr, flush=True)
    T = host_threads()
    v, rows, dt = out[T]
    v1, rows1, dt1 = out[1]
    return {"value": round(v, 6), "unit": "Lanczos iter/s", "cores": T, "kind": "port",
            "sample": f"one H*v of the same workload over {rows} of its {D} rows (8 equally spaced blocks, all "
                      f"states) in {dt:.1f} s, scaled to all rows; FP64 oracle, OpenMP over contiguous row blocks",
            "one_core": {"value": round(v1, 6), "cores": 1,
                         "sample": f"{rows1} rows in {dt1:.1f} s, scaled"},
            "host": f"{lscpu_model()}, {os.cpu_count()} CPUs visible, {T} used"}


def headline_job(eng, prec, args, sync, niter):
    """One whole job in slices: W warmup slices, K timed slices, the rest of
    the job with the first slice's H*v launches bracketed by HIP events.
    Returns (timed seconds, in-job H*v ms and launches, eigen-solution)."""
    S, W, K = args.slice, args.warmup, args.steps
    eng.lanczos_init(niter=niter, reorthog=0, prec=prec)
    it = 1
    for k in range(W):
        eng.run(it, it + S - 1)
        it += S
        sync()
        print(f"# {prec} warmup slice {k + 1}/{W} (iterations to {it - 1})", file=sys.stderr, flush=True)
    sync()
    t0 = time.perf_counter()
    for k in range(K):
        eng.run(it, it + S - 1)
        it += S
        sync()
        print(f"# {prec} slice {k + 1}/{K} done at {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    hv_ms, hv_n = 0.0, 0
    if it <= niter:
        i1 = min(niter, it + S - 1)
        eng.profile(True)
        eng.profile_read()
        eng.run(it, i1)
        hv_ms, hv_n = eng.profile_read()
        eng.profile(False)
        it = i1 + 1
    while it <= niter:
        i1 = min(niter, it + 499)
        eng.run(it, i1)
        sync()
        print(f"# {prec} rest of the job: {i1} iterations", file=sys.stderr, flush=True)
        it = i1 + 1
    a, b, _ = eng.coefficients()
    if not (np.all(np.isfinite(a)) and np.all(np.isfinite(b))):
        raise RuntimeError(f"{prec} job produced non-finite coefficients")
    return dt, hv_ms, hv_n, eng.ritz(a, b)[:3], a


def cfg2_jobs(args, sync):
    """The 1.38M-row workload's whole 2000-iteration jobs with partial
    re-orthogonalisation (reorthog 1, chkorthog 100), mixed against FP64:
    job time, PRO work, H*v alone, and the converged-root error; the same jobs
    without re-orthogonalisation; BF16 input against FP32 on the first 10
    coefficients."""
    from nadvibs_amd.engine import Engine
    from nadvibs_amd.synth import workload
    prob = workload("cfg2")
    sub = argparse.Namespace(niter=prob.niter, reorthog=prob.reorthog, chkorthog=prob.chkorthog, seed=2024,
                             steps=args.cfg2_jobs, warmup=1)
    N = prob.n
    res = {"workload": f"nstates={prob.nstates} nmodes={prob.nmodes} basis={prob.nfunc} ordr={prob.ordr} N={N}",
           "niter": prob.niter, "reorthog": prob.reorthog, "chkorthog": prob.chkorthog, "step": "one whole job"}
    with Engine(prob, device=0) as eng:
        res["hv_kernels"] = {p: eng.hv_kernel(p) for p in ("f32", "f64")}
        dt32, _, _, out32, pro32 = timed(eng, sub, "f32", sync)
        dt64, _, _, out64, pro64 = timed(eng, sub, "f64", sync)
        burst = {}
        for p in ("f32", "f64"):
            eng.lanczos_init(niter=4, reorthog=0, prec=p)
            eng.update_lanczos(1)
            burst[p] = eng.profile_hv(200)
        plain = {}
        for p in ("f32", "f64"):
            run_job(eng, sub.niter, p, 0, sub.chkorthog, sub.seed)
            sync()
            t0 = time.perf_counter()
            run_job(eng, sub.niter, p, 0, sub.chkorthog, sub.seed)
            sync()
            plain[p] = sub.niter / (time.perf_counter() - t0)
        first = {}
        for p in ("f64", "f32", "bf16"):
            eng.lanczos_init(niter=10, reorthog=0, prec=p, seed=sub.seed)
            eng.run(1, 10)
            first[p] = eng.coefficients()[0]
    dE, dI, nconv = compare(out64[2], out64[3], out64[4], out32[2], out32[3], out32[4])
    v32, v64 = sub.steps * sub.niter / dt32, sub.steps * sub.niter / dt64
    res.update({
        "f32_iter_per_s": round(v32, 1), "f64_iter_per_s": round(v64, 1), "speedup_vs_fp64": round(v32 / v64, 3),
        "job_ms": {"f32": round(dt32 / sub.steps * 1e3, 1), "f64": round(dt64 / sub.steps * 1e3, 1)},
        "max_abs_dE_cm1": float(f"{dE:.3e}"), "max_abs_dI": float(f"{dI:.3e}"), "n_converged_fp64": nconv,
        "error_target_met": bool(dE <= 1e-3),
        "pro": {"f32": {"events_per_job": pro32[0], "vectors_touched_per_job": pro32[1]},
                "f64": {"events_per_job": pro64[0], "vectors_touched_per_job": pro64[1]}},
        "hv_avg_launch_us": {p: round(burst[p], 2) for p in burst},
        "hv_gbps": {"f32": round(N * 8 / burst["f32"] / 1e3, 1), "f64": round(N * 16 / burst["f64"] / 1e3, 1)},
        "reorthog0_iter_per_s": {p: round(plain[p], 1) for p in plain},
        "alpha_max_rel_diff_first_10": {p: float(f"{np.max(np.abs(first[p] - first['f64']) / np.abs(first['f64'])):.3e}")
                                        for p in ("f32", "bf16")},
    })
    print(f"# 1.38M-row jobs: {res['job_ms']} ms, dE {dE:.2e}", file=sys.stderr, flush=True)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed slices")
    ap.add_argument("--warmup", type=int, default=5, help="untimed slices before them")
    ap.add_argument("--slice", type=int, default=100, help="Lanczos iterations per step")
    ap.add_argument("--workload", default="cfg3")
    ap.add_argument("--niter", type=int, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=16.0)
    ap.add_argument("--cfg2-jobs", type=int, default=3, help="timed whole jobs of the 1.38M-row workload")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cfg2", action="store_true", help="skip the 1.38M-row re-orthogonalised jobs")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-problem batched workload")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the spectrum-gradient workload")
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

    import torch
    from nadvibs_amd.engine import Engine
    from nadvibs_amd.synth import workload

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    prob = workload(args.workload)
    niter = args.niter or prob.niter
    if (args.warmup + args.steps) * args.slice > niter:
        args.slice = max(1, niter // (args.warmup + args.steps + 1))
    sync = torch.cuda.synchronize
    N = prob.n
    print(f"# workload {args.workload}: N={N} niter={niter}, slices of {args.slice}", file=sys.stderr, flush=True)
    runs = {}
    with Engine(prob, device=0) as eng:
        kinfo = {p: eng.hv_kernel(p) for p in ("f32", "f64")}
        print(f"# H*v kernels: {kinfo}", file=sys.stderr, flush=True)
        for prec in ("f32",) if args.no_fp64 else ("f32", "f64"):
            runs[prec] = headline_job(eng, prec, args, sync, niter)
            print(f"# {prec}: {args.steps * args.slice / runs[prec][0]:.2f} iter/s", file=sys.stderr, flush=True)
        # the H*v kernel alone: back-to-back launches between one pair of HIP events on the library's stream
        burst = {}
        for prec in ("f32", "f64"):
            eng.lanczos_init(niter=4, reorthog=0, prec=prec)
            eng.update_lanczos(1)
            burst[prec] = eng.profile_hv(20)
        print(f"# H*v back-to-back: {burst} us", file=sys.stderr, flush=True)
        hv_flops = eng.hv_flops()
    iters = args.steps * args.slice
    dt32, hv32_ms, hv32_n, sol32, _ = runs["f32"]
    value = iters / dt32
    hv_avg_s = burst["f32"] / 1e6
    hv_bytes = N * (4 + 4)  # x in FP32 read, off-diagonal y in FP32 written
    achieved = hv_bytes / hv_avg_s / 1e9
    rec = {
        "metric": METRIC, "value": round(value, 3), "unit": "Lanczos iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt32 / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: nstates={prob.nstates} nmodes={prob.nmodes} "
                               f"basis={prob.nfunc} ordr={prob.ordr} N={N}",
                   "niter": niter, "reorthog": 0, "hv_precision": "f32", "recurrence_precision": "f64",
                   "step": f"{args.slice} Lanczos iterations of the {niter}-iteration job"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": load_traffic(os.path.join(ROOT, "profiles", f"hv_f32_{args.workload}_traffic.txt")),
                     "kernel": "iteration H*v, FP32 (" + kinfo["f32"] + ")",
                     "avg_launch_us": round(hv_avg_s * 1e6, 1), "bytes_per_launch": hv_bytes, "launches": 20,
                     "in_job_avg_launch_us": round(hv32_ms / max(hv32_n, 1) * 1e3, 1)},
        "hv_compute": {"flops_per_launch": hv_flops, "achieved_tflops": round(hv_flops / hv_avg_s / 1e12, 2),
                       "peak_tflops_fp32": FP32_PEAK_TFLOPS,
                       "frac": round(hv_flops / hv_avg_s / 1e12 / FP32_PEAK_TFLOPS, 4),
                       "note": "flops = twice the multiply-adds of the operator's nonzero structure"},
    }
    if "f64" in runs:
        dt64, hv64_ms, hv64_n, sol64, _ = runs["f64"]
        v64 = iters / dt64
        ev64, it64, bj64 = sol64
        ev32, it32, bj32 = sol32
        win = compare_windows(ev64, bj64, ev32, bj32)
        dE, nboth = win["0_6000_both"]
        rec["fp64"] = {"value": round(v64, 3), "ms_per_step": round(dt64 / args.steps * 1e3, 3),
                       "hv_avg_launch_us": round(burst["f64"], 1),
                       "hv_gbps": round(N * 16 / burst["f64"] / 1e3, 2),
                       "in_job_hv_avg_launch_us": round(hv64_ms / max(hv64_n, 1) * 1e3, 1)}
        rec["speedup_vs_fp64"] = round(value / v64, 3)
        rec["max_abs_dE_cm1"] = float(f"{dE:.3e}")
        rec["n_converged_both"] = nboth
        rec["n_converged_fp64"] = win["0_6000"][1]
        rec["error_target_met"] = bool(dE <= 1e-3)
        rec["error_windows"] = {k: {"max_abs_dE_cm1": float(f"{v[0]:.3e}"), "roots": v[1]} for k, v in win.items()}
        rec["error_note"] = (
            "max_abs_dE_cm1: roots within 0-6000 cm^-1 above the lowest converged (bji < 1e-9 hartree) in both the "
            "FP64 and the mixed run, each FP64 root matched to the nearest mixed-run root.  n_converged_fp64 - "
            "n_converged_both window roots converged in FP64 are still converging in the mixed run after the same "
            "iterations (error_windows.0_6000 includes them: their gap is the Ritz residual, not a converged shift).  "
            "No re-orthogonalisation, so converged roots have ghost copies in both runs and intensities are not "
            "compared")
    if not args.no_cfg2:
        rec["reorthogonalised_1p38M"] = cfg2_jobs(args, sync)
    if not args.no_batch:
        rec["batched_cfg5"] = batched_cfg5(sync, 0.0 if args.no_cpu else args.cpu_seconds)
    if not args.no_adjoint:
        rec["adjoint_cfg4"] = adjoint_cfg4(sync)
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline_rows(prob, args.cpu_seconds)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
