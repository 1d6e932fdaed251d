#!/usr/bin/env python3
"""Benchmark of the MH-RE engine on one MI355X.

Headline workload (one "step" = one MH-RE iteration: nprm = 9 single-parameter
proposals for every chain followed by the replica exchange):
    plain MH-RE fit of the 9-parameter ring+Gaussian visibility model,
    32 temperatures x 128 sequences (4096 chains), 2000 visibility amplitudes
    + 600 closure phases (8 stations, 6 h tracks at 230 GHz), 5000 burn-in +
    20000 sampling iterations.  --warmup + --steps iterations are run
    (default 100 + 24900 = the whole fit); only --steps are timed.
value = visibility model evaluations performed per second (G/s), inputs
resident in HBM, whole job (one GPU).

Also reported:
  roofline      the algorithmic FP32 rate of the timed iterations (170 flop
                per visibility-eval, 15 per closure term, 40 per chain setup)
                over their device time (HIP events around the window, which
                hold the step launches of the three sequence groups,
                overlapping on three streams; each step launch also runs the
                replica exchange of the iteration before), against the
                157.3 TFLOP/s peak; also the average duration of one step
                launch and of the exchange launches per iteration (HIP events
                around each);
  cpu_baseline  the float64 C/OpenMP oracle port on the host cores;
  pcie_inclusive the same whole fit through host buffers (upload, run,
                download of the recorded samples), host wall time;
  batched       64 uneven fits (500-20000 visibilities, 16-48 x 64-256
                chains) in one session, launch form picked by the
                autotuner, vs the same fits one
                session at a time (fused native kernel, and the per-datum
                model.cuh contract path);
  adjoint       chi^2 + gradient of 4096 parameter points against 20000
                visibilities + 6000 closure phases (fp32, HIP events), then
                64 short fits of the plain-fit shape sampled in one batched
                session and a Gauss-Newton polish + Laplace covariance of
                each fit's best chain;
  quadsys       1024 systems of 8 quadratic equations, 16 x 32 chains each
                (524288 chains), 3000 + 10000 iterations in one session, then
                Gauss-Newton polish; success = residual norm < 1e-5.

Prints exactly one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "G visibility-evals/s in MH-RE fit (batched vs plain), % of FP32 VALU peak"
PEAK_FP32_TFLOPS = 157.3
F_VIS, F_CPH, F_CHAIN = 170.0, 15.0, 40.0
# quadratic system of 8 equations in 8 unknowns, monomial form: 36 products
# + 8 * (45 FMA + 2) = 772 flop per chi^2 evaluation
F_QUAD8 = 772.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def eht_fit(d, nbeta, nseq, nburn, niter, seed):
    from cuda_mcmc_amd import synthdata
    from cuda_mcmc_amd.engine import Fit
    idat = np.array([d["nvis"], nbeta * nseq * (d["nvis"] + d["ncph"]), 0], np.int32)
    return Fit(synthdata.EHT_PDESCR, synthdata.EHT_PMIN, synthdata.EHT_PMAX, d["coor"], d["dat"], d["std2r"],
               nbeta=nbeta, nseq=nseq, nburn=nburn, niter=niter, model="sgra9", imodel=1,
               icoor=d["icoor"], idat=idat, npass=2, seed=seed)


def traffic_from_profiles():
    """HBM bytes per step launch from the committed PMC summary, if present."""
    p = os.path.join(ROOT, "profiles", "sgra9_step_4096chains_2000vis_pmc.txt")
    if not os.path.isfile(p):
        return None
    for line in open(p):
        if line.startswith("HBM_BYTES_PER_LAUNCH"):
            try:
                return float(line.split()[1])
            except (IndexError, ValueError):
                return None
    return None


def headline_options(args):
    """Launch form of the plain fit's session: 4 waves per work-group, one
    wave per chain, the 128 sequences in three groups on three streams
    (--streams changes the group count, e.g. 1 for hardware-counter runs)."""
    return dict(waves_per_block=4, persistent=0, streams=args.streams or 3)


def leg_streams(args):
    """Sequence groups of the other legs' sessions (two unless --streams)."""
    return args.streams or 2


def host_threads():
    """Threads for the CPU baseline: the process's OpenMP allotment (the GPU
    box gives each GPU 16 host CPUs and sets OMP_NUM_THREADS to that), else
    every CPU the process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def cpu_baseline(d, state, it0, iters, nthreads):
    """The float64 C/OpenMP oracle port run for `iters` iterations of the same
    fit from the device's state at iteration it0 (sampling stage, recording
    on), timed on the host."""
    from oracle import cport
    cport.load()
    fit = eht_fit(d, 32, 128, nburn=it0, niter=iters, seed=0)
    for k in ("pcur", "chi2c", "pstp", "n_cnt", "n_acpt", "n_hist", "n_exch", "tcur", "flag", "ptent", "ptentn"):
        getattr(fit, k)[...] = getattr(state, k)
    t = time.time()
    nev = 0
    for it in range(it0, it0 + iters):
        nev += cport.eht_iteration(fit, d, it, nthreads)
    dt = time.time() - t
    return dict(value=nev * d["nvis"] / dt / 1e9, unit="G visibility-evals/s", cores=int(nthreads),
                host_cpus=int(os.cpu_count() or 0), kind="port",
                sample="%d MH-RE iterations (%d..%d, sampling stage, recording on; all 4096 chains, %d chi^2 "
                       "evaluations of 2000 vis + 600 closures) of the same fit from the device's state, float64 "
                       "C/OpenMP port of the oracle on %d threads, %.1f s"
                       % (iters, it0, it0 + iters - 1, nev, nthreads, dt))


def batched_vs_plain(iters, nfits, log_every=16, session_only=False, streams=2):
    """64 uneven fits (500-20000 visibilities, 5-10 stations, 16-48 x 64-256
    chains) in one session vs one session per fit (session_only: the batched
    session alone, for profiling)."""
    from cuda_mcmc_amd import synthdata
    from cuda_mcmc_amd.engine import Session
    fits, datas = [], []
    for f in range(nfits):
        rng = np.random.default_rng(300 + f)
        nsta = int(rng.integers(5, 11))
        nvis = int(np.exp(rng.uniform(np.log(500), np.log(20000))))
        per_time = nsta * (nsta - 1) // 2
        ntri = nsta * (nsta - 1) * (nsta - 2) // 6
        ncph = min(int(0.3 * nvis), int(nvis / per_time * ntri * 0.8))
        d = synthdata.ring_gauss_dataset(nvis=nvis, ncph=ncph, nsta=nsta, seed=300 + f)
        nb = int(rng.integers(16, 49))
        ns = int(rng.integers(64, 257))
        fits.append((d, nb, ns))
        if (f + 1) % log_every == 0:
            log("  generated %d/%d fits" % (f + 1, nfits))
    tune_iters = 2
    mk = lambda d, nb, ns, f: eht_fit(d, nb, ns, nburn=iters + 2 + 6 * tune_iters, niter=0, seed=300 + f)
    total_vis = 0.0
    with Session([mk(d, nb, ns, f) for f, (d, nb, ns) in enumerate(fits)], persistent=0, streams=streams) as s:
        s.run(1)
        wpb, pers, _ = s.autotune(tune_iters)
        s.sync()
        e0 = s.stats()
        t0 = time.time()
        s.run(iters)
        s.sync()
        tb = time.time() - t0
        e1 = s.stats()
        total_vis = e1["vis_evals"] - e0["vis_evals"]
        total_ev = e1["chi2_evals"] - e0["chi2_evals"]
        total_cph = e1["cph_evals"] - e0["cph_evals"]
        tb_dev = e1["run_ms"] * 1e-3
    batched = total_vis / tb
    if session_only:
        return dict(fits=nfits, iterations=iters, streams=streams,
                    autotuned={"waves_per_block": wpb, "persistent": pers},
                    batched_gvis_per_s=batched / 1e9, batched_device_gvis_per_s=total_vis / tb_dev / 1e9,
                    iteration_ms=tb_dev * 1e3 / iters, vis_evals=total_vis, cph_evals=total_cph,
                    chi2_evals=total_ev)
    tp = 0.0
    plain_vis = 0.0
    for f, (d, nb, ns) in enumerate(fits):
        with Session([mk(d, nb, ns, f)], streams=streams) as s:
            s.run(1)
            s.sync()
            e0 = s.stats()
            t0 = time.time()
            s.run(iters)
            s.sync()
            tp += time.time() - t0
            plain_vis += s.stats()["vis_evals"] - e0["vis_evals"]
    plain = plain_vis / tp
    # the same fits one at a time through the reference's per-datum model()
    # contract (generic model.cuh path, one lane per chain)
    log("  one session per fit done; now the model.cuh path")
    tg = 0.0
    gen_vis = 0.0
    for f, (d, nb, ns) in enumerate(fits):
        if f % log_every == 0:
            log("  model.cuh path: fit %d/%d" % (f, nfits))
        with Session([mk(d, nb, ns, f)], force_generic=True, streams=streams) as s:
            s.run(1)
            s.sync()
            e0 = s.stats()
            t0 = time.time()
            s.run(iters)
            s.sync()
            tg += time.time() - t0
            gen_vis += s.stats()["vis_evals"] - e0["vis_evals"]
    generic = gen_vis / tg
    return dict(fits=nfits, iterations=iters, streams=streams, autotuned={"waves_per_block": wpb, "persistent": pers},
                batched_device_gvis_per_s=total_vis / tb_dev / 1e9,
                plain_generic_sequential_gvis_per_s=generic / 1e9, ratio_vs_generic=batched / generic,
                chains=int(sum(nb * ns for _, nb, ns in fits)),
                vis_total=int(sum(d["nvis"] for d, _, _ in fits)),
                batched_gvis_per_s=batched / 1e9, plain_sequential_gvis_per_s=plain / 1e9,
                ratio=batched / plain,
                pct_fp32_peak_batched=100.0 * batched * F_VIS / (PEAK_FP32_TFLOPS * 1e12))


def adjoint_leg(npts=4096, nvis=20000, ncph=6000, reps=5, gn_fits=64, gn_burn=1000, gn_iter=250, streams=2):
    """chi^2 + gradient of 4096 points against 20000 visibilities + 6000
    closures, then Gauss-Newton polish + Laplace covariance of 64 fits."""
    import torch
    from cuda_mcmc_amd import adjoint, synthdata
    from cuda_mcmc_amd.engine import Session
    d = synthdata.ring_gauss_dataset(nvis=nvis, ncph=ncph, nsta=8, seed=4)
    data = adjoint.EhtData.from_dataset(d)
    rng = np.random.default_rng(7)
    lo, hi = synthdata.EHT_PMIN.astype(np.float64), synthdata.EHT_PMAX.astype(np.float64)
    P = torch.as_tensor((lo + (hi - lo) * rng.uniform(size=(npts, 9))).astype(np.float32), device="cuda")
    data.chi2_grad(P)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        data.chi2_grad(P)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    e0.record()
    for _ in range(reps):
        data.chi2_grad(P, need_grad=False)
    e1.record()
    e1.synchronize()
    ms_fwd = e0.elapsed_time(e1) / reps
    out = dict(points=npts, nvis=nvis, ncph=ncph, ms_per_call=round(ms, 4), ms_chi2_only=round(ms_fwd, 4),
               vis_grad_evals_per_s=npts * nvis / (ms * 1e-3),
               model_evals_per_s=npts * (nvis + 3 * ncph) / (ms * 1e-3))
    # 64 short fits of the plain-fit shape sampled in one batched session,
    # then Gauss-Newton on the best chain of each and its Laplace covariance
    before, after, ratios = [], [], []
    dsets = [synthdata.ring_gauss_dataset(nvis=2000, ncph=600, nsta=8, seed=500 + f) for f in range(gn_fits)]
    fits = [eht_fit(dd, 32, 128, nburn=gn_burn, niter=gn_iter, seed=500 + f) for f, dd in enumerate(dsets)]
    t0 = time.time()
    with Session(fits, streams=streams) as s:
        s.run(gn_burn + gn_iter)
        s.download()
    t_sample = time.time() - t0
    t0 = time.time()
    for dd, fit in zip(dsets, fits):
        i = np.unravel_index(np.argmin(fit.chi2c), fit.chi2c.shape)
        dat = adjoint.EhtData.from_dataset(dd)
        p, c = adjoint.gauss_newton(dat, fit.pcur[i][None, :], synthdata.EHT_PMIN, synthdata.EHT_PMAX,
                                    synthdata.EHT_PDESCR, iters=20)
        before.append(float(fit.chi2c[i]))
        after.append(float(c[0]))
        cov, vv = adjoint.laplace_covariance(dat, p, synthdata.EHT_PDESCR)
        smp = fit.pout[:, 0].reshape(9, -1).T.astype(np.float64)
        smp[:, 8] = np.angle(np.exp(1j * (smp[:, 8] - p[0, 8]))) + p[0, 8]
        sd_mc = smp[:, vv].std(axis=0)
        ratios.append(float(np.median(sd_mc / np.sqrt(np.diag(cov[0])))))
        dat.close()
    t_polish = time.time() - t0
    out["gauss_newton"] = dict(fits=gn_fits, burn_in=gn_burn, samples=gn_iter, sampling_s=round(t_sample, 2),
                               polish_and_laplace_s=round(t_polish, 2), iterations=20,
                               chi2_best_chain_median=float(np.median(before)),
                               chi2_polished_median=float(np.median(after)),
                               chi2_drop_median=float(np.median(np.array(before) - np.array(after))),
                               laplace_vs_chain_sd_ratio_median=float(np.median(ratios)))
    return out


def quadsys_leg(nsys=1024, nburn=3000, niter=10000, streams=2):
    """1024 systems of 8 quadratic equations, every system one fit of one
    session, then Gauss-Newton polish."""
    from cuda_mcmc_amd import quadsys, synthdata
    from cuda_mcmc_amd.engine import Session
    ds = synthdata.quadsys_dataset(nsys=nsys, seed=5)
    fits = [quadsys.quad_fit(ds["Q"][s], ds["b"][s], ds["c"][s], nbeta=16, nseq=32, nburn=nburn, niter=niter,
                             seed=7000 + s, record=False) for s in range(nsys)]
    log("  %d systems packed; sampling" % nsys)
    with Session(fits, streams=streams) as s:
        s.run(1)
        s.sync()
        e0 = s.stats()
        t0 = time.time()
        left = nburn + niter - 1
        while left > 0:
            n = min(1000, left)
            s.run(n)
            s.sync()
            left -= n
            log("  %d iterations left (%.1f s)" % (left, time.time() - t0))
        dt = time.time() - t0
        e1 = s.stats()
        s.download()
    ev = e1["chi2_evals"] - e0["chi2_evals"]
    X0, c2 = quadsys.best_rows(fits)
    t1 = time.time()
    X, resid = quadsys.gauss_newton(ds["Q"], ds["b"], -ds["c"], X0, iters=20)
    tgn = time.time() - t1
    # the same polish from the best chain of each of the 4 best sequences of
    # every system (independent runs of the sampler), keeping the best result
    Xs, _, sysv = quadsys.best_sequence_rows(fits, k=4)
    t1 = time.time()
    _, resid4 = quadsys.polish_best(ds["Q"], ds["b"], -ds["c"], Xs, sysv, nsys, iters=20)
    tgn4 = time.time() - t1
    return dict(systems=nsys, chains=nsys * 16 * 32, iterations=nburn + niter - 1, seconds=round(dt, 3),
                chi2_evals_per_s=ev / dt, tflops=ev * F_QUAD8 / dt / 1e12,
                pct_fp32_peak=100.0 * ev * F_QUAD8 / dt / (PEAK_FP32_TFLOPS * 1e12),
                gn_seconds=round(tgn, 3), solved=int((resid < 1e-5).sum()),
                solved_fraction=float((resid < 1e-5).mean()),
                gn4_seconds=round(tgn4, 3), solved_best_of_4_sequences=int((resid4 < 1e-5).sum()),
                best_chain_resid_median=float(np.median(np.sqrt(c2 / 1e4))))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=25000)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-fits", type=int, default=64)
    ap.add_argument("--batched-iters", type=int, default=16)
    ap.add_argument("--batched-session-only", action="store_true",
                    help="batched leg: only the one-session run (no per-fit comparison sessions)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-iters", type=int, default=8)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-quadsys", action="store_true")
    ap.add_argument("--no-pcie", action="store_true")
    ap.add_argument("--streams", type=int, default=0,
                    help="sequence groups of every session (0: three for the plain fit, two for the "
                         "other legs; 1 for hardware-counter runs)")
    ap.add_argument("--launch-probe", type=int, default=20,
                    help="iterations of the per-launch timing probe (0: no probe session)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: every fit and sequence is independent, so the job "
              "shards by fit or sequence with no exchange during sampling (see the design notes); "
              "no multi-GPU figure is produced.", file=sys.stderr)
        return 2

    if not args.no_adjoint:
        # torch's HIP runtime is brought up before the engine library's, as a
        # user of the autograd interface would do
        import torch
        if not torch.cuda.is_available():
            log("torch sees no GPU: the adjoint leg is skipped")
            args.no_adjoint = True
        else:
            torch.zeros(1, device="cuda")
    from cuda_mcmc_amd import synthdata, mcmc_interf
    from cuda_mcmc_amd.engine import Session

    log(mcmc_interf.list_gpus().strip())
    d = synthdata.ring_gauss_dataset(nvis=2000, ncph=600, nsta=8, seed=2)
    nburn_cfg, niter_cfg = 5000, 20000
    total_cfg = nburn_cfg + niter_cfg
    # The timed window: the whole fit when --warmup + --steps cover it (the
    # default); otherwise --steps iterations of the sampling stage, after the
    # burn-in and --warmup more untimed iterations, so that a short window
    # sees adapted step sizes and the sampling-stage recording.
    if args.warmup + args.steps >= total_cfg:
        w0 = min(args.warmup, total_cfg)
    else:
        w0 = nburn_cfg + args.warmup
    w1 = min(w0 + args.steps, total_cfg)
    steps = w1 - w0
    fit = eht_fit(d, 32, 128, nburn=nburn_cfg, niter=niter_cfg, seed=0)
    opts = dict(headline_options(args))
    # sequence groups of the session (the engine caps them at the 64 exchange waves)
    ngroups = max(1, min(opts.get("streams", 1) or 2, 8, 64))
    with Session([fit], **opts) as ses:
        if w0 > 0:
            ses.run(w0)
        ses.sync()
        e1 = ses.stats()
        log("untimed iterations 0..%d done" % w0)
        # timed region: inputs and chain state resident in HBM
        t0 = time.time()
        ses.run(steps)
        ses.sync()
        wall = time.time() - t0
        e2 = ses.stats()
        gpu_ms = e2["run_ms"]
        log("timed iterations %d..%d: %.3f s wall, %.3f s device" % (w0, w1 - 1, wall, gpu_ms / 1e3))
        vis = e2["vis_evals"] - e1["vis_evals"]
        cph = e2["cph_evals"] - e1["cph_evals"]
        ev = e2["chi2_evals"] - e1["chi2_evals"]
        props = e2["proposals"] - e1["proposals"]
    # A second session of the same fit (no recording) run to the first
    # sampling iteration (or the window's start, if later): its state there
    # seeds the CPU baseline, and 20 more iterations with HIP events around
    # every launch give the step and exchange kernels' per-launch durations.
    p0 = max(w0, nburn_cfg)
    nprobe = max(0, args.launch_probe)
    step_ms = exch_ms = None
    if nprobe or not args.no_cpu_baseline:
        state = eht_fit(d, 32, 128, nburn=nburn_cfg, niter=niter_cfg, seed=0)
        state.pout = state.tout = state.chi2 = None
        with Session([state], **opts) as ses:
            ses.run(p0)
            ses.download()
            if nprobe:
                step_ms, exch_ms = ses.time_kernels(nprobe)
                log("per launch at iterations %d..%d: step %.4f ms, exchange %.4f ms"
                    % (p0, p0 + nprobe - 1, step_ms, exch_ms))
    value = vis / wall / 1e9
    flop = vis * F_VIS + cph * F_CPH + ev * F_CHAIN
    achieved = flop / (gpu_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "G visibility-evals/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / max(steps, 1), 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: ring+Gaussian source, 8 stations, 2000 amplitudes + 600 closure phases, data seed 2",
        "config": {"workload": "plain MH-RE fit, 9-param ring+Gaussian, 32 temperatures x 128 sequences, "
                               "2000 vis + 600 closures, 5000 burn-in + 20000 sampling iterations",
                   "nbeta": 32, "nseq": 128, "nvis": 2000, "ncph": 600, "nprm": 9,
                   "nburn": nburn_cfg, "niter": niter_cfg, "philox_seed": 0,
                   "timed_iterations": [w0, w1], "iterations_run": w1,
                   "timed_stage": "whole fit" if w0 == 0 and w1 == total_cfg else
                                  ("sampling (recording on)" if w0 >= nburn_cfg else "burn-in and sampling"),
                   "launch_form": opts},
        "pct_fp32_peak": round(100.0 * flop / wall / 1e12 / PEAK_FP32_TFLOPS, 2),
        "proposals_per_s": props / wall,
        "chi2_evals_per_s": ev / wall,
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": traffic_from_profiles(),
                     "kernel": "one MH-RE iteration: k_step_eht<Sgra9> on each of %d sequence-group streams "
                               "(overlapping), each also running the exchange of the iteration before" % ngroups,
                     "iteration_ms": round(gpu_ms / max(steps, 1), 5),
                     "step_launches_per_iteration": ngroups,
                     "step_launch_ms": None if step_ms is None else round(step_ms, 5),
                     "exchange_launch_ms": None if exch_ms is None else round(exch_ms, 5),
                     "launch_probe_iterations": [p0, p0 + nprobe] if nprobe else None,
                     "flop_per_iteration": flop / max(steps, 1)},
    }
    if not args.no_pcie:
        # the same whole fit through the host-buffer boundary: session
        # creation (upload), every iteration, and the download of the
        # recorded samples, timed on the host
        from cuda_mcmc_amd.engine import run_fits
        hf = eht_fit(d, 32, 128, nburn=nburn_cfg, niter=niter_cfg, seed=0)
        t0 = time.time()
        st = run_fits([hf], **opts)
        tw = time.time() - t0
        nbytes = sum(getattr(hf, k).nbytes for k in ("pout", "tout", "chi2", "pcur", "n_hist"))
        out["pcie_inclusive"] = {"value": round(st["vis_evals"] / tw / 1e9, 3), "unit": "G visibility-evals/s",
                                 "seconds": round(tw, 3), "iterations": total_cfg,
                                 "download_bytes": int(nbytes)}
        log("host buffers in and out: %.3f s for the whole fit" % tw)
    if not args.no_batched:
        log("batched vs plain (%d fits)..." % args.batched_fits)
        out["batched"] = batched_vs_plain(args.batched_iters, args.batched_fits,
                                          session_only=args.batched_session_only, streams=leg_streams(args))
    if not args.no_adjoint:
        log("adjoint gradient + Gauss-Newton...")
        out["adjoint"] = adjoint_leg(streams=leg_streams(args))
    if not args.no_quadsys:
        log("quadratic systems (1024 x 8 equations)...")
        out["quadsys"] = quadsys_leg(streams=leg_streams(args))
    if not args.no_cpu_baseline:
        nth = args.cpu_threads or host_threads()
        log("cpu baseline on %d threads..." % nth)
        out["cpu_baseline"] = cpu_baseline(d, state, p0, args.cpu_iters, nth)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
