#!/usr/bin/env python3
"""PIP ESS/s and MCMC iterations/s of the device PMV sampler: learned+batched vs plain.

Workload: n=10,000 individuals, p=500,000 SNPs, AR(1) LD blocks (rho ~
U(0.7, 0.95), block length U{20..100}), MAF ~ U(0.02, 0.5), 100 causal SNPs,
h2 = 0.5, additive coding, generated on the GPU from a fixed seed.

A "step" is one `Batch.run(iters_per_step)`: every chain of the batch does
`iters_per_step` MCMC iterations (moves, tau2/alpha refreshes, thinning and
the Rao-Blackwellised all-SNP scans of its schedule).  Batches run back to
back on the same dataset, each after `--warmup` untimed steps (which cover
the burn-in of 4 RB blocks, so adaptation and the learned table are frozen
before the timed steps):

* learned_batched: `--chains` chains (16) in one kernel with both learned
  proposals: the learned swap and the learned per-SNP add/remove table
  (both networks trained online during burn-in, frozen at its end);
* batched:         `--chains` chains with the plain moves only;
* learned_swap:    `--chains` chains, plain moves plus the learned swap only;
* learned_proposal: `--chains` chains, the learned per-SNP table only;
* plain:           1 chain, plain moves (the reference's sampler on the GPU),
                   `--plain-steps` timed steps (60: long enough for its own ESS);
* batched_full:    `--full-chains` chains (256: one per CU) with both learned
  proposals (`--full-proposal`), to show the whole GPU (reported beside the
  headline) and to carry the PIP check.

PIP ESS (median over the monitored SNPs: those whose PIP in the long
plain-method run lies in [0.05, 0.95]) is the multi-chain ESS
(`bmagwa_amd.ess.ess_multichain`: pooled within/between-chain variance,
Geyer's initial monotone sequence) of the thinned inclusion indicators of
the timed steps of all chains of a run, per wall second of the timed steps.
`value` is the learned_batched figure (the batched one with --no-learned).

Each run also reports its PIP ESS per iteration per chain.  The 1-chain CPU
oracle runs the same chain as the device (bit for bit), but a bounded CPU
sample is far shorter than the monitored SNPs' mixing time, so
`cpu_baseline.value` is its measured iterations/s times the plain method's
ESS per iteration per chain (from the plain batch); `headline_vs_plain_ess`
divides by the plain GPU chain's own measured ESS/s and
`headline_vs_plain_ess_rate` by the same rate estimate.

`pip_check` compares each run's Rao-Blackwellised PIPs with a long
plain-method run (`--pip-ref-chains` chains, 256 by default, plain moves,
`--pip-ref-steps` timed steps, its own seeds).

`roofline` covers the all-SNP X^T r scan kernel (HBM-bound) at the launch
shape the batch uses (3 residual vectors x 5 digit planes), timed with HIP
events on the stream it is launched on.  `cpu_baseline` is the CPU oracle's
plain sampler (one chain, one core) on the same dataset and settings, timed
over a bounded sample after an untimed warm part.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

N, P, N_CAUSAL, H2 = 10000, 500000, 100, 0.5
DELAY_REJECTION = 0      # --delay-rejection: the sampler option of every run
MAX_MODEL_SIZE = 0       # model-size limit of every run (0: none)
# --config large: n=50,000, p=1,000,000, MAF ~ U(0.01, 0.5), the same LD
# blocks, 300 causal SNPs, h2 = 0.5, 128 chains, models of at most 200 SNPs
LARGE = dict(N=50000, P=1000000, N_CAUSAL=300, H2=0.5, MAF=(0.01, 0.5), SEED=20261202 + 5000,
             chains=128, max_model_size=200)
MAF, LD = (0.02, 0.5), (0.7, 0.95, 20, 100)
SEED = 20261202 + 4000
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the tuned scan at this workload, from rocprofv3
# FETCH_SIZE / WRITE_SIZE passes scaled by a copy kernel of known traffic
# (profiles/kernels_scan_band_copy_n10000_p500000.md): fetch 2 x 626,386.4
# KiB, write 31,264 KiB
SCAN_TRAFFIC = dict(n=10000, p=500000, nvec=3, digits=5,
                    fetch_bytes=int(2 * 626737.6 * 1024), write_bytes=int(31269.7 * 1024),
                    source="profiles/kernels_scan_band_copy_n10000_p500000.md")
INT8_PEAK_TOPS = 5000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def indicator_rows(th, snps, first):
    """Thinned samples first.. of one chain -> bool (len(snps), T)."""
    from bmagwa_amd import ess
    return ess.indicator_matrix(th["sizes"], th["loci"], snps, burn=first)


def pooled_counts(th, first, p):
    sizes = th["sizes"].astype(np.int64)
    off = int(sizes[:first].sum())
    return np.bincount(th["loci"][off:].astype(np.int64), minlength=p), len(sizes) - first


def run_batch(bmg, synth, ctx, n_chains, steps, warmup, ips, seed0, batch_cfg=None, thin=1):
    total = (warmup + steps) * ips + 1000     # sizes the trace buffers
    cfgs = [synth.sampler_cfg(N, P, N_CAUSAL, types="A", seed=seed0 + c, chain=c,
                              do_n_iter=total, n_rao_burnin=4, thin=thin,
                              delay_rejection=DELAY_REJECTION)
            for c in range(n_chains)]
    bc = dict(batch_cfg or {})
    bc["trace_capacity"] = (total // thin + 16) * 200    # thinned loci: up to 200 per sample
    bc["max_model_size"] = MAX_MODEL_SIZE
    bc.setdefault("use_graph", 1)     # after burn-in: segments replayed from a captured hipGraph
    t0 = time.perf_counter()
    b = bmg.Batch([ctx], [0] * n_chains, cfgs, bc)
    b.setup()
    t_setup = time.perf_counter() - t0
    # burn-in (adaptation, and for the learned proposal its training rounds)
    # runs in the warm-up steps; its wall time is reported, not hidden
    t0 = time.perf_counter()
    for w in range(warmup):
        b.run(ips)
        log("%d chains warmup step %d/%d size %d" % (n_chains, w + 1, warmup, b.state(0)["size"]))
    b.sync()
    t_burnin = time.perf_counter() - t0
    first = [len(b.thinned(c)["sizes"]) for c in range(n_chains)]
    prop = (batch_cfg or {}).get("proposal", 0)
    learn = b.learn_state(0) if prop else None
    tm0 = b.timing()
    b.sync()
    t0 = time.perf_counter()
    for s in range(steps):
        b.run(ips)
        if (s + 1) % 10 == 0 and s + 1 < steps:
            log("%d chains timed step %d/%d" % (n_chains, s + 1, steps))
    b.sync()
    wall = time.perf_counter() - t0
    tm1 = b.timing()
    gs = b.graph_stats()
    log("%d chains: %d timed steps in %.2f s (%d graph replays)" % (n_chains, steps, wall, gs["replays"]))
    ths = [b.thinned(c) for c in range(n_chains)]
    sizes = [b.state(c)["size"] for c in range(n_chains)]
    # Rao-Blackwellised PIPs averaged over the post-burn-in RB scans, mean of chains
    p_rao_chains = np.array([b.p_rao(c) for c in range(n_chains)])
    p_rao = p_rao_chains.mean(axis=0)
    # split-half Monte-Carlo noise of the PIPs: max |difference| between the
    # means of the even and the odd chains
    half = (float(np.abs(p_rao_chains[0::2].mean(axis=0) - p_rao_chains[1::2].mean(axis=0)).max())
            if n_chains >= 2 else None)
    del p_rao_chains
    acc = [b.state(c)["n_accepted"] / max(b.state(c)["n_iter"], 1) for c in range(n_chains)]
    if learn is not None and not prop & 1:
        learn = dict(rounds=learn["rounds"], frozen=learn["frozen"],
                     table="per-SNP network (features: marginal score, RB running means, LD to the model)")
    elif learn is not None:
        tr = b.trace(0)
        mv, jd = tr["move"], tr["jd"]
        k0 = len(mv) - steps * ips      # first timed iteration
        learned = mv[k0:] == 3
        learn = dict(rounds=learn["rounds"], frozen=learn["frozen"], labels=learn["labels"],
                     window=learn["H"],
                     learned_swap_share=round(float(learned.mean()), 3) if len(learned) else None,
                     learned_swap_accept=round(float((jd[k0:][learned] > 0).mean()), 3)
                     if learned.any() else None)
        if prop & 2:
            learn["table"] = "per-SNP network (features: marginal score, RB running means, LD to the model)"
    b.close()
    return dict(wall=wall, steps=steps, setup=t_setup, burnin=t_burnin, half_split=half, ths=ths, first=first, sizes=sizes, acc=acc, learn=learn,
                graph_replays=gs["replays"], graph_ms=gs["graph_ms"],
                p_rao=p_rao,
                scan_ms=tm1["scan_ms"] - tm0["scan_ms"], chain_ms=tm1["chain_ms"] - tm0["chain_ms"],
                n_scans=tm1["n_scans"] - tm0["n_scans"])


def ess_of(run, monitored, wall):
    """Multi-chain PIP ESS per wall second: (median, min) over monitored SNPs."""
    from bmagwa_amd import ess
    mats = [indicator_rows(th, monitored, f) for th, f in zip(run["ths"], run["first"])]
    T = min(M.shape[1] for M in mats)
    if T < 4:
        return 0.0, 0.0
    e = ess.ess_multichain(np.stack([M[:, :T] for M in mats]))
    return float(np.median(e)) / wall, float(np.min(e)) / wall


def scan_roofline(bmg, ctx):
    """k_scan_mfma at the batch's launch shape, HIP events on its stream."""
    n_pad = ctx.n_pad()
    nvec, digits = 3, 5
    rng = np.random.default_rng(1)
    r = np.zeros((nvec, n_pad))
    r[:, :N] = rng.normal(size=(nvec, N))
    r_dev = bmg.DevBuf(r.nbytes)
    r_dev.upload(r)
    out_dev = bmg.DevBuf(nvec * P * 8)
    ms = ctx.scan_timed(r_dev.ptr, nvec, out_dev.ptr, 0, digits, 30)
    r_dev.free()
    out_dev.free()
    bytes_per_launch = P * ((N + 3) // 4) + nvec * (N * 8 + P * 8)
    gbs = bytes_per_launch / (ms * 1e-3) / 1e9
    return dict(ms=ms, bytes=bytes_per_launch, gbs=gbs, nvec=nvec, digits=digits,
                shape=ctx.scan_shape(nvec, digits))


def band_mfma(bmg, ctx, W=500):
    """k_band (int8 MFMA, exact int32) over all p SNPs at window W."""
    out = bmg.DevBuf(P * W * 4)
    ctx.band_timed(W, 0, 0, out.ptr, reps=1)  # first launch
    ms = ctx.band_timed(W, 0, 0, out.ptr, reps=3)
    out.free()
    ops = 2.0 * N * (P * W - W * (W - 1) / 2)  # 2n * sum_j min(W, p - j)
    return dict(ms=ms, W=W, tops=ops / (ms * 1e-3) / 1e12)


def copy_calibration(bmg, nbytes=1 << 30):
    """A plain copy of known traffic: scales the profiler's HBM counters."""
    a, b = bmg.DevBuf(nbytes), bmg.DevBuf(nbytes)
    ms = bmg.copy_timed(a, b, nbytes, reps=5)
    a.free()
    b.free()
    return dict(bytes=2 * nbytes, ms=ms, gbs=2 * nbytes / (ms * 1e-3) / 1e9)


def read_calibration(bmg, nbytes=1 << 31):
    """Streaming read of 2 GiB: the read rate the device reaches."""
    a, sink = bmg.DevBuf(nbytes), bmg.DevBuf(256)
    ms = bmg.read_timed(a, nbytes, sink, reps=5)
    a.free()
    sink.free()
    return dict(bytes=nbytes, ms=ms, gbs=nbytes / (ms * 1e-3) / 1e9)


def phase_profile(bmg, synth, ctx, args):
    """Per-chain device clocks (Batch.profile) over the timed steps: us per
    iteration by phase, and per add / per removal."""
    ips, n = args.iters_per_step, args.chains
    total = (args.warmup + args.steps) * ips + 1000
    cfgs = [synth.sampler_cfg(N, P, N_CAUSAL, types="A", seed=100 + c, chain=c, do_n_iter=total,
                              n_rao_burnin=4, thin=args.thin, delay_rejection=DELAY_REJECTION)
            for c in range(n)]
    b = bmg.Batch([ctx], [0] * n, cfgs, dict(proposal=args.phase_proposal, use_graph=0,
                                              max_model_size=MAX_MODEL_SIZE,
                                              trace_capacity=(total // args.thin + 16) * 200))
    b.setup()
    b.run(args.warmup * ips)
    p0 = [b.profile(c) for c in range(n)]
    t0 = b.timing()
    b.run(args.steps * ips)
    p1 = [b.profile(c) for c in range(n)]
    t1 = b.timing()
    it = args.steps * ips
    busy = []
    for c in range(n):
        d = {k: p1[c][k] - p0[c][k] for k in p0[c]}
        na, nr = max(d["n_add"], 1), max(d["n_rem"], 1)
        us = lambda k, per: round(d[k] / per * 1e3, 2)
        busy.append(d["total"] / it * 1e3)
        print(json.dumps(dict(chain=c, size=b.state(c)["size"], us_per_iter=us("total", it),
                              adds_per_iter=round(d["n_add"] / it, 2), removes_per_iter=round(d["n_rem"] / it, 2),
                              add_xy=us("add_xy", na), add_xtx=us("add_xtx", na), add_solve=us("add_trsv", na),
                              add_rest=us("add_rest", na), remove=us("rem", nr),
                              remove_downdate=us("rem_downdate", nr), remove_solve=us("rem_solve", nr),
                              refresh=us("tau2", it), refresh_draws=us("refresh_draws", it),
                              beta_sigma2=us("beta_sigma2", it), pve_thin=us("pve_thin", it),
                              addrem_moves=us("addrem", it), statechange=us("statechange", it),
                              swap=us("swap", it), second_stage=us("dr", it),
                              second_stage_scoring=us("dr_score", it), addrem_picks=us("ar_picks", it),
                              addrem_apply=us("ar_apply", it), addrem_reverse=us("ar_reverse", it),
                              addrem_decide=us("ar_decide", it))), flush=True)
    # a launch lasts as long as its slowest chain: the share of the chains'
    # workgroup-time spent waiting for it is the launch's idle fraction
    print(json.dumps(dict(kernel_us_per_iter=round((t1["chain_ms"] - t0["chain_ms"]) / it * 1e3, 1),
                          slowest_chain_us_per_iter=round(max(busy), 1),
                          mean_chain_us_per_iter=round(float(np.mean(busy)), 1),
                          idle_fraction=round(1.0 - float(np.mean(busy)) / max(busy), 3))), flush=True)
    b.close()


def cpu_baseline(ctx, y, monitored, seconds, thin):
    import oracle
    oracle.load()
    body = ctx.read_body()
    t0 = time.perf_counter()
    d = oracle.Data(N, P, body, y)
    del body
    from bmagwa_amd import synth
    cfg = synth.sampler_cfg(N, P, N_CAUSAL, types="A", seed=9001, chain=0, do_n_iter=10 ** 8,
                            n_rao_burnin=4, thin=thin, delay_rejection=DELAY_REJECTION)
    s = oracle.Sampler(cfg, d)
    s.init_flat()
    s.setup()
    t_setup = time.perf_counter() - t0
    log("cpu oracle setup %.1f s" % t_setup)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds / 3:
        s.iterate(100)
    first = len(s.thinned()["sizes"])
    it0 = s.state()["n_iter"]
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        s.iterate(100)
    wall = time.perf_counter() - t0
    its = s.state()["n_iter"] - it0
    from bmagwa_amd import ess
    th = s.thinned()
    M = ess.indicator_matrix(th["sizes"], th["loci"], monitored, burn=first)
    e = ess.ess_multichain(M[None]) if M.shape[1] >= 4 else np.zeros(len(monitored))
    log("cpu oracle: %d iterations in %.1f s" % (its, wall))
    return dict(value=float(np.median(e)) / wall, iters_per_s=its / wall, iters=int(its),
                wall=wall, samples=int(M.shape[1]), setup_s=t_setup)


def cpu_worker(body_path, y_path, seed, seconds, thin, ready_dir, n_workers, out_path, n, p, n_causal, dr):
    """One oracle chain on one core (a child process of the all-cores leg):
    set-up and an untimed warm part, then a barrier on the other workers, then
    `seconds` of iterations; writes its iteration count and wall time."""
    import oracle
    from bmagwa_amd import synth
    oracle.load()
    body = np.memmap(body_path, dtype=np.uint8, mode="r")
    y = np.fromfile(y_path, dtype=np.float64)
    d = oracle.Data(n, p, body, y)
    del body
    cfg = synth.sampler_cfg(n, p, n_causal, types="A", seed=seed, chain=0, do_n_iter=10 ** 8,
                            n_rao_burnin=4, thin=thin, delay_rejection=dr)
    s = oracle.Sampler(cfg, d)
    s.init_flat()
    s.setup()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds / 3:
        s.iterate(100)
    open(os.path.join(ready_dir, "ready.%d" % seed), "w").close()
    t0 = time.perf_counter()
    while (len([f for f in os.listdir(ready_dir) if f.startswith("ready.")]) < n_workers
           and time.perf_counter() - t0 < 120):
        time.sleep(0.01)
    it0 = s.state()["n_iter"]
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        s.iterate(100)
    wall = time.perf_counter() - t0
    with open(out_path, "w") as f:
        json.dump(dict(iters=int(s.state()["n_iter"] - it0), wall=wall), f)


def cpu_all_cores(ctx, y, seconds, thin):
    """One oracle chain per host core, each in its own process (the reference
    runs one chain per thread, src/main.cpp:79-95), on the same dataset:
    aggregate iterations/s over a common timed window."""
    import shutil
    import subprocess
    import tempfile
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    # the host's CPU share for one GPU is 16 cores
    cores = max(1, min(avail, int(os.environ.get("OMP_NUM_THREADS", "16")), 16))
    base = "/dev/shm" if os.path.isdir("/dev/shm") else None
    tmp = tempfile.mkdtemp(prefix="bmg_cpu_", dir=base)
    try:
        body_path, y_path = os.path.join(tmp, "body.u8"), os.path.join(tmp, "y.f64")
        ctx.read_body().tofile(body_path)
        np.asarray(y, dtype=np.float64).tofile(y_path)
        procs = []
        for w in range(cores):
            out = os.path.join(tmp, "out.%d.json" % w)
            procs.append((subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker",
                                            body_path, y_path, str(9001 + w), str(seconds), str(thin),
                                            tmp, str(cores), out, str(N), str(P), str(N_CAUSAL),
                                            str(DELAY_REJECTION)]), out))
        res = []
        for pr, out in procs:
            if pr.wait() != 0:
                raise RuntimeError("cpu worker failed with status %d" % pr.returncode)
            with open(out) as f:
                res.append(json.load(f))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    rate = sum(r["iters"] / r["wall"] for r in res)
    log("cpu oracle, %d cores: %.1f iterations/s aggregate" % (cores, rate))
    return dict(cores=cores, iters_per_s=rate, iters=sum(r["iters"] for r in res),
                wall=max(r["wall"] for r in res))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-worker":
        a = sys.argv[2:]
        cpu_worker(a[0], a[1], int(a[2]), float(a[3]), int(a[4]), a[5], int(a[6]), a[7], int(a[8]),
                   int(a[9]), int(a[10]), int(a[11]))
        return
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--iters-per-step", type=int, default=1000)
    ap.add_argument("--chains", type=int, default=None,
                    help="chains of the batched runs (16; 128 with --config large)")
    ap.add_argument("--config", choices=("default", "large"), default="default",
                    help="default: n=10,000, p=500,000, 100 causal, 16 chains (the headline "
                         "workload); large: n=50,000, p=1,000,000, 300 causal, 128 chains, models of "
                         "at most 200 SNPs")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the band and scan timings")
    ap.add_argument("--no-learned", dest="learned", action="store_false",
                    help="skip the learned-proposal batches (the headline is then the plain batched run)")
    ap.add_argument("--full-chains", type=int, default=256,
                    help="also run the learned batch with this many chains (e.g. 256: one per CU)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="time only the scan, band and copy kernels (for profiler runs)")
    ap.add_argument("--phase-profile", action="store_true",
                    help="print each chain's per-phase device clocks over the timed steps of a batch "
                         "(eager, so no graph hides the segments) and exit")
    ap.add_argument("--phase-proposal", type=int, default=0, choices=(0, 1, 2, 3),
                    help="proposal of the --phase-profile batch (0: plain; 3: both learned proposals)")
    ap.add_argument("--thin", type=int, default=10,
                    help="thinning (part of the sampler's schedule: beta and sigma2 are drawn at "
                         "every thinned iteration, as in the reference)")
    ap.add_argument("--delay-rejection", type=int, default=0,
                    help="the sampler's delay_rejection option in every run (0: none; the "
                         "reference example uses 10)")
    ap.add_argument("--plain-steps", type=int, default=60,
                    help="timed steps of the 1-chain plain run (long enough for its own ESS)")
    ap.add_argument("--full-proposal", type=int, default=3, choices=(0, 1, 2, 3),
                    help="proposal of the --full-chains run (1: learned swap, 2: learned per-SNP "
                         "proposal, 3: both; 0: plain)")
    ap.add_argument("--full-steps", type=int, default=80,
                    help="timed steps of the --full-chains run")
    ap.add_argument("--pip-ref-chains", type=int, default=256,
                    help="chains of the long plain-moves reference run for the PIP check (0: skip)")
    ap.add_argument("--pip-ref-steps", type=int, default=160,
                    help="timed steps of the PIP reference run")
    args = ap.parse_args()
    global DELAY_REJECTION, MAX_MODEL_SIZE, N, P, N_CAUSAL, H2, MAF, SEED
    DELAY_REJECTION = args.delay_rejection
    if args.config == "large":
        N, P, N_CAUSAL, H2 = LARGE["N"], LARGE["P"], LARGE["N_CAUSAL"], LARGE["H2"]
        MAF, SEED, MAX_MODEL_SIZE = LARGE["MAF"], LARGE["SEED"], LARGE["max_model_size"]
    if args.chains is None:
        args.chains = LARGE["chains"] if args.config == "large" else 16
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: chains and phenotypes are independent, so more "
              "GPUs would each run their own share of the chains; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import bmagwa_amd as bmg
    from bmagwa_amd import synth

    t0 = time.perf_counter()
    thresh, rho, _ = synth.snp_layout(P, MAF, LD, SEED)
    ctx = bmg.Ctx(N, P, synth=(SEED, thresh, rho), types=("A",))
    y, causal, _ = synth.phenotype(lambda s: ctx.decode(np.asarray(s), 0), N, P, N_CAUSAL, H2, SEED)
    ctx.set_y(y)
    t_data = time.perf_counter() - t0
    log("dataset on device in %.1f s" % t_data)

    if args.phase_profile:
        phase_profile(bmg, synth, ctx, args)
        return

    if args.kernels_only:
        sc = scan_roofline(bmg, ctx)
        bd = band_mfma(bmg, ctx)
        cp = copy_calibration(bmg)
        rd = read_calibration(bmg)
        print(json.dumps(dict(scan=sc, band=bd, copy=cp, read=rd)), flush=True)
        return

    ips = args.iters_per_step
    runs = {}
    th = args.thin
    if args.learned:
        runs["learned_batched"] = run_batch(bmg, synth, ctx, args.chains, args.steps, args.warmup, ips,
                                            100, dict(proposal=3, trace_every_iter=1), th)
    runs["batched"] = run_batch(bmg, synth, ctx, args.chains, args.steps, args.warmup, ips, 100,
                                dict(proposal=0), th)
    if args.learned:
        # each learned proposal alone (same seeds as the batched run): the
        # learned swap, and the per-SNP add/remove table set at the end of
        # burn-in
        runs["learned_swap"] = run_batch(bmg, synth, ctx, args.chains, args.steps, args.warmup, ips, 100,
                                         dict(proposal=1, trace_every_iter=1), th)
        runs["learned_proposal"] = run_batch(bmg, synth, ctx, args.chains, args.steps, args.warmup, ips, 100,
                                             dict(proposal=2), th)
    if not args.no_plain:
        runs["plain"] = run_batch(bmg, synth, ctx, 1, args.plain_steps, args.warmup, ips, 100,
                                  dict(proposal=0), th)
    if args.full_chains > 0:
        runs["batched_full"] = run_batch(bmg, synth, ctx, args.full_chains, args.full_steps,
                                         args.warmup, ips,
                                         100, dict(proposal=args.full_proposal if args.learned else 0,
                                                   trace_every_iter=1),
                                         th)
    # the long plain-method run the PIPs are checked against (its own seeds;
    # RB PIPs only, so its samples are thinned)
    pip_ref = None
    if args.pip_ref_chains > 0:
        pip_ref = run_batch(bmg, synth, ctx, args.pip_ref_chains, args.pip_ref_steps, args.warmup, ips,
                            50000, dict(proposal=0), 10)
        pip_ref["ths"] = None

    if pip_ref is not None:
        # SNPs whose PIP is uncertain according to the long plain-method run
        monitored = np.nonzero((pip_ref["p_rao"] >= 0.05) & (pip_ref["p_rao"] <= 0.95))[0]
    else:
        counts = np.zeros(P, np.int64)
        total = 0
        for run in runs.values():
            for th, f in zip(run["ths"], run["first"]):
                c, t = pooled_counts(th, f, P)
                counts += c
                total += t
        pip = counts / max(total, 1)
        monitored = np.nonzero((pip >= 0.05) & (pip <= 0.95))[0]
    summary = {}
    for name, run in runs.items():
        med, mn = ess_of(run, monitored, run["wall"])
        nch = len(run["sizes"])
        n_it = ips * run["steps"]
        summary[name] = dict(chains=nch, ess_per_s_median=round(med, 3), ess_per_s_min=round(mn, 3),
                             ess_per_iter_per_chain=float("%.4g" % (med * run["wall"] / (nch * n_it))),
                             iters_per_s=round(n_it * nch / run["wall"], 1),
                             timed_iters_per_chain=n_it,
                             ms_per_step=round(run["wall"] / run["steps"] * 1e3, 2),
                             model_size_mean=round(float(np.mean(run["sizes"])), 1),
                             accept_rate_mean=round(float(np.mean(run["acc"])), 4),
                             rb_block_ms=round(run["scan_ms"], 1), chain_kernel_ms=round(run["chain_ms"], 1),
                             graph_segments=run["graph_replays"], graph_ms=round(run["graph_ms"], 1),
                             burnin_s=round(run["burnin"], 2), setup_s=round(run["setup"], 2))
        if run.get("learn"):
            summary[name]["learner"] = run["learn"]
    # the plain method's PIP ESS per iteration per chain, from its 16-chain
    # batch (every chain is the single-chain sampler, bit for bit)
    e_plain = summary["batched"]["ess_per_iter_per_chain"]
    head = "learned_batched" if args.learned else "batched"
    lb = summary[head]
    b_med = lb["ess_per_s_median"]
    # Rao-Blackwellised PIPs of each run against the long plain-method run
    pip_check = {}
    if pip_ref is not None:
        n_ref = args.pip_ref_chains * (args.warmup + args.pip_ref_steps) * ips
        pip_check["reference"] = ("plain moves, %d chains x %d iterations (%d after burn-in), RB PIPs"
                                  % (args.pip_ref_chains, (args.warmup + args.pip_ref_steps) * ips,
                                     args.pip_ref_steps * ips))
        pip_check["reference_iterations"] = n_ref
        # the reference's own Monte-Carlo noise: its even chains against its
        # odd chains (each half has twice the variance of the whole run, so a
        # run as long as the reference that samples the same posterior
        # differs from it by about 1/sqrt(2) of this figure)
        pip_check["reference_half_split_max_abs_diff"] = round(pip_ref["half_split"], 4)
        for name, run in runs.items():
            dpip = np.abs(run["p_rao"] - pip_ref["p_rao"])
            pip_check[name] = dict(max_abs_diff=round(float(dpip.max()), 4),
                                   max_abs_diff_monitored=round(float(dpip[monitored].max()), 4),
                                   argmax_snp=int(np.argmax(dpip)),
                                   chains=len(run["sizes"]), timed_iters_per_chain=ips * run["steps"],
                                   half_split_max_abs_diff=(round(run["half_split"], 4)
                                                            if run["half_split"] is not None else None))
        # the target: a learned + batched run within 0.02 of the long plain
        # run; checked on the longest learned run (the whole-GPU batch)
        tgt = "batched_full" if "batched_full" in pip_check else head
        pip_check["target"] = dict(run=tgt, max_abs_diff=pip_check[tgt]["max_abs_diff"], bound=0.02,
                                   passed=bool(pip_check[tgt]["max_abs_diff"] <= 0.02))
    roof = None
    extra = {}
    if not args.no_extra:
        sc = scan_roofline(bmg, ctx)
        t = SCAN_TRAFFIC
        traffic = (t["fetch_bytes"] + t["write_bytes"]
                   if (t["n"], t["p"], t["nvec"], t["digits"]) == (N, P, sc["nvec"], sc["digits"]) else None)
        roof = dict(bound="hbm", achieved=round(sc["gbs"], 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(sc["gbs"] / HBM_PEAK_GBS, 4), traffic=traffic,
                    traffic_source=t["source"] if traffic else None,
                    kernel="k_scan_mfma<%d> (3 vectors x 5 digit planes)" % sc["shape"]["tiles_per_wave"],
                    autotuned_shape=sc["shape"],
                    ms_per_launch=round(sc["ms"], 4), algorithmic_bytes=sc["bytes"])
        bd = band_mfma(bmg, ctx)
        extra["band_xtx_int8"] = dict(W=bd["W"], ms=round(bd["ms"], 3), tops=round(bd["tops"], 1),
                                      frac_of_int8_peak=round(bd["tops"] / INT8_PEAK_TOPS, 4))
        cp = copy_calibration(bmg)
        extra["copy_calibration"] = dict(bytes=cp["bytes"], ms=round(cp["ms"], 4),
                                         gbs=round(cp["gbs"], 1))
        rd = read_calibration(bmg)
        extra["read_calibration"] = dict(bytes=rd["bytes"], ms=round(rd["ms"], 4), gbs=round(rd["gbs"], 1),
                                         scan_frac_of_read=round(sc["gbs"] / rd["gbs"], 4))

    cpu = None
    if not args.no_cpu:
        c = cpu_baseline(ctx, y, monitored, args.cpu_seconds, args.thin)
        ca = cpu_all_cores(ctx, y, args.cpu_seconds, args.thin)
        cpu = dict(value=round(e_plain * c["iters_per_s"], 4), unit="PIP ESS/s (median)", cores=1,
                   kind="port", iters_per_s=round(c["iters_per_s"], 1),
                   ess_per_s_own_sample=round(c["value"], 4),
                   sample="oracle plain PMV, 1 chain on 1 core, same dataset and settings: %d iterations "
                          "timed over %.1f s after an untimed warm part (setup %.1f s excluded); value = "
                          "iterations/s x the plain method's PIP ESS per iteration per chain (%.3g, from "
                          "the plain 16-chain GPU batch, whose chains are this sampler bit for bit): a "
                          "20 s sample is far shorter than the monitored SNPs' mixing time"
                          % (c["iters"], c["wall"], c["setup_s"], e_plain),
                   all_cores=dict(cores=ca["cores"], iters_per_s=round(ca["iters_per_s"], 1),
                                  value=round(e_plain * ca["iters_per_s"], 4),
                                  sample="one oracle chain per core in %d processes (own seeds), "
                                         "%d iterations over a common %.1f s window"
                                         % (ca["cores"], ca["iters"], ca["wall"])))

    wall = runs[head]["wall"]
    pl = summary.get("plain")
    line = {
        "metric": "PIP ESS/sec + MCMC iters/sec, learned+batched vs plain, 1 MI355X",
        "value": round(b_med, 3),
        "unit": "PIP ESS/s (median over monitored SNPs; multi-chain ESS of all chains of the run)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (generated on device, seeded)",
        "config": {"workload": "PMV MCMC over SNP inclusion, n=%d p=%d AR(1) LD, MAF U(%g, %g), "
                               "%d causal, h2=%g, additive" % (N, P, MAF[0], MAF[1], N_CAUSAL, H2),
                   "chains": args.chains, "iters_per_step": ips,
                   "max_model_size": MAX_MODEL_SIZE or None,
                   "delay_rejection": DELAY_REJECTION,
                   "proposal": ("plain moves + learned swap + learned per-SNP add/remove table "
                                "(two MLPs 16-256-256-1, bf16)" if args.learned
                                else "plain PMV moves (learned proposals not run at this size)")},
        "iters_per_s": lb["iters_per_s"],
        "runs": summary,
        "headline_run": head,
        "headline_vs_plain_ess": (round(b_med / pl["ess_per_s_median"], 2)
                                  if pl and pl["ess_per_s_median"] > 0 else None),
        "headline_vs_plain_ess_rate": (round(b_med / (e_plain * pl["iters_per_s"]), 2)
                                       if pl and e_plain > 0 else None),
        "headline_vs_plain_iters": (round(lb["iters_per_s"] / pl["iters_per_s"], 2)
                                    if pl and pl["iters_per_s"] > 0 else None),
        "monitored_snps": int(len(monitored)),
        "pip_check": pip_check,
        "setup_s": dict(dataset=round(t_data, 2),
                        batch=round(runs[head]["setup"], 2)),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
