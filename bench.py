"""Benchmark of the HOPS propagation hot path on one MI355X.

Workload (the batched non-adaptive configuration of the headline metric):
10-site chain, V = 50 cm^-1 nearest neighbour, E_n ~ N(0, 100 cm^-1) (seed 0),
2 Drude-Lorentz modes per site (lambda 35, gamma 50 cm^-1, 295 K, high-T mode +
1 Matsubara mode), LINEAR EOM, depth 4 (10626 auxiliaries x 10 states),
256 trajectories (noise seeds 0..255, FFT_FILTER, TAU 1 fs, TLEN 2048 fs),
dt = 2 fs.  A step is one RK4 step of all 256 trajectories; psi of every
trajectory is recorded on the device after every step.

Prints one JSON line.  value = trajectories x RK4 steps / s over the timed
region (inputs, operator and noise already in HBM).  The line also carries a
short measurement of the adaptive batch ("adaptive_c3", below).

--workload c3: adaptive NORMALIZED NONLINEAR batch — 100-site chain (same
bath), depth 10, delta_a = delta_s = 1e-3, psi0 on site 50, dt 1 fs, TAU
0.5 fs, TLEN 512 fs, noise seeds 0..B-1; every trajectory keeps its own basis
(ragged batch); complex64 stages, FP64 state carry, norms and error bounds.
--workload c5: adaptive NORMALIZED NONLINEAR batch on a 30 x 30 lattice,
V = 40 cm^-1, E_n ~ N(0, 150 cm^-1) (seed 0); per site the DL bath (lambda
35, gamma 50 cm^-1, 77 K) with 4 Matsubara terms: the high-T mode and the
first Matsubara mode in the hierarchy, the other three in the low-temperature
correction (sum g/w) and all five in the noise; depth 8, delta 1e-3, psi0 on
site (15, 15), dt 1 fs, TAU 0.5 fs, TLEN 1024 fs, 1024 trajectories.
--workload c4: serving stream — Poisson arrivals (200/s, seed 0) of adaptive
requests with chain length U{2..200}, depth U{2..10}, delta U[1e-4, 1e-2],
t_max U[200, 2000] fs, dt 1 fs; continuous batching; p50/p99 latency.  With
--c4-window W: an open stream at 200/s for W seconds, then streams at the
--c4-rates rates, each reporting offered against completed load, the queue
over time and whether it stayed bounded.  The default run ends with a 20 s
200/s stream ("serving_c4").
--workload dense: the dense system block on the matrix cores — a fully
coupled 100-site aggregate (point-dipole couplings V / |i - j|^3 between every
pair, V = 100 cm^-1, E_n ~ N(0, 100 cm^-1), one high-T DL mode per site),
LINEAR EOM, depth 2 (5151 auxiliaries x 100 states), B trajectories (default
64), dt 2 fs; -iH Y is a complex GEMM on the 16x16x4 MFMA (f32 with complex64
stages), reported against the 157.3 TFLOP/s dense fp32 peak.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "traj·RK4-steps/s at ≤1e-4 pop. error vs FP64; HBM GB/s; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# HBM bytes per stage launch of this workload, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes (scaled x2.00 and x1.00 by a 512 MiB device copy counted
# in the same runs), by stage kernel and interleave
PROFILED_TRAFFIC = {
    ("k_stage_lds<float, 4, 4, 2, 2>", 8): (1.250e9, "profiles/k_stage_lds_float_v2_nt2_chain10_d4_b256_lanes8.md"),
    ("k_stage_shared<float, 4, 4, 2, 1>", 4): (1.436e9, "profiles/k_stage_shared_float_nt2_chain10_d4_b256_lanes4.md"),
}


def c2_system():
    from mesohops_amd import bcf_convert_dl_to_exp, bcf_exp
    import scipy.sparse as sp
    nsite = 10
    rng = np.random.default_rng(0)
    H = np.diag(rng.normal(0.0, 100.0, nsite)).astype(np.complex128)
    for i in range(nsite - 1):
        H[i, i + 1] = H[i + 1, i] = 50.0
    modes = bcf_convert_dl_to_exp(35.0, 50.0, 295.0, k_matsubara=1)
    gw, ll = [], []
    for i in range(nsite):
        L = sp.coo_matrix(([1.0], ([i], [i])), shape=(nsite, nsite))
        for j in range(0, len(modes), 2):
            gw.append([modes[j], modes[j + 1]])
            ll.append(L)
    return dict(HAMILTONIAN=H, GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll, ALPHA_NOISE1=bcf_exp, PARAM_NOISE1=gw)


NOISE = {"MODEL": "FFT_FILTER", "TLEN": 2048.0, "TAU": 1.0}
EOM = {"EQUATION_OF_MOTION": "LINEAR"}
HIER = {"MAXHIER": 4}
DT = 2.0


def psi0():
    p = np.zeros(10, dtype=np.complex128)
    p[0] = 1.0
    return p


def _cpu_worker(seed, steps):
    """One trajectory of the workload on the FP64 oracle (1 thread); returns
    the seconds its propagation took."""
    from oracle import hops_oracle as O
    sysp = c2_system()
    sysp["ALPHA_NOISE1"] = O.bcf_exp
    tr = O.Trajectory(sysp, EOM, dict(NOISE, SEED=seed), hierarchy_param=HIER)
    tr.initialize(psi0())
    _ = tr.noise1.noise
    t0 = time.perf_counter()
    tr.propagate(steps * DT, DT)
    return time.perf_counter() - t0


def cpu_baseline(steps_per_proc, procs):
    """procs child processes (this script with --cpu-worker), one trajectory
    each with single-threaded BLAS, all waited for before returning."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    t0 = time.perf_counter()
    kids = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", str(s), str(steps_per_proc)],
                             stdout=subprocess.PIPE, env=env, text=True) for s in range(procs)]
    per = []
    for k in kids:
        out, _ = k.communicate()
        if k.returncode != 0:
            raise RuntimeError("CPU baseline worker failed (%d)" % k.returncode)
        per.append(float(out.strip().split()[-1]))
    wall = time.perf_counter() - t0
    # throughput over the propagation parts, all processes in parallel
    value = procs * steps_per_proc / max(per)
    return {"value": value, "unit": "traj·RK4-steps/s", "cores": procs, "kind": "port",
            "sample": "%d trajectories x %d RK4 steps of the workload on the FP64 NumPy/SciPy oracle, one "
                      "process per core (OMP_NUM_THREADS=1); slowest process %.1f s, wall %.1f s"
                      % (procs, steps_per_proc, max(per), wall)}


def algorithmic_bytes_per_traj_step(tables, B):
    """96 B per hierarchy element (complex128 state read + write, 4 stages x
    complex64 stage input read + derivative write), noise reads, and the
    shared operator amortised over the batch."""
    n_phi = tables.n_state * tables.n_hier
    s_op = tables.ell_w * n_phi * 12 + tables.n_hier * 16
    return 96 * n_phi + 3 * tables.n_act * 8 * 2 + s_op / B


def adaptive_workload(kind):
    """(system, psi0, noise, hierarchy) of the adaptive batch workloads: c3 the
    100-site chain at depth 10, c5 the 30 x 30 lattice at depth 8 with the
    low-temperature correction."""
    from mesohops_amd.workloads import chain_system, lattice_system
    if kind == "c5":
        sysp = lattice_system(30, 30)
        psi = np.zeros(900, np.complex128)
        psi[15 * 30 + 15] = 1.0
        return sysp, psi, {"MODEL": "FFT_FILTER", "TLEN": 1024.0, "TAU": 0.5}, {"MAXHIER": 8}
    sysp = chain_system(100)
    psi = np.zeros(100, np.complex128)
    psi[50] = 1.0
    return sysp, psi, {"MODEL": "FFT_FILTER", "TLEN": 512.0, "TAU": 0.5}, {"MAXHIER": 10}


def run_c3(B, steps, prec="mixed", single_steps=20, ref_traj=4, chunk=10, arena_bytes=None, log=None, kind="c3",
           baselines=True):
    """The adaptive batch: B trajectories, each with its own basis.  With
    baselines, also the plain per-trajectory FP64 path and one trajectory in
    the same runtime."""
    import torch
    from mesohops_amd import HopsTrajectory, _lib
    from mesohops_amd.ragged import AdaptiveBatch
    sysp, psi, noise, hier = adaptive_workload(kind)
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}

    def make(nb, p):
        ab = AdaptiveBatch(sysp, eom, dict(noise, SEED=0), None, hier, seeds=list(range(nb)), precision_mode=p,
                           arena_bytes=arena_bytes)
        ab.make_adaptive(1e-3, 1e-3)
        ab.initialize(psi, tau=1.0)
        return ab

    t0 = time.perf_counter()
    if log:
        log("%s: preparing %d trajectories (noise, first bases)" % (kind, B))
    ab = make(B, prec)
    torch.cuda.synchronize()
    t_init = time.perf_counter() - t0
    if log:
        log("%s: prepared in %.1f s" % (kind, t_init))
    ab.rt.profile_enable(True)
    chunks = []
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        c = min(chunk, steps - done)
        t1 = time.perf_counter()
        ab.propagate(float(c), 1.0)
        torch.cuda.synchronize()
        infos = [ab.rt.info(sl) for sl in ab.slots]
        chunks.append({"t_fs": done + c, "s": time.perf_counter() - t1,
                       "mean_n_phi": float(np.mean([i["n_state"] * i["n_hier"] for i in infos])),
                       "max_n_hier": int(max(i["n_hier"] for i in infos))})
        done += c
        if log:
            log("%s t=%d fs: %.0f traj*steps/s, mean N_phi %.0f" % (kind, done, B * c / chunks[-1]["s"],
                                                                  chunks[-1]["mean_n_phi"]))
    elapsed = time.perf_counter() - t0
    ms, n = ab.rt.profile_read()
    ab.rt.profile_enable(False)
    timers = ab.rt.timers()
    es = 16 if prec == "fp64" else 8
    stage_bytes = timers["rows_stepped"] * (8 * es + 32)
    achieved = stage_bytes / (ms[1] * 1e-3) / 1e9 if ms[1] > 0 else 0.0
    out = {"value": B * steps / elapsed, "trajectories": B, "sub_batches": ab.K, "threads_per_sub_batch": ab.n_threads,
           "steps": steps, "elapsed_s": elapsed, "init_s": t_init,
           "ms_per_step": elapsed / steps * 1e3, "chunks": chunks,
           "host_phase_s": {k: v for k, v in timers.items() if k != "rows_stepped"},
           "stage_kernel": {"avg_ms": ms[1] / max(n[1], 1), "launches": n[1],
                            "algorithmic_bytes": stage_bytes, "achieved_GBs": achieved,
                            "frac_of_8TBs": achieved / HBM_PEAK_GBS}}
    if not baselines:
        ab.rt.close()
        return out
    # plain per-trajectory FP64 GPU path over the first single_steps steps
    if log:
        log("%s: plain per-trajectory FP64 path, %d steps" % (kind, single_steps))
    tr = HopsTrajectory(sysp, eom, dict(noise, SEED=0), hierarchy_param=hier)
    tr.make_adaptive(1e-3, 1e-3)
    tr.initialize(psi)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    import contextlib
    with contextlib.redirect_stdout(sys.stderr):  # the reference's progress print stays off the JSON line
        tr.propagate(float(single_steps), 1.0)
    torch.cuda.synchronize()
    single_rate = single_steps / (time.perf_counter() - t1)
    # the batch's rate over the same first steps (whole chunks)
    win, prev = [], 0
    for c in chunks:
        if c["t_fs"] > single_steps and win:
            break
        win.append((c["t_fs"] - prev, c["s"]))
        prev = c["t_fs"]
    batch_win = B * sum(w[0] for w in win) / sum(w[1] for w in win)
    # one trajectory alone in the same native runtime, FP64, every step
    if log:
        log("%s: one trajectory in the same runtime, FP64, %d steps" % (kind, steps))
    one = make(1, "fp64")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for t_done in range(0, steps, chunk):
        one.propagate(float(min(chunk, steps - t_done)), 1.0)
    torch.cuda.synchronize()
    one_rate = steps / (time.perf_counter() - t1)
    one.rt.close()
    out.update({"gpu_fp64_single_traj_steps_per_s": single_rate, "single_window_steps": single_steps,
                "batch_rate_same_window": batch_win,
                "speedup_vs_gpu_fp64_single_same_window": batch_win / single_rate,
                "same_runtime_fp64_b1_steps_per_s": one_rate,
                "speedup_vs_same_runtime_fp64_b1": (B * steps / elapsed) / one_rate})
    if ref_traj > 0:
        ref = make(ref_traj, "fp64")
        for t_done in range(0, steps, chunk):
            ref.propagate(float(min(chunk, steps - t_done)), 1.0)
            if log:
                log("%s: FP64 reference, t=%d fs" % (kind, min(steps, t_done + chunk)))
        err = 0.0
        for b in range(ref_traj):
            pa = np.abs(ab.results(b)[0]) ** 2
            pb = np.abs(ref.results(b)[0]) ** 2
            err = max(err, float(np.max(np.abs(pa - pb))))
        out["pop_err_vs_fp64"] = err
        out["pop_err_trajectories"] = ref_traj
        ref.rt.close()
    ab.rt.close()
    return out


def run_c4(n_req, prec="mixed", time_limit=600.0, log=None):
    from mesohops_amd.serving import ServingLoop
    from mesohops_amd.workloads import c4_requests
    reqs = c4_requests(n_req)
    loop = ServingLoop(precision_mode=prec, max_slots=512, record=False)
    done, stats = loop.run(reqs, time_limit=time_limit, log=log)
    stats["host_phase_s"] = {k: v for k, v in loop.rt.timers().items() if k != "rows_stepped"}
    loop.close()
    return done, stats


def main_c3(args):
    c5 = args.workload == "c5"
    B = args.batch if args.batch != 256 else (1024 if c5 else 256)
    out3 = run_c3(B, args.steps, args.precision, args.single_steps, args.ref_traj,
                  log=lambda m: print(m, file=sys.stderr, flush=True), kind=args.workload)
    sk = out3["stage_kernel"]
    if c5:
        wl = ("adaptive NORMALIZED NONLINEAR batch: 30x30 lattice (V 40, E ~ N(0,150)), per site 2 DL modes "
              "(high-T + 1 Matsubara, 77 K) in the hierarchy and 3 more Matsubara modes in the low-temperature "
              "correction and the noise, depth 8, delta_a = delta_s = 1e-3, %d trajectories with their own bases, "
              "dt 1 fs, TLEN 1024 fs, t = 0..%d fs" % (out3["trajectories"], args.steps))
    else:
        wl = ("adaptive NORMALIZED NONLINEAR batch: 100-site chain, 2 DL modes/site, depth 10, "
              "delta_a = delta_s = 1e-3, %d trajectories with their own bases, dt 1 fs, "
              "t = 0..%d fs" % (out3["trajectories"], args.steps))
    out = {"metric": METRIC, "value": out3["value"], "unit": "traj·RK4-steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": 0, "ms_per_step": out3["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "c64" if args.precision == "mixed" else "c128",
           "data": "synthetic",
           "config": {"workload": wl},
           "roofline": {"bound": "hbm", "achieved": sk["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": sk["frac_of_8TBs"], "traffic": None, "kernel": "k_stage_slots",
                        "kernel_avg_ms": sk["avg_ms"]},
           "adaptive": out3}
    print(json.dumps(out), flush=True)


def run_c4_stream(rate, window, prec="mixed", log=None):
    """Requests of the serving distribution arriving at ``rate`` per second
    for ``window`` seconds (the loop stops at the end of the window).
    Returns the loop's statistics plus offered and completed load and whether
    the number of requests in the system (waiting + running) stopped growing
    over the second half of the window."""
    from mesohops_amd.serving import ServingLoop
    from mesohops_amd.workloads import c4_requests
    n = int(math.ceil(rate * window * 1.2)) + 8
    reqs = [r for r in c4_requests(n, rate=rate) if r.arrival <= window]
    loop = ServingLoop(precision_mode=prec, max_slots=512, record=False)
    done, stats = loop.run(reqs, time_limit=window, log=log)
    loop.close()
    tl = stats["timeline"]
    in_sys = [(t, w + a) for (t, _, w, a, _) in tl]
    mid = min(in_sys, key=lambda x: abs(x[0] - window / 2))
    end = in_sys[-1]
    stats["in_system_mid"], stats["in_system_end"] = mid[1], end[1]
    wait2 = max((w for (t, _, w, _, _) in tl if t >= mid[0]), default=0)
    stats["max_waiting_second_half"] = wait2
    # bounded: no request waits for admission in the second half, and the
    # second half completes at least 80 % of its arrival rate
    done_mid = min(tl, key=lambda x: abs(x[0] - window / 2))[4]
    rate2 = (tl[-1][4] - done_mid) / max(end[0] - mid[0], 1e-9)
    stats["completed_rate_second_half"] = rate2
    stats["bounded"] = wait2 <= max(2, 0.02 * stats["arrived"]) and rate2 >= 0.8 * rate
    stats["offered_rate"] = rate
    stats["completed_rate"] = stats["completed"] / stats["wall_s"]
    stats["timeline"] = tl[:: max(1, len(tl) // 12)] + ([tl[-1]] if len(tl) > 12 else [])
    return stats


MFMA_PEAK_TFLOPS = {"mixed": 157.3, "fp64": 78.6}  # dense fp32 / fp64 matrix peaks of MI355X


def main_dense(args):
    """Batched LINEAR propagation of a fully coupled aggregate: the dense
    system block (k_dense_h, MFMA) next to the stage kernel.  Algorithmic
    flops of one k_dense_h launch: 8 n_state^2 n_hier B (complex GEMM of
    -iH over every auxiliary block of every trajectory)."""
    import torch
    from mesohops_amd import HopsBatch, _lib
    from mesohops_amd.workloads import aggregate_system
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    nsite, depth = args.dense_sites, args.dense_depth
    B = args.batch if args.batch != 256 else 64
    sysp = aggregate_system(nsite)
    noise = {"MODEL": "FFT_FILTER", "TLEN": max(64.0, (args.steps + 2) * DT + 8.0), "TAU": 1.0}
    psi = np.zeros(nsite, np.complex128)
    psi[nsite // 2] = 1.0
    dev = torch.device("cuda:0")
    ns = int(round(DT * 0.5 / noise["TAU"]))

    def make(prec):
        b = HopsBatch(sysp, EOM, dict(noise, SEED=0), None, {"MAXHIER": depth}, seeds=list(range(B)),
                      precision_mode=prec, device="cuda:0", store_psi=True, lanes=args.lanes or None)
        b.initialize(psi)
        return b

    def run_on(state, steps):
        if state.psi_store is None or state.store_slots < steps:
            state.store_slots = steps
            state.psi_store = torch.zeros(B * steps * state_n, dtype=torch.complex128, device=dev)
            state.bt.store_stride = steps
        state.bt.psi_store_d = state.psi_store.data_ptr()
        state.propagate(0, steps, DT, ns, 0)

    out = {}
    choice = None
    from mesohops_amd import autotune
    for prec in (args.precision, "fp64") if args.precision != "fp64" else ("fp64",):
        log("dense: %d-site aggregate, depth %d, %d trajectories, %s" % (nsite, depth, B, prec))
        batch = make(prec)
        tb = batch.tables
        if not tb.dense_h:
            sys.exit("dense: the system block did not take the dense path")
        state_n = tb.n_state
        if args.lanes == 0:
            # the layout and stage-kernel form tuned on the first batch, kept
            # for the second
            if choice is None:
                choice, tune_t = autotune.tune(batch, psi, DT, verbose=True)
                out["layout"] = {"lanes": choice[0], "stationary_stage": bool(choice[1]), "nt_stream": choice[3],
                                 "vec_lanes": choice[6], "lds_stage": bool(choice[7])}
            else:
                batch.relayout(choice[0])
                autotune.apply(choice, batch)
                batch.rewind(psi)
        st = batch.state
        run_on(st, args.warmup)
        torch.cuda.synchronize()
        batch.rewind(psi)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run_on(st, args.steps)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        psi_t = st.psi_store_host(args.steps)
        batch.rewind(psi)
        torch.cuda.synchronize()
        _lib.profile_enable(True)
        run_on(st, args.steps)
        torch.cuda.synchronize()
        ms, n = _lib.profile_read()
        _lib.profile_enable(False)
        flops = 8.0 * tb.n_state ** 2 * tb.n_hier * B
        dense_ms = ms[3] / max(n[3], 1)
        achieved = flops / (dense_ms * 1e-3) / 1e12 if n[3] else 0.0
        out[prec] = {"value": B * args.steps / elapsed, "elapsed_s": elapsed, "ms_per_step": elapsed / args.steps * 1e3,
                     "dense_kernel": {"avg_ms": dense_ms, "launches": n[3], "flops_per_launch": flops,
                                      "achieved_TFLOPs": achieved, "peak_TFLOPs": MFMA_PEAK_TFLOPS[prec],
                                      "frac": achieved / MFMA_PEAK_TFLOPS[prec]},
                     "stage_kernel_avg_ms": ms[1] / max(n[1], 1), "prologue_avg_ms": ms[0] / max(n[0], 1),
                     "psi": psi_t}
        del batch, st
        torch.cuda.empty_cache()
    layout = out.pop("layout", None)
    main = out[args.precision]
    err = None
    if args.precision != "fp64":
        pm, pr = np.abs(out[args.precision]["psi"]) ** 2, np.abs(out["fp64"]["psi"]) ** 2
        err = float(np.max(np.abs(pm - pr) / pr.sum(-1, keepdims=True)))
    for v in out.values():
        v.pop("psi")
    dk = main["dense_kernel"]
    line = {"metric": METRIC, "value": main["value"], "unit": "traj·RK4-steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": main["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "c64" if args.precision == "mixed" else "c128", "data": "synthetic",
            "config": {"workload": "batched LINEAR HOPS of a fully coupled %d-site aggregate (dense H), 1 DL mode/site, "
                                   "depth %d (%d aux x %d states), %d trajectories, dt 2 fs"
                                   % (nsite, depth, int(round(dk["flops_per_launch"] / (8.0 * nsite ** 2 * B))),
                                      nsite, B)},
            "roofline": {"bound": "mfma", "achieved": dk["achieved_TFLOPs"], "peak": dk["peak_TFLOPs"],
                         "unit": "TFLOP/s", "frac": dk["frac"], "traffic": None, "kernel": "k_dense_h",
                         "kernel_avg_ms": dk["avg_ms"]},
            "pop_err_rel_vs_fp64": err, "layout": layout, "legs": out}
    print(json.dumps(line), flush=True)


def main_c4(args):
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if args.c4_window > 0:
        # an open stream at the configured rate, then lower rates, to find the
        # largest rate at which the system keeps up
        main = run_c4_stream(200.0, args.c4_window, args.precision, log=log)
        sweep = {}
        for r in [float(x) for x in args.c4_rates.split(",") if x]:
            log("c4: stream at %.2f requests/s for %.0f s" % (r, args.c4_window))
            sweep[r] = run_c4_stream(r, args.c4_window, args.precision, log=log)
        bounded = [r for r, st in sweep.items() if st["bounded"]]
        out = {"metric": METRIC, "value": main["traj_steps_per_s"], "unit": "traj·RK4-steps/s", "n_gpus": 1,
               "steps": main["iterations"], "warmup": 0,
               "ms_per_step": main["wall_s"] / max(main["iterations"], 1) * 1e3, "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "c64" if args.precision == "mixed" else "c128",
               "data": "synthetic",
               "config": {"workload": "serving stream: adaptive requests arriving as a Poisson stream at 200/s for "
                                      "%.0f s, chain 2-200 sites, depth 2-10, delta 1e-4..1e-2, t_max 200-2000 fs, "
                                      "dt 1 fs" % args.c4_window},
               "p50_latency_s": main["p50_latency_s"], "p99_latency_s": main["p99_latency_s"],
               "offered": {"rate": 200.0, "arrived": main["arrived"], "completed": main["completed"],
                           "running_at_end": main["admitted_unfinished"], "waiting_at_end": main["waiting_at_end"]},
               "serving": main,
               "rate_sweep": {str(r): {k: st[k] for k in ("arrived", "completed", "completed_rate",
                                                            "admitted_unfinished", "waiting_at_end", "in_system_mid",
                                                            "in_system_end", "max_waiting_second_half",
                                                            "completed_rate_second_half", "bounded", "p50_latency_s",
                                                            "p99_latency_s", "traj_steps_per_s", "timeline")}
                              for r, st in sweep.items()},
               "largest_bounded_rate": max(bounded) if bounded else None,
               "latency_note": "p50/p99 over the requests completed inside the window; at 200/s the stream "
                               "exceeds capacity, so these are the short requests"}
        print(json.dumps(out), flush=True)
        return
    done, stats = run_c4(args.requests, args.precision, log=log)
    out = {"metric": METRIC, "value": stats["traj_steps_per_s"], "unit": "traj·RK4-steps/s", "n_gpus": 1,
           "steps": stats["iterations"], "warmup": 0, "ms_per_step": stats["wall_s"] / max(stats["iterations"], 1) * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "c64" if args.precision == "mixed" else "c128", "data": "synthetic",
           "config": {"workload": "serving stream: %d adaptive requests, Poisson arrivals 200/s, chain 2-200 sites, "
                                  "depth 2-10, delta 1e-4..1e-2, t_max 200-2000 fs, dt 1 fs" % args.requests},
           "p50_latency_s": stats["p50_latency_s"], "p99_latency_s": stats["p99_latency_s"], "serving": stats}
    print(json.dumps(out), flush=True)


def main():
    if len(sys.argv) == 4 and sys.argv[1] == "--cpu-worker":
        print(_cpu_worker(int(sys.argv[2]), int(sys.argv[3])), flush=True)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--ref-traj", type=int, default=32,
                    help="adaptive workloads: trajectories re-run in FP64 for the population error")
    ap.add_argument("--single-steps", type=int, default=200, help="steps of the plain B=1 FP64 GPU path")
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lanes", type=int, default=0, help="trajectories interleaved per row (0: autotune)")
    ap.add_argument("--chunk", type=int, default=0, help="trajectories per cache-resident chunk with --lanes")
    ap.add_argument("--nt", type=int, default=0, help="nontemporal hint on streamed arrays with --lanes")
    ap.add_argument("--groups-q", type=int, default=0, help="trajectory-group split with --lanes (0: automatic)")
    ap.add_argument("--tile-order", type=int, default=0, help="locality tile order with --lanes (0 or 1)")
    ap.add_argument("--vec", type=int, default=1, help="trajectory lanes per thread with --lanes (1, 2 or 4)")
    ap.add_argument("--lds", type=int, default=0, help="LDS-staged stationary kernel with --lanes (0 or 1)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a plain 512 MiB device copy (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "c5", "dense"])
    ap.add_argument("--dense-sites", type=int, default=100, help="dense: sites of the fully coupled aggregate")
    ap.add_argument("--dense-depth", type=int, default=2, help="dense: hierarchy depth")
    ap.add_argument("--requests", type=int, default=64, help="c4: requests in the stream")
    ap.add_argument("--c4-window", type=float, default=0.0,
                    help="c4: serve an open 200/s stream for this many seconds, then the --c4-rates streams "
                         "(0: the --requests burst instead)")
    ap.add_argument("--c4-rates", default="2,4,6,8", help="c4: lower arrival rates (requests/s) for --c4-window")
    ap.add_argument("--c4-leg", type=float, default=20.0,
                    help="seconds of the 200/s serving stream in the default run (0: off)")
    ap.add_argument("--c3-steps", type=int, default=500, help="steps of the adaptive leg in the c2 run (0: off)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (trajectories would shard "
              "one batch per GPU, see DESIGN.md); got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload in ("c3", "c5"):
        if args.workload == "c3" and args.steps > 500:
            sys.exit("c3 runs at most 500 steps of 1 fs (TLEN 512 fs)")
        if args.workload == "c5" and args.steps > 1000:
            sys.exit("c5 runs at most 1000 steps of 1 fs (TLEN 1024 fs)")
        return main_c3(args)
    if args.workload == "c4":
        return main_c4(args)
    if args.workload == "dense":
        return main_dense(args)
    if (args.steps + 1) * DT > NOISE["TLEN"] or (args.warmup + 1) * DT > NOISE["TLEN"]:
        sys.exit("steps x dt exceeds the noise length %.0f fs" % NOISE["TLEN"])
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    cpu = None
    if not args.no_cpu:
        # child processes started before this process touches the GPU
        procs = min(16, os.cpu_count() or 1)
        log("cpu baseline: %d processes x %d steps of the oracle" % (procs, args.cpu_steps))
        cpu = cpu_baseline(args.cpu_steps, procs)
        log("cpu baseline: %.1f traj*steps/s" % cpu["value"])

    import torch
    from mesohops_amd import HopsBatch, _lib
    from mesohops_amd.engine import _stream_ptr

    dev = torch.device("cuda:0")
    sysp = c2_system()
    B = args.batch
    t_init = time.perf_counter()
    batch = HopsBatch(sysp, EOM, dict(NOISE, SEED=0), None, HIER, seeds=list(range(B)),
                      precision_mode=args.precision, device="cuda:0", store_psi=True)
    batch.initialize(psi0())
    torch.cuda.synchronize()
    t_init = time.perf_counter() - t_init
    tb = batch.tables
    ns = int(round(DT * 0.5 / NOISE["TAU"]))
    from mesohops_amd import autotune
    if args.lanes == 0:
        choice, tune_t = autotune.tune(batch, psi0(), DT, verbose=True)
    else:
        batch.relayout(args.lanes)
        choice, tune_t = (args.lanes, True, args.chunk, args.nt, args.groups_q, args.tile_order, args.vec, args.lds), {}
        autotune.apply(choice, batch)
    st = batch.state

    def run_on(state, steps):
        """steps RK4 steps from t = 0, psi stored on the device (no download)."""
        if state.psi_store is None or state.store_slots < steps:
            state.store_slots = steps
            state.psi_store = torch.zeros(B * steps * tb.n_state, dtype=torch.complex128, device=dev)
            state.bt.store_stride = steps
        state.bt.psi_store_d = state.psi_store.data_ptr()
        state.propagate(0, steps, DT, ns, 0)

    def run(steps):
        run_on(st, steps)

    # warmup, then rewind to t = 0 with the psi store sized for the timed region
    run(args.warmup)
    torch.cuda.synchronize()
    batch.rewind(psi0())
    st.psi_store = None
    st.psi_store = torch.zeros(B * args.steps * tb.n_state, dtype=torch.complex128, device=dev)
    st.store_slots = args.steps
    st.bt.store_stride = args.steps
    torch.cuda.synchronize()

    # timed region: the default dispatch (graph-replayed steps), no profiler
    log("c2: timed region, %d steps" % args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = B * args.steps / elapsed
    psi_mixed = st.psi_store_host(args.steps)
    # the same steps again with a HIP event after every kernel (launched
    # eagerly): per-kernel times for the roofline
    batch.rewind(psi0())
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    t0 = time.perf_counter()
    run(args.steps)
    torch.cuda.synchronize()
    elapsed_prof = time.perf_counter() - t0
    ms, n = _lib.profile_read()
    _lib.profile_enable(False)
    prof_equal = bool(np.array_equal(st.psi_store_host(args.steps), psi_mixed))

    # the same batch with complex128 stages: its rate, and the FP64 reference
    # of the population error on every trajectory (identical noise)
    log("c2: complex128 leg, %d trajectories x %d steps" % (B, args.steps))
    ref = HopsBatch(sysp, EOM, dict(NOISE, SEED=0), None, HIER, seeds=list(range(B)),
                    precision_mode="fp64", device="cuda:0", store_psi=True)
    ref.initialize(psi0())
    if ref.state.lanes != st.lanes:
        ref.relayout(st.lanes)
    run_on(ref.state, 2)   # warm-up, then rewind
    ref.rewind(psi0())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_on(ref.state, args.steps)
    torch.cuda.synchronize()
    c128_elapsed = time.perf_counter() - t0
    pref = ref.state.psi_store_host(args.steps)
    pm = np.abs(psi_mixed) ** 2
    pr = np.abs(pref) ** 2
    pop_err_abs = float(np.max(np.abs(pm - pr)))
    # LINEAR trajectories are not normalised: also the error of the
    # normalised populations |psi_n|^2 / |psi|^2
    pop_err_norm = float(np.max(np.abs(pm / pm.sum(-1, keepdims=True) - pr / pr.sum(-1, keepdims=True))))
    # the same difference relative to the trajectory's |psi|^2 at that time
    pop_err_rel = float(np.max(np.abs(pm - pr) / pr.sum(-1, keepdims=True)))
    max_pop = float(np.max(pr))
    nref = B
    del ref, pref
    torch.cuda.empty_cache()

    # plain per-trajectory FP64 GPU path (B = 1, no batching, every kernel
    # launched eagerly)
    _lib.set_option("graph_steps", 0)
    single = HopsBatch(sysp, EOM, dict(NOISE, SEED=0), None, HIER, seeds=[0], precision_mode="fp64",
                       device="cuda:0", store_psi=True)
    single.initialize(psi0())
    ss = min(args.single_steps, args.steps)
    single.propagate(2 * DT, DT)
    single.rewind(psi0())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    single.propagate(ss * DT, DT)
    torch.cuda.synchronize()
    single_rate = ss / (time.perf_counter() - t1)
    del single
    _lib.set_option("graph_steps", 16)

    if args.calib_copy:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device=dev)
        c = torch.empty_like(a)
        for _ in range(3):
            c.copy_(a)
        torch.cuda.synchronize()
        del a, c

    stage_ms = ms[1] / max(n[1], 1)
    bytes_step = algorithmic_bytes_per_traj_step(tb, B)
    # launches may cover a chunk of the batch: spread the timed region's
    # algorithmic bytes over the stage launches that moved them
    bytes_launch = bytes_step * B * args.steps / max(n[1], 1)
    achieved = bytes_launch / (stage_ms * 1e-3) / 1e9
    rt = "double" if args.precision == "fp64" else "float"
    pat_w = int(np.max(np.diff(np.asarray(tb.tables["pat_ptr"])))) if len(tb.tables["pat_ptr"]) > 1 else 0
    if choice[1] and 0 < pat_w <= 8 and tb.ell_w <= 16:
        vec = choice[6]
        while vec > 1 and (vec > choice[0] or B % vec != 0):
            vec //= 2
        kernel_name = "%s<%s, %d, %d, %d, %d>" % ("k_stage_lds" if choice[7] else "k_stage_shared", rt, 4 if tb.ell_w <= 4 else (8 if tb.ell_w <= 8 else 16),
                                                              4 if pat_w <= 4 else 8, choice[3], vec)
    else:
        kernel_name = "k_stage<%s>" % rt
    traffic, traffic_src = PROFILED_TRAFFIC.get((kernel_name, choice[0]), (None, None))
    if B != 256 or args.precision != "mixed" or choice[2]:
        traffic, traffic_src = None, None  # profiled for the default workload only
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "traj·RK4-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64" if args.precision == "mixed" else "c128",
        "data": "synthetic",
        "config": {"workload": "batched non-adaptive LINEAR HOPS: 10-site chain, 2 DL modes/site, depth 4 "
                               "(10626 aux x 10 states), %d trajectories, dt 2 fs; complex64 stages with a "
                               "complex128 state carry" % B if args.precision == "mixed" else
                               "same workload in complex128",
                   "trajectories": B, "n_state": tb.n_state, "n_hier": tb.n_hier, "dt_fs": DT,
                   "precision": args.precision},
        "pop_err_vs_fp64": pop_err_abs,
        "pop_err_definition": "max over t, n, trajectory of | |psi_n|^2 (this run) - |psi_n|^2 (complex128 run of "
                              "the same batch, identical noise) |",
        "pop_err_normalised_vs_fp64": pop_err_norm,
        "max_unnormalised_population": max_pop,
        "pop_err_gate": {
            "value": pop_err_rel, "bound": 1e-4, "met": pop_err_rel <= 1e-4,
            "definition": "max over t, n, trajectory of | |psi_n|^2 (this run) - |psi_n|^2 (complex128 run) | / "
                          "|psi|^2 (complex128 run, same trajectory and time)",
            "why": "LINEAR HOPS does not normalise: over these steps the populations reach %.3g, so the absolute "
                   "difference above scales with them; the 1e-4 bound is applied to the difference relative to "
                   "the trajectory's |psi|^2" % max_pop},
        "pop_err_trajectories": nref,
        "profiled_ms_per_step": elapsed_prof / args.steps * 1e3,
        "profiled_pass_bitwise_equal": prof_equal,
        "c128_leg": {"value": B * args.steps / c128_elapsed, "ms_per_step": c128_elapsed / args.steps * 1e3,
                     "dtype": "c128", "note": "same batch and steps with complex128 stages"},
        "gpu_fp64_single_traj_steps_per_s": single_rate,
        "speedup_vs_gpu_fp64_single": value / single_rate,
        "init_s": t_init,
        "layout": {"lanes": choice[0], "stationary_stage": bool(choice[1]), "chunk_traj": choice[2],
                   "nt_stream": choice[3], "groups_q": choice[4], "tile_order": bool(choice[5]),
                   "vec_lanes": choice[6], "lds_stage": bool(choice[7]),
                   "autotune_ms_per_step (lanes/stationary/chunk/nt/q/order/vec/lds)":
                       {"%d/%d/%d/%d/%d/%d/%d/%d" % k: v * 1e3 for k, v in tune_t.items()}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": kernel_name, "kernel_avg_ms": stage_ms,
                     "algorithmic_bytes_per_launch": bytes_launch,
                     "prologue_avg_ms": ms[0] / max(n[0], 1)},
    }
    if args.c3_steps > 0:
        del batch, st
        torch.cuda.empty_cache()
        c3 = run_c3(256, args.c3_steps, "mixed", min(20, args.c3_steps), args.ref_traj, chunk=25, log=log)
        c3["chunks"] = [{"t_fs": c["t_fs"], "rate": round(256 * 25 / c["s"], 1), "mean_n_phi": round(c["mean_n_phi"])}
                        for c in c3["chunks"]]
        out["adaptive_c3"] = dict(c3, config="adaptive NORMALIZED NONLINEAR batch: 100-site chain, depth 10, "
                                             "delta 1e-3, 256 trajectories with their own bases, t = 0..%d fs"
                                             % args.c3_steps)
        # the same batch with complex128 stages: the configuration that meets
        # the 1e-4 population bound over the whole run (the complex64-stage
        # run above flips some near-threshold basis decisions late in the run)
        torch.cuda.empty_cache()
        c3f = run_c3(256, args.c3_steps, "fp64", 0, 0, chunk=25, log=log, baselines=False)
        w = [c for c in c3f["chunks"] if c["t_fs"] <= max(25, min(20, args.c3_steps))]
        win_rate = 256 * sum(c["t_fs"] - p for c, p in zip(w, [0] + [x["t_fs"] for x in w[:-1]])) / \
            sum(c["s"] for c in w)
        c3f["chunks"] = [{"t_fs": c["t_fs"], "rate": round(256 * 25 / c["s"], 1),
                          "mean_n_phi": round(c["mean_n_phi"])} for c in c3f["chunks"]]
        out["adaptive_c3_fp64"] = dict(
            c3f, config="the same batch with complex128 stages (FP64 throughout)",
            pop_err_vs_fp64=0.0,
            pop_err_note="this run is the FP64 computation; the FP64 path matches the oracle to 1e-10 in the tests",
            batch_rate_first_window=win_rate,
            speedup_vs_gpu_fp64_single_same_window=win_rate / c3["gpu_fp64_single_traj_steps_per_s"],
            speedup_vs_same_runtime_fp64_b1=c3f["value"] / c3["same_runtime_fp64_b1_steps_per_s"])
        # the north-star figure: the leg that meets the 1e-4 population bound
        mixed_ok = c3.get("pop_err_vs_fp64") is not None and c3["pop_err_vs_fp64"] <= 1e-4
        out["adaptive_c3"]["pop_err_bound_met"] = mixed_ok
        leg = "adaptive_c3" if mixed_ok else "adaptive_c3_fp64"
        ns_leg = out[leg]
        out["north_star_c3"] = {
            "leg": leg, "value": ns_leg["value"], "unit": "traj·RK4-steps/s", "pop_err_vs_fp64": ns_leg["pop_err_vs_fp64"],
            "pop_err_bound": 1e-4, "bound_met": True,
            "speedup_vs_gpu_fp64_single_same_window": ns_leg["speedup_vs_gpu_fp64_single_same_window"],
            "speedup_vs_same_runtime_fp64_b1": ns_leg["speedup_vs_same_runtime_fp64_b1"],
            "note": "the 256-trajectory 100-site adaptive batch over t = 0..%d fs with complex64 stages if their "
                    "population error against FP64 stays within 1e-4, else with complex128 stages (complex64 "
                    "stages: %.2e on %s trajectories)"
                    % (args.c3_steps, c3.get("pop_err_vs_fp64", float("nan")), c3.get("pop_err_trajectories"))}
    if args.c4_leg > 0:
        torch.cuda.empty_cache()
        log("c4: open stream at 200 requests/s for %.0f s" % args.c4_leg)
        st = run_c4_stream(200.0, args.c4_leg, "mixed", log=log)
        out["serving_c4"] = {
            "config": "serving stream: adaptive requests arriving as a Poisson stream at 200/s for %.0f s, chain "
                      "2-200 sites, depth 2-10, delta 1e-4..1e-2, t_max 200-2000 fs, dt 1 fs, up to 512 "
                      "requests on the device at once" % args.c4_leg,
            "traj_steps_per_s": st["traj_steps_per_s"], "arrived": st["arrived"], "completed": st["completed"],
            "running_at_end": st["admitted_unfinished"], "waiting_at_end": st["waiting_at_end"],
            "p50_latency_s": st["p50_latency_s"], "p99_latency_s": st["p99_latency_s"],
            "runtime_batches": st["runtime_batches"], "threads_per_batch": st["threads_per_batch"],
            "timeline (t, arrived, waiting, running, done)": st["timeline"],
            "note": "200/s exceeds what one GPU completes (the rate sweep of --workload c4 --c4-window 60 "
                    "finds the largest rate it keeps up with); latencies are those of the requests completed "
                    "inside the window, i.e. the short ones"}
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
