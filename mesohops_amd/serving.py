"""Serving loop: continuous batching of adaptive HOPS requests on one GPU.

Requests (mesohops_amd.workloads.Request: chain length, depth, delta, t_max,
arrival time) arrive over time.  A request is admitted into a free slot of
one ragged batch (include/hops_runtime.h) as soon as it has arrived, its
noise is prepared and the arena has room for its footprint; every loop
iteration advances all admitted requests by one RK4 step (each at its own
time), and a finished request leaves the batch at once, freeing its slot and
memory for the next.  Host preparation (system tables, raw noise draws) runs
in worker threads ahead of admission; the correlated-noise filter runs on the
device at admission.

Results do not depend on which requests share the batch: every slot's
kernels and basis bookkeeping touch only that slot's data.  The slots are
spread over several runtime batches (ragged.RaggedGroup), each with its own
share of the arena, stream and worker threads, stepped concurrently, so one
batch's host-side selection overlaps the others' device work; a request goes
to the batch with a free slot and the most free memory.

Back-pressure: admission needs room for twice a request's estimated
footprint.  When bases grow until a step cannot place some slots' new bases,
those slots are preempted (removed, their requests queued again at the front
to restart from t = 0) and the rest of the batch carries on; a preempted
request returns only once the arena has room for the footprint it had
reached.  A request that cannot fit even alone raises UnsupportedRequest.
"""
from __future__ import annotations

import itertools
import math
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import _lib
from .basis import SystemInfo
from .engine import _stream_ptr
from .errors import ArenaFull, UnsupportedRequest
from .noise import HopsNoise, device_noise
from .ragged import RaggedGroup, RtSystem, traj_param
from .workloads import chain_system

_EOM = {"LINEAR": _lib.HOPS_EOM_LINEAR, "NONLINEAR": _lib.HOPS_EOM_NONLINEAR,
        "NORMALIZED NONLINEAR": _lib.HOPS_EOM_NORMALIZED_NONLINEAR}


class ServingLoop:
    def __init__(self, precision_mode="mixed", arena_bytes=None, max_slots=256, n_threads=None, prep_workers=4,
                 eom="NORMALIZED NONLINEAR", record=False, device="cuda:0", system_fn=chain_system,
                 reserve_states=12, reserve_aux=4000, sub_batches=None):
        """``sub_batches``: runtime batches stepped concurrently
        (ragged.RaggedGroup; default one per 64 slots, at most 8), each
        with its share of the arena and of the slots; ``n_threads``: worker
        threads per batch (default 16 for one batch, else 32 / sub_batches)."""
        K = max(1, min(8, max_slots // 64)) if sub_batches is None else int(sub_batches)
        if K < 1:
            raise UnsupportedRequest("sub_batches < 1", "ServingLoop")
        nthr = (16 if K == 1 else max(2, 32 // K)) if n_threads is None else int(n_threads)
        self.K = K
        self.per_slots = -(-max_slots // K)
        self.rt = RaggedGroup(K, precision_mode, arena_bytes, max_slots=self.per_slots, n_threads=nthr,
                              device=device)
        self.device = self.rt.device
        self.max_slots = max_slots
        self.eom = eom
        self.record = record
        self.system_fn = system_fn
        self.reserve = (reserve_states, reserve_aux)
        self._systems = {}
        self._prep = ThreadPoolExecutor(max_workers=prep_workers)

    def _system(self, nsite):
        if nsite not in self._systems:
            sysp = self.system_fn(nsite)
            info = SystemInfo(sysp)
            self._systems[nsite] = (sysp, info, RtSystem(info))
        return self._systems[nsite]

    def _prepare(self, req):
        """Host part of a request: its system and the raw noise draws."""
        sysp, info, rts = self._system(req.nsite)
        nz = HopsNoise({"SEED": int(req.rid), "MODEL": "FFT_FILTER", "TLEN": req.tlen, "TAU": req.tau_noise},
                       info, sysp, 1)
        return sysp, info, rts, nz, nz.raw()

    def _footprint(self, rts, req):
        return self.rt.estimate(rts, min(req.nsite, self.reserve[0]), self.reserve[1])

    def _admit(self, req, prep, k):
        sysp, info, rts, nz, raw = prep
        z1 = device_noise([nz], list(range(info.n_l2)), self.device, self.rt.parts[k].stream, raws=[raw])[0]
        psi0 = np.zeros(info.nstates, np.complex128)
        psi0[info.nstates // 2] = 1.0
        ns = int(round(req.dt * 0.5 / req.tau_noise))
        prm = traj_param(_EOM[self.eom], req.maxhier, req.delta, req.delta, 1, 0.01, req.dt, ns, nz.n_t,
                         nz.time_index(0.0), record=self.record)
        slot = self.rt.add(k, rts, prm, psi0, z1)
        return slot, (z1, prm)

    def run(self, requests, time_limit=None, prep_ahead=16, log=None):
        """Serve ``requests`` (arrival times relative to the start).  Returns
        a dict of per-request results and the serving statistics."""
        reqs = deque(sorted(requests, key=lambda r: r.arrival))
        arrivals = np.array([r.arrival for r in reqs])
        timeline = []        # (t, arrived, waiting, active, done) about once a second
        last_sample = -1.0
        futures = {}
        active = {}          # (batch, slot) -> [req, steps_done, keep, t_admit, peak device bytes]
        n_in = [0] * self.K  # active slots per runtime batch
        done = {}
        start = time.perf_counter()
        steps_total = 0
        n_iter = 0
        n_preempt = 0
        need = {}            # rid -> device bytes a preempted request had reached
        lone_retry = set()   # requests restarted once after outgrowing the arena alone
        last_log = start
        while reqs or active:
            now = time.perf_counter() - start
            if now - last_sample >= 1.0:
                last_sample = now
                arrived = int(np.searchsorted(arrivals, now, side="right"))
                waiting = sum(1 for r in reqs if r.arrival <= now)
                timeline.append((round(now, 2), arrived, waiting, len(active), len(done)))
            if time_limit is not None and now > time_limit:
                break
            # host preparation of the next requests, in arrival order
            for r in itertools.islice(reqs, prep_ahead):
                if r.rid not in futures:
                    futures[r.rid] = self._prep.submit(self._prepare, r)
            # admission: arrived, prepared, a free slot and room in the arena
            while reqs and reqs[0].arrival <= now and len(active) < self.max_slots:
                r = reqs[0]
                fut = futures.get(r.rid)
                if fut is None:
                    futures[r.rid] = self._prep.submit(self._prepare, r)
                    break
                if not fut.done():
                    break
                prep = fut.result()
                # the runtime batch with a free slot and the most free memory
                room = [q for q in range(self.K) if n_in[q] < self.per_slots]
                if not room:
                    break
                free = {q: self.rt.parts[q].free_bytes() for q in room}
                k = max(room, key=lambda q: free[q])
                if free[k] < max(2 * self._footprint(prep[2], r), need.get(r.rid, 0)):
                    if not active:
                        raise UnsupportedRequest("request %d (%d sites, depth %d) needs more device memory than "
                                                 "a runtime batch's arena holds" % (r.rid, r.nsite, r.maxhier),
                                                 "ServingLoop")
                    break
                reqs.popleft()
                del futures[r.rid]
                slot, keep = self._admit(r, prep, k)
                n_in[k] += 1
                active[slot] = [r, 0, keep, time.perf_counter() - start, 0]
            if not active:
                if reqs:
                    wait = reqs[0].arrival - (time.perf_counter() - start)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                    else:
                        fut = futures.get(reqs[0].rid)
                        if fut is None or not fut.done():
                            time.sleep(0.001)
                continue
            slots = list(active)
            try:
                self.rt.step(1, slots=slots)
            except ArenaFull:
                # back-pressure: the slots whose grown bases did not fit are
                # preempted and their requests queued again (they restart from
                # t = 0 once memory frees up); the others completed the step
                oom = [s for s in slots if self.rt.info(s)["oom"]]
                for s in oom:
                    if n_in[s[0]] == 1:
                        # alone in its batch and out of room: the arena may
                        # only have been fragmented by earlier requests, so
                        # the request restarts once before it is refused
                        r = active[s][0]
                        if r.rid in lone_retry:
                            raise UnsupportedRequest("request %d (%d sites, depth %d) outgrowing the arena while "
                                                     "running alone" % (r.rid, r.nsite, r.maxhier), "ServingLoop")
                        lone_retry.add(r.rid)
                for s in oom:
                    r = active.pop(s)[0]
                    need[r.rid] = 2 * self.rt.info(s)["bytes"]
                    self.rt.remove(s)
                    n_in[s[0]] -= 1
                    reqs.appendleft(r)
                    n_preempt += 1
                slots = [s for s in slots if s not in oom]
            steps_total += len(slots)
            n_iter += 1
            t_now = time.perf_counter() - start
            for s in slots:
                a = active[s]
                a[1] += 1
                if self.record:
                    a[4] = max(a[4], self.rt.info(s)["bytes"])
                if a[1] >= a[0].n_steps:
                    r = a[0]
                    res = {"rid": r.rid, "arrival": r.arrival, "admitted": a[3], "finished": t_now,
                           "latency": t_now - r.arrival, "steps": a[1], "nsite": r.nsite, "maxhier": r.maxhier,
                           "delta": r.delta, "t_max": r.t_max}
                    inf = self.rt.info(s)
                    res["n_state_final"], res["n_hier_final"] = inf["n_state"], inf["n_hier"]
                    if self.record:
                        res["psi"], res["states"], res["nhier"] = self.rt.records(s, r.nsite)
                        res["peak_bytes"] = a[4]
                    done[r.rid] = res
                    self.rt.remove(s)
                    n_in[s[0]] -= 1
                    del active[s]
            if log is not None and time.perf_counter() - last_log > 5.0:
                last_log = time.perf_counter()
                log("t=%.1fs active=%d done=%d queued=%d steps=%d" % (t_now, len(active), len(done), len(reqs),
                                                                      steps_total))
        wall = time.perf_counter() - start
        lat = np.array([d["latency"] for d in done.values()]) if done else np.zeros(0)
        arrived = int(np.searchsorted(arrivals, wall, side="right"))
        timeline.append((round(wall, 2), arrived, sum(1 for r in reqs if r.arrival <= wall), len(active), len(done)))
        stats = {"completed": len(done), "admitted_unfinished": len(active), "not_admitted": len(reqs),
                 "arrived": arrived, "waiting_at_end": timeline[-1][2],
                 "timeline": timeline,
                 "wall_s": wall, "traj_steps": steps_total, "traj_steps_per_s": steps_total / wall if wall else 0.0,
                 "iterations": n_iter, "preempted": n_preempt,
                 "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
                 "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None}
        return done, stats

    def close(self):
        self._prep.shutdown(wait=True)
        self.rt.close()
