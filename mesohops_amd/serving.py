"""Serving loop: continuous batching of adaptive HOPS requests on one GPU.

Requests (mesohops_amd.workloads.Request: chain length, depth, delta, t_max,
arrival time) arrive over time.  A request is admitted into a free slot of
one ragged batch (include/hops_runtime.h) as soon as it has arrived, its
noise is prepared and the arena has room for its footprint; each step of the
batch advances all its admitted requests by one RK4 step (each at its own
time), and a finished request leaves the batch at once, freeing its slot and
memory for the next.  Host preparation (system tables, raw noise draws) runs
in worker threads ahead of admission; the correlated-noise filter runs on the
device at admission.

Results do not depend on which requests share the batch: every slot's
kernels and basis bookkeeping touch only that slot's data.  The slots are
spread over several runtime batches (ragged.RaggedGroup), each with its own
share of the arena, stream and worker threads.  Each batch is served by a
host thread of its own that admits, steps and retires its requests without
waiting for the other batches, so one batch's host-side selection overlaps
the others' device work and a slow step holds up only its own batch; the
next request goes to the first batch with a free slot and room for it.

Back-pressure: admission needs room for twice a request's estimated
footprint.  When bases grow until a step cannot place some slots' new bases,
those slots are preempted (removed, their requests queued again at the front
to restart from t = 0) and the rest of the batch carries on; a preempted
request returns only once a batch has room for twice the footprint it had
reached, or an empty batch room for that footprint.  A request that cannot
fit even alone raises UnsupportedRequest.
"""
from __future__ import annotations

import itertools
import threading
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import _lib
from .basis import SystemInfo
from .errors import ArenaFull, UnsupportedRequest
from .noise import HopsNoise, device_noise
from .ragged import RaggedGroup, RtSystem, traj_param
from .workloads import chain_system

_EOM = {"LINEAR": _lib.HOPS_EOM_LINEAR, "NONLINEAR": _lib.HOPS_EOM_NONLINEAR,
        "NORMALIZED NONLINEAR": _lib.HOPS_EOM_NORMALIZED_NONLINEAR}


class ServingLoop:
    def __init__(self, precision_mode="mixed", arena_bytes=None, max_slots=256, n_threads=None, prep_workers=8,
                 eom="NORMALIZED NONLINEAR", record=False, device="cuda:0", system_fn=chain_system,
                 reserve_states=12, reserve_aux=4000, sub_batches=None):
        """``sub_batches``: runtime batches, each served by its own host
        thread (ragged.RaggedGroup; default one per 32 slots, at most 16),
        each with its share of the arena and of the slots; ``n_threads``:
        worker threads per batch (default 16 / sub_batches, at least 1)."""
        K = max(1, min(16, max_slots // 32)) if sub_batches is None else int(sub_batches)
        if K < 1:
            raise UnsupportedRequest("sub_batches < 1", "ServingLoop")
        K = min(K, max(1, max_slots))  # no batch without a slot
        nthr = max(1, 16 // K) if n_threads is None else int(n_threads)
        self.K = K
        self.n_threads = nthr
        self.slot_cap = self.split_slots(max_slots, K)
        self.rt = RaggedGroup(K, precision_mode, arena_bytes, max_slots=max(1, max(self.slot_cap)), n_threads=nthr,
                              device=device)
        self.device = self.rt.device
        self.max_slots = max_slots
        self.eom = eom
        self.record = record
        self.system_fn = system_fn
        self.reserve = (reserve_states, reserve_aux)
        self._systems = {}
        self._prep = ThreadPoolExecutor(max_workers=prep_workers)

    @staticmethod
    def split_slots(max_slots, K):
        """Slots of each of the K runtime batches: max_slots split as evenly
        as it goes, the first batches taking the remainder."""
        return [max_slots // K + (1 if k < max_slots % K else 0) for k in range(K)]

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

    def _footprint(self, rts, req, k=0):
        return self.rt.parts[k].estimate(rts, min(req.nsite, self.reserve[0]), self.reserve[1])

    def _admit(self, req, prep, k):
        sysp, info, rts, nz, raw = prep
        part = self.rt.parts[k]
        # the filter and torch's own fills and copies on the batch's stream
        ts = part._torch_stream if part._torch_stream is not None else torch.cuda.current_stream(self.device)
        with torch.cuda.stream(ts):
            z1 = device_noise([nz], list(range(info.n_l2)), self.device, part.stream, raws=[raw])[0]
        psi0 = np.zeros(info.nstates, np.complex128)
        psi0[info.nstates // 2] = 1.0
        ns = int(round(req.dt * 0.5 / req.tau_noise))
        prm = traj_param(_EOM[self.eom], req.maxhier, req.delta, req.delta, 1, 0.01, req.dt, ns, nz.n_t,
                         nz.time_index(0.0), record=self.record)
        slot = self.rt.add(k, rts, prm, psi0, z1)
        return slot, (z1, prm)

    def run(self, requests, time_limit=None, prep_ahead=16, log=None):
        """Serve ``requests`` (arrival times relative to the start).  Returns
        a dict of per-request results and the serving statistics.

        Runtime batch k is served by host thread k, which admits the next
        arrived, prepared request whenever its batch has a free slot and
        room, steps its own slots and retires them; the batches do not wait
        for each other between steps.  The calling thread samples the
        timeline and joins the batch threads."""
        reqs = deque(sorted(requests, key=lambda r: r.arrival))
        arrivals = np.array([r.arrival for r in reqs])
        timeline = []        # (t, arrived, waiting, active, done) about once a second
        futures = {}
        active = {}          # (batch, slot) -> request, for the statistics
        done = {}
        need = {}            # rid -> device bytes a preempted request had reached
        lone_retry = set()   # requests restarted once after outgrowing the arena alone
        cnt = {"steps": 0, "iter": 0, "preempt": 0}
        lock = threading.Lock()
        stop = threading.Event()
        errors = []
        start = time.perf_counter()

        def clock():
            return time.perf_counter() - start

        def next_request(k, part, n_mine, now):
            """Under the lock: the head request if it has arrived, is prepared
            and fits batch k (taken off the queue), else None."""
            for r in itertools.islice(reqs, prep_ahead):  # host preparation, in arrival order
                if r.rid not in futures:
                    futures[r.rid] = self._prep.submit(self._prepare, r)
            if not reqs or reqs[0].arrival > now or n_mine >= self.slot_cap[k]:
                return None
            r = reqs[0]
            fut = futures[r.rid]
            if not fut.done():
                return None
            prep = fut.result()
            fp = self._footprint(prep[2], r, k)
            reached = need.get(r.rid, 0)
            free = part.free_bytes()
            if free < 2 * max(fp, reached):
                if n_mine > 0:
                    return None
                # an empty batch holds no less than any other: a request
                # alone in it needs room for what it had reached before
                if free < max(2 * fp, reached):
                    raise UnsupportedRequest("request %d (%d sites, depth %d) needs more device memory than "
                                             "a runtime batch's arena holds" % (r.rid, r.nsite, r.maxhier),
                                             "ServingLoop")
            reqs.popleft()
            del futures[r.rid]
            return r, prep

        def serve(k):
            part = self.rt.parts[k]
            mine = {}        # slot -> [req, steps_done, keep, t_admit, peak device bytes]
            try:
                serve_batch(k, part, mine)
            finally:
                # requests still running when the loop stops (time limit or
                # error) leave the batch, which is then ready for another run
                for s in list(mine):
                    try:
                        part.remove(s)
                    except Exception:  # the first error is the one reported
                        pass

        def serve_batch(k, part, mine):
            while not stop.is_set():
                now = clock()
                if time_limit is not None and now > time_limit:
                    return
                # admission: arrived, prepared, a free slot and room in this batch
                while True:
                    with lock:
                        if not reqs and not mine:
                            return
                        got = next_request(k, part, len(mine), now)
                    if got is None:
                        break
                    r, prep = got
                    (_, slot), keep = self._admit(r, prep, k)
                    mine[slot] = [r, 0, keep, clock(), 0]
                    with lock:
                        active[(k, slot)] = r
                if not mine:
                    with lock:
                        head = reqs[0] if reqs else None
                    if head is not None:
                        time.sleep(min(max(head.arrival - clock(), 0.001), 0.01))
                    continue
                slots = list(mine)
                oom = []
                try:
                    part.step(1, slots=slots)
                except ArenaFull:
                    # back-pressure: the slots whose grown bases did not fit
                    # are preempted and their requests queued again (they
                    # restart from t = 0 once memory frees up); the others
                    # completed the step
                    oom = [s for s in slots if part.info(s)["oom"]]
                    if len(mine) == 1:
                        # alone in its batch and out of room: the arena may
                        # only have been fragmented by earlier requests, so
                        # the request restarts once before it is refused
                        r = mine[oom[0]][0]
                        with lock:
                            if r.rid in lone_retry:
                                raise UnsupportedRequest("request %d (%d sites, depth %d) outgrowing the arena "
                                                         "while running alone" % (r.rid, r.nsite, r.maxhier),
                                                         "ServingLoop")
                            lone_retry.add(r.rid)
                    for s in oom:
                        r = mine.pop(s)[0]
                        reached = part.info(s)["bytes"]
                        part.remove(s)
                        with lock:
                            need[r.rid] = reached
                            del active[(k, s)]
                            reqs.appendleft(r)
                            cnt["preempt"] += 1
                    slots = [s for s in slots if s not in oom]
                t_now = clock()
                finished = []
                for s in slots:
                    a = mine[s]
                    a[1] += 1
                    if self.record:
                        a[4] = max(a[4], part.info(s)["bytes"])
                    if a[1] >= a[0].n_steps:
                        r = a[0]
                        res = {"rid": r.rid, "arrival": r.arrival, "admitted": a[3], "finished": t_now,
                               "latency": t_now - r.arrival, "steps": a[1], "nsite": r.nsite,
                               "maxhier": r.maxhier, "delta": r.delta, "t_max": r.t_max}
                        inf = part.info(s)
                        res["n_state_final"], res["n_hier_final"] = inf["n_state"], inf["n_hier"]
                        if self.record:
                            res["psi"], res["states"], res["nhier"] = part.records(s, r.nsite)
                            res["peak_bytes"] = a[4]
                        part.remove(s)
                        del mine[s]
                        finished.append((s, res))
                with lock:
                    cnt["steps"] += len(slots)
                    cnt["iter"] += 1
                    for s, res in finished:
                        done[res["rid"]] = res
                        del active[(k, s)]

        def worker(k):
            try:
                with torch.cuda.device(self.device):
                    serve(k)
            except BaseException as e:  # re-raised on the calling thread
                with lock:
                    errors.append(e)
                stop.set()

        threads = [threading.Thread(target=worker, args=(k,), name="hops-serve-%d" % k) for k in range(self.K)]
        for t in threads:
            t.start()
        last_sample = -1.0
        last_log = 0.0
        while any(t.is_alive() for t in threads):
            now = clock()
            if now - last_sample >= 1.0:
                last_sample = now
                arrived = int(np.searchsorted(arrivals, now, side="right"))
                with lock:
                    waiting = sum(1 for r in reqs if r.arrival <= now)
                    timeline.append((round(now, 2), arrived, waiting, len(active), len(done)))
            if log is not None and now - last_log > 5.0:
                last_log = now
                with lock:
                    msg = "t=%.1fs active=%d done=%d queued=%d steps=%d" % (now, len(active), len(done), len(reqs),
                                                                             cnt["steps"])
                log(msg)
            threads[0].join(timeout=0.02)
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        wall = clock()
        lat = np.array([d["latency"] for d in done.values()]) if done else np.zeros(0)
        arrived = int(np.searchsorted(arrivals, wall, side="right"))
        timeline.append((round(wall, 2), arrived, sum(1 for r in reqs if r.arrival <= wall), len(active), len(done)))
        stats = {"completed": len(done), "admitted_unfinished": len(active), "not_admitted": len(reqs),
                 "arrived": arrived, "waiting_at_end": timeline[-1][2],
                 "timeline": timeline,
                 "wall_s": wall, "traj_steps": cnt["steps"],
                 "traj_steps_per_s": cnt["steps"] / wall if wall else 0.0,
                 "iterations": cnt["iter"], "preempted": cnt["preempt"],
                 "runtime_batches": self.K, "threads_per_batch": self.n_threads,
                 "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
                 "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None}
        return done, stats

    def close(self):
        self._prep.shutdown(wait=True)
        self.rt.close()
