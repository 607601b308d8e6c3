"""Ragged batches: many adaptive trajectories, each with its own basis (and,
if wanted, its own system), advanced together on one GPU by the native
runtime of include/hops_runtime.h.

``AdaptiveBatch`` is the batched form of ``HopsTrajectory`` + ``make_adaptive``
(trajectory/hops_trajectory.py:84-538, :338-399): B trajectories of one
system with different noise, every one keeping its own adaptive basis.  The
per-step basis bookkeeping that the reference runs in Python per trajectory
runs in C++ worker threads for all trajectories at once; the GPU runs one
launch per RK4 stage for all of them.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np
import torch

from . import _lib
from ._lib import RtSystemDesc, RtTrajParam, check, lib
from .basis import SystemInfo, apply_static_filters, triangular_aux_list
from .engine import _require_gpu, _stream_ptr
from .errors import EngineError, TrajectoryError, UnsupportedRequest
from .noise import HopsNoise, device_noise
from .params import precision as _precision


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class RtSystem:
    """A system registered with the native runtime (basis/system_functions.py:10-171)."""

    def __init__(self, info: SystemInfo):
        self.info = info
        n = info.nstates
        self._H = np.ascontiguousarray(info.H, dtype=np.complex128)
        self._g = np.ascontiguousarray(info.G, dtype=np.complex128)
        self._w = np.ascontiguousarray(info.W, dtype=np.complex128)
        self._lm = np.ascontiguousarray(info.l_index_by_hmode, dtype=np.int32)
        rows, cols, vals, ptr = [], [], [], [0]
        for L in info.l_ops:
            rows += L.row.tolist()
            cols += L.col.tolist()
            vals += L.data.tolist()
            ptr.append(len(rows))
        self._lp = np.array(ptr, dtype=np.int32)
        self._lr = np.array(rows, dtype=np.int32)
        self._lc = np.array(cols, dtype=np.int32)
        self._lv = np.array(vals, dtype=np.complex128)
        self._lt = np.ascontiguousarray(info.lt_param, dtype=np.complex128)
        d = RtSystemDesc()
        d.n_state, d.n_hmode, d.n_lop = n, info.n_hmodes, info.n_l2
        d.H, d.g, d.w = _ptr(self._H), _ptr(self._g), _ptr(self._w)
        d.l_of_mode, d.l_ptr, d.l_row, d.l_col = _ptr(self._lm), _ptr(self._lp), _ptr(self._lr), _ptr(self._lc)
        d.l_val = _ptr(self._lv)
        d.lt = _ptr(self._lt) if np.any(self._lt) else None
        self.handle = lib().hops_rt_system_create(ctypes.byref(d))
        if not self.handle:
            raise EngineError(lib().hops_last_error().decode())

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _lib._lib is not None:
            try:
                _lib._lib.hops_rt_system_destroy(h)
            except Exception:
                pass
        self.handle = None


def _aux_csr(aux_list):
    ptr = [0]
    mode, val = [], []
    for a in aux_list:
        for (m, v) in a:
            mode.append(m)
            val.append(v)
        ptr.append(len(mode))
    return (np.array(ptr, dtype=np.int32), np.array(mode if mode else [0], dtype=np.int32),
            np.array(val if val else [0], dtype=np.int32))


def _aux_from_csr(ptr, mode, val):
    return [tuple((int(mode[q]), int(val[q])) for q in range(ptr[i], ptr[i + 1])) for i in range(len(ptr) - 1)]


def traj_param(eom_kind, maxhier, delta_a, delta_s, update_step, f_discard, dt, noise_step, n_t, it0,
               early_integrator="INCH_WORM", early_steps=5, inchworm_cap=5, markov=None, aux0=None,
               record=True, filters=None, noise_c128=False):
    """hops_rt_traj_param_t and the arrays it points to (kept alive on the
    returned object)."""
    p = RtTrajParam()
    p.eom, p.maxhier = int(eom_kind), int(min(maxhier, 255))
    p.update_step = int(update_step) if update_step else 1
    codes = {"INCH_WORM": 0, "STATIC": 1, "STATIC_STATE_INCHWORM_HIERARCHY": 2}
    if early_integrator not in codes:
        raise UnsupportedRequest(early_integrator, "early time integrator clause of the propagate")
    p.early_integrator = codes[early_integrator]
    p.early_steps, p.inchworm_cap = int(early_steps), int(inchworm_cap)
    p.noise_step, p.n_t, p.it0, p.record = int(noise_step), int(n_t), int(it0), int(bool(record))
    p.noise_c128 = int(bool(noise_c128))
    p.delta_a, p.delta_s, p.f_discard, p.dt = float(delta_a), float(delta_s), float(f_discard), float(dt)
    keep = []
    if markov is not None:
        mk = np.ascontiguousarray(np.asarray(markov, dtype=np.uint8))
        keep.append(mk)
        p.markov = _ptr(mk)
    if filters:
        kinds = {"Triangular": 0, "LongEdge": 1}
        fk = np.ascontiguousarray([kinds[name] for name, _ in filters], dtype=np.int32)
        fm = np.ascontiguousarray([int(prm[1]) for _, prm in filters], dtype=np.int32)
        fmask = np.ascontiguousarray(np.array([np.asarray(prm[0], bool) for _, prm in filters], dtype=np.uint8))
        keep += [fk, fm, fmask]
        p.n_filters = len(filters)
        p.filter_kind, p.filter_kmax, p.filter_mask = _ptr(fk), _ptr(fm), _ptr(fmask)
    if aux0 is not None:
        a_ptr, a_mode, a_val = _aux_csr(aux0)
        keep += [a_ptr, a_mode, a_val]
        p.n_aux0 = len(aux0)
        p.aux0_ptr, p.aux0_mode, p.aux0_val = _ptr(a_ptr), _ptr(a_mode), _ptr(a_val)
    p._keep = keep
    return p


class DebugBasis:
    """Host-only view of the runtime's basis tables and selection (tests)."""

    def __init__(self, system: RtSystem, states, aux_list, eom_kind=2):
        st = np.ascontiguousarray(np.asarray(states, dtype=np.int32))
        a_ptr, a_mode, a_val = _aux_csr(aux_list)
        self.h = lib().hops_rt_debug_basis(system.handle, int(eom_kind), len(st), _ptr(st), len(aux_list),
                                           _ptr(a_ptr), _ptr(a_mode), _ptr(a_val))
        if not self.h:
            raise EngineError(lib().hops_last_error().decode())

    def table(self, name, dtype):
        n = lib().hops_rt_debug_table(self.h, name.encode(), None, 0)
        if n < 0:
            raise EngineError(lib().hops_last_error().decode())
        out = np.zeros(max(n, 1), dtype=np.uint8)
        lib().hops_rt_debug_table(self.h, name.encode(), _ptr(out), n)
        return out[:n].view(dtype)

    def select(self, prm, pn, e_aux, dt):
        pn = np.ascontiguousarray(pn, dtype=np.complex128)
        e = np.ascontiguousarray(e_aux, dtype=np.float64)
        nh = int(self.table("sizes", np.int32)[1])
        mask = np.zeros(nh, dtype=np.uint8)
        check(lib().hops_rt_debug_select(self.h, ctypes.byref(prm), _ptr(pn), _ptr(e), float(dt), _ptr(mask)))
        return mask.astype(bool)

    def select_states(self, e_state):
        e = None if e_state is None else np.ascontiguousarray(e_state, dtype=np.float64)
        check(lib().hops_rt_debug_select_states(self.h, _ptr(e)))

    def set_noise(self, z1, z2, zmem):
        """Noise of the step (complex64 z1 / z2 of the basis L-operators, z2
        may be None; complex128 noise memory by absolute mode) for bases with
        off-diagonal L-operators."""
        z1 = np.ascontiguousarray(z1, dtype=np.complex64)
        z2 = None if z2 is None else np.ascontiguousarray(z2, dtype=np.complex64)
        zm = np.ascontiguousarray(zmem, dtype=np.complex128)
        check(lib().hops_rt_debug_noise(self.h, _ptr(z1), _ptr(z2), _ptr(zm)))

    def errors(self, which):
        """Error terms of the last selection: "aux", "state" or "outside"
        ({state: error})."""
        w = {"aux": 0, "state": 1, "outside": 2}[which]
        n = lib().hops_rt_debug_errors(self.h, w, None, 0)
        if n < 0:
            raise EngineError(lib().hops_last_error().decode())
        out = np.zeros(max(n, 1))
        lib().hops_rt_debug_errors(self.h, w, _ptr(out), n)
        out = out[:n]
        if w == 2:
            return {int(out[i]): float(out[i + 1]) for i in range(0, n, 2)}
        return out

    def result(self, cap=1 << 20):
        ns, na = ctypes.c_int32(), ctypes.c_int32()
        st = np.zeros(cap, np.int32)
        ap, am, av = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        rc = lib().hops_rt_debug_result(self.h, ctypes.byref(ns), _ptr(st), ctypes.byref(na), _ptr(ap), _ptr(am),
                                        _ptr(av), cap)
        if rc != 0:
            raise EngineError("result buffer too small")
        return [int(s) for s in st[:ns.value]], _aux_from_csr(ap[:na.value + 1], am, av)

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib._lib is not None:
            try:
                _lib._lib.hops_rt_debug_free(h)
            except Exception:
                pass
        self.h = None


class RaggedBatch:
    """Thin owner of a runtime batch: arena, stream, slots."""

    def __init__(self, precision_mode="fp64", arena_bytes=None, max_slots=1024, n_threads=16, device="cuda:0",
                 stream=None):
        """``stream``: a torch.cuda.Stream the batch's work is queued on (the
        current stream when None)."""
        _require_gpu()
        self.device = torch.device(device)
        self.precision = {"fp64": _lib.HOPS_PREC_FP64, "mixed": _lib.HOPS_PREC_MIXED}[precision_mode]
        if arena_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            arena_bytes = int(free * 0.6)
        arena_bytes = (int(arena_bytes) // 256) * 256
        self.arena = torch.empty(arena_bytes, dtype=torch.uint8, device=self.device)
        self._torch_stream = stream
        self.stream = _stream_ptr(self.device) if stream is None else ctypes.c_void_p(stream.cuda_stream)
        self.h = lib().hops_rt_batch_create(self.precision, ctypes.c_void_p(self.arena.data_ptr()), arena_bytes,
                                            int(max_slots), int(n_threads), self.stream)
        if not self.h:
            raise EngineError(lib().hops_last_error().decode())
        self._keep = {}

    def set_option(self, name, value):
        """Runtime option: "device_select" 1 selects the adaptive bases on the
        device (experimental: verified in a CPU emulation of its kernels, not
        yet on the GPU), 0 (default) on the host threads."""
        check(lib().hops_rt_batch_option(self.h, name.encode(), int(value)))

    def add(self, system: RtSystem, prm: RtTrajParam, psi0, z1, z2=None):
        psi = np.ascontiguousarray(np.asarray(psi0, dtype=np.complex128))
        rc = lib().hops_rt_batch_add(self.h, system.handle, ctypes.byref(prm), _ptr(psi),
                                     ctypes.c_void_p(z1.data_ptr()),
                                     None if z2 is None else ctypes.c_void_p(z2.data_ptr()))
        if rc < 0:
            raise EngineError(lib().hops_last_error().decode())
        self._keep[rc] = (system, prm, z1, z2)
        return rc

    def remove(self, slot):
        check(lib().hops_rt_batch_remove(self.h, int(slot)))
        self._keep.pop(slot, None)

    def step(self, n_steps=1, slots=None):
        if slots is None:
            check(lib().hops_rt_batch_step(self.h, int(n_steps), None, 0))
        else:
            s = np.ascontiguousarray(np.asarray(slots, dtype=np.int32))
            check(lib().hops_rt_batch_step(self.h, int(n_steps), _ptr(s), len(s)))

    def info(self, slot):
        a = np.zeros(9, dtype=np.int64)
        check(lib().hops_rt_slot_info(self.h, int(slot), _ptr(a)))
        return {"n_state": int(a[0]), "n_hier": int(a[1]), "steps": int(a[2]), "it": int(a[3]),
                "records": int(a[4]), "bytes": int(a[5]), "early_left": int(a[6]), "used": bool(a[7]),
                "oom": bool(a[8])}

    def states(self, slot):
        n = self.info(slot)["n_state"]
        out = np.zeros(n, np.int32)
        check(lib().hops_rt_slot_states(self.h, int(slot), _ptr(out)))
        return out

    def aux_list(self, slot):
        inf = self.info(slot)
        nh = inf["n_hier"]
        ptr = np.zeros(nh + 1, np.int32)
        cap = nh * 256 + 1
        mode, val = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        check(lib().hops_rt_slot_aux(self.h, int(slot), _ptr(ptr), _ptr(mode), _ptr(val)))
        return _aux_from_csr(ptr, mode, val)

    def phi(self, slot):
        inf = self.info(slot)
        out = np.zeros(inf["n_state"] * inf["n_hier"], np.complex128)
        check(lib().hops_rt_slot_phi(self.h, int(slot), _ptr(out)))
        return out

    def zmem(self, slot, n_hmode):
        out = np.zeros(n_hmode, np.complex128)
        check(lib().hops_rt_slot_zmem(self.h, int(slot), _ptr(out)))
        return out

    def slot_noise(self, slot, z1, z2, n_t, it, noise_step, c128):
        """Point a slot at new device noise arrays (hops_rt_slot_noise)."""
        check(lib().hops_rt_slot_noise(self.h, int(slot), ctypes.c_void_p(z1.data_ptr()),
                                       None if z2 is None else ctypes.c_void_p(z2.data_ptr()), int(n_t), int(it),
                                       int(noise_step), int(bool(c128))))

    def stage_noise(self, slot, z1, z2):
        """Noise the slot's RK stages read (hops_rt_slot_stage_noise)."""
        check(lib().hops_rt_slot_stage_noise(self.h, int(slot), ctypes.c_void_p(z1.data_ptr()),
                                             None if z2 is None else ctypes.c_void_p(z2.data_ptr())))

    def static_basis(self, slot, states, aux_list):
        """STATIC_BASIS of a slot before its first step (hops_rt_slot_static_basis)."""
        st = np.ascontiguousarray(np.asarray(states, dtype=np.int32))
        a_ptr, a_mode, a_val = _aux_csr(aux_list)
        check(lib().hops_rt_slot_static_basis(self.h, int(slot), len(st), _ptr(st), len(aux_list), _ptr(a_ptr),
                                              _ptr(a_mode), _ptr(a_val)))

    def records(self, slot, nstates, clear=False):
        """Dense psi(t) [n_rec, nstates], state lists and n_hier per record."""
        inf = self.info(slot)
        n_rec = inf["records"]
        psi = np.zeros((n_rec, nstates), np.complex128)
        states, nhier = [], []
        buf = np.zeros(nstates, np.complex128)
        sbuf = np.zeros(nstates, np.int32)
        ns, nh = ctypes.c_int32(), ctypes.c_int32()
        for i in range(n_rec):
            check(lib().hops_rt_slot_record(self.h, int(slot), i, _ptr(buf), _ptr(sbuf), ctypes.byref(ns),
                                            ctypes.byref(nh)))
            st = sbuf[:ns.value].copy()
            psi[i, st] = buf[:ns.value]
            states.append(st)
            nhier.append(nh.value)
        if clear:
            check(lib().hops_rt_slot_clear_records(self.h, int(slot)))
        return psi, states, nhier

    def estimate(self, system: RtSystem, n_state, n_hier):
        return int(lib().hops_rt_slot_estimate(self.h, system.handle, int(n_state), int(n_hier)))

    def free_bytes(self):
        return int(lib().hops_rt_arena_free(self.h))

    def synchronize(self):
        torch.cuda.synchronize(self.device)

    def timers(self):
        """Accumulated seconds per runtime phase (see hops_rt_batch_timers)."""
        out = np.zeros(8, np.float64)
        check(lib().hops_rt_batch_timers(self.h, _ptr(out)))
        names = ["rk4_step", "errors_device", "select_hier", "select_states", "build_tables", "update_device",
                 "fetch_psi", "rows_stepped"]
        return dict(zip(names, out.tolist()))

    def close(self):
        h = getattr(self, "h", None)
        if h and _lib._lib is not None:
            try:
                _lib._lib.hops_rt_batch_destroy(h)
            except Exception:
                pass
        self.h = None

    def __del__(self):
        self.close()


class RaggedGroup:
    """K runtime batches, each with its share of the arena, its own stream
    and its own worker threads, stepped concurrently from K host threads.
    A batch's host work (basis selection, table building) runs while the
    others wait on their device work and transfers, so the CPU cores and the
    GPU stay busy together.  Trajectories never interact, so a trajectory's
    result is the same in whichever batch it sits.  Slots are addressed as
    (batch, slot) pairs."""

    def __init__(self, k, precision_mode="fp64", arena_bytes=None, max_slots=1024, n_threads=16,
                 device="cuda:0"):
        if k < 1:
            raise UnsupportedRequest("sub_batches < 1", "RaggedGroup")
        self.device = torch.device(device)
        if arena_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            arena_bytes = int(free * 0.6)
        # one batch keeps the caller's current stream, as a plain RaggedBatch
        self.streams = [None] if k == 1 else [torch.cuda.Stream(device=self.device) for _ in range(k)]
        self.parts = [RaggedBatch(precision_mode, int(arena_bytes) // k, max_slots, n_threads, device, stream=s)
                      for s in self.streams]
        self._lock = threading.Lock()
        self._profiling = False
        self._prof = ([0.0] * 4, [0] * 4)

    @property
    def stream(self):
        return self.parts[0].stream

    def stream_of(self, ref):
        return self.parts[ref[0]].stream

    def set_option(self, name, value):
        for p in self.parts:
            p.set_option(name, value)

    def add(self, k, system, prm, psi0, z1, z2=None):
        return (k, self.parts[k].add(system, prm, psi0, z1, z2))

    def remove(self, ref):
        self.parts[ref[0]].remove(ref[1])

    def info(self, ref):
        return self.parts[ref[0]].info(ref[1])

    def states(self, ref):
        return self.parts[ref[0]].states(ref[1])

    def aux_list(self, ref):
        return self.parts[ref[0]].aux_list(ref[1])

    def phi(self, ref):
        return self.parts[ref[0]].phi(ref[1])

    def zmem(self, ref, n_hmode):
        return self.parts[ref[0]].zmem(ref[1], n_hmode)

    def slot_noise(self, ref, z1, z2, n_t, it, noise_step, c128):
        self.parts[ref[0]].slot_noise(ref[1], z1, z2, n_t, it, noise_step, c128)

    def stage_noise(self, ref, z1, z2):
        self.parts[ref[0]].stage_noise(ref[1], z1, z2)

    def static_basis(self, ref, states, aux_list):
        self.parts[ref[0]].static_basis(ref[1], states, aux_list)

    def records(self, ref, nstates, clear=False):
        return self.parts[ref[0]].records(ref[1], nstates, clear)

    def free_bytes(self):
        return sum(p.free_bytes() for p in self.parts)

    def synchronize(self):
        torch.cuda.synchronize(self.device)

    def profile_enable(self, on=True):
        """Time the engine's launches of the following steps with HIP events
        (hops_profile_*), summed over the batches' host threads."""
        with self._lock:
            self._profiling = bool(on)
            self._prof = ([0.0] * 4, [0] * 4)

    def profile_read(self):
        """(ms, launches) for [prologue, stage, other, dense system block]."""
        with self._lock:
            return list(self._prof[0]), list(self._prof[1])

    def estimate(self, system, n_state, n_hier):
        return self.parts[0].estimate(system, n_state, n_hier)

    def _step_one(self, k, n_steps, slots=None):
        prof = self._profiling
        if prof:
            _lib.profile_enable(True)  # per host thread
        try:
            self.parts[k].step(n_steps, slots)
        finally:
            if prof:
                ms, n = _lib.profile_read()
                _lib.profile_enable(False)
                with self._lock:
                    for i in range(4):
                        self._prof[0][i] += ms[i]
                        self._prof[1][i] += n[i]

    def step(self, n_steps=1, slots=None):
        """Advance every slot of every batch (or the listed (batch, slot)
        pairs) by n_steps; batch k's steps run on host thread k (the runtime
        call releases the interpreter lock).  A batch's error is raised after
        every batch has finished its steps."""
        per = None
        if slots is not None:
            per = [[] for _ in self.parts]
            for k, sl in slots:
                per[k].append(sl)
        work = [k for k in range(len(self.parts)) if per is None or per[k]]
        if len(self.parts) == 1:
            if work:
                self._step_one(0, n_steps, None if per is None else per[0])
            return
        errs = [None] * len(self.parts)

        def run(k):
            try:
                self._step_one(k, n_steps, None if per is None else per[k])
            except BaseException as e:  # re-raised on the calling thread
                errs[k] = e

        th = [threading.Thread(target=run, args=(k,), name="hops-batch-%d" % k) for k in work]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for e in errs:
            if e is not None:
                raise e
        # every batch's work complete: arrays the caller drops after this
        # call (per-call noise) are read by no stream any more
        torch.cuda.synchronize(self.device)

    def timers(self):
        """Seconds per runtime phase summed over the batches (their phases
        overlap in time), and rows stepped."""
        out = {}
        for p in self.parts:
            for key, v in p.timers().items():
                out[key] = out.get(key, 0.0) + v
        return out

    def close(self):
        for p in self.parts:
            p.close()


class AdaptiveBatch:
    """B adaptive trajectories of one system that differ in their noise,
    each with its own basis, on one GPU.

    Same dictionaries as HopsTrajectory; ``make_adaptive`` as the
    reference's; ``initialize(psi_0)``; ``propagate(t_advance, tau)``;
    ``psi_traj[b]`` (dense [n_t, n_state]), ``state_traj[b]``,
    ``nhier_traj[b]`` after propagate.

    ``sub_batches``: the trajectories are split over this many runtime
    batches stepped concurrently (``RaggedGroup``); default one per 32
    trajectories, at most 8.  ``n_threads``: worker threads per runtime
    batch (default 16 for one batch, else 32 / sub_batches, at least 2).
    Neither changes any result."""

    def __init__(self, system_param, eom_param=None, noise_param=None, noise2_param=None,
                 hierarchy_param=None, integration_param=None, seeds=None, noise_params=None,
                 precision_mode="fp64", device="cuda:0", arena_bytes=None, n_threads=None, record=True,
                 device_select=False, sub_batches=None):
        from .trajectory import _Setup
        self.setup = _Setup(system_param, eom_param, noise_param or {"TLEN": 1000.0, "TAU": 1.0}, noise2_param,
                            hierarchy_param, integration_param)
        self.static_basis = None
        sb = self.setup.integ["STATIC_BASIS"]
        if sb is not None:
            from .adaptive import aux_tuple
            self.static_basis = ([int(x) for x in sb[0]], [aux_tuple(a) for a in sb[1]])
        if self.setup.integ["EARLY_ADAPTIVE_INTEGRATOR"] == "STATIC_STATE_INCHWORM_HIERARCHY" and sb is None:
            raise UnsupportedRequest("STATIC_STATE_INCHWORM_HIERARCHY without a STATIC_BASIS",
                                     "early time integrator clause of the propagate")
        self.effective = bool(self.setup.integ["EFFECTIVE_NOISE_INTEGRATION"])
        if noise_params is None:
            if seeds is None:
                raise TrajectoryError("AdaptiveBatch needs seeds or noise_params")
            noise_params = []
            for s in seeds:
                p = dict(self.setup.noise_param)
                p["SEED"] = s
                noise_params.append(p)
        self.noise_params = noise_params
        self.B = len(noise_params)
        info = self.setup.info
        self.noise1 = [HopsNoise(p, info, system_param, 1) for p in noise_params]
        self.noise2 = [HopsNoise(self.setup.noise2_param, info, system_param, 2) for _ in noise_params]
        self.precision_mode = precision_mode
        self.device = torch.device(device)
        self.arena_bytes = arena_bytes
        K = max(1, min(8, self.B // 32)) if sub_batches is None else int(sub_batches)
        if K < 1:
            raise UnsupportedRequest("sub_batches < 1", "AdaptiveBatch")
        self.K = min(K, max(self.B, 1))
        self.n_threads = (16 if self.K == 1 else max(2, 32 // self.K)) if n_threads is None else int(n_threads)
        self.record = record
        self.device_select = bool(device_select)
        self.delta_a = self.delta_s = 0.0
        self.update_step = 1
        self.f_discard = 0.0
        self.initialized = False
        self.t = 0.0

    def make_adaptive(self, delta_a=1e-4, delta_s=1e-4, update_step=1, f_discard=0.01, adaptive_noise=True):
        if float(f_discard) < 0 or float(f_discard) > 1:
            raise UnsupportedRequest("make_adaptive", "f_discard not in acceptable range of [0,1]")
        if self.initialized:
            raise TrajectoryError("Calling make_adaptive on an initialized trajectory")
        self.delta_a = float(delta_a) if delta_a > 0 else 0.0
        self.delta_s = float(delta_s) if delta_s > 0 else 0.0
        self.update_step = update_step
        self.f_discard = float(f_discard)

    def initialize(self, psi_0, tau=None):
        """Build every trajectory's first basis.  ``tau`` is the time step the
        batch will be propagated with (the slots are sized for it)."""
        if self.initialized:
            from .errors import LockedException
            raise LockedException("AdaptiveBatch.initialize()")
        info = self.setup.info
        self.system = RtSystem(info)
        self.per = -(-self.B // self.K)  # trajectories per runtime batch
        self.rt = RaggedGroup(self.K, self.precision_mode, self.arena_bytes, max_slots=max(self.per, 1),
                              n_threads=self.n_threads, device=self.device)
        self.rt.set_option("device_select", self.device_select)
        sp = self.rt.stream
        n1, n2 = self.noise1[0], self.noise2[0]
        if n2.model != "ZERO" and n2.interpolate != n1.interpolate:
            raise UnsupportedRequest("INTERPOLATE set for one noise and not the other", "AdaptiveBatch")
        # INTERPOLATE: the grid noise stays in HBM in complex128 and each
        # propagate call samples its cubic interpolation on the call's own
        # half-step grid (noise/hops_noise.py:242-243, 379-385); complex128
        # PRE_CALCULATED arrays are used as given
        self.interp = n1.interpolate
        self.c128 = any(nz.needs_c128() for nz in self.noise1 + self.noise2)
        if self.effective and self.c128:
            raise UnsupportedRequest("EFFECTIVE_NOISE_INTEGRATION with complex128 (interpolated or "
                                     "PRE_CALCULATED complex128) noise", "AdaptiveBatch")
        nl = info.n_l2
        self.z1 = device_noise(self.noise1, list(range(nl)), self.device, sp, c128=self.c128)
        self.z2 = None
        if n2.model != "ZERO":
            self.z2 = device_noise(self.noise2, list(range(nl)), self.device, sp, c128=self.c128)
        self._zcall = []
        torch.cuda.synchronize(self.device)
        markov = None
        filters = []
        for name, prm in self.setup.hier["STATIC_FILTERS"]:
            if name == "Markovian":
                mk = np.array(prm, bool)
                markov = mk if markov is None else markov | mk
            elif name in ("Triangular", "LongEdge"):
                filters.append((name, [list(prm[0]), int(prm[1])]))
            elif self.delta_a > 0:
                raise UnsupportedRequest("static filter %s with an adaptive hierarchy" % name, "make_adaptive")
        self._filters = filters
        aux0 = None
        if self.delta_a <= 0:
            aux0 = triangular_aux_list(info.n_hmodes, self.setup.hier["MAXHIER"])
            aux0 = apply_static_filters(aux0, self.setup.hier["STATIC_FILTERS"], info.n_hmodes)
        self._psi0 = np.asarray(psi_0, dtype=np.complex128)
        self._markov, self._aux0 = markov, aux0
        self._tau = tau
        self.slots = []
        self._params = []
        self._zeff = []
        self.initialized = True
        self.t = 0.0
        if tau is not None:
            self._admit(tau)

    def _admit(self, tau):
        info = self.setup.info
        nz1 = self.noise1[0]
        # interpolated noise: the slots read each call's own grid
        # (propagate), so the grid noise given here is never sampled
        ns = max(1, int(round(tau * 0.5 / nz1.param["TAU"]))) if self.interp else \
            int(round(tau * 0.5 / nz1.param["TAU"]))
        integ = self.setup.integ
        for b in range(self.B):
            prm = traj_param(self.setup.eom_kind, self.setup.hier["MAXHIER"], self.delta_a, self.delta_s,
                             self.update_step, self.f_discard, tau, ns, nz1.n_t, nz1.time_index(0.0),
                             integ["EARLY_ADAPTIVE_INTEGRATOR"], integ["EARLY_INTEGRATOR_STEPS"],
                             integ["INCHWORM_CAP"], self._markov, self._aux0, self.record, filters=self._filters,
                             noise_c128=self.c128)
            z1 = self.z1[b]
            z2 = None if self.z2 is None else self.z2[b]
            slot = self.rt.add(b // self.per, self.system, prm, self._psi0, z1, z2)
            if self.static_basis is not None and (self.delta_a > 0 or self.delta_s > 0):
                self.rt.static_basis(slot, *self.static_basis)
            if self.effective:
                # the RK stages read window means over the step's noise
                # samples, the selection the point samples
                # (integrator/integrator_rk.py:115-128)
                from .noise import window_mean
                sp = self.rt.stream_of(slot)
                z1e = window_mean(z1, ns, sp)
                z2e = None if z2 is None else window_mean(z2, ns, sp)
                self.rt.stage_noise(slot, z1e, z2e)
                self._zeff.append((z1e, z2e))
            self.slots.append(slot)
            self._params.append(prm)
        self._tau = tau

    def propagate(self, t_advance, tau):
        if not self.initialized:
            raise TrajectoryError("propagate called before initialize")
        from .trajectory import _check_tau
        nz1, nz2 = self.noise1[0], self.noise2[0]
        if not _check_tau(tau, nz1, nz2):
            raise TrajectoryError("Timesteps(" + str(tau * 0.5) + ") that do not match noise.param['TAU'] ("
                                  + str(nz1.param["TAU"]) + ")")
        if (self.t + t_advance + tau) > nz1.param["TLEN"]:
            raise TrajectoryError("Trajectory times longer than noise.param['TLEN'] =" + str(nz1.param["TLEN"]))
        ns = int(round(tau * 0.5 / nz1.param["TAU"]))
        if not self.interp and abs(tau * 0.5 - ns * nz1.param["TAU"]) >= _precision:
            raise UnsupportedRequest("Off axis t-samples when INTERPOLATE = False", "NoiseModel.get_noise()")
        if not self.slots:
            self._admit(tau)
        elif abs(tau - self._tau) > 1e-12:
            raise UnsupportedRequest("changing tau between propagate calls", "AdaptiveBatch.propagate")
        n_steps = int(np.ceil(t_advance / tau))
        if self.interp:
            from .noise import spline_noise
            n_out = 2 * n_steps + 1
            zcall = []
            for b, slot in enumerate(self.slots):
                sp = self.rt.stream_of(slot)
                z1 = spline_noise(self.z1[b], nz1.param["TAU"], self.t, 0.5 * tau, n_out, sp,
                                  nz1.param["FLAG_REAL"])
                z2 = None if self.z2 is None else spline_noise(self.z2[b], nz2.param["TAU"], self.t, 0.5 * tau,
                                                               n_out, sp, nz2.param["FLAG_REAL"])
                self.rt.slot_noise(slot, z1, z2, n_out, 0, 1, True)
                zcall.append((z1, z2))
            self._zcall = zcall  # the previous call's arrays are no longer read
        self.rt.step(n_steps)
        self.t += n_steps * tau

    # -- read-backs ---------------------------------------------------------
    def results(self, b):
        """(psi_traj dense [n_rec, nstates], state lists, n_hier list)."""
        return self.rt.records(self.slots[b], self.setup.info.nstates)

    def phi(self, b):
        return self.rt.phi(self.slots[b])

    def state_list(self, b):
        return self.rt.states(self.slots[b])

    def auxiliary_list(self, b):
        return self.rt.aux_list(self.slots[b])

    def z_mem(self, b):
        """Noise memory over the absolute modes, len(L_NOISE1) entries as the
        reference sizes it (hops_trajectory.py:290-292); the modes beyond the
        hierarchy modes stay zero."""
        info = self.setup.info
        zm = self.rt.zmem(self.slots[b], info.n_hmodes)
        out = np.zeros(max(info.n_noise_modes, len(zm)), dtype=np.complex128)
        out[: len(zm)] = zm
        return out
