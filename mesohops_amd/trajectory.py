"""HopsTrajectory / HopsBatch: the reference's user API on top of the HIP engine.

HopsTrajectory mirrors trajectory/hops_trajectory.py:84-538 (same constructor
dicts, defaults, validation and exceptions; initialize(); propagate();
readbacks psi/phi/t/z_mem/state_list/auxiliary_list/storage).  HopsBatch is
the batched form: B trajectories of one system that differ only in their noise
run through one device operator.
"""
from __future__ import annotations

import time
import warnings

import numpy as np
import torch

from . import _lib
from .basis import BasisTables, SystemInfo, apply_static_filters, aux_order_key, triangular_aux_list
from .engine import BatchState, DeviceOperator, _stream_ptr
from .errors import LockedException, TrajectoryError, UnsupportedRequest
from .noise import HopsNoise, device_noise
from .params import (EOM_DICT_DEFAULT, EOM_DICT_TYPES, HIERARCHY_DICT_DEFAULT, HIERARCHY_DICT_TYPES, NOISE_DICT_DEFAULT,
                     INTEGRATION_DICT_DEFAULT, INTEGRATION_DICT_TYPES, initialize_dict, precision)

_EOM_KIND = {"LINEAR": _lib.HOPS_EOM_LINEAR, "NONLINEAR": _lib.HOPS_EOM_NONLINEAR,
             "NORMALIZED NONLINEAR": _lib.HOPS_EOM_NORMALIZED_NONLINEAR}
_PREC = {"fp64": _lib.HOPS_PREC_FP64, "mixed": _lib.HOPS_PREC_MIXED}


def _save_psi_traj(phi_new, state_list, **kw):
    return np.array(phi_new[: len(state_list)], dtype=np.complex128)


def _save_aux_norm(phi_new, aux_list, **kw):
    n_aux = len(aux_list)
    return np.sqrt(np.sum(np.abs(np.asarray(phi_new)).reshape(n_aux, -1) ** 2, axis=1))


# storage/storage_functions.py:8-196
STORAGE_DEFAULT_FUNC = {
    "psi_traj": _save_psi_traj,
    "phi_traj": lambda phi_new, **kw: np.array(phi_new),
    "phi_norm": lambda phi_new, **kw: np.linalg.norm(phi_new),
    "t_axis": lambda t_new, **kw: t_new,
    "aux_list": lambda aux_list, **kw: [np.array([[m, v] for (m, v) in a], dtype=int).reshape(-1, 2)
                                        if len(a) else np.array([], dtype=int) for a in aux_list],
    "state_list": lambda state_list, **kw: list(state_list),
    "list_nstate": lambda state_list, **kw: len(state_list),
    "list_nhier": lambda aux_list, **kw: len(aux_list),
    "list_aux_norm": _save_aux_norm,
    "z_mem": lambda z_mem_new, **kw: z_mem_new,
}


class HopsStorage:
    """Per-step record of a trajectory (storage/hops_storage.py:13-209): the
    same keys, defaults (adaptive runs add aux_list, state_list, list_nhier,
    list_nstate, list_aux_norm), callables as custom savers, STORAGE_TIME, and
    the dense reconstruction of psi_traj for adaptive runs."""

    def __init__(self, adaptive, storage_dic):
        self.storage_dic = dict(storage_dic or {})
        self.storage_time = self.storage_dic.pop("STORAGE_TIME", True)
        self._adaptive = False
        self.n_dim = 0
        self.metadata = {"INITIALIZATION_TIME": 0, "LIST_PROPAGATION_TIME": []}
        self.adaptive = adaptive

    @property
    def adaptive(self):
        return self._adaptive

    @adaptive.setter
    def adaptive(self, new):
        self._adaptive = new
        self.dic_save = {}
        self.data = {}
        for k, v in (("phi_traj", False), ("psi_traj", True), ("t_axis", True), ("z_mem", False)):
            self.storage_dic.setdefault(k, v)
        if new:
            for k in ("aux_list", "state_list", "list_nhier", "list_nstate", "list_aux_norm"):
                self.storage_dic.setdefault(k, True)
        for key, value in self.storage_dic.items():
            if value is not False:
                if isinstance(value, bool):
                    if key not in STORAGE_DEFAULT_FUNC:
                        raise UnsupportedRequest("this value", "Storage.adaptive")
                    self.dic_save[key] = STORAGE_DEFAULT_FUNC[key]
                elif callable(value):
                    self.dic_save[key] = value
                else:
                    raise UnsupportedRequest("this value", "Storage.adaptive")
        for key, value in self.storage_dic.items():
            if value:
                self.data[key] = []

    @property
    def needs_full_state(self):
        """True when a saver needs more than psi and t of each step."""
        return any(k not in ("psi_traj", "t_axis", "state_list", "list_nstate") for k in self.dic_save)

    def store_step(self, **kwargs):
        for key, fn in self.dic_save.items():
            self.data[key].append(fn(**kwargs))

    def check_storage_time(self, t):
        st = self.storage_time
        if isinstance(st, (list, np.ndarray)):
            return t in st
        if isinstance(st, (int, float)) and not isinstance(st, bool):
            return t % st == 0
        if isinstance(st, bool):
            return st
        raise UnsupportedRequest("Type {} is not allowed in the STORAGE_TIME key of storage parameters for "
                                 "checking whether the current time should be saved".format(type(st)),
                                 "check_storage_time", override=True)

    def __getitem__(self, item):
        if self._adaptive and item == "psi_traj":
            full = np.zeros([len(self["t_axis"]), self.n_dim], dtype=np.complex128)
            for i, psi in enumerate(self.data["psi_traj"]):
                full[i, np.array(self.data["state_list"][i])] = psi
            return full
        if self._adaptive and item == "psi_traj_sparse":
            import scipy.sparse as sps
            dat, row, col = [], [], []
            for i, psi in enumerate(self.data["psi_traj"]):
                dat.extend(list(psi))
                row.extend([i] * len(psi))
                col.extend(self.data["state_list"][i])
            return sps.csr_array((dat, (row, col)), shape=(len(self["t_axis"]), self.n_dim))
        if item in self.data:
            return self.data[item]
        print("{} not a key of storage.data".format(item))


class _Setup:
    """Validated parameter dicts shared by HopsTrajectory and HopsBatch."""

    def __init__(self, system_param, eom_param, noise_param, noise2_param, hierarchy_param,
                 integration_param):
        self.system_param = system_param
        self.eom = initialize_dict(eom_param or {}, EOM_DICT_DEFAULT, EOM_DICT_TYPES, "HopsEOM")
        if self.eom["EQUATION_OF_MOTION"] not in _EOM_KIND:
            raise UnsupportedRequest("EQUATION_OF_MOTION =" + self.eom["EQUATION_OF_MOTION"], "HopsEOM")
        self.eom["ADAPTIVE"] = bool(self.eom["ADAPTIVE_H"] or self.eom["ADAPTIVE_S"])
        self.hier = initialize_dict(hierarchy_param or {}, HIERARCHY_DICT_DEFAULT, HIERARCHY_DICT_TYPES,
                                    "HopsHierarchy")
        if self.hier["MAXHIER"] > 255:
            print("Warning: using a hierarchy depth greater than 255 can cause integer overflow issues "
                  "when calculating derivative error. Resetting hierarchy depth to 255.")
            self.hier["MAXHIER"] = 255
        if integration_param is None:
            integration_param = INTEGRATION_DICT_DEFAULT
        self.integ = initialize_dict(integration_param, INTEGRATION_DICT_DEFAULT, INTEGRATION_DICT_TYPES,
                                     "integration_param in the HopsTrajectory initialization")
        if self.integ["INTEGRATOR"] != "RUNGE_KUTTA":
            raise UnsupportedRequest("Integrator of type " + str(self.integ["INTEGRATOR"]), "HopsTrajectory")
        self.info = SystemInfo(system_param)
        self.noise_param = dict(noise_param or {})
        if noise2_param is None:
            noise2_param = {"TLEN": self.noise_param["TLEN"], "TAU": self.noise_param["TAU"],
                            "MODEL": "ZERO", "SEED": None}
        else:
            if ("SEED" in self.noise_param and "SEED" in noise2_param
                    and noise2_param["SEED"] is not None
                    and not isinstance(noise2_param["SEED"], np.ndarray)
                    and noise2_param["SEED"] == self.noise_param["SEED"]):
                warnings.warn("Using the same seed for both noise 1 and noise 2 may introduce unphysical "
                              "correlations between independent noise terms.")
        self.noise2_param = dict(noise2_param)
        self.eom_kind = _EOM_KIND[self.eom["EQUATION_OF_MOTION"]]

    def full_basis(self):
        info = self.info
        if self.eom["ADAPTIVE"]:
            raise UnsupportedRequest("adaptive propagation (make_adaptive)", "HopsTrajectory.initialize")
        aux = triangular_aux_list(info.n_hmodes, self.hier["MAXHIER"])
        aux = apply_static_filters(aux, self.hier["STATIC_FILTERS"], info.n_hmodes)
        return BasisTables(info, np.arange(info.nstates), aux, self.eom_kind)


def _check_tau(tau, noise1, noise2, integrator_step=0.5):
    """trajectory/hops_trajectory.py:579-614."""
    ts = tau * integrator_step
    ok = True
    for nz in (noise1, noise2):
        T = nz.param["TAU"]
        ok = ok and (ts / T - int(round(ts / T)) < precision) and int(round(tau / T)) >= 1
    return ok


class HopsBatch:
    """B trajectories of one system, one hierarchy and one EOM, differing in
    their noise seeds, propagated together on the device.

    precision: "fp64" (complex128 everywhere) or "mixed" (complex64 stages and
    super-operator, complex128 state carry and FP64 norm/<L>/z-memory)."""

    def __init__(self, system_param, eom_param=None, noise_param=None, noise2_param=None,
                 hierarchy_param=None, integration_param=None, seeds=None, noise_params=None,
                 precision_mode="fp64", device="cuda:0", store_psi=True, lanes=None):
        self.lanes = lanes
        self.setup = _Setup(system_param, eom_param, noise_param, noise2_param, hierarchy_param,
                            integration_param)
        if noise_params is None:
            if seeds is None:
                raise TrajectoryError("HopsBatch needs seeds or noise_params")
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
        self.store_psi = store_psi
        self.t = 0.0
        self.initialized = False

    def initialize(self, psi_0):
        if self.initialized:
            raise LockedException("HopsTrajectory.initialize()")
        t0 = time.time()
        self.tables = self.setup.full_basis()
        tb = self.tables
        self.dop = DeviceOperator(tb, _PREC[self.precision_mode], self.device)
        sp = _stream_ptr(self.device)
        nz1, nz2 = self.noise1[0], self.noise2[0]
        # noise the reference uses in complex128 (interpolated, or a wider
        # PRE_CALCULATED array) stays complex128 here; noise 2 follows noise 1
        c128 = any(n.needs_c128() for n in self.noise1 + self.noise2)
        self._interp = nz1.interpolate
        if nz2.model != "ZERO" and nz2.interpolate != self._interp:
            raise UnsupportedRequest("INTERPOLATE set for one noise and not the other", "HopsTrajectory")
        if c128 and self.setup.integ["EFFECTIVE_NOISE_INTEGRATION"]:
            raise UnsupportedRequest("EFFECTIVE_NOISE_INTEGRATION with complex128 (interpolated or "
                                     "PRE_CALCULATED complex128) noise", "HopsTrajectory")
        z1 = device_noise(self.noise1, tb.absL_active, self.device, sp, c128=c128)
        z2 = None
        if nz2.model != "ZERO":
            z2 = device_noise(self.noise2, tb.absL_active, self.device, sp, c128=c128)
        self._z_raw = (z1, z2)
        self._eff_w = None
        self.state = BatchState(self.dop, self.B, self.setup.eom_kind, z1, z2, self.noise1[0].n_t,
                                store_slots=0, lanes=self.lanes)
        psi0 = np.asarray(psi_0, dtype=np.complex128)
        phi = np.zeros(tb.n_hier * tb.n_state, dtype=np.complex128)
        phi[:tb.n_state] = psi0[tb.state_list]
        self.state.set_phi(phi)
        self.psi_traj = [[phi[:tb.n_state].copy()] for _ in range(self.B)]
        self.t_axis = [0.0]
        self.t = 0.0
        self.initialized = True
        self.init_time = time.time() - t0

    def _plan(self, t_advance, tau):
        nz1, nz2 = self.noise1[0], self.noise2[0]
        if getattr(self, "_interp", False):
            # cubic noise: any stage time passing the reference's step check
            # (trajectory/hops_trajectory.py:432-446); the noise of the call
            # is evaluated on t + j tau/2 (noise index j, one point per half step)
            if not _check_tau(tau, nz1, nz2):
                raise TrajectoryError("Timesteps(" + str(tau * 0.5) + ") that do not match noise.param['TAU'] ("
                                      + str(nz1.param["TAU"]) + ")")
            if (self.t + t_advance + tau) > nz1.param["TLEN"]:
                raise TrajectoryError("Trajectory times longer than noise.param['TLEN'] ="
                                      + str(nz1.param["TLEN"]))
            return int(np.ceil(t_advance / tau)), 0, 1
        if not _check_tau(tau, nz1, nz2):
            raise TrajectoryError("Timesteps(" + str(tau * 0.5) + ") that do not match noise.param['TAU'] ("
                                  + str(nz1.param["TAU"]) + ")")
        if (self.t + t_advance + tau) > nz1.param["TLEN"]:
            raise TrajectoryError("Trajectory times longer than noise.param['TLEN'] =" + str(nz1.param["TLEN"]))
        n_steps = int(np.ceil(t_advance / tau))
        it0 = nz1.time_index(self.t)
        ns = int(round(tau * 0.5 / nz1.param["TAU"]))
        # the signed check above lets a half step between grid points through;
        # the reference then fails on the off-grid noise lookup
        # (noise/hops_noise.py:439-447), and so does this engine
        for nz in (nz1, nz2):
            if abs(tau * 0.5 - ns * nz.param["TAU"]) >= precision or nz.param["TAU"] != nz1.param["TAU"]:
                raise UnsupportedRequest("Off axis t-samples when INTERPOLATE = False", "NoiseModel.get_noise()")
        return n_steps, it0, ns

    def propagate(self, t_advance, tau, store=None):
        """Advance all trajectories; psi after every step is recorded when
        ``store`` (default: store_psi) is true."""
        if not self.initialized:
            raise TrajectoryError("propagate called before initialize")
        n_steps, it0, ns = self._plan(t_advance, tau)
        store = self.store_psi if store is None else store
        st = self.state
        if self.setup.integ["EFFECTIVE_NOISE_INTEGRATION"]:
            self._use_effective_noise(ns)
        if store and (st.psi_store is None or st.store_slots < n_steps):
            st.store_slots = n_steps
            st.psi_store = torch.zeros(self.B * n_steps * self.tables.n_state, dtype=torch.complex128,
                                       device=self.device)
            st.bt.store_stride = n_steps
        st.bt.psi_store_d = st.psi_store.data_ptr() if store else None
        if self._interp and n_steps > 0:
            self._interpolated_noise(tau, n_steps)
        st.propagate(it0, n_steps, tau, ns, 0)
        t_axis = self.t + np.arange(1, n_steps + 1) * tau
        if store:
            psi = st.psi_store_host(n_steps)
            for b in range(self.B):
                self.psi_traj[b].extend(list(psi[b]))
            self.t_axis.extend([float(t) for t in t_axis])
        self.t = float(t_axis[-1]) if n_steps else self.t

    def _interpolated_noise(self, tau, n_steps):
        """The cubic noise at t + j tau/2, j = 0..2 n_steps, for one call."""
        from .noise import spline_noise
        sp = _stream_ptr(self.device)
        z1g, z2g = self._z_raw
        n_out = 2 * n_steps + 1
        nz1, nz2 = self.noise1[0], self.noise2[0]
        z1 = spline_noise(z1g, nz1.param["TAU"], self.t, 0.5 * tau, n_out, sp, nz1.param["FLAG_REAL"])
        z2 = None if z2g is None else spline_noise(z2g, nz2.param["TAU"], self.t, 0.5 * tau, n_out, sp,
                                                   nz2.param["FLAG_REAL"])
        self.state.set_noise(z1, z2, n_out)

    def _use_effective_noise(self, w):
        """Stage noise = moving averages over w samples (integrator/integrator_rk.py:115-128)."""
        if self._eff_w == w:
            return
        from .noise import window_mean
        sp = _stream_ptr(self.device)
        z1, z2 = self._z_raw
        st = self.state
        st.z1 = window_mean(z1, w, sp)
        st.bt.z1_d = st.z1.data_ptr()
        if z2 is not None:
            st.z2 = window_mean(z2, w, sp)
            st.bt.z2_d = st.z2.data_ptr()
        self._eff_w = w

    def rewind(self, psi_0):
        """Reset every trajectory to t = 0 with the initial state psi_0 and
        zero noise memory (the device operator and noise are kept)."""
        tb = self.tables
        phi = np.zeros(tb.n_hier * tb.n_state, dtype=np.complex128)
        phi[:tb.n_state] = np.asarray(psi_0, dtype=np.complex128)[tb.state_list]
        self.state.set_phi(phi)
        self.state.zmem.zero_()
        self.psi_traj = [[phi[:tb.n_state].copy()] for _ in range(self.B)]
        self.t_axis = [0.0]
        self.t = 0.0

    def relayout(self, lanes):
        """Rebuild the device pools with ``lanes`` interleaved trajectories per
        row, keeping the state, the noise memory and the noise."""
        old = self.state
        phi = old.phi_host()
        zmem = old.zmem.clone()
        self.state = None
        del old.phi, old.y0, old.ya, old.yb, old.acc
        st = BatchState(self.dop, self.B, self.setup.eom_kind, old.z1, old.z2, int(old.z1.shape[-1]),
                        store_slots=0, lanes=lanes)
        st.set_phi(phi)
        st.zmem.copy_(zmem)
        self.state = st
        self.lanes = lanes

    def synchronize(self):
        torch.cuda.synchronize(self.device)

    @property
    def phi(self):
        return self.state.phi_host()

    @property
    def psi(self):
        return self.state.phi_host()[:, :self.tables.n_state]


def write_checkpoint(filepath, phi, z_mem, t, state_list, aux_list, early_counter, storage_data, storage_meta,
                     params, compress=True):
    """The reference's checkpoint file (trajectory/hops_trajectory.py:884-936):
    keys phi, z_mem (the sparse (len(L_NOISE1), 1) noise memory), t,
    state_list, aux_list (an object array of (k, 2) [mode, depth] int arrays),
    early_counter, storage_data, storage_meta and params."""
    aux = np.empty(len(aux_list), dtype=object)
    for i, a in enumerate(aux_list):
        aux[i] = np.array([[m, v] for (m, v) in a], dtype=int).reshape(-1, 2)
    ck = {"phi": np.asarray(phi, dtype=np.complex128), "z_mem": z_mem, "t": float(t),
          "state_list": np.array(state_list, dtype=int), "aux_list": aux, "early_counter": int(early_counter),
          "storage_data": storage_data, "storage_meta": storage_meta, "params": params}
    (np.savez_compressed if compress else np.savez)(filepath, **ck)


def read_checkpoint(filename):
    """Parse a checkpoint in the reference's layout (as written by its
    save_checkpoint or by write_checkpoint; trajectory/hops_trajectory.py:
    939-1034).  The auxiliary list may come back as a 1-D object array of
    (k, 2) arrays or, when every auxiliary array has the same shape, stacked
    by NumPy into one [n_aux, k, 2] array; both give the same (mode, depth)
    tuples.  The noise memory comes back as a dense vector.  Reads pickled
    parameter objects (functions, sparse matrices), as the reference does:
    load only files you trust."""
    data = np.load(filename, allow_pickle=True)
    aux = [tuple((int(m), int(v)) for m, v in np.asarray(a).reshape(-1, 2)) for a in data["aux_list"]]
    zm = data["z_mem"].item()
    zm = np.asarray(zm.toarray() if hasattr(zm, "toarray") else zm, dtype=np.complex128).reshape(-1)
    return {"params": data["params"].item(), "phi": np.asarray(data["phi"], dtype=np.complex128),
            "z_mem": zm, "t": float(data["t"]), "state_list": [int(s) for s in data["state_list"]],
            "aux_list": aux, "early_counter": int(data["early_counter"]),
            "storage_data": data["storage_data"].item(), "storage_meta": data["storage_meta"].item()}


class HopsTrajectory:
    """One HOPS trajectory with the reference's constructor and methods
    (trajectory/hops_trajectory.py:84-538, :1071-1125).  make_adaptive()
    switches to the adaptive basis (mesohops_amd/adaptive.py).  Extra keyword
    arguments select the precision mode ("fp64" by default) and the device."""

    def __init__(self, system_param, eom_param=None, noise_param=None, noise2_param=None,
                 hierarchy_param=None, storage_param=None, integration_param=None,
                 precision_mode="fp64", device="cuda:0"):
        if noise_param is None:
            noise_param = {}
        self._batch = HopsBatch(system_param, eom_param, noise_param, noise2_param, hierarchy_param,
                                integration_param, noise_params=[dict(noise_param)],
                                precision_mode=precision_mode, device=device)
        self.noise_param = noise_param
        self.noise1 = self._batch.noise1[0]
        self.noise2 = self._batch.noise2[0]
        self.integration_param = self._batch.setup.integ
        self.storage = HopsStorage(self._batch.setup.eom["ADAPTIVE"], storage_param or {})
        self._adaptive = None
        self.__initialized__ = False

    # -- reference API ----------------------------------------------------
    def make_adaptive(self, delta_a=1e-4, delta_s=1e-4, update_step=1, f_discard=0.01, adaptive_noise=True):
        """trajectory/hops_trajectory.py:338-399.  adaptive_noise changes no
        value: the noise of an L-operator does not depend on when it is first
        generated (tests/test_hops_noise.py:297-362)."""
        if float(f_discard) < 0 or float(f_discard) > 1:
            raise UnsupportedRequest("make_adaptive", "f_discard not in acceptable range of [0,1]")
        if self.__initialized__:
            raise TrajectoryError("Calling make_adaptive on an initialized trajectory")
        eom = self._batch.setup.eom
        if delta_a > 0:
            eom.update(ADAPTIVE=True, ADAPTIVE_H=True, DELTA_A=delta_a, UPDATE_STEP=update_step,
                       F_DISCARD=float(f_discard))
            self.storage.adaptive = True
        if delta_s > 0:
            eom.update(ADAPTIVE=True, ADAPTIVE_S=True, DELTA_S=delta_s, UPDATE_STEP=update_step,
                       F_DISCARD=float(f_discard))
            self.storage.adaptive = True

    @property
    def adaptive(self):
        return bool(self._batch.setup.eom["ADAPTIVE"])

    def _store(self, t, phi, aux_list, state_list, z_mem):
        self.storage.store_step(phi_new=phi, aux_list=aux_list, state_list=state_list, t_new=t,
                                z_mem_new=z_mem)

    def initialize(self, psi_0, timer_checkpoint=None):
        if timer_checkpoint is None:
            timer_checkpoint = time.time()
        if self.__initialized__:
            raise LockedException("HopsTrajectory.initialize()")
        b = self._batch
        self.storage.n_dim = b.setup.info.nstates
        if self.adaptive:
            from .adaptive import AdaptiveTrajectoryEngine
            e = b.setup.eom
            ip = self.integration_param
            self._adaptive = AdaptiveTrajectoryEngine(
                b.setup, self.noise1, self.noise2, b.precision_mode, b.device,
                e["DELTA_A"] if e["ADAPTIVE_H"] else 0.0, e["DELTA_S"] if e["ADAPTIVE_S"] else 0.0,
                e["UPDATE_STEP"], e["F_DISCARD"], ip["EARLY_ADAPTIVE_INTEGRATOR"],
                ip["EARLY_INTEGRATOR_STEPS"], ip["INCHWORM_CAP"], ip["STATIC_BASIS"])
            self._adaptive.initialize(psi_0, self.noise1.time_index(0.0))
            self._t = 0.0
        else:
            b.initialize(psi_0)
        self._store(0, self.phi, self.auxiliary_list, list(self.state_list), self.z_mem)
        self.storage.metadata["INITIALIZATION_TIME"] = time.time() - timer_checkpoint
        self.__initialized__ = True

    def propagate(self, t_advance, tau, timer_checkpoint=None):
        if timer_checkpoint is None:
            timer_checkpoint = time.time()
        if not self.__initialized__:
            raise TrajectoryError("propagate called before initialize")
        b = self._batch
        n_steps, it0, ns = b._plan(t_advance, tau) if not self.adaptive else self._plan_adaptive(t_advance, tau)
        t_axis = self.t + np.arange(1, 1 + n_steps) * tau
        print("Integration from ", self.t, " to ", np.max(t_axis) if len(t_axis) else self.t)
        if self.adaptive:
            ad = self._adaptive
            ad.begin_call(self.t, tau, n_steps)
            for index_t, t in enumerate(t_axis):
                z_lag = self.z_mem
                ad.step(index_t, it0 + 2 * ns * index_t, tau, ns)
                if self.storage.check_storage_time(t):
                    self._store(float(t), ad.phi_host(), self.auxiliary_list, list(self.state_list), z_lag)
                self._t = float(t)
        elif self.storage.needs_full_state:
            # one step per call: every saver sees the whole state
            for t in t_axis:
                z_lag = self.z_mem
                b.propagate(tau, tau, store=False)
                if self.storage.check_storage_time(t):
                    self._store(float(t), self.phi, self.auxiliary_list, list(self.state_list), z_lag)
        else:
            n0 = len(b.t_axis)
            b.propagate(t_advance, tau, store=True)
            for t, psi in zip(b.t_axis[n0:], b.psi_traj[0][n0:]):
                if self.storage.check_storage_time(t):
                    self._store(t, psi, self.auxiliary_list, list(self.state_list), None)
        self.storage.metadata["LIST_PROPAGATION_TIME"].append(time.time() - timer_checkpoint)

    def _operator(self, op):
        """Acts with a system operator (dense or SciPy sparse, over all
        states) on every auxiliary of the hierarchy vector
        (trajectory/hops_trajectory.py:540-577).  With state adaptivity the
        states op populates are added to the basis first and the early-time
        integrator starts over."""
        if not self.__initialized__:
            raise TrajectoryError("_operator called before initialize")
        import scipy.sparse as sps
        if self.adaptive:
            self._adaptive.apply_operator(op)
            return
        full = op.toarray() if sps.issparse(op) else np.asarray(op)
        self._batch.state.apply_state_op(full)

    def reset_early_time_integrator(self):
        """The next propagate starts with EARLY_INTEGRATOR_STEPS early-time
        steps again (trajectory/hops_trajectory.py:761-772)."""
        if self.early_integrator != "INCH_WORM":
            raise UnsupportedRequest("Early type integrators of type other than INCH_WORM",
                                     "reset_early_time_integrator", override=True)
        if self._adaptive is not None:
            self._adaptive.early_counter = 0

    @property
    def early_integrator(self):
        return self.integration_param["EARLY_ADAPTIVE_INTEGRATOR"]

    @property
    def static_basis(self):
        return self.integration_param["STATIC_BASIS"]

    def _plan_adaptive(self, t_advance, tau):
        b = self._batch
        saved, saved_interp = b.t, getattr(b, "_interp", False)
        b.t = self._t
        b._interp = self._adaptive.interp
        try:
            return b._plan(t_advance, tau)
        finally:
            b.t, b._interp = saved, saved_interp

    # -- checkpoints (trajectory/hops_trajectory.py:853-1034) ------------
    _SYSTEM_KEYS = ["HAMILTONIAN", "GW_SYSBATH", "L_HIER", "L_NOISE1", "ALPHA_NOISE1", "PARAM_NOISE1",
                    "L_NOISE2", "ALPHA_NOISE2", "PARAM_NOISE2", "L_LT_CORR", "PARAM_LT_CORR"]

    def save_checkpoint(self, filepath, compress=True, drop_seed=False):
        """Parameters, the current hierarchy vector, state and auxiliary
        lists, noise memory, time, early-step counter and the storage, in the
        reference's .npz layout (keys phi, z_mem, t, state_list, aux_list,
        early_counter, storage_data, storage_meta, params)."""
        if not self.__initialized__:
            raise TrajectoryError("save_checkpoint before initialize")
        b = self._batch
        setup = b.setup
        eom_param = {k: v for k, v in setup.eom.items() if k != "ADAPTIVE"}
        for key in ("DELTA_A", "DELTA_S", "F_DISCARD"):
            if key in eom_param:
                eom_param[key] = float(eom_param[key])
        if eom_param.get("UPDATE_STEP") is None:
            eom_param.pop("UPDATE_STEP", None)

        def noise_dict(nz):
            return {k: nz.param[k] for k in NOISE_DICT_DEFAULT if k in nz.param and not (k == "SEED" and drop_seed)}

        sysp = setup.system_param
        params = {
            "system_param": {k: sysp[k] for k in self._SYSTEM_KEYS if k in sysp},
            "hierarchy_param": {k: setup.hier[k] for k in HIERARCHY_DICT_DEFAULT if k in setup.hier},
            "eom_param": eom_param,
            "noise1_param": noise_dict(self.noise1),
            "noise2_param": noise_dict(self.noise2),
            "integration_param": dict(self.integration_param),
            "storage_param": dict(self.storage.storage_dic),
        }
        if not drop_seed and self.noise1.param["SEED"] is None:
            warnings.warn("Noise1 seed is None but drop_seed is False; this may cause unphysical behavior.",
                          UserWarning)
        early = self._adaptive.early_counter if self._adaptive is not None else 0
        write_checkpoint(filepath, self.phi, self.z_mem, self.t, self.state_list, self.auxiliary_list, early,
                         self.storage.data, self.storage.metadata, params, compress)

    @classmethod
    def load_checkpoint(cls, filename, add_seed1=None, add_seed2=None, add_system_param=None,
                        precision_mode="fp64", device="cuda:0"):
        """Rebuild a trajectory from a file written by save_checkpoint: the
        same parameters (seeds and system optionally replaced), then the
        stored basis, hierarchy vector, noise memory, time and storage, so
        that propagation continues exactly where it stopped.  Reads pickled
        parameter objects (functions, sparse matrices), as the reference's
        checkpoints do: load only files you trust."""
        ck = read_checkpoint(filename)
        params = ck["params"]
        if add_seed1 is not None:
            params["noise1_param"]["SEED"] = add_seed1
        if add_seed2 is not None:
            params["noise2_param"]["SEED"] = add_seed2
        if add_system_param is not None:
            params["system_param"] = add_system_param
        traj = cls(params["system_param"], eom_param=params["eom_param"], noise_param=params["noise1_param"],
                   noise2_param=params["noise2_param"], hierarchy_param=params["hierarchy_param"],
                   storage_param=params["storage_param"], integration_param=params["integration_param"],
                   precision_mode=precision_mode, device=device)
        states, phi, aux, zm = ck["state_list"], ck["phi"], ck["aux_list"], ck["z_mem"]
        psi_0 = np.zeros(traj._batch.setup.info.nstates, dtype=np.complex128)
        psi_0[states] = phi[:len(states)]
        traj.initialize(psi_0)
        info = traj._batch.setup.info
        zdev = torch.zeros(info.n_hmodes + 1, dtype=torch.complex128)
        zdev[: info.n_hmodes] = torch.from_numpy(zm[: info.n_hmodes])
        if traj._adaptive is not None:
            ad = traj._adaptive
            ad.update_basis(ad.phi_host(), states, aux)
            if [int(s) for s in ad.tb.state_list] != sorted(states) or set(ad.tb.aux_list) != set(aux):
                raise TrajectoryError("checkpoint basis could not be restored")
            ad.st.set_phi(phi)
            ad.st.zmem.copy_(zdev.to(ad.device))
            ad._sync_zmem_from_state()
            ad.early_counter = ck["early_counter"]
            traj._t = ck["t"]
        else:
            b = traj._batch
            if [tuple(a) for a in b.tables.aux_list] != sorted(aux, key=aux_order_key):
                raise TrajectoryError("checkpoint hierarchy does not match the parameters")
            b.state.set_phi(phi)
            b.state.zmem[: info.n_hmodes + 1].copy_(zdev.to(b.device)[: info.n_hmodes + 1])
            b.t = ck["t"]
            b.t_axis = [b.t]
            b.psi_traj = [[phi[: b.tables.n_state].copy()]]
        traj.storage.data = ck["storage_data"]
        traj.storage.metadata = ck["storage_meta"]
        return traj

    # -- readbacks (trajectory/hops_trajectory.py:1071-1125) -------------
    @property
    def phi(self):
        if self._adaptive is not None:
            return self._adaptive.phi_host()
        return self._batch.phi[0]

    @phi.setter
    def phi(self, phi):
        """Replaces the whole hierarchy vector (in the current basis) on the
        device (trajectory/hops_trajectory.py:1099-1101).  The next step takes
        it as given, like a state an operator has just produced."""
        phi = np.asarray(phi, dtype=np.complex128).reshape(-1)
        if len(phi) != self.n_state * self.n_hier:
            raise TrajectoryError("phi of length %d does not match the basis (%d states x %d auxiliaries)"
                                  % (len(phi), self.n_state, self.n_hier))
        if self._adaptive is not None:
            self._adaptive.st.set_phi(phi)
            self._adaptive.raw_pending = True
        else:
            self._batch.state.set_phi(phi)
            self._batch.state.raw_first = True
        torch.cuda.synchronize(self._batch.device)

    @property
    def psi(self):
        return self.phi[:self.n_state]

    @property
    def t(self):
        return self._t if self._adaptive is not None else self._batch.t

    @property
    def z_mem(self):
        """Noise memory over the absolute modes as a (len(L_NOISE1), 1) CSR
        array, as the reference keeps it."""
        import scipy.sparse as sps
        if self._adaptive is not None:
            zm = self._adaptive.st.zmem.cpu().numpy()[: self._batch.setup.info.n_hmodes]
        else:
            zm = self._batch.state.zmem_host()[0]
        n = max(self._batch.setup.info.n_noise_modes, len(zm))
        full = np.zeros((n, 1), dtype=np.complex128)
        full[: len(zm), 0] = zm
        return sps.csr_array(full)

    @property
    def _tables(self):
        return self._adaptive.tb if self._adaptive is not None else self._batch.tables

    @property
    def n_hier(self):
        return self._tables.n_hier

    @property
    def n_state(self):
        return self._tables.n_state

    @property
    def state_list(self):
        return self._tables.state_list

    @property
    def auxiliary_list(self):
        return self._tables.aux_list

    @property
    def update_step(self):
        return self._batch.setup.eom["UPDATE_STEP"]

    @property
    def early_steps(self):
        return self.integration_param["EARLY_INTEGRATOR_STEPS"]

    @property
    def inchworm_cap(self):
        return self.integration_param["INCHWORM_CAP"]
