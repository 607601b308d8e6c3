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
from .basis import BasisTables, SystemInfo, apply_static_filters, triangular_aux_list
from .engine import BatchState, DeviceOperator, _stream_ptr
from .errors import LockedException, TrajectoryError, UnsupportedRequest
from .noise import HopsNoise, device_noise
from .params import (EOM_DICT_DEFAULT, EOM_DICT_TYPES, HIERARCHY_DICT_DEFAULT, HIERARCHY_DICT_TYPES,
                     INTEGRATION_DICT_DEFAULT, INTEGRATION_DICT_TYPES, initialize_dict, precision)

_EOM_KIND = {"LINEAR": _lib.HOPS_EOM_LINEAR, "NONLINEAR": _lib.HOPS_EOM_NONLINEAR,
             "NORMALIZED NONLINEAR": _lib.HOPS_EOM_NORMALIZED_NONLINEAR}
_PREC = {"fp64": _lib.HOPS_PREC_FP64, "mixed": _lib.HOPS_PREC_MIXED}


class HopsStorage:
    """Per-step record of psi and t (storage/hops_storage.py:33-209); psi is
    downloaded from the device at the end of each propagate call."""

    def __init__(self, storage_param):
        p = dict(storage_param or {})
        self.storage_time = p.pop("STORAGE_TIME", True)
        self.storage_dic = {"phi_traj": False, "psi_traj": True, "t_axis": True, "z_mem": False}
        for k, v in p.items():
            if k not in ("phi_traj", "psi_traj", "t_axis", "z_mem"):
                raise UnsupportedRequest("storage key {}".format(k), "HopsStorage")
            self.storage_dic[k] = v
        self.data = {k: [] for k, v in self.storage_dic.items() if v}
        self.metadata = {"INITIALIZATION_TIME": 0, "LIST_PROPAGATION_TIME": []}

    def check_storage_time(self, t):
        st = self.storage_time
        if isinstance(st, (list, np.ndarray)):
            return t in st
        if isinstance(st, (int, float)) and not isinstance(st, bool):
            return t % st == 0
        if isinstance(st, bool):
            return st
        raise UnsupportedRequest("Type {} is not allowed in the STORAGE_TIME key".format(type(st)),
                                 "check_storage_time", override=True)

    def store(self, t, psi, phi=None, z_mem=None):
        if "psi_traj" in self.data:
            self.data["psi_traj"].append(psi)
        if "t_axis" in self.data:
            self.data["t_axis"].append(t)
        if "phi_traj" in self.data and phi is not None:
            self.data["phi_traj"].append(phi)
        if "z_mem" in self.data and z_mem is not None:
            self.data["z_mem"].append(z_mem)

    def __getitem__(self, key):
        return self.data[key]


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
        if self.integ["EFFECTIVE_NOISE_INTEGRATION"]:
            raise UnsupportedRequest("EFFECTIVE_NOISE_INTEGRATION", "HopsTrajectory")
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
        z1 = device_noise(self.noise1, tb.absL_active, self.device, sp)
        z2 = None
        if self.noise2[0].model != "ZERO":
            z2 = device_noise(self.noise2, tb.absL_active, self.device, sp)
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
        if store and (st.psi_store is None or st.store_slots < n_steps):
            st.store_slots = n_steps
            st.psi_store = torch.zeros(self.B * n_steps * self.tables.n_state, dtype=torch.complex128,
                                       device=self.device)
            st.bt.store_stride = n_steps
        st.bt.psi_store_d = st.psi_store.data_ptr() if store else None
        st.propagate(it0, n_steps, tau, ns, 0)
        t_axis = self.t + np.arange(1, n_steps + 1) * tau
        if store:
            psi = st.psi_store_host(n_steps)
            for b in range(self.B):
                self.psi_traj[b].extend(list(psi[b]))
            self.t_axis.extend([float(t) for t in t_axis])
        self.t = float(t_axis[-1]) if n_steps else self.t

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
        st = BatchState(self.dop, self.B, self.setup.eom_kind, old.z1, old.z2, self.noise1[0].n_t,
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


class HopsTrajectory:
    """One HOPS trajectory with the reference's constructor and methods
    (trajectory/hops_trajectory.py:84-538).  Extra keyword arguments select
    the precision mode ("fp64" by default) and the device."""

    def __init__(self, system_param, eom_param=None, noise_param=None, noise2_param=None,
                 hierarchy_param=None, storage_param=None, integration_param=None,
                 precision_mode="fp64", device="cuda:0"):
        if noise_param is None:
            noise_param = {}
        self._batch = HopsBatch(system_param, eom_param, noise_param, noise2_param, hierarchy_param,
                                integration_param, noise_params=[dict(noise_param)],
                                precision_mode=precision_mode, device=device)
        self.noise_param = noise_param
        self.storage = HopsStorage(storage_param)
        self.__initialized__ = False

    # -- reference API ----------------------------------------------------
    def make_adaptive(self, delta_a=1e-4, delta_s=1e-4, update_step=1, f_discard=0.01, adaptive_noise=True):
        if float(f_discard) < 0 or float(f_discard) > 1:
            raise UnsupportedRequest("f_discard", "f_discard not in acceptable range of [0,1]")
        if self.__initialized__:
            raise TrajectoryError("Calling make_adaptive on an initialized trajectory")
        eom = self._batch.setup.eom
        if delta_a > 0:
            eom.update(ADAPTIVE=True, ADAPTIVE_H=True, DELTA_A=delta_a, UPDATE_STEP=update_step,
                       F_DISCARD=float(f_discard))
        if delta_s > 0:
            eom.update(ADAPTIVE=True, ADAPTIVE_S=True, DELTA_S=delta_s, UPDATE_STEP=update_step,
                       F_DISCARD=float(f_discard))

    def initialize(self, psi_0, timer_checkpoint=None):
        if timer_checkpoint is None:
            timer_checkpoint = time.time()
        if self.__initialized__:
            raise LockedException("HopsTrajectory.initialize()")
        self._batch.initialize(psi_0)
        b = self._batch
        self.storage.store(0, b.psi_traj[0][0])
        self.storage.metadata["INITIALIZATION_TIME"] = time.time() - timer_checkpoint
        self.__initialized__ = True

    def propagate(self, t_advance, tau, timer_checkpoint=None):
        if timer_checkpoint is None:
            timer_checkpoint = time.time()
        b = self._batch
        n0 = len(b.t_axis)
        t_end = b.t + np.arange(1, 1 + int(np.ceil(t_advance / tau))) * tau
        print("Integration from ", b.t, " to ", np.max(t_end) if len(t_end) else b.t)
        b.propagate(t_advance, tau, store=True)
        for t, psi in zip(b.t_axis[n0:], b.psi_traj[0][n0:]):
            if self.storage.check_storage_time(t):
                self.storage.store(t, psi)
        self.storage.metadata["LIST_PROPAGATION_TIME"].append(time.time() - timer_checkpoint)

    # -- readbacks (trajectory/hops_trajectory.py:1071-1125) -------------
    @property
    def phi(self):
        return self._batch.phi[0]

    @property
    def psi(self):
        return self.phi[:self.n_state]

    @property
    def t(self):
        return self._batch.t

    @property
    def z_mem(self):
        return self._batch.state.zmem_host()[0]

    @property
    def n_hier(self):
        return self._batch.tables.n_hier

    @property
    def n_state(self):
        return self._batch.tables.n_state

    @property
    def state_list(self):
        return self._batch.tables.state_list

    @property
    def auxiliary_list(self):
        return self._batch.tables.aux_list
