"""Dyadic HOPS on the MI355X engine.

A dyadic trajectory propagates a ket and a bra together as one wave function
of the doubled space (ket block: states 0..N-1, bra block: N..2N-1) under the
block-diagonal Hamiltonian and L-operators H + H, L + L
(trajectory/hops_dyadic.py:15-246).  Field interactions are system operators
acting on one side, applied on the device by HopsTrajectory._operator; the
response function is read from the stored wave functions
(util/spectroscopy_analysis.py:67-113).  DyadicSpectra is the linear
spectroscopy driver (absorption, fluorescence) of trajectory/dyadic_spectra.py.

Propagation, the operator products and the noise filter run on the GPU
through the engine; this module only builds the dyadic system, sequences
the interactions and evaluates the response from the stored psi(t).
"""
import time

import numpy as np
import scipy.sparse as sps

from .errors import UnsupportedRequest
from .params import bcf_exp, c, h
from .trajectory import HopsTrajectory

__all__ = ["DyadicTrajectory", "DyadicSpectra", "dyad_matrix", "response_function_calc", "zeropadded_fft",
           "prepare_spectroscopy_input_dict", "prepare_chromophore_input_dict",
           "prepare_convergence_parameter_dict"]


def dyad_matrix(m):
    """(N, N) -> (2N, 2N) block-diagonal [[M, 0], [0, M]]: a SciPy COO matrix
    for a sparse input, a dense complex128 array otherwise
    (hops_dyadic.py:113-143)."""
    n = m.shape[0]
    if sps.issparse(m):
        q = m.tocoo()
        return sps.coo_matrix((np.concatenate((q.data, q.data)),
                               (np.concatenate((q.row, q.row + n)), np.concatenate((q.col, q.col + n)))),
                              shape=(2 * n, 2 * n))
    out = np.zeros((2 * n, 2 * n), dtype=np.complex128)
    out[:n, :n] = m
    out[n:, n:] = m
    return out


def response_function_calc(F_op, index_t, traj, list_response_norm_sq):
    """Normalised response <psi(t)|F|psi(t)> prod(norms) / |psi(t)|^2 for
    every stored time after index_t (util/spectroscopy_analysis.py:67-113).
    traj: [n_t, 2N], dense or sparse."""
    F = F_op.tocsr() if sps.issparse(F_op) else sps.csr_array(F_op)
    scale = np.prod(list_response_norm_sq)
    if sps.issparse(traj):
        rows = sps.csr_array(traj)
        out = []
        for t in range(int(index_t) + 1, rows.shape[0]):
            row = rows[[t], :]
            col = row.T
            out.append(scale / np.linalg.norm(row.data) ** 2 * (np.conj(row) @ (F @ col)).todense()[0, 0])
        return np.asarray(out)
    tr = np.asarray(traj)
    return np.array([scale / np.linalg.norm(tr[t]) ** 2 * (np.conj(tr[t]) @ (F @ tr[t]))
                     for t in range(int(index_t) + 1, tr.shape[0])])


def zeropadded_fft(R_t, t_step, w_res=c * (10 ** -7)):
    """Real part of the FFT of R(t) zero-padded to 1 / (w_res t_step) points
    (frequency resolution w_res, default 1 cm^-1), scaled by the padded
    length, on the frequency axis -fftshift(h fftfreq)
    (util/spectroscopy_analysis.py:5-64)."""
    n_data = len(R_t)
    n_pad = int(1 / (w_res * t_step))
    if n_pad < n_data:
        n_pad = n_data
        print("WARNING: Defined frequency resolution is less than inherent resolution from the time step. "
              "FFT will continue with resolution provided by time step.")
    padded = np.zeros(n_pad, dtype=complex)
    padded[:n_data] = R_t[:]
    R_w = np.real(np.fft.fftshift(np.fft.fft(padded))) / n_pad
    w_axis = -np.fft.fftshift(h * np.fft.fftfreq(n_pad, t_step))
    return R_w, w_axis


class DyadicTrajectory(HopsTrajectory):
    """A HOPS trajectory in the doubled (ket + bra) space
    (trajectory/hops_dyadic.py:15-246).  The system dictionary's Hamiltonian
    and L-operators (L_HIER, L_NOISE1, L_NOISE2, L_LT_CORR) are replaced in
    place by their dyadic forms, as the reference does."""

    def __init__(self, system_param, eom_param=None, noise_param=None, noise2_param=None,
                 hierarchy_param=None, storage_param=None, integration_param=None, **engine_kw):
        if system_param is None:
            system_param = {}
        system_param.update({"HAMILTONIAN": dyad_matrix(system_param["HAMILTONIAN"])})
        for key in ("L_HIER", "L_NOISE1", "L_NOISE2", "L_LT_CORR"):
            if key in system_param:
                system_param.update({key: [dyad_matrix(m) for m in system_param[key]]})
        super().__init__(system_param, eom_param or {}, noise_param or {}, noise2_param, hierarchy_param or {},
                         storage_param or {}, integration_param, **engine_kw)
        self._norm_sq = []

    def _M2_dyad_conversion(self, M2_hilbert):
        return dyad_matrix(M2_hilbert)

    def initialize(self, psi_ket, psi_bra, timer_checkpoint=None):
        """Starts from (psi_ket, psi_bra) normalised as one vector; its squared
        norm is the first response normalisation factor
        (hops_dyadic.py:83-111)."""
        if timer_checkpoint is None:
            timer_checkpoint = time.time()
        psi = np.concatenate((psi_ket, psi_bra))
        self._norm_sq = [np.linalg.norm(psi) ** 2]
        super().initialize(psi / np.sqrt(self._norm_sq[0]), timer_checkpoint=timer_checkpoint)

    @property
    def _normalized(self):
        return self._batch.setup.eom["EQUATION_OF_MOTION"] == "NORMALIZED NONLINEAR"

    def _dyad_operator(self, op_hilbert, side):
        """Acts with op_hilbert on the ket (side 'ket': [[op, 0], [0, 1]]) or
        the bra (side 'bra': [[1, 0], [0, op]]) through _operator, the
        identity covering the other side's basis states, and records the
        response normalisation factor; normalised EOMs then renormalise the
        hierarchy to a unit physical wave function (hops_dyadic.py:145-214)."""
        n = op_hilbert.shape[0]
        states = [int(s) for s in self.state_list]
        before = np.linalg.norm(self.phi[:len(states)]) ** 2
        if side not in ("ket", "bra"):
            raise UnsupportedRequest('sides other than "ket" or "bra"', "_dyad_operator")
        if sps.issparse(op_hilbert):
            q = op_hilbert.tocoo()
            other = [s for s in states if (s < n if side == "bra" else s >= n)]
            shift = n if side == "bra" else 0
            rows = np.concatenate((other, q.row + shift))
            cols = np.concatenate((other, q.col + shift))
            vals = np.concatenate((np.ones(len(other)), q.data))
            op = sps.coo_matrix((vals, (rows.astype(int), cols.astype(int))), shape=(2 * n, 2 * n),
                                dtype=np.complex128)
        else:
            op = np.zeros((2 * n, 2 * n))
            op[states, states] = 1
            block = slice(n, 2 * n) if side == "bra" else slice(0, n)
            op[block, block] = op_hilbert
        self._operator(op)
        after = np.linalg.norm(self.phi[:len(self.state_list)]) ** 2  # the basis after the operator
        if self._normalized:
            self._norm_sq.append(after)
            phi = self.phi
            self.phi = phi / np.sqrt(np.sum(np.abs(phi[:self.n_state]) ** 2))
        else:
            self._norm_sq.append(after / before)

    def _response_function_comp(self, F_op, index_t):
        """Response function of the stored trajectory after index_t
        (hops_dyadic.py:216-242)."""
        traj = self.storage["psi_traj_sparse"] if self.adaptive else self.storage["psi_traj"]
        return response_function_calc(F_op, index_t, traj, self._norm_sq)

    @property
    def list_response_norm_sq(self):
        return self._norm_sq


class DyadicSpectra(DyadicTrajectory):
    """Linear spectra of one noise realisation by dyadic HOPS
    (trajectory/dyadic_spectra.py:12-360): the ground state on both sides,
    field interactions as raising / lowering operators on the ket or bra, a
    NORMALIZED NONLINEAR propagation between them, and the response of the
    final signal operator.  ABSORPTION and FLUORESCENCE pathways."""

    def __init__(self, spectroscopy_dict, chromophore_dict, convergence_dict, seed, **engine_kw):
        self._initialized = False
        sd, cd, vd = spectroscopy_dict, chromophore_dict, convergence_dict
        self.spectrum_type = sd["spectrum_type"]
        self.t_1, self.t_2, self.t_3 = sd["t_1"], sd["t_2"], sd["t_3"]
        self.list_t = [self.t_1, self.t_2, self.t_3]
        self.E_1, self.E_2, self.E_3, self.E_sig = sd["E_1"], sd.get("E_2"), sd.get("E_3"), sd["E_sig"]
        self.list_ket_sites = sd["list_ket_sites"]
        self.list_bra_sites = sd.get("list_bra_sites", None)
        self.M2_mu_ge = cd["M2_mu_ge"]
        self.n_chromophore = cd["n_chromophore"]
        self.H2_sys_hamiltonian = cd["H2_sys_hamiltonian"]
        self.lop_list_hier, self.gw_sysbath_hier = cd["lop_list_hier"], cd["gw_sysbath_hier"]
        self.lop_list_noise, self.gw_sysbath_noise = cd["lop_list_noise"], cd["gw_sysbath_noise"]
        self.lop_list_ltc, self.ltc_param = cd["lop_list_ltc"], cd["ltc_param"]
        self.static_filter_list = cd.get("static_filter_list", None)
        self.t_step, self.max_hier = vd["t_step"], vd["max_hier"]
        self.delta_a, self.delta_s = vd["delta_a"], vd["delta_s"]
        self.set_update_step, self.set_f_discard = vd["set_update_step"], vd["set_f_discard"]
        self.n_state_hilb = self.n_chromophore + 1
        self.n_state_dyad = 2 * self.n_state_hilb
        if np.shape(self.H2_sys_hamiltonian) != (self.n_state_hilb, self.n_state_hilb):
            raise ValueError("H2_sys_hamiltonian must be ((n_chrom + 1) x (n_chrom + 1)) to account for each "
                             "chromophore and the ground state.")
        system_param = {"HAMILTONIAN": self.H2_sys_hamiltonian, "GW_SYSBATH": self.gw_sysbath_hier,
                        "L_HIER": self.lop_list_hier, "L_NOISE1": self.lop_list_noise, "ALPHA_NOISE1": bcf_exp,
                        "PARAM_NOISE1": self.gw_sysbath_noise, "L_LT_CORR": self.lop_list_ltc,
                        "PARAM_LT_CORR": self.ltc_param}
        # the noise covers all propagation windows plus 1000 fs; TAU is half a
        # step for fractional steps, else 0.5 fs
        self.noise_param = {"SEED": seed, "MODEL": "FFT_FILTER", "TLEN": float(1000 + np.sum(self.list_t)),
                            "TAU": 0.5 if self.t_step % 1 == 0 else self.t_step / 2}
        self.noise2_param = None
        hierarchy_param = {"MAXHIER": self.max_hier}
        if self.static_filter_list:
            hierarchy_param["STATIC_FILTERS"] = self.static_filter_list
        super().__init__(system_param, {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}, self.noise_param,
                         self.noise2_param, hierarchy_param, {}, **engine_kw)

    def initialize(self):
        """Ground state on ket and bra; adaptive when delta_a or delta_s > 0
        (dyadic_spectra.py:161-196)."""
        if self._initialized:
            print("WARNING: DyadicTrajectory has already been initialized.")
            return
        t0 = time.time()
        ground = np.zeros(self.n_state_hilb)
        ground[0] = 1
        if self.delta_a > 0 or self.delta_s > 0:
            self.make_adaptive(self.delta_a, self.delta_s, self.set_update_step, self.set_f_discard)
        super().initialize(ground, ground.copy(), timer_checkpoint=t0)
        self._initialized = True

    def _hilb_operator(self, action_type, field, list_sites):
        """Raising |site><0| (or lowering |0><site|) operator weighted by
        mu_site . field, sites counted from 1 (dyadic_spectra.py:198-238)."""
        list_sites = np.asarray(list_sites)
        weights = np.dot(self.M2_mu_ge[list_sites - 1], field)
        ground = np.zeros_like(list_sites)
        shape = (self.n_state_hilb, self.n_state_hilb)
        if action_type == "raise":
            return sps.coo_matrix((weights, (list_sites, ground)), shape=shape, dtype=np.float64)
        if action_type == "lower":
            return sps.coo_matrix((weights, (ground, list_sites)), shape=shape, dtype=np.float64)
        raise ValueError("action_type must be either 'raise' or 'lower'.")

    def _final_dyad_operator(self):
        """Signal operator sum_n (mu_n . E_sig) |0_bra><n_ket| of the doubled
        space and the time index the response starts after
        (dyadic_spectra.py:240-266)."""
        weights = np.dot(self.M2_mu_ge, self.E_sig)
        index = int(self.t_2 / self.t_step)
        op = sps.coo_matrix((weights, ([self.n_state_hilb] * self.n_chromophore, np.arange(1, self.n_state_hilb))),
                            shape=(self.n_state_dyad, self.n_state_dyad), dtype=np.float64)
        return op, index

    def calculate_spectrum(self):
        """Time-domain response of this trajectory for the configured pathway,
        scaled for its degenerate and conjugate pathways: absorption x 2,
        fluorescence x 4 (dyadic_spectra.py:268-356)."""
        self.initialize()
        scale = 1
        final_op, final_index = self._final_dyad_operator()
        if self.spectrum_type == "ABSORPTION":
            t0 = time.time()
            self._dyad_operator(self._hilb_operator("raise", self.E_1, self.list_ket_sites), "ket")
            self.propagate(self.t_1, self.t_step, t0)
            scale = 2
        elif self.spectrum_type == "FLUORESCENCE":
            t0 = time.time()
            self._dyad_operator(self._hilb_operator("raise", self.E_1, self.list_ket_sites), "ket")
            self._dyad_operator(self._hilb_operator("raise", self.E_2, self.list_bra_sites), "bra")
            self.propagate(self.t_2, self.t_step, t0)
            t0 = time.time()
            self._dyad_operator(self._hilb_operator("lower", self.E_3, np.arange(1, self.n_state_hilb)), "bra")
            self.propagate(self.t_3, self.t_step, t0)
            scale = 4
        return scale * self._response_function_comp(final_op, final_index)

    @property
    def initialized(self):
        return self._initialized


def prepare_spectroscopy_input_dict(spectrum_type, propagation_time_dict, field_dict, site_dict):
    """spectroscopy_dict for DyadicSpectra from the spectrum type, the
    propagation times (t_1, t_2, t_3; entries <= 0 or None are dropped), the
    field vectors (3-vectors) and the excited sites (indexed from 1)
    (dyadic_spectra.py:362-491)."""
    kinds = ["ABSORPTION", "FLUORESCENCE"]
    if "list_ket_sites" not in site_dict:
        raise ValueError("list_ket_sites must be defined.")
    if not isinstance(site_dict["list_ket_sites"], np.ndarray):
        site_dict["list_ket_sites"] = np.array(site_dict["list_ket_sites"])
    for sites in site_dict.values():
        if 0 in sites:
            raise ValueError("Ket and Bra sites must be indexed starting from 1.")
    for vec in field_dict.values():
        if not isinstance(vec, np.ndarray):
            raise ValueError("All field entries should be numpy arrays.")
        if vec.shape != (3,):
            raise ValueError("All field entries should be numpy arrays with exactly 3 entries.")
    for key in [k for k, v in propagation_time_dict.items() if v is None or v <= 0]:
        del propagation_time_dict[key]
    if spectrum_type == "ABSORPTION":
        if "t_1" not in propagation_time_dict:
            raise ValueError("Propagation time after first field interaction (t_1) must be defined as > 0 for "
                             "absorption.")
        if "E_1" not in field_dict:
            raise ValueError("E_1 must be defined for absorption.")
        if len(propagation_time_dict) > 1:
            print("WARNING: Only t_1 is necessary for absorption. Setting all other propagation times to zero.")
        if len(field_dict) > 1:
            print("WARNING: Only E_1 is necessary for absorption. E_sig is set to E_1. All other field "
                  "definitions will be discarded")
        return {"spectrum_type": spectrum_type, "E_1": field_dict["E_1"], "E_sig": field_dict["E_1"],
                "t_1": propagation_time_dict["t_1"], "t_2": 0, "t_3": 0,
                "list_ket_sites": site_dict["list_ket_sites"]}
    if spectrum_type == "FLUORESCENCE":
        if "list_bra_sites" not in site_dict:
            raise ValueError("list_bra_sites must be defined for fluorescence.")
        if not isinstance(site_dict["list_bra_sites"], np.ndarray):
            site_dict["list_bra_sites"] = np.array(site_dict["list_bra_sites"])
        if "t_2" not in propagation_time_dict or "t_3" not in propagation_time_dict:
            raise ValueError("Propagation times after second and third field interactions (t_2, t_3) must be "
                             "defined as > 0 for fluorescence.")
        if "E_1" not in field_dict:
            raise ValueError("E_1 must be defined for fluorescence.")
        if "E_sig" not in field_dict:
            print("WARNING: E_sig is not defined. Setting E_sig to default, [0, 0, 1].")
        if len(propagation_time_dict) > 2:
            print("WARNING: Only t_2 and t_3 are necessary for fluorescence. Setting all other propagation times "
                  "to zero.")
        if len(field_dict) > 2:
            print("WARNING: Only E_1 and E_sig are necessary for fluorescence. All other field definitions will "
                  "be discarded.")
        e_sig = field_dict.get("E_sig", np.array([0, 0, 1]))
        return {"spectrum_type": spectrum_type, "E_1": field_dict["E_1"], "E_2": field_dict["E_1"],
                "E_3": e_sig, "E_sig": e_sig, "t_1": 0, "t_2": propagation_time_dict["t_2"],
                "t_3": propagation_time_dict["t_3"], "list_ket_sites": site_dict["list_ket_sites"],
                "list_bra_sites": site_dict["list_bra_sites"]}
    raise ValueError(f"spectrum_type must be one of the following: {kinds}")


def _check_static_filters(bath_dict):
    filters = bath_dict["static_filter_list"]
    if not isinstance(filters, list):
        raise ValueError("static_filter_list must be a list.")
    for i, item in enumerate(filters):
        where = f"Error in filter {i}: "
        if not isinstance(item, list):
            raise ValueError(where + "static_filter_list must be a list.")
        if len(item) != 2:
            raise ValueError(where + "each filter in static_filter_list must be a 2-element list of the form: "
                             "[filter_name, filter_params].")
        name, params = item
        if name not in ["Markovian", "Triangular", "LongEdge"]:
            raise ValueError(where + "Filter names must be 'Markovian', 'Triangular', or 'LongEdge'.")
        if name == "Markovian":
            if not all(isinstance(x, bool) for x in params):
                raise ValueError(where + "filter_params for Markovian filters must be a list of booleans.")
            flags = params
        else:
            if len(params) != 2:
                raise ValueError(where + "Triangular/LongEdge filter_params must be a list containing a list of "
                                 "booleans and an integer.")
            if not isinstance(params[1], int):
                raise ValueError(where + "The second entry in filter_params for Triangular/LongEdge filters must "
                                 "be an integer.")
            if not all(isinstance(x, bool) for x in params[0]):
                raise ValueError(where + "The first entry in filter_params for Triangular/LongEdge filters must "
                                 "be a list of booleans.")
            if params[1] < 0:
                raise ValueError(where + "Triangular/LongEdge filter_params must have a positive integer as the "
                                 "second element.")
            flags = params[0]
        if "list_modes" in bath_dict:
            if len(flags) != len(bath_dict["list_modes"]) // 2:
                raise ValueError(where + "The list of booleans in filter_params must have the same length as the "
                                 "number of modes in each bath.")
            expanded = flags * len(bath_dict["list_lop"])
            if name == "Markovian":
                item[1] = expanded
            else:
                item[1][0] = expanded
        elif len(flags) != sum(len(m) for m in bath_dict["list_modes_by_bath"]) // 2:
            raise ValueError(where + "The list of booleans in filter_params must have the same length as the "
                             "number of modes in all baths combined.")


def prepare_chromophore_input_dict(M2_mu_ge, H2_sys_hamiltonian, bath_dict):
    """chromophore_dict for DyadicSpectra (dyadic_spectra.py:493-848): one
    bath per L-operator (site projectors |n><n|, n = 1..N, by default), each
    with its exponential modes (list_modes shared by all baths, or
    list_modes_by_bath); the last nmodes_LTC modes of a bath go to the
    low-temperature correction (sum of g/w) instead of the hierarchy, and
    every mode enters the noise."""
    n_chrom = len(M2_mu_ge)
    M2_mu_ge = np.array(M2_mu_ge)
    if M2_mu_ge.shape[1] != 3:
        raise ValueError("M2_mu_ge must be a numpy array with shape (n_chromophore, 3).")
    for key, value in list(bath_dict.items()):
        if type(value) == np.ndarray:
            bath_dict[key] = list(value)
        elif key != "nmodes_LTC" and (value is None or value == 0):
            del bath_dict[key]
    if bath_dict.get("nmodes_LTC") is None:
        bath_dict["nmodes_LTC"] = 0
    elif not isinstance(bath_dict["nmodes_LTC"], int):
        raise ValueError("nmodes_LTC must be an integer or None.")
    elif bath_dict["nmodes_LTC"] < 0:
        raise ValueError("nmodes_LTC must be >= 0 or set to None.")
    if bath_dict["nmodes_LTC"] > 0 and "static_filter_list" in bath_dict:
        raise ValueError("The use of static hierarchy filters with low-temperature correction is not currently "
                         "supported.")
    if "list_modes_by_bath" in bath_dict and "list_modes" in bath_dict:
        raise ValueError("list_modes_by_bath and list_modes should not both be defined.")
    if "list_lop" not in bath_dict:
        bath_dict["list_lop"] = [sps.coo_matrix(([1], ([i + 1], [i + 1])), shape=(n_chrom + 1, n_chrom + 1))
                                 for i in range(n_chrom)]
    if "list_modes" in bath_dict:
        if len(bath_dict["list_modes"]) % 2 != 0:
            raise ValueError("list_modes should contain paired Gs and Ws, which guarantees an even number of "
                             "elements.")
        bath_dict["list_modes_by_bath"] = [bath_dict["list_modes"] for _ in range(len(bath_dict["list_lop"]))]
    elif "list_modes_by_bath" in bath_dict:
        if len(bath_dict["list_modes_by_bath"]) != len(bath_dict["list_lop"]):
            raise ValueError("list_modes_by_bath and list_lop must have the same length.")
        for modes in bath_dict["list_modes_by_bath"]:
            if not isinstance(modes, list):
                raise ValueError("list_modes_by_bath must be a list of lists.")
            if len(modes) % 2 != 0:
                raise ValueError("sublists within list_modes_by_bath should contain paired Gs and Ws, which "
                                 "guarantees an even number of elements in each sublist.")
    else:
        raise ValueError("Either list_modes_by_bath or list_modes must be defined.")
    if "static_filter_list" in bath_dict:
        _check_static_filters(bath_dict)
    gw_hier, l_hier, gw_noise, l_noise, l_ltc, ltc = [], [], [], [], [], []
    n_ltc = bath_dict["nmodes_LTC"]
    for lop, flat in zip(bath_dict["list_lop"], bath_dict["list_modes_by_bath"]):
        modes = [(flat[i], flat[i + 1]) for i in range(0, len(flat), 2)]
        if n_ltc >= len(modes):
            raise ValueError("nmodes_LTC must be less than the number of modes in each bath.")
        l_ltc.append(lop)
        corr = 0
        for i, mode in enumerate(modes):
            gw_noise.append(mode)
            l_noise.append(lop)
            if len(modes) - i > n_ltc:
                gw_hier.append(mode)
                l_hier.append(lop)
            else:
                corr += mode[0] / mode[1]
        ltc.append(corr)
    return {"M2_mu_ge": M2_mu_ge, "n_chromophore": n_chrom, "H2_sys_hamiltonian": H2_sys_hamiltonian,
            "lop_list_hier": l_hier, "gw_sysbath_hier": gw_hier, "lop_list_noise": l_noise,
            "gw_sysbath_noise": gw_noise, "lop_list_ltc": l_ltc, "ltc_param": ltc,
            "static_filter_list": bath_dict.get("static_filter_list", None)}


def prepare_convergence_parameter_dict(t_step, max_hier, delta_a=0, delta_s=0, set_update_step=1,
                                       set_f_discard=0):
    """convergence_dict for DyadicSpectra (dyadic_spectra.py:851-885)."""
    return {"t_step": t_step, "max_hier": max_hier, "delta_a": delta_a, "delta_s": delta_s,
            "set_update_step": set_update_step, "set_f_discard": set_f_discard}
