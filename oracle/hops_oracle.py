"""FP64 NumPy/SciPy restatement of the reference HOPS propagation path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  The functions below
follow the reference step by step so that their outputs can be compared with
the HIP engine; where the reference's arithmetic order fixes the last bits of
a golden fixture (the super-operator values, the noise filter), the same NumPy
operations are used in the same order and dtype.

Citations are to MesoHOPS v1.6.0 ``src/mesohops/...``.

Pins (tests/test_oracle_pins*.py):
- golden fixtures of the reference: super-operators, raw noise, correlation
  function, filtered noise, a whole trajectory, the raw noise rows of the
  basis tests;
- the reference's hand-computed answers: low-temperature-correction terms
  (matrices, norm correction, linear term) and their effect on the three
  equations of motion; every adaptive error term (deletion flux, state flux
  out of auxiliaries and states, boundary-state flux, flux up / down for
  site-projector, Peierls and mixed L-operators, by destination state, the
  mode-from-state and expectation tables, the noise and low-temperature
  operators of off-diagonal couplings), the boundary-auxiliary binning with
  f_discard, and the stable / boundary auxiliaries and states chosen on the
  reference's noise;
- brute force and closed forms (derivative, RK4, z-memory,
  determine_error_thresh, flux terms by explicit loops, cubic interpolation
  of cubics);
- invariants (adaptive with vanishing thresholds equals the full hierarchy;
  derivative error below delta, for site projectors and for off-diagonal
  L-operators with and without the low-temperature correction; _operator
  equal to a block-by-block product; STATIC_BASIS kept).
The dyadic part (tests/test_dyadic_host.py): the reference's dyadic-matrix
and side-operator answers, and absorption / fluorescence without a bath
against exact RK4 matrix powers.

Parity unpinned: the raw draws of L-operator rows beyond those the fixtures
hold (index >= 2 for SUM_GAUSSIAN, >= 1 for BOX_MULLER) are NumPy's own
PCG64 output in the reference's call sequence, with no fixture of their
values.
"""
from __future__ import annotations

import copy
import itertools
from collections import Counter

import numpy as np
import scipy.sparse as sp

# util/physical_constants.py:21-28
H_PLANCK = 33357.179
HBAR = H_PLANCK / (2 * np.pi)
KB = 0.69503
PRECISION = 1e-8


# ---------------------------------------------------------------------------
# Bath correlation functions
# ---------------------------------------------------------------------------
def bcf_exp(t_axis, g, w):
    """g * exp(-w t / hbar)  (trajectory/exp_noise.py:5-25)."""
    return g * np.exp(-w * t_axis / HBAR)


def bcf_convert_dl_to_exp(lambda_dl, gamma_dl, temp, k_matsubara=0):
    """High-temperature Drude-Lorentz mode plus Matsubara modes
    (util/bath_corr_functions.py:70-112)."""
    beta = 1 / (KB * temp)
    g_exp = 2 * lambda_dl / beta - 1j * lambda_dl * gamma_dl
    w_exp = gamma_dl
    nu = 2 * np.pi / beta

    def spectral(w):
        return 2 * lambda_dl * gamma_dl * w / (w ** 2 + gamma_dl ** 2)

    extra = []
    for k in np.arange(k_matsubara) + 1:
        w_k = k * nu
        g_k = 2j * spectral(1j * w_k) / beta
        g_exp += 2 * lambda_dl * (1 / beta) * (2 * gamma_dl ** 2) / (gamma_dl ** 2 - w_k ** 2)
        extra += [g_k, w_k]
    return [g_exp, w_exp] + extra


# ---------------------------------------------------------------------------
# System dictionary (basis/system_functions.py:10-171)
# ---------------------------------------------------------------------------
def l_signature(op):
    """Nonzero pattern of an operator as a hashable tuple
    (util/helper_functions.py:8-32)."""
    if sp.issparse(op):
        if op.getnnz() == 0:
            return tuple()
        return tuple(tuple(x) for x in op.nonzero())
    arr = np.asarray(op)
    if len(arr) == 0:
        return tuple()
    return tuple(tuple(x) for x in np.nonzero(arr))


def _states_of(op):
    """States touched by an L-operator (util/helper_functions.py:50-66)."""
    if sp.issparse(op):
        r, c = op.nonzero()
    else:
        r, c = np.nonzero(np.asarray(op))
    return tuple(sorted(set(r.tolist()) | set(c.tolist())))


def derive_system(system_param):
    """Derived system dictionary; restates initialize_system_dict
    (basis/system_functions.py:10-171) for the keys the propagation uses."""
    p = dict(system_param)
    ham = p["HAMILTONIAN"]
    H = ham.toarray() if sp.issparse(ham) else np.asarray(ham)
    p["NSTATES"] = H.shape[0]
    p["H_DENSE"] = np.array(H, dtype=np.complex128)
    p["N_HMODES"] = len(p["GW_SYSBATH"])
    p["G"] = np.array([g for (g, w) in p["GW_SYSBATH"]])
    p["W"] = np.array([w for (g, w) in p["GW_SYSBATH"]])
    p["LIST_STATE_INDICES_BY_HMODE"] = [_states_of(L) for L in p["L_HIER"]]
    p["LIST_HMODE_INDICES_BY_STATE"] = [[] for _ in range(p["NSTATES"])]
    for hm, sts in enumerate(p["LIST_STATE_INDICES_BY_HMODE"]):
        for s in sts:
            p["LIST_HMODE_INDICES_BY_STATE"][s].append(hm)
    if "L_LT_CORR" not in p:
        p["L_LT_CORR"] = []
        p["PARAM_LT_CORR"] = []
    sigs = [l_signature(L) for L in p["L_HIER"]]
    uniq = {}
    p["LIST_INDEX_L2_BY_HMODE"] = []
    p["LIST_L2_COO"] = []
    p["LIST_STATE_INDICES_BY_INDEX_L2"] = []
    for i, s in enumerate(sigs):
        if s not in uniq:
            uniq[s] = len(uniq)
            tmp = sp.coo_matrix(p["L_HIER"][i])
            tmp.eliminate_zeros()
            p["LIST_L2_COO"].append(tmp)
            p["LIST_STATE_INDICES_BY_INDEX_L2"].append(p["LIST_STATE_INDICES_BY_HMODE"][i])
        p["LIST_INDEX_L2_BY_HMODE"].append(uniq[s])
    p["N_L2"] = len(uniq)
    p["LIST_INDEX_L2_BY_STATE_INDICES"] = [[] for _ in range(p["NSTATES"])]
    for il, sts in enumerate(p["LIST_STATE_INDICES_BY_INDEX_L2"]):
        for s in sts:
            p["LIST_INDEX_L2_BY_STATE_INDICES"][s].append(il)
    p["LIST_LT_PARAM"] = [0] * p["N_L2"]
    for i, L in enumerate(p["L_LT_CORR"]):
        s = l_signature(L)
        if s in uniq:
            p["LIST_LT_PARAM"][uniq[s]] += p["PARAM_LT_CORR"][i]
    for tag in ("1", "2"):
        key = "L_NOISE" + tag
        if key not in p:
            continue
        p["LIST_NMODE%s_BY_INDEX_L2" % tag] = [[] for _ in range(len(sigs))]
        p["LIST_INDEX_L2_BY_NMODE%s" % tag] = [None] * len(p["PARAM_NOISE" + tag])
        for i, L in enumerate(p[key]):
            s = l_signature(L)
            if s in uniq:
                p["LIST_INDEX_L2_BY_NMODE%s" % tag][i] = uniq[s]
                p["LIST_NMODE%s_BY_INDEX_L2" % tag][uniq[s]].append(i)
        p["N_L2_NOISE" + tag] = len(set(l_signature(L) for L in p[key]))
    # Hamiltonian couplings used by the adaptive state basis
    # (basis/system_functions.py:41-46)
    hc = sp.coo_matrix(H)
    p["COUPLED_STATES"] = [[] for _ in range(p["NSTATES"])]
    for r, c in zip(hc.row, hc.col):
        p["COUPLED_STATES"][r].append(c)
    p["COUPLED_STATES"] = [sorted(x) for x in p["COUPLED_STATES"]]
    dest = [[] for _ in range(p["NSTATES"])]
    for s in uniq:
        if len(s) == 0:
            continue
        for a, b in zip(s[0], s[1]):
            dest[a].append(b)
    p["LIST_DESTINATION_STATES_BY_STATE_INDEX"] = [sorted(set(x)) for x in dest]
    return p


# ---------------------------------------------------------------------------
# Hierarchy (basis/hops_hierarchy.py:340-368, basis/hops_aux.py:171-201,460-480)
# ---------------------------------------------------------------------------
def aux_sort_key(aux):
    """Reference ordering: by ID-string length (= depth), then by the ID string,
    which for fixed-width mode fields is the lexicographic order of the
    expanded (sorted) mode sequence (basis/hops_aux.py:171-201,460-480)."""
    expanded = tuple(m for (m, v) in aux for _ in range(v))
    return (len(expanded), expanded)


def triangular_hierarchy(n_modes, maxhier):
    """All multi-indices of depth <= maxhier as sparse (mode, value) tuples,
    sorted as the reference sorts them (hops_hierarchy.py:340-368,582)."""
    out = []
    for k in range(maxhier + 1):
        for raw in itertools.combinations_with_replacement(range(n_modes), k):
            c = Counter(raw)
            out.append(tuple((m, c[m]) for m in sorted(c)))
    out.sort(key=aux_sort_key)
    return out


def aux_dict(aux):
    return {m: v for (m, v) in aux}


def aux_step(aux, mode, step):
    d = aux_dict(aux)
    d[mode] = d.get(mode, 0) + step
    if d[mode] < 0:
        return None
    if d[mode] == 0:
        del d[mode]
    return tuple((m, d[m]) for m in sorted(d))


# ---------------------------------------------------------------------------
# Basis bookkeeping (basis/hops_system.py:206-286, basis/hops_modes.py:111-171,
# basis/hops_basis.py:90-115)
# ---------------------------------------------------------------------------
class Basis:
    """State list + auxiliary list + derived relative indexing."""

    def __init__(self, sysd, state_list, aux_list):
        self.sysd = sysd
        self.state_list = np.array(sorted(int(s) for s in state_list), dtype=int)
        self.aux_list = sorted(aux_list, key=aux_sort_key)
        assert len(self.aux_list[0]) == 0, "zero auxiliary must be present"
        self.aux_index = {a: i for i, a in enumerate(self.aux_list)}
        st = self.state_list
        n = len(st)
        self.n_state = n
        self.n_hier = len(self.aux_list)
        modes_in_use = sorted({m for a in self.aux_list for (m, v) in a})
        state_modes = sorted({m for s in st for m in sysd["LIST_HMODE_INDICES_BY_STATE"][s]})
        self.state_modes = np.array(state_modes, dtype=int)
        self.modes = np.array(sorted(set(modes_in_use) | set(state_modes)), dtype=int)
        self.rel_mode = {int(m): i for i, m in enumerate(self.modes)}
        l_by_hmode = sysd["LIST_INDEX_L2_BY_HMODE"]
        self.absL = np.array(sorted({l_by_hmode[m] for m in self.modes}), dtype=int)
        self.relL_by_mode = np.array([list(self.absL).index(l_by_hmode[m]) for m in self.modes], dtype=int)
        self.absL_active = np.array(
            sorted({l for s in st for l in sysd["LIST_INDEX_L2_BY_STATE_INDICES"][s]}), dtype=int)
        self.relL_active = np.array([list(self.absL).index(l) for l in self.absL_active], dtype=int)
        self.g = np.array([sysd["G"][m] for m in self.modes])
        self.w = np.array([sysd["W"][m] for m in self.modes])
        # reduced L-operators (basis/hops_system.py:318-354)
        pos = {int(s): i for i, s in enumerate(st)}
        self.L_coo = []
        for l in self.absL:
            c = sysd["LIST_L2_COO"][l]
            keep = [(pos[int(i)], pos[int(j)], d) for (i, j, d) in zip(c.row, c.col, c.data)
                    if int(i) in pos and int(j) in pos]
            if keep:
                r = [k[0] for k in keep]
                cc = [k[1] for k in keep]
                dd = np.array([k[2] for k in keep])
                self.L_coo.append(sp.coo_matrix((dd, (r, cc)), shape=(n, n)))
            else:
                self.L_coo.append(sp.coo_matrix((n, n)))
        self.L_csr = [sp.csr_array(L) for L in self.L_coo]
        self.L_sq_csr = [L @ L for L in self.L_csr]
        self.H = sysd["H_DENSE"][np.ix_(st, st)]
        lt = np.array(sysd["LIST_LT_PARAM"], dtype=np.complex128)
        self.lt_active = lt[self.absL_active] if len(self.absL_active) else np.zeros(0, complex)
        self._build_super()

    # -- hierarchy super-operators: eom/eom_hops_ksuper.py:575-668, :36-230 --
    def _build_super(self):
        n, nh = self.n_state, self.n_hier
        W = self.sysd["W"]
        K0 = np.zeros(nh, dtype=np.complex128)
        for i, a in enumerate(self.aux_list):
            if a:
                arr = np.array(a, dtype=int)
                K0[i] = -np.dot(arr[:, 1], W[arr[:, 0]])
        self.K0 = K0
        kr, kc, kd = [], [], []
        mr, md = [], []
        zr = [[] for _ in self.absL]
        zc = [[] for _ in self.absL]
        zd = [[] for _ in self.absL]
        # connections (lower aux, upper aux, depth of upper in the mode)
        self.conn = []
        for lm, m in enumerate(self.modes):
            lows, ups, vals = [], [], []
            for i, a in enumerate(self.aux_list):
                up = aux_step(a, int(m), 1)
                j = self.aux_index.get(up)
                if j is not None:
                    lows.append(i)
                    ups.append(j)
                    vals.append(aux_dict(up)[int(m)])
                    self.conn.append((i, j, lm, vals[-1]))
            if not lows:
                continue
            il = self.relL_by_mode[lm]
            L = self.L_coo[il]
            gw = self.g[lm] / self.w[lm]
            lows = np.array(lows)
            ups = np.array(ups)
            nd = len(L.data)
            kd.extend(np.repeat(np.ones(len(lows)) * -gw, nd) * np.tile(L.data, len(lows)))
            kr.extend(np.repeat(lows, nd) * n + np.tile(L.row, len(lows)))
            kc.extend(np.repeat(ups, nd) * n + np.tile(L.col, len(lows)))
            md.extend(np.repeat(np.array(vals) * self.w[lm], nd) * np.tile(np.conj(L.data), len(lows)))
            zd[il].extend(np.ones(len(lows)) * gw)
            zr[il].extend(lows)
            zc[il].extend(ups)
        N = n * nh
        self.K2_k = sp.coo_matrix((K0, (np.arange(nh), np.arange(nh))), shape=(nh, nh),
                                  dtype=np.complex128).tocsr()
        self.K2_k.eliminate_zeros()
        self.K2_kp1 = sp.coo_matrix((kd, (kr, kc)), shape=(N, N), dtype=np.complex128).tocsr()
        self.K2_km1 = sp.coo_matrix((md, (kc, kr)), shape=(N, N), dtype=np.complex128).tocsr()
        self.Z2_kp1 = [sp.coo_matrix((zd[i], (zr[i], zc[i])), shape=(nh, nh),
                                     dtype=np.complex128).tocsr() for i in range(len(self.absL))]
        self.K2_stable = self.K2_kp1 + self.K2_km1
        # first-level auxiliaries e_m used by the normalisation correction
        # (eom/hops_eom.py:178-193)
        act = list(self.relL_active)
        self.first_level = []
        for lm, m in enumerate(self.modes):
            j = self.aux_index.get(((int(m), 1),))
            if j is None:
                continue
            il = self.relL_by_mode[lm]
            if il in act:
                self.first_level.append((j, act.index(il), lm))


# ---------------------------------------------------------------------------
# Derivative (eom/hops_eom.py:118-552, eom/eom_functions.py:11-309)
# ---------------------------------------------------------------------------
def expectation(op, vec):
    """<vec|op|vec>/<vec|vec>  (eom/eom_functions.py:11-29)."""
    return (np.conj(vec) @ (op @ vec)) / (np.conj(vec) @ vec)


def _dense(op):
    return op.toarray() if sp.issparse(op) else np.asarray(op)


def lt_corr(lt, L_list, avg, Lsq_list):
    """Low-temperature correction operators of the nonlinear equations
    (eom/eom_functions.py:215-256): the physical-wave-function term
    sum_n (c_n <L_n> L_n - c_n L_n^2) and the all-auxiliary term
    sum_n conj(c_n) <L_n> L_n, summed in list order.  Operators may be dense
    or sparse; with no operators both terms are the scalar 0."""
    c_phys, c_hier = 0, 0
    for c, L, a, L2 in zip(lt, L_list, avg, Lsq_list):
        c_phys = c_phys + ((c * a) * L - c * L2)
        c_hier = c_hier + (np.conj(c) * a) * L
    return c_phys, c_hier


def lt_corr_to_norm_corr(lt, avg, avg_sq):
    """Low-temperature part of the normalisation correction,
    sum_n Re(c_n) (2 <L_n>^2 - <L_n^2>)  (eom/eom_functions.py:258-284)."""
    if len(lt) == 0:
        return 0.0
    terms = np.real(np.asarray(lt)) * (2 * np.asarray(avg) ** 2 - np.asarray(avg_sq))
    return np.sum(terms)


def lt_corr_linear(lt, Lsq_list):
    """Low-temperature correction of the linear equation, -sum_n c_n L_n^2,
    for the physical wave function only (eom/eom_functions.py:286-309).  The
    coefficients and squared operators must pair up one to one: the reference
    raises on a length mismatch and so does this."""
    if len(lt) != len(Lsq_list):
        raise ValueError("%d correction coefficients for %d operators" % (len(lt), len(Lsq_list)))
    if len(lt) == 0:
        return 0
    acc = lt[0] * Lsq_list[0]
    for c, L2 in zip(lt[1:], Lsq_list[1:]):
        acc = acc + c * L2
    return -acc


def make_derivative(basis, eom_type):
    """Returns dsystem_dt(phi, z_mem, z1, z2) -> (hbar dphi/dt, hbar dz_mem/dt).

    z_mem is a dense complex vector over the absolute hierarchy modes; z1 and
    z2 are the noise values of the L-operators of the basis in relative order
    (mode.list_absindex_L2).  Restates the closures at eom/hops_eom.py:196-441
    (NONLINEAR / NORMALIZED NONLINEAR) and :445-542 (LINEAR)."""
    b = basis
    n, nh = b.n_state, b.n_hier
    H = b.H
    act = list(b.relL_active)
    normalized = eom_type == "NORMALIZED NONLINEAR"
    sysd = b.sysd

    def rows_apply(op, phi):
        # op (n x n) applied to every auxiliary block of phi
        return (op @ phi.reshape(nh, n).T).T.reshape(-1)

    if eom_type == "LINEAR":
        def dsystem_dt(phi, z_mem, z1, z2):
            d = b.K2_stable @ phi
            d += (b.K0[:, None] * phi.reshape(nh, n)).reshape(-1)
            d += rows_apply(-1j * H, phi)
            if np.any(b.lt_active):
                d[:n] += lt_corr_linear(list(b.lt_active), list(b.L_sq_csr)) @ phi[:n]
            zhat = np.conj(z1)
            for j in range(len(b.L_csr)):
                d += (zhat[j] - 1j * z2[j]) * rows_apply(b.L_csr[j], phi)
            return d, 0 * z_mem
        return dsystem_dt

    if eom_type not in ("NONLINEAR", "NORMALIZED NONLINEAR"):
        raise ValueError(eom_type)

    def dsystem_dt(phi, z_mem, z1, z2):
        # z-hat: conj(noise) + memory compressed by L-operator (eom_functions.py:32-62)
        comp = np.zeros(len(b.absL), dtype=np.complex128)
        for i, m in enumerate(b.modes):
            comp[b.relL_by_mode[i]] += z_mem[m]
        zhat = (np.conj(z1) + comp)[act]
        ztmp2 = np.asarray(z2)[act]
        psi = phi[:n]
        avg = [expectation(b.L_csr[j], psi) for j in act]
        norm_corr = 0
        if normalized:
            # eom_functions.py:163-213
            delta = np.dot(zhat, avg)
            for (i_aux, l_ind, lm) in b.first_level:
                phi1 = (b.g[lm] / b.w[lm]) * phi[n * i_aux:n * (i_aux + 1)]
                delta -= np.conj(psi) @ (b.L_csr[act[l_ind]] @ phi1)
                delta += (np.conj(psi) @ phi1) * avg[l_ind]
            norm_corr = np.real(delta)
        lt = b.lt_active
        has_lt = bool(np.any(lt))
        if has_lt:
            avg_sq = [expectation(b.L_sq_csr[j], psi) for j in act]
            L_act = [b.L_csr[j] for j in act]
            Lsq_act = [b.L_sq_csr[j] for j in act]
            c_phys, c_hier = lt_corr(lt, L_act, avg, Lsq_act)
            lt_norm = lt_corr_to_norm_corr(lt, avg, avg_sq)
        d = b.K2_stable @ phi
        d += (b.K0[:, None] * phi.reshape(nh, n)).reshape(-1)
        d += rows_apply(-1j * H, phi)
        if has_lt:
            d += rows_apply(c_hier, phi)
            d[:n] += c_phys @ phi[:n]
            norm_corr += lt_norm
        if normalized:
            d -= norm_corr * phi
        dv = d.reshape(nh, n)
        pv = phi.reshape(nh, n)
        for k, j in enumerate(act):
            dv += ((zhat[k] - 1j * ztmp2[k]) * (b.L_csr[j] @ pv.T)).T
            dv += np.conj(avg[k]) * (b.Z2_kp1[j] @ pv)
        # memory drift (eom_functions.py:65-160)
        dz = np.zeros(len(z_mem), dtype=np.complex128)
        absL_by_hmode = sysd["LIST_INDEX_L2_BY_HMODE"]
        act_abs = list(b.absL_active)
        for m in b.modes:
            la = absL_by_hmode[m]
            l_avg = avg[act_abs.index(la)] if la in act_abs else 0
            dz[m] = l_avg * np.conj(sysd["G"][m]) - np.conj(sysd["W"][m]) * z_mem[m]
        in_basis = set(int(m) for m in b.modes)
        for m in np.where(z_mem > PRECISION)[0]:
            if int(m) not in in_basis:
                dz[m] -= np.conj(sysd["W"][m]) * z_mem[m]
        return d, dz

    return dsystem_dt


# ---------------------------------------------------------------------------
# RK4 (integrator/integrator_rk.py:10-75)
# ---------------------------------------------------------------------------
def rk4_step(dsystem_dt, phi, z_mem, z1, z2, tau):
    """z1, z2: noise at [t, t+tau/2, t+tau] (columns); stage noise map [0,1,1,2]."""
    c = [0.0, 0.5, 0.5, 1.0]
    iz = [0, 1, 1, 2]
    k = [None] * 4
    kz = [None] * 4
    for i in range(4):
        if i == 0:
            pt, zt = phi.copy(), z_mem.copy()
        else:
            zt = z_mem + c[i] * kz[i - 1] * tau / HBAR
            pt = phi + c[i] * k[i - 1] * tau / HBAR
        k[i], kz[i] = dsystem_dt(pt, zt, z1[:, iz[i]], z2[:, iz[i]])
    phi = phi + tau / HBAR * (k[0] + 2.0 * k[1] + 2.0 * k[2] + k[3]) / 6.0
    z_mem = z_mem + tau / HBAR * (kz[0] + 2.0 * kz[1] + 2.0 * kz[2] + kz[3]) / 6.0
    return phi, z_mem


# ---------------------------------------------------------------------------
# Noise (noise/hops_noise.py, noise/prepare_functions.py)
# ---------------------------------------------------------------------------
NOISE_DEFAULTS = {
    "SEED": None, "MODEL": "FFT_FILTER", "TLEN": 1000.0, "TAU": 1.0,
    "INTERPOLATE": False, "RAND_MODEL": "SUM_GAUSSIAN", "STORE_RAW_NOISE": False,
    "NOISE_WINDOW": None, "ADAPTIVE": False, "FLAG_REAL": False,
}


def noise_time_axis(tlen, tau):
    """hops_noise.py:161-163."""
    nstep = int(np.ceil(tlen / tau)) + 1
    return np.arange(nstep) * tau


def raw_index(ntaus):
    """Assignment of raw draws to re/im (or radius/phase) slots so that a
    prefix is independent of the time-axis length (hops_noise.py:592-628)."""
    m1 = ntaus - 1
    a = [0]
    a.extend(np.arange(2, 4 * (ntaus - 2), 4))
    a.append(4 * m1 - 2)
    a.extend(np.arange(4 * (ntaus - 2), 3, -4))
    b = [1]
    b.extend(np.arange(3, 4 * (ntaus - 1) - 1, 4))
    b.append(4 * m1 - 1)
    b.extend(np.arange(4 * (m1 - 1) + 1, 4, -4))
    return a, b


def raw_draws(seed, lops, ntaus, rand_model):
    """Per-L stream PCG64(seed).jumped(lop) (hops_noise.py:558-590)."""
    out = np.zeros((len(lops), 4 * (ntaus - 1)))
    for i, lop in enumerate(lops):
        rng = np.random.default_rng(np.random.PCG64(seed=seed).jumped(int(lop)))
        if rand_model == "BOX_MULLER":
            out[i, :] = rng.random(size=4 * (ntaus - 1))
        elif rand_model == "SUM_GAUSSIAN":
            out[i, :] = rng.normal(loc=0, size=4 * (ntaus - 1))
        else:
            raise ValueError(rand_model)
    return out


def uncorrelated_noise(seed, lops, ntaus, rand_model):
    """Complex64 white noise of shape [n_lop, 2(ntaus-1)] (hops_noise.py:453-556)."""
    r = raw_draws(seed, lops, ntaus, rand_model)
    ia, ib = raw_index(ntaus)
    if rand_model == "BOX_MULLER":
        g, ph = r[:, np.array(ia)], r[:, np.array(ib)]
        return np.complex64(np.sqrt(-2.0 * np.log(g)) * np.exp(2.0j * np.pi * ph))
    return np.complex64(r[:, np.array(ia)] + 1j * r[:, np.array(ib)])


def correlated_noise(c_t, z_t):
    """Circulant-embedding FFT filter (Percival 2006), in the dtypes the
    reference uses (hops_noise.py:630-716)."""
    n = len(c_t[0, :])
    ang = np.array(np.angle(c_t), dtype=np.float64)
    nu = ang[:, -1] / ((n - 1.0) * 2.0 * np.pi)
    E2 = np.exp(-1.0j * 2.0 * np.pi * np.outer(nu, np.arange(n)))
    tc = np.abs(c_t) * np.exp(1.0j * ang) * E2
    s_w = np.zeros([len(tc[:, 0]), 2 * n - 2])
    for l in range(len(tc[:, 0])):
        s_w[l, :] = np.real(np.fft.fft(np.concatenate([tc[l, :], np.conj(np.flip(tc[l, 1:-1]))])))
    if np.min(s_w) < 0:
        s_w[np.where(s_w < 0)] = 0
    z_w = np.zeros([len(z_t[:, 0]), len(z_t[0, :])], dtype=np.complex128)
    for l in range(len(z_t[:, 0])):
        z_w[l, :] = np.fft.fft(z_t[l, :])
    ty = np.zeros_like(E2)
    for l in range(len(z_w[:, 0])):
        ty[l, :] = np.fft.ifft(np.abs(z_w[l, :]) * np.exp(1.0j * np.angle(z_w[l, :]))
                               * np.sqrt(s_w[l, :] / 2.0))[:n]
    return np.abs(ty) * np.exp(1.0j * np.angle(ty)) * np.conj(E2)


class Noise:
    """Dense (non-windowed) restatement of HopsNoise for the FFT_FILTER, ZERO
    and PRE_CALCULATED models (hops_noise.py:85-451)."""

    def __init__(self, noise_param, sysd, noise_type=1):
        p = copy.deepcopy(NOISE_DEFAULTS)
        p.update(noise_param)
        tag = str(noise_type)
        if noise_type == 1 or p["MODEL"] == "ZERO":
            tag = "1"
            if noise_type == 1:
                p["FLAG_REAL"] = False
        elif "FLAG_REAL" not in noise_param:
            p["FLAG_REAL"] = True
        self.p = p
        self.t_axis = noise_time_axis(p["TLEN"], p["TAU"])
        self.n_l2 = len(set(l_signature(L) for L in sysd["L_NOISE" + tag]))
        self.corr = sysd["ALPHA_NOISE" + tag]
        self.corr_param = sysd["PARAM_NOISE" + tag]
        self.nmode_by_lind = sysd["LIST_NMODE%s_BY_INDEX_L2" % tag]
        self._noise = None

    def alpha(self, lops):
        """hops_noise.py:167-192 (complex128; the caller casts)."""
        a = np.zeros([len(lops), len(self.t_axis)], dtype=np.complex128)
        for r, l in enumerate(sorted(set(lops))):
            for i in self.nmode_by_lind[l]:
                a[r, :] += self.corr(self.t_axis, *self.corr_param[i])
        return a

    def prepare(self):
        p = self.p
        nt = len(self.t_axis)
        lops = list(range(self.n_l2))
        if p["MODEL"] == "ZERO":
            self._noise = np.zeros((self.n_l2, nt), dtype=np.complex64)
        elif p["MODEL"] == "FFT_FILTER":
            seed = p["SEED"]
            if isinstance(seed, np.ndarray):
                z_unc = seed
            else:
                z_unc = uncorrelated_noise(seed, lops, nt, p["RAND_MODEL"])
            self.z_unc = z_unc
            a = np.complex64(self.alpha(lops))
            corr = correlated_noise(a, z_unc)
            self._noise = np.complex64(corr)
            if p["INTERPOLATE"]:  # hops_noise.py:242-243: cubic spline of the unrounded noise
                from scipy.interpolate import interp1d
                self._spline = interp1d(self.t_axis, corr, kind="cubic", axis=1)
        elif p["MODEL"] == "PRE_CALCULATED":
            self._noise = np.asarray(p["SEED"])  # used as given (hops_noise.py:255-263)
            if p["INTERPOLATE"]:
                from scipy.interpolate import interp1d
                self._spline = interp1d(self.t_axis, p["SEED"], kind="cubic", axis=1)
        else:
            raise ValueError(p["MODEL"])

    @property
    def noise(self):
        if self._noise is None:
            self.prepare()
        return self._noise

    def get_noise(self, t_list, lops):
        """hops_noise.py:339-451: grid lookup, or the cubic spline with
        INTERPOLATE (:379-385)."""
        if self.p["MODEL"] == "ZERO":
            return np.zeros([len(lops), len(t_list)], dtype=np.complex64)
        Z = self.noise
        if self.p["INTERPOLATE"]:
            vals = self._spline(np.asarray(t_list, dtype=float))[list(lops), :]
            return np.real(vals) if self.p["FLAG_REAL"] else vals
        it = []
        for t in t_list:
            hit = np.abs(self.t_axis - t) < PRECISION
            if np.sum(hit) != 1:
                raise ValueError("off-grid noise time %r" % t)
            it.append(int(np.where(hit)[0][0]))
        out = np.zeros((len(lops), len(t_list)), dtype=Z.dtype)
        for i, l in enumerate(lops):
            for j, t in enumerate(it):
                out[i, j] = Z[l, t]
        if self.p["FLAG_REAL"]:
            return np.real(out)
        return out


def effective_noise(noise, t, tau, lops):
    """Moving averages of the noise over the three RK sub-intervals
    [t, t+tau/2), [t+tau/2, t+tau), [t+tau, t+3tau/2)
    (integrator/integrator_rk.py:115-128)."""
    r = round(tau / noise.p["TAU"])
    raw = noise.get_noise([t + (i / r) * tau for i in range(round(r * 1.5))], lops)
    return np.array([np.mean(raw[:, :round(r / 2)], axis=1), np.mean(raw[:, round(r / 2):r], axis=1),
                     np.mean(raw[:, r:], axis=1)]).T


# ---------------------------------------------------------------------------
# Trajectory driver, non-adaptive (trajectory/hops_trajectory.py:253-538,616-634)
# ---------------------------------------------------------------------------
class Trajectory:
    """Plain (non-adaptive) HOPS trajectory restated from HopsTrajectory."""

    def __init__(self, system_param, eom_param=None, noise_param=None, noise2_param=None,
                 hierarchy_param=None, effective_noise_integration=False):
        eom_param = dict(eom_param or {})
        hierarchy_param = dict(hierarchy_param or {})
        self.effective = bool(effective_noise_integration)
        self.eom_type = eom_param.get("EQUATION_OF_MOTION", "NORMALIZED NONLINEAR")
        self.maxhier = min(int(hierarchy_param.get("MAXHIER", 3)), 255)
        self.sysd = derive_system(system_param)
        self.noise_param = dict(noise_param)
        if noise2_param is None:
            noise2_param = {"TLEN": noise_param["TLEN"], "TAU": noise_param["TAU"],
                            "MODEL": "ZERO", "SEED": None}
        self.noise1 = Noise(noise_param, self.sysd, 1)
        self.noise2 = Noise(noise2_param, self.sysd, 2)
        self.t = 0.0
        self.psi_traj = []
        self.t_axis = []
        self.z_mem_traj = []

    def initialize(self, psi_0):
        sysd = self.sysd
        aux = triangular_hierarchy(sysd["N_HMODES"], self.maxhier)
        self.basis = Basis(sysd, np.arange(sysd["NSTATES"]), aux)
        self.dsystem_dt = make_derivative(self.basis, self.eom_type)
        b = self.basis
        phi = np.zeros(b.n_hier * b.n_state, dtype=np.complex128)
        phi[:b.n_state] = np.asarray(psi_0, dtype=np.complex128)[b.state_list]
        self.phi = phi
        self.z_mem = np.zeros(len(sysd["L_NOISE1"]), dtype=np.complex128)
        self._store(0.0)
        self.t = 0.0

    def _store(self, t):
        self.psi_traj.append(self.phi[:self.basis.n_state].copy())
        self.t_axis.append(t)
        self.z_mem_traj.append(self.z_mem.copy())

    def operator(self, op):
        """HopsTrajectory._operator (trajectory/hops_trajectory.py:540-577):
        populated states join the basis (state adaptivity), op restricted to
        the state basis multiplies the (n_state x n_hier) hierarchy matrix, and
        the early-time integrator is reset; the reference's define_basis /
        update_basis pair after the product leaves the basis unchanged."""
        import scipy.sparse as sps
        if sps.issparse(op):
            op = op.tocsr()
        adaptive_s = getattr(self, "ad", None) is not None and self.ad.adaptive_s
        if adaptive_s:
            sl = [int(s) for s in self.basis.state_list]
            new = list(sl) + list(np.nonzero(op[:, sl])[0])
            self.phi = self.update_basis(self.phi, list(set(int(s) for s in new)), list(self.basis.aux_list))
        sl = [int(s) for s in self.basis.state_list]
        opr = op[np.ix_(sl, sl)]
        phi_mat = np.reshape(self.phi, [self.basis.n_state, self.basis.n_hier], order="F")
        self.phi = np.reshape(opr @ phi_mat, len(self.phi), order="F")
        if adaptive_s:
            self.early_counter = 0

    def noise_block(self, t, tau):
        lops = list(self.basis.absL)
        if self.effective:
            return (effective_noise(self.noise1, t, tau, lops), effective_noise(self.noise2, t, tau, lops))
        z1 = self.noise1.get_noise([t, t + tau * 0.5, t + tau], lops)
        z2 = self.noise2.get_noise([t, t + tau * 0.5, t + tau], lops)
        return z1, z2

    def propagate(self, t_advance, tau):
        t0 = self.t
        t_axis = t0 + np.arange(1, 1 + int(np.ceil(t_advance / tau))) * tau
        if (t0 + t_advance + tau) > self.noise1.p["TLEN"]:
            raise ValueError("trajectory longer than TLEN")
        for t in t_axis:
            z1, z2 = self.noise_block(self.t, tau)
            phi, z_mem = rk4_step(self.dsystem_dt, self.phi, self.z_mem, z1, z2, tau)
            if self.eom_type == "NORMALIZED NONLINEAR":
                phi = phi / np.sqrt(np.sum(np.abs(phi[:self.basis.n_state]) ** 2))
            self.phi = phi
            # the stored z_mem lags one step (hops_trajectory.py:526-532)
            self.psi_traj.append(self.phi[:self.basis.n_state].copy())
            self.t_axis.append(float(t))
            self.z_mem_traj.append(self.z_mem.copy())
            self.z_mem = z_mem
            self.t = float(t)

    @property
    def psi(self):
        return self.phi[:self.basis.n_state]


# ---------------------------------------------------------------------------
# Adaptive basis (basis/hops_basis.py:117-292, :294-998, :1000-1151;
# basis/basis_functions_adaptive.py; basis/basis_functions.py;
# basis/hops_fluxfilters.py).  General L-operators: the diagonal parts enter
# through M2_diag [mode, state], the off-diagonal parts through one
# [mode, state] table per destination state, plus the noise and
# low-temperature terms that off-diagonal operators add to the state flux
# (hops_basis.py:914-998).  Arrays are dense; Phi is aux-major
# (row k*n_state + s), so C2 = Phi.reshape(n_hier, n_state).T is [s, k].
# ---------------------------------------------------------------------------
def determine_error_thresh(sorted_error, max_error, offset=0.0):
    """basis/basis_functions.py:8-41."""
    thresh = 0.0
    if len(sorted_error) > 0:
        idx = np.argmax(np.cumsum(sorted_error) + offset >= max_error)
        if idx > 0:
            thresh = sorted_error[idx - 1]
        elif sorted_error[0] + offset < max_error:
            thresh = np.inf
    return thresh


def basis_from_list(error, max_error, members):
    """hops_basis.py:1000-1028."""
    thresh = determine_error_thresh(np.sort(error), max_error)
    idx = np.where(error > thresh)[0]
    return idx, [members[i] for i in idx]


def _aux_depth(aux):
    return sum(v for (m, v) in aux)


def _C2(phi, n_state, n_hier):
    return np.asarray(phi).reshape(n_hier, n_state).T


def _abs2(x):
    x = x.toarray() if sp.issparse(x) else np.asarray(x)
    return np.abs(x) ** 2


def _off_super(M2_off, shape):
    """sum over destinations d of |M_d|^2."""
    out = np.zeros(shape)
    for M in M2_off:
        out = out + _abs2(M)
    return out


def mode_from_state_tables(L_by_mode, states, dests):
    """(M2_diag, [M_d for d in dests]) for the L-operators of the modes given
    on the full state space: M2_diag[m, i] = L_m[s_i, s_i] over the basis
    states s_i, and M_d[m, i] = L_m[d, s_i] for s_i != d
    (hops_basis.py:1030-1100)."""
    L_by_mode = [_dense(L) for L in L_by_mode]
    nm, n = len(L_by_mode), len(states)
    diag = np.zeros((nm, n), complex)
    for m, L in enumerate(L_by_mode):
        for i, s in enumerate(states):
            diag[m, i] = L[s, s]
    off = []
    for d in dests:
        M = np.zeros((nm, n), complex)
        for m, L in enumerate(L_by_mode):
            for i, s in enumerate(states):
                if s != d:
                    M[m, i] = L[d, s]
        off.append(M)
    return diag, off


def expectation_table(avg_by_mode, n_state):
    """X2[m, s] = <L_m> for every basis state s (hops_basis.py:1118-1129)."""
    return np.repeat(np.asarray(avg_by_mode, dtype=complex).reshape(-1, 1), n_state, axis=1)


def error_deriv(dsystem_dt, phi, z_step, n_state, n_hier, dt, stable_idx=None):
    """|dPhi/dt / hbar + Phi / dt|^2 as [s, k]: the deletion flux, optionally
    out of the stable auxiliaries only (basis_functions_adaptive.py:10-74)."""
    if stable_idx is not None:
        P = np.zeros(n_state * n_hier, complex)
        Pv = P.reshape(n_hier, n_state)
        Pv[stable_idx] = np.asarray(phi).reshape(n_hier, n_state)[stable_idx]
    else:
        P = np.asarray(phi)
    d = dsystem_dt(P, z_step[2], z_step[0], z_step[1])[0] / HBAR
    d = d + P / dt
    D = _C2(d, n_state, n_hier)
    if stable_idx is not None:
        D = D[:, stable_idx]
    return np.abs(D) ** 2


def error_sflux_hier(phi, s0, sc, n_state, n_hier, H2, T_phys=None, T_hier=None):
    """Per auxiliary k: sum_d |((H + T_k)[d, s] @ psi_k)[d]|^2 / hbar^2 over the
    states d outside the basis; H2 is the full-space (-iH + noise) operator and
    T_k the low-temperature operator (T_phys for k = 0, T_hier otherwise)
    (basis_functions_adaptive.py:76-163)."""
    C2 = _C2(phi, n_state, n_hier)
    s0, sc = list(s0), list(sc)
    H2 = _dense(H2)
    if T_phys is None:
        Hsc = H2[np.ix_(sc, s0)]
        return np.sum(np.abs(Hsc @ C2 / HBAR) ** 2, axis=0)
    Hp = (H2 + _dense(T_phys))[np.ix_(sc, s0)]
    Hh = (H2 + _dense(T_hier))[np.ix_(sc, s0)]
    Caux = C2.copy()
    Caux[:, 0] = 0
    out = np.sum(np.abs(Hh @ Caux / HBAR) ** 2, axis=0)
    out[0] = np.sum(np.abs(Hp @ C2[:, :1] / HBAR) ** 2)
    return out


def error_flux_up_hier_stable(phi, n_state, n_hier, n_modes, w, K2, M2_diag, M2_off):
    """E[m, k] = |w_m (1 + k_m)|^2 (sum_d |M_d|^2 @ |Phi|^2)[m, k] / hbar^2,
    the diagonal part counted as one more destination
    (basis_functions_adaptive.py:165-266)."""
    KW = np.abs(np.asarray(w)).reshape(n_modes, 1) * (1 + np.asarray(K2))
    Msup = _abs2(M2_diag)
    if len(M2_off):
        Msup = Msup + _off_super(M2_off, Msup.shape)
    return np.abs(KW) ** 2 * (Msup @ (np.abs(_C2(phi, n_state, n_hier)) ** 2)) / HBAR ** 2


def error_flux_up_state_stable(phi, n_state, n_hier, n_modes, w, K2, M2_diag, M2_off,
                               F_diag=None, F_off=None):
    """E[s, k] = (|M|^2.T @ |F * w (1 + k)|^2)[s, k] |Phi[s, k]|^2 / hbar^2, for
    the diagonal table with filter F_diag plus the off-diagonal tables with
    F_off (basis_functions_adaptive.py:268-384)."""
    KW = np.abs(np.asarray(w)).reshape(n_modes, 1) * (1 + np.asarray(K2))
    C2sq = np.abs(_C2(phi, n_state, n_hier)) ** 2
    Fd = 1 if F_diag is None else F_diag
    E = (_abs2(M2_diag).T @ (np.abs(Fd * KW) ** 2)) * C2sq / HBAR ** 2
    Fo = 1 if F_off is None else F_off
    Moff = _off_super(M2_off, np.shape(M2_diag))
    if np.sum(Moff) > 0:
        E = E + (Moff.T @ (np.abs(Fo * KW) ** 2)) * C2sq / HBAR ** 2
    return E


def error_flux_up_by_dest_state(phi, n_state, n_hier, n_modes, w, K2, M2_off_dest,
                                stable_idx=None, F_off=None, stable_states=None):
    """Total off-diagonal flux up into each destination state, out of the
    stable auxiliaries and stable states (basis_functions_adaptive.py:386-515)."""
    KW = np.abs(np.asarray(w)).reshape(n_modes, 1) * (1 + np.asarray(K2))
    C2 = _C2(phi, n_state, n_hier)
    if stable_states is not None:
        keep = np.zeros(n_state, bool)
        keep[list(stable_states)] = True
        C2 = np.where(keep[:, None], C2, 0)
    idx = list(range(n_hier)) if stable_idx is None else list(stable_idx)
    Fo = 1 if F_off is None else np.asarray(F_off)[:, idx]
    KWo = np.abs(Fo * KW[:, idx]) ** 2
    C2sq = np.abs(C2[:, idx]) ** 2
    return np.array([np.sum((_abs2(M).T @ KWo) * C2sq / HBAR ** 2) for M in M2_off_dest])


def error_flux_down_hier_stable(phi, n_state, n_hier, n_modes, g, w, M2_diag, M2_off, X2):
    """E[m, k] = |g_m / w_m|^2 ((|M_diag - X2|^2 + sum_d |M_d|^2) @ |Phi|^2)[m, k]
    / hbar^2 (basis_functions_adaptive.py:517-624)."""
    gw2 = (np.abs(np.asarray(g) / np.asarray(w)) ** 2).reshape(n_modes, 1)
    D = np.abs(_dense(M2_diag) - _dense(X2)) ** 2
    if len(M2_off):
        D = D + _off_super(M2_off, D.shape)
    return gw2 * (D @ (np.abs(_C2(phi, n_state, n_hier)) ** 2)) / HBAR ** 2


def error_flux_down_state_stable(phi, n_state, n_hier, n_modes, g, w, M2_diag, M2_off, X2,
                                 F_diag=None, F_off=None):
    """E[s, k] of the flux down, off-diagonal tables (filter F_off) then the
    diagonal one minus <L> (filter F_diag) (basis_functions_adaptive.py:626-758)."""
    gw2 = (np.abs(np.asarray(g) / np.asarray(w)) ** 2).reshape(n_modes, 1)
    C2sq = np.abs(_C2(phi, n_state, n_hier)) ** 2
    E = 0
    Fo = 1 if F_off is None else F_off
    Moff = _off_super(M2_off, np.shape(M2_diag))
    if np.sum(Moff) != 0:
        E = (Moff.T @ (Fo * gw2)) * C2sq / HBAR ** 2
    Fd = 1 if F_diag is None else F_diag
    D = np.abs(_dense(M2_diag) - _dense(X2)) ** 2
    return E + (D.T @ (Fd * gw2)) * C2sq / HBAR ** 2


def error_flux_down_by_dest_state(phi, n_state, n_hier, n_modes, g, w, M2_off_dest,
                                  stable_idx=None, F_off=None, stable_states=None):
    """Total off-diagonal flux down into each destination state
    (basis_functions_adaptive.py:760-891)."""
    gw2 = (np.abs(np.asarray(g) / np.asarray(w)) ** 2).reshape(n_modes, 1)
    C2sq = np.abs(_C2(phi, n_state, n_hier)) ** 2
    if stable_states is not None:
        keep = np.zeros(n_state, bool)
        keep[list(stable_states)] = True
        C2sq = np.where(keep[:, None], C2sq, 0.0)
    idx = list(range(n_hier)) if stable_idx is None else list(stable_idx)
    Fo = 1 if F_off is None else np.asarray(F_off)[:, idx]
    G = Fo * gw2
    return np.array([np.sum((_abs2(M).T @ G) * C2sq[:, idx] / HBAR ** 2) for M in M2_off_dest])


def error_sflux_stable_state(phi, n_state, n_hier, H2, stable_idx, states, T_phys=None, T_hier=None):
    """Per basis state s: sum_d |H2[d, s]|^2 (d != s, over the full space) times
    sum_k |Phi[s, k]|^2 over the stable auxiliaries, with the low-temperature
    couplings split between k = 0 and k > 0 (basis_functions_adaptive.py:893-995)."""
    C2 = _C2(phi, n_state, n_hier)[:, list(stable_idx)]
    H2 = _dense(H2)
    Hc = H2 - np.diag(np.diag(H2))
    if T_phys is None:
        V1 = np.sum(np.abs(Hc[:, list(states)]) ** 2, axis=0)
        return V1 * np.sum(np.abs(C2) ** 2, axis=1) / HBAR ** 2
    Tp, Th = _dense(T_phys), _dense(T_hier)
    Vp = np.sum(np.abs((Hc + Tp - np.diag(np.diag(Tp)))[:, list(states)]) ** 2, axis=0)
    Vh = np.sum(np.abs((Hc + Th - np.diag(np.diag(Th)))[:, list(states)]) ** 2, axis=0)
    Caux = C2.copy()
    Caux[:, 0] = 0
    return (Vp * np.abs(C2[:, 0]) ** 2 + Vh * np.sum(np.abs(Caux) ** 2, axis=1)) / HBAR ** 2


def error_sflux_boundary_state(phi, s0, sc, n_state, n_hier, H2, state_stable_idx, stable_idx,
                               sc_dest, flux_updown, T_phys=None, T_hier=None):
    """Flux from the stable states and auxiliaries into each state outside the
    basis, plus the given flux up / down into some of them; returns the states
    with nonzero flux and their errors (basis_functions_adaptive.py:997-1132)."""
    H2 = _dense(H2)
    if not len(state_stable_idx) < H2.shape[0]:
        return [], []
    C2 = _C2(phi, n_state, n_hier)[np.ix_(list(state_stable_idx), list(stable_idx))]
    s0, sc = list(s0), list(sc)
    if T_phys is None:
        E = np.sum(np.abs(H2[np.ix_(sc, s0)] @ (C2 / HBAR)) ** 2, axis=1)
    else:
        E = (np.abs((H2 + _dense(T_phys))[np.ix_(sc, s0)] @ (C2[:, :1] / HBAR)) ** 2)[:, 0]
        if C2.shape[1] > 1:
            E = E + np.sum(np.abs((H2 + _dense(T_hier))[np.ix_(sc, s0)] @ (C2[:, 1:] / HBAR)) ** 2, axis=1)
    E = np.array(E, dtype=float)
    for i, f in zip(sc_dest, flux_updown):
        E[i] += f
    nz = np.nonzero(E)[0]
    return np.array(sc, dtype=int)[nz], E[nz]


class AdaptiveState:
    """The pieces of HopsBasis used by define_basis/update_basis for one
    trajectory, built on Basis (which is rebuilt for every new basis, the
    invariant of eom/eom_hops_ksuper.py:387-572)."""

    def __init__(self, sysd, eom_type, maxhier, delta_a, delta_s, f_discard, markov=None, filters=None):
        self.markov = None if markov is None else np.array(markov, bool)
        # Triangular / LongEdge static filters: (name, bool mask over the
        # absolute modes, kmax2)
        self.filters = [(n, np.array(p[0], bool), int(p[1])) for n, p in (filters or [])]
        self.sysd = sysd
        self.eom_type = eom_type
        self.maxhier = maxhier
        self.delta_a = delta_a
        self.delta_s = delta_s
        self.f_discard = f_discard
        self.adaptive_h = delta_a > 0
        self.adaptive_s = delta_s > 0
        self.L_full = [L.toarray() for L in sysd["LIST_L2_COO"]]
        self.off_diag = np.array([bool(np.any(L.row != L.col)) for L in sysd["LIST_L2_COO"]], dtype=bool)

    # -- per-basis arrays (hops_basis.py:1030-1151) --
    def off_couplings(self, b):
        """Any off-diagonal L-operator among the basis modes
        (hops_basis.py:1223-1227)."""
        return bool(np.any(self.off_diag[b.absL]))

    def dest_states(self, b):
        """States reachable by the L-operators of the basis states
        (hops_system.py:218-226)."""
        dest = set()
        for s in b.state_list:
            dest |= set(self.sysd["LIST_DESTINATION_STATES_BY_STATE_INDEX"][int(s)])
        return sorted(dest)

    def tables(self, b):
        """M2_diag over the basis states (diagonal of each mode's reduced L) and,
        with off-diagonal couplings, the per-destination tables."""
        nm, n = len(b.modes), b.n_state
        M2 = np.zeros((nm, n), complex)
        for i in range(nm):
            L = b.L_coo[b.relL_by_mode[i]]
            for r, c, d in zip(L.row, L.col, L.data):
                if r == c:
                    M2[i, c] += d
        if not self.off_couplings(b):
            return M2, [], []
        dests = self.dest_states(b)
        L_by_mode = [self.L_full[b.absL[b.relL_by_mode[i]]] for i in range(nm)]
        _, off = mode_from_state_tables(L_by_mode, [int(s) for s in b.state_list], dests)
        return M2, off, dests

    def expectations(self, b, psi):
        return np.array([expectation(L, psi) for L in b.L_csr])

    def noise_operator(self, b, z_step):
        """Full-space noise and memory operator of the off-diagonal L-operators,
        added to -iH in the state-flux terms (hops_basis.py:914-953); None
        without off-diagonal couplings."""
        if not self.off_couplings(b):
            return None
        z1, z2, zm = z_step
        comp = np.zeros(len(b.absL), dtype=np.complex128)
        for i, m in enumerate(b.modes):
            comp[b.relL_by_mode[i]] += zm[m]
        coef = np.conj(np.asarray(z1)) - 1j * np.asarray(z2) + comp
        N = self.sysd["NSTATES"]
        Z = np.zeros((N, N), complex)
        for j, l in enumerate(b.absL):
            if self.off_diag[l]:
                Z = Z + coef[j] * self.L_full[l]
        return Z

    def ltc_operators(self, b, avg):
        """(T_phys, T_hier) of the off-diagonal L-operators with a
        low-temperature coefficient, or (None, None) (hops_basis.py:955-998):
        T_hier = sum conj(c) <L> L, T_phys = T_hier + T_hier^dagger - sum c L^2."""
        lt = np.array(self.sysd["LIST_LT_PARAM"], dtype=np.complex128)[b.absL]
        if not (np.any(lt) and self.off_couplings(b)):
            return None, None
        N = self.sysd["NSTATES"]
        Th = np.zeros((N, N), complex)
        Lsq = np.zeros((N, N), complex)
        for j, l in enumerate(b.absL):
            if self.off_diag[l]:
                Th = Th + np.conj(lt[j]) * avg[j] * self.L_full[l]
                Lsq = Lsq + lt[j] * (self.L_full[l] @ self.L_full[l])
        return Th + Th.conj().T - Lsq, Th

    def K2_aux_by_mode(self, b):
        K = np.zeros((len(b.modes), b.n_hier), dtype=np.int64)
        for j, a in enumerate(b.aux_list):
            for (m, v) in a:
                if m in b.rel_mode:
                    K[b.rel_mode[m], j] = v
        return K

    def markov_up(self, b):
        """No flux up along Markovian-filtered modes except from the physical
        wave function, and none out of their first-level auxiliaries
        (hops_fluxfilters.py:354-397)."""
        nm, nh = len(b.modes), b.n_hier
        F = np.ones((nm, nh), bool)
        if self.markov is None:
            return F
        mask = self.markov[b.modes]
        F[mask, 1:] = False
        for lm, m in enumerate(b.modes):
            if mask[lm]:
                j = b.aux_index.get(((int(m), 1),))
                if j is not None:
                    F[:, j] = False
        return F

    def up_filter(self, b):
        """Allowed flux up along (mode, aux): the Markovian mask, times the
        Triangular mask (no step up along a filtered mode from an auxiliary
        whose depth in the filtered modes is already >= kmax2), times the
        LongEdge mask (from depth >= kmax2: an auxiliary with depth in a
        filtered mode steps up only along its own mode when it is an edge
        term, one without steps up along no filtered mode); the filters of one
        kind share one mask, applied in order (hops_fluxfilters.py:313-501)."""
        F = self.markov_up(b)
        nm, nh = len(b.modes), b.n_hier
        modes = np.array(b.modes)
        for kind in ("Triangular", "LongEdge"):
            G = np.ones((nm, nh), bool)
            for name, mask_abs, kmax2 in self.filters:
                if name != kind:
                    continue
                mask = mask_abs[modes]
                filt = set(int(m) for m in modes[mask])
                for j, a in enumerate(b.aux_list):
                    if kind == "Triangular":
                        if sum(v for m, v in a if m in filt) >= kmax2:
                            G[mask, j] = False
                    elif sum(v for _, v in a) >= kmax2:
                        if any(m in filt for m, _ in a):
                            G[:, j] = False
                            if len(a) == 1:
                                G[b.rel_mode[a[0][0]], j] = True
                        else:
                            G[mask, j] = False
            F = F & G
        return F

    def state_filters(self, b, aux_set):
        """(F_down, F_up) over [mode, aux] of the current basis: F_down marks
        auxiliaries k with k - e_m in aux_set, F_up those with k + e_m in
        aux_set (hops_fluxfilters.py:292-352)."""
        nm, nh = len(b.modes), b.n_hier
        F_dn = np.zeros((nm, nh), bool)
        F_up = np.zeros((nm, nh), bool)
        for a in aux_set:
            for lm, m in enumerate(b.modes):
                j = b.aux_index.get(aux_step(a, int(m), 1))
                if j is not None:
                    F_dn[lm, j] = True
            for (m, v) in a:
                j = b.aux_index.get(aux_step(a, m, -1))
                if j is not None and m in b.rel_mode:
                    F_up[b.rel_mode[m], j] = True
        return F_dn, F_up

    def list_sc(self, b):
        """States outside the basis reachable by L or H (hops_system.py:191-204)."""
        st = set(int(s) for s in b.state_list)
        dest = set()
        for s in st:
            dest |= set(self.sysd["LIST_DESTINATION_STATES_BY_STATE_INDEX"][s])
        coupled = set()
        for s in st:
            coupled |= set(self.sysd["COUPLED_STATES"][s])
        return sorted((dest - st) | (coupled - st))

    def _H2(self, b, z_step):
        H2 = -1j * self.sysd["H_DENSE"]
        Z = self.noise_operator(b, z_step)
        return H2 if Z is None else H2 + Z

    def hier_stable_error(self, dsys, b, Phi, dt, z_step):
        """hops_basis.py:560-665."""
        n, nh, nm = b.n_state, b.n_hier, len(b.modes)
        E1 = np.sum(error_deriv(dsys, Phi, z_step, n, nh, dt), axis=0)
        sc = self.list_sc(b)
        avg = self.expectations(b, Phi[:n])
        T_phys, T_hier = self.ltc_operators(b, avg)
        if len(sc):
            E1 = E1 + error_sflux_hier(Phi, b.state_list, sc, n, nh, self._H2(b, z_step), T_phys, T_hier)
        M2, M2_off, _ = self.tables(b)
        K2 = self.K2_aux_by_mode(b)
        E_up = error_flux_up_hier_stable(Phi, n, nh, nm, b.w, K2, M2, M2_off)
        F_up = self.up_filter(b)
        for j, a in enumerate(b.aux_list):
            if _aux_depth(a) == self.maxhier:
                F_up[:, j] = False
        E1 = E1 + np.sum(F_up * E_up, axis=0)
        X2 = expectation_table(avg[b.relL_by_mode], n)
        E_down = error_flux_down_hier_stable(Phi, n, nh, nm, b.g, b.w, M2, M2_off, X2)
        F_down = K2 > 0
        E1 = E1 + np.sum(F_down * E_down, axis=0)
        return E1, [E_up, E_down]

    def boundary_hier(self, b, E_updown, stable_idx, dbound_sq):
        """hops_basis.py:667-793."""
        nm, nh = len(b.modes), b.n_hier
        E_up, E_down = E_updown[0].copy(), E_updown[1].copy()
        F_up = np.ones((nm, nh), bool)
        F_down = np.ones((nm, nh), bool)
        for j, a in enumerate(b.aux_list):
            d = dict(a)
            if _aux_depth(a) < self.maxhier:
                for lm, m in enumerate(b.modes):
                    if aux_step(a, int(m), 1) in b.aux_index:
                        F_up[lm, j] = False
            else:
                F_up[:, j] = False
            for lm, m in enumerate(b.modes):
                m = int(m)
                if m not in d or aux_step(a, m, -1) in b.aux_index:
                    F_down[lm, j] = False
        F_up &= self.up_filter(b)
        E_up = (E_up * F_up)[:, stable_idx]
        E_down = (E_down * F_down)[:, stable_idx]
        srt = np.sort(np.append(E_up[E_up != 0], E_down[E_down != 0]))
        thr = determine_error_thresh(srt, dbound_sq * self.f_discard * self.f_discard)
        discarded = np.sum(E_up[E_up <= thr]) + np.sum(E_down[E_down <= thr])
        E_up[E_up <= thr] = 0
        E_down[E_down <= thr] = 0
        acc = {}
        for col, j in enumerate(stable_idx):
            a = b.aux_list[j]
            for lm in np.nonzero(E_up[:, col])[0]:
                key = aux_step(a, int(b.modes[lm]), 1)
                acc[key] = acc.get(key, 0.0) + E_up[lm, col]
            d = dict(a)
            for lm in np.nonzero(E_down[:, col])[0]:
                m = int(b.modes[lm])
                if m in d:
                    key = aux_step(a, m, -1)
                    acc[key] = acc.get(key, 0.0) + E_down[lm, col]
        thr = determine_error_thresh(np.sort(np.array(list(acc.values()))), dbound_sq, offset=discarded)
        return [k for k, v in acc.items() if v > thr]

    def state_stable_error(self, dsys, b, Phi, dt, z_step, stable_idx, aux_bound, aux_new):
        """hops_basis.py:795-912."""
        n, nh, nm = b.n_state, b.n_hier, len(b.modes)
        E1 = np.sum(error_deriv(dsys, Phi, z_step, n, nh, dt, stable_idx), axis=1)
        avg = self.expectations(b, Phi[:n])
        T_phys, T_hier = self.ltc_operators(b, avg)
        E1 = E1 + error_sflux_stable_state(Phi, n, nh, self._H2(b, z_step), stable_idx,
                                           [int(s) for s in b.state_list], T_phys, T_hier)
        M2, M2_off, _ = self.tables(b)
        X2 = expectation_table(avg[b.relL_by_mode], n)
        F_dn, F_up = self.state_filters(b, aux_bound)
        if len(M2_off):
            F_dn_off, F_up_off = self.state_filters(b, aux_new)
        else:
            F_dn_off = F_up_off = None
        E_down = error_flux_down_state_stable(Phi, n, nh, nm, b.g, b.w, M2, M2_off, X2, F_dn, F_dn_off)
        E1 = E1 + np.sum(E_down[:, stable_idx], axis=1)
        K2 = self.K2_aux_by_mode(b)
        E_up = error_flux_up_state_stable(Phi, n, nh, nm, b.w, K2, M2, M2_off, F_up, F_up_off)
        E1 = E1 + np.sum(E_up[:, stable_idx], axis=1)
        return E1

    def state_basis(self, dsys, b, Phi, dt, z_step, stable_idx, aux_bound, aux_new):
        """hops_basis.py:294-482."""
        n, nh, nm = b.n_state, b.n_hier, len(b.modes)
        dsq = self.delta_s ** 2
        err = self.state_stable_error(dsys, b, Phi, dt, z_step, stable_idx, aux_bound, aux_new)
        rel_stable, st_stable = basis_from_list(err, dsq / 2.0, list(b.state_list))
        unstable = np.setdiff1d(np.arange(n), rel_stable)
        dbound = dsq - np.sum(err[unstable])
        sc = self.list_sc(b)
        avg = self.expectations(b, Phi[:n])
        T_phys, T_hier = self.ltc_operators(b, avg)
        H2 = self._H2(b, z_step)
        M2, M2_off, dests = self.tables(b)
        sc_dest, flux = [], []
        if len(M2_off):
            M2_sc = []
            for i, d in enumerate(dests):
                if d in sc:
                    sc_dest.append(sc.index(d))
                    M2_sc.append(M2_off[i])
            F_dn_off, F_up_off = self.state_filters(b, aux_new)
            K2 = self.K2_aux_by_mode(b)
            E_dn = error_flux_down_by_dest_state(Phi, n, nh, nm, b.g, b.w, M2_sc, stable_idx, F_dn_off, rel_stable)
            E_up = error_flux_up_by_dest_state(Phi, n, nh, nm, b.w, K2, M2_sc, stable_idx, F_up_off, rel_stable)
            flux = list(E_up + E_dn)
        idx, E = error_sflux_boundary_state(Phi, [int(s) for s in st_stable], sc, n, nh, H2, rel_stable,
                                            stable_idx, sc_dest, flux, T_phys, T_hier)
        bound = []
        if len(E):
            _, bound = basis_from_list(np.asarray(E), dbound, [int(s) for s in idx])
        bound = sorted(set(bound) - set(int(s) for s in b.state_list))
        return [int(s) for s in st_stable], bound

    def hierarchy_basis(self, dsys, b, Pn, dt, z_step):
        """Stable and boundary auxiliaries of a normalised Pn
        (hops_basis.py:484-558)."""
        dsq = self.delta_a ** 2
        err, E_updown = self.hier_stable_error(dsys, b, Pn, dt, z_step)
        idx_stable, aux_stable = basis_from_list(err, dsq / 2.0, b.aux_list)
        unstable = np.setdiff1d(np.arange(b.n_hier), idx_stable)
        dbound = dsq - np.sum(err[unstable])
        aux_bound = self.boundary_hier(b, E_updown, idx_stable, dbound)
        return aux_stable, aux_bound

    def define_basis(self, dsys, b, Phi, dt, z_step):
        """hops_basis.py:117-182."""
        n = b.n_state
        Pn = Phi / np.linalg.norm(Phi[:n])
        if self.adaptive_h:
            aux_stable, aux_bound = self.hierarchy_basis(dsys, b, Pn, dt, z_step)
            aux_new = list(set(aux_stable) | set(aux_bound))
            stable_idx = sorted(b.aux_index[a] for a in aux_stable)
        else:
            aux_new = list(b.aux_list)
            aux_bound = []
            stable_idx = list(range(b.n_hier))
        if self.adaptive_s:
            st, sb = self.state_basis(dsys, b, Pn, dt, z_step, stable_idx, aux_bound, aux_new)
            states_new = sorted(set(st) | set(sb))
        else:
            states_new = [int(s) for s in b.state_list]
        return states_new, aux_new


def permute_phi(b_old, b_new, Phi):
    """Copy the entries of the stable (aux, state) pairs into the new basis and
    rescale to the old physical norm (hops_basis.py:214-274)."""
    n_old, n_new = b_old.n_state, b_new.n_state
    old_pos = {int(s): i for i, s in enumerate(b_old.state_list)}
    new_pos = {int(s): i for i, s in enumerate(b_new.state_list)}
    stable_states = sorted(set(old_pos) & set(new_pos))
    out = np.zeros(b_new.n_hier * n_new, complex)
    for a in b_old.aux_list:
        j = b_new.aux_index.get(a)
        if j is None:
            continue
        i = b_old.aux_index[a]
        for s in stable_states:
            out[j * n_new + new_pos[s]] = Phi[i * n_old + old_pos[s]]
    norm_old = np.linalg.norm(Phi[:n_old])
    return norm_old * out / np.linalg.norm(out[:n_new])


class AdaptiveTrajectory(Trajectory):
    """adHOPS trajectory (trajectory/hops_trajectory.py:253-336, :338-399,
    :456-533, :636-714) for site-projector L-operators."""

    def __init__(self, *args, delta_a=1e-4, delta_s=1e-4, update_step=1, f_discard=0.01,
                 early_steps=5, inchworm_cap=5, markov=None, early_integrator="INCH_WORM", static_basis=None,
                 filters=None, **kw):
        super().__init__(*args, **kw)
        self.ad = AdaptiveState(self.sysd, self.eom_type, self.maxhier, delta_a, delta_s, f_discard,
                                markov=markov, filters=filters)
        if filters and not self.ad.adaptive_h:
            raise NotImplementedError("Triangular / LongEdge filters need an adaptive hierarchy here")
        self.update_step = update_step
        self.early_steps = early_steps
        self.inchworm_cap = inchworm_cap
        self.early_integrator = early_integrator  # INCH_WORM, STATIC_STATE_INCHWORM_HIERARCHY or STATIC
        self.static_basis = static_basis          # [states, aux] (hops_trajectory.py:303-317)
        self.early_counter = 0
        self.state_traj = []
        self.aux_traj = []

    def _basis(self, states, aux):
        b = Basis(self.sysd, states, aux)
        return b, make_derivative(b, self.eom_type)

    def z_step(self, z_mem):
        lops = list(self.basis.absL)
        z1 = self.noise1.get_noise([self.t], lops)[:, 0]
        z2 = self.noise2.get_noise([self.t], lops)[:, 0]
        return [z1, z2, z_mem]

    def update_basis(self, phi, states, aux):
        b_old = self.basis
        if set(states) == set(int(s) for s in b_old.state_list) and set(aux) == set(b_old.aux_list):
            return phi
        b_new, d = self._basis(states, aux)
        phi = permute_phi(b_old, b_new, phi)
        self.basis, self.dsystem_dt = b_new, d
        return phi

    def initialize(self, psi_0):
        psi_0 = np.asarray(psi_0, dtype=np.complex128)
        states = [int(s) for s in np.where(np.abs(psi_0) > 0)[0]]
        if not self.ad.adaptive_s:
            states = list(range(self.sysd["NSTATES"]))
        if self.ad.adaptive_h:
            aux = [()]
        else:
            aux = triangular_hierarchy(self.sysd["N_HMODES"], self.maxhier)
            if self.ad.markov is not None:
                # Markovian static filter (basis/hierarchy_functions.py:95-140)
                mk = self.ad.markov
                aux = [a for a in aux if sum(v for _, v in a) <= 1 or not any(mk[m] for m, _ in a)]
        self.basis, self.dsystem_dt = self._basis(states, aux)
        b = self.basis
        phi = np.zeros(b.n_hier * b.n_state, complex)
        phi[:b.n_state] = psi_0[b.state_list]
        self.z_mem = np.zeros(len(self.sysd["L_NOISE1"]), complex)
        self.t = 0.0
        if self.static_basis is None:
            st, ax = self.ad.define_basis(self.dsystem_dt, b, phi, 1, self.z_step(self.z_mem))
        else:  # hops_trajectory.py:303-317
            st = list(set(int(s) for s in b.state_list) | set(int(s) for s in self.static_basis[0]))
            ax = list(set(b.aux_list) | set(tuple(a) for a in self.static_basis[1]))
        self.phi = self.update_basis(phi, st, ax)
        self._store(0.0)

    def _store(self, t):
        b = self.basis
        full = np.zeros(self.sysd["NSTATES"], complex)
        full[b.state_list] = self.phi[:b.n_state]
        self.psi_traj.append(full)
        self.t_axis.append(t)
        self.z_mem_traj.append(self.z_mem.copy())
        self.state_traj.append(list(int(s) for s in b.state_list))
        self.aux_traj.append(list(b.aux_list))

    def _rk(self, phi, z_mem, tau):
        z1, z2 = self.noise_block(self.t, tau)
        phi, z_mem = rk4_step(self.dsystem_dt, phi, z_mem, z1, z2, tau)
        if self.eom_type == "NORMALIZED NONLINEAR":
            phi = phi / np.sqrt(np.sum(np.abs(phi[:self.basis.n_state]) ** 2))
        return phi, z_mem

    def propagate(self, t_advance, tau):
        t0 = self.t
        t_axis = t0 + np.arange(1, 1 + int(np.ceil(t_advance / tau))) * tau
        for index_t, t in enumerate(t_axis):
            phi, z_mem = self._rk(self.phi, self.z_mem, tau)
            if self.early_counter < self.early_steps:
                # hops_trajectory.py:469-509
                if self.early_integrator in ("INCH_WORM", "STATIC_STATE_INCHWORM_HIERARCHY"):
                    static_states = self.early_integrator == "STATIC_STATE_INCHWORM_HIERARCHY"
                    st, ax = self.ad.define_basis(self.dsystem_dt, self.basis, phi, tau, self.z_step(z_mem))
                    if static_states:
                        st = list(self.static_basis[0])
                    k = 0
                    while (set(st) != set(int(s) for s in self.basis.state_list)
                           or set(ax) != set(self.basis.aux_list)):
                        # inchworm: union basis, redo the step from the old phi
                        st_u = sorted(set(st) | set(int(s) for s in self.basis.state_list))
                        ax_u = list(set(ax) | set(self.basis.aux_list))
                        self.phi = self.update_basis(self.phi, st_u, ax_u)
                        phi, z_mem = self._rk(self.phi, self.z_mem, tau)
                        st, ax = self.ad.define_basis(self.dsystem_dt, self.basis, phi, tau, self.z_step(z_mem))
                        if static_states:
                            st = list(self.static_basis[0])
                        k += 1
                        if k >= self.inchworm_cap:
                            break
                    phi = self.update_basis(phi, st, ax)
                self.early_counter += 1
            elif (index_t + 1) % self.update_step == 0:
                st, ax = self.ad.define_basis(self.dsystem_dt, self.basis, phi, tau, self.z_step(z_mem))
                phi = self.update_basis(phi, st, ax)
            self.phi = phi
            b = self.basis
            full = np.zeros(self.sysd["NSTATES"], complex)
            full[b.state_list] = phi[:b.n_state]
            self.psi_traj.append(full)
            self.t_axis.append(float(t))
            self.z_mem_traj.append(self.z_mem.copy())
            self.state_traj.append(list(int(s) for s in b.state_list))
            self.aux_traj.append(list(b.aux_list))
            self.z_mem = z_mem
            self.t = float(t)

    def reset_early_time_integrator(self):
        """hops_trajectory.py:761-772."""
        self.early_counter = 0


# ---------------------------------------------------------------------------
# Dyadic HOPS (trajectory/hops_dyadic.py:15-246), the linear-spectroscopy
# pathways (trajectory/dyadic_spectra.py:198-356) and the response function
# (util/spectroscopy_analysis.py:67-113), on the trajectories above.
# ---------------------------------------------------------------------------
def dyad(m):
    """[[M, 0], [0, M]] (hops_dyadic.py:113-143); dense input -> complex128."""
    n = m.shape[0]
    if sp.issparse(m):
        return sp.block_diag((m, m), format="coo")
    out = np.zeros((2 * n, 2 * n), dtype=np.complex128)
    out[:n, :n] = m
    out[n:, n:] = m
    return out


def dyadic_system(system_param):
    """A copy of system_param with the Hamiltonian and every L-operator list
    in dyadic form (hops_dyadic.py:63-78)."""
    sysp = dict(system_param)
    sysp["HAMILTONIAN"] = dyad(system_param["HAMILTONIAN"])
    for key in ("L_HIER", "L_NOISE1", "L_NOISE2", "L_LT_CORR"):
        if key in sysp:
            sysp[key] = [dyad(m) for m in system_param[key]]
    return sysp


class Dyadic:
    """A dyadic trajectory driving an oracle Trajectory (or AdaptiveTrajectory)
    built on dyadic_system(...)."""

    def __init__(self, traj):
        self.tr = traj
        self.norm_sq = []

    def initialize(self, psi_ket, psi_bra):
        """hops_dyadic.py:83-111."""
        psi = np.concatenate((np.asarray(psi_ket, complex), np.asarray(psi_bra, complex)))
        n2 = float(np.sum(np.abs(psi) ** 2))
        self.norm_sq = [n2]
        self.tr.initialize(psi / np.sqrt(n2))

    def side_operator(self, op_h, side):
        """hops_dyadic.py:145-214: op on the chosen side's block, the identity
        on the other side's basis states; the response normalisation factor;
        renormalisation for the normalised EOM."""
        op_h = op_h.toarray() if sp.issparse(op_h) else np.asarray(op_h)
        n = op_h.shape[0]
        states = [int(s) for s in self.tr.basis.state_list]
        before = float(np.sum(np.abs(self.tr.phi[:len(states)]) ** 2))
        full = np.zeros((2 * n, 2 * n), dtype=np.complex128)
        for s in states:
            if (s < n) == (side == "bra"):
                full[s, s] = 1.0
        blk = slice(n, 2 * n) if side == "bra" else slice(0, n)
        full[blk, blk] = op_h
        self.tr.operator(sp.csr_matrix(full))
        ns = self.tr.basis.n_state
        after = float(np.sum(np.abs(self.tr.phi[:ns]) ** 2))
        if self.tr.eom_type == "NORMALIZED NONLINEAR":
            self.norm_sq.append(after)
            self.tr.phi = self.tr.phi / np.sqrt(after)
        else:
            self.norm_sq.append(after / before)

    def full_traj(self):
        """Stored wave functions over the whole dyadic state space."""
        tr = self.tr
        if hasattr(tr, "state_traj"):
            return np.array(tr.psi_traj)
        out = np.zeros((len(tr.psi_traj), tr.sysd["NSTATES"]), complex)
        out[:, tr.basis.state_list] = np.array(tr.psi_traj)
        return out

    def response(self, F, index_t):
        """spectroscopy_analysis.py:67-113 on the stored trajectory."""
        F = F.toarray() if sp.issparse(F) else np.asarray(F)
        scale = float(np.prod(self.norm_sq))
        return np.array([scale / np.sum(np.abs(p) ** 2) * (np.conj(p) @ (F @ p))
                         for p in self.full_traj()[int(index_t) + 1:]])


def hilbert_field_operator(mu, field, sites, n_hilb, lower=False):
    """Raising sum_s (mu_s . E) |s><0| or lowering |0><s| over sites
    counted from 1 (dyadic_spectra.py:198-238)."""
    op = np.zeros((n_hilb, n_hilb))
    for s in np.asarray(sites):
        w = float(np.dot(mu[s - 1], field))
        if lower:
            op[0, s] += w
        else:
            op[s, 0] += w
    return op


def signal_operator(mu, e_sig, n_chrom):
    """sum_n (mu_n . E_sig) |0_bra><n_ket| in the doubled space
    (dyadic_spectra.py:240-266)."""
    n = n_chrom + 1
    F = np.zeros((2 * n, 2 * n))
    for k in range(n_chrom):
        F[n, k + 1] = float(np.dot(mu[k], e_sig))
    return F


def dyadic_spectrum(dy, kind, mu, fields, ket_sites, bra_sites, t1, t2, t3, dt):
    """The ABSORPTION / FLUORESCENCE pathways of calculate_spectrum
    (dyadic_spectra.py:268-356) on an initialised-from-ground Dyadic."""
    n_chrom = len(mu)
    n = n_chrom + 1
    g = np.zeros(n)
    g[0] = 1.0
    dy.initialize(g, g)
    F = signal_operator(mu, fields["E_sig"], n_chrom)
    if kind == "ABSORPTION":
        dy.side_operator(hilbert_field_operator(mu, fields["E_1"], ket_sites, n), "ket")
        dy.tr.propagate(t1, dt)
        scale = 2
    else:
        dy.side_operator(hilbert_field_operator(mu, fields["E_1"], ket_sites, n), "ket")
        dy.side_operator(hilbert_field_operator(mu, fields["E_2"], bra_sites, n), "bra")
        dy.tr.propagate(t2, dt)
        dy.side_operator(hilbert_field_operator(mu, fields["E_3"], np.arange(1, n), n, lower=True), "bra")
        dy.tr.propagate(t3, dt)
        scale = 4
    return scale * dy.response(F, int(t2 / dt))
