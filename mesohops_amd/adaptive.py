"""Adaptive HOPS (adHOPS) on the device engine.

Per RK4 step of one trajectory: the step runs on the device in the current
basis; the error bound of every auxiliary and state is then assembled from two
device derivative evaluations (hops_derivative: the full hierarchy, and the
hierarchy restricted to the stable auxiliaries) and the flux terms; the host
selects the stable and boundary sets (sorting and cumulative thresholds, the
reference's set logic); the new basis gets a device-built super-operator and
the hierarchy vector is moved into it on the device (hops_permute_c128).

Follows basis/hops_basis.py:117-292 (define_basis / update_basis), :294-998
(error bounds), basis/basis_functions_adaptive.py (error terms),
basis/basis_functions.py:8-65 (thresholds), basis/hops_fluxfilters.py (flux
filters) and trajectory/hops_trajectory.py:456-533, :636-714 (adaptive time
loop and INCH_WORM early integrator).  L-operators with entries off the
diagonal add their destination-state flux terms, their noise operator and
their low-temperature operators to the bounds (hops_basis.py:914-998,
:1030-1100).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .basis import BasisTables
from .engine import BatchState, DeviceOperator, _stream_ptr
from .errors import UnsupportedRequest
from .params import hbar


def error_threshold(sorted_err, max_err, offset=0.0):
    """Largest error value that can be dropped with the cumulative dropped
    error staying below max_err (basis/basis_functions.py:8-41)."""
    if len(sorted_err) == 0:
        return 0.0
    cum = np.cumsum(sorted_err) + offset
    i = int(np.argmax(cum >= max_err))
    if i > 0:
        return sorted_err[i - 1]
    return np.inf if sorted_err[0] + offset < max_err else 0.0


def keep_above(err, max_err, members):
    """hops_basis.py:1000-1028."""
    thr = error_threshold(np.sort(err), max_err)
    idx = np.where(err > thr)[0]
    return idx, [members[i] for i in idx]


def _step_aux(aux, mode, step):
    d = dict(aux)
    v = d.get(mode, 0) + step
    if v < 0:
        return None
    if v == 0:
        d.pop(mode, None)
    else:
        d[mode] = v
    return tuple((m, d[m]) for m in sorted(d))


def aux_tuple(a):
    """An auxiliary as this package's sorted ((mode, value), ...) tuple, from
    such a tuple or list of pairs, a {mode: value} dict, or an object with the
    reference AuxiliaryVector's dict_aux_vec (basis/hops_aux.py:60-66)."""
    if hasattr(a, "dict_aux_vec"):
        a = a.dict_aux_vec
    if isinstance(a, dict):
        a = a.items()
    return tuple(sorted((int(m), int(v)) for m, v in a if int(v) != 0))


def _depth(aux):
    return sum(v for _, v in aux)


class BasisSelector:
    """Host half of the adaptive basis update: from the current basis tables,
    the normalised hierarchy vector and two derivative evaluations (a callable
    returning hbar dPhi/dt for a vector in this basis), the new state and
    auxiliary lists (basis/hops_basis.py:117-182, :294-998).  It touches no
    device memory; the trajectory engine supplies the derivative."""

    def __init__(self, info, delta_a, delta_s, maxhier, f_discard, markov=None, filters=None):
        import scipy.sparse as sps
        self.info = info
        # L-operators with entries off the diagonal couple a basis state to
        # other states (Peierls-type couplings): their flux terms sum over the
        # destination states and their noise enters the state flux
        # (hops_basis.py:1030-1100, :914-998)
        self.off_diag = np.array([bool(np.any(L.row != L.col)) for L in info.l_ops], dtype=bool)
        self.L_csc = []
        for L in info.l_ops:
            c = sps.csc_array(L)
            c.sum_duplicates()
            c.sort_indices()
            self.L_csc.append(c)
        dest = [set() for _ in range(info.nstates)]
        for L in info.l_ops:
            for r, c in zip(L.row, L.col):
                dest[int(r)].add(int(c))
        self.dest = [sorted(d) for d in dest]  # hops_system.py:218-226
        self.delta_a, self.delta_s = float(delta_a), float(delta_s)
        self.maxhier = int(maxhier)
        self.f_discard = float(f_discard)
        self.markov = None if markov is None else np.array(markov, bool)
        # Triangular / LongEdge static filters: (name, mask over the absolute
        # modes, kmax2), in the order given
        self.filters = [(n, np.array(p[0], bool), int(p[1])) for n, p in (filters or [])]

    def _bind(self, tb):
        self.tb = tb
        self.aux_index = {a: i for i, a in enumerate(tb.aux_list)}
        self.rel_mode = {int(m): i for i, m in enumerate(tb.modes)}

    # -- error terms ---------------------------------------------------------
    def _mode_tables(self):
        tb = self.tb
        nm, n = tb.n_mode, tb.n_state
        M2 = np.zeros((nm, n), complex)
        for i in range(nm):
            L = tb.L[tb.mode_lrel[i]].tocoo()
            for r, c, v in zip(L.row, L.col, L.data):
                if r == c:
                    M2[i, c] += v
        K2 = np.zeros((nm, tb.n_hier), dtype=np.int64)
        for j, a in enumerate(tb.aux_list):
            for (m, v) in a:
                K2[self.rel_mode[m], j] = v
        g = self.info.G[tb.modes]
        w = self.info.W[tb.modes]
        psi = self._pn[: n]
        nrm = np.vdot(psi, psi)
        avg = np.array([np.vdot(psi, tb.L[j] @ psi) / nrm for j in range(tb.n_lop)])
        self._avg = avg
        X2 = np.repeat(avg[tb.mode_lrel].reshape(nm, 1), n, axis=1)
        return M2, K2, g, w, X2

    # -- off-diagonal L-operators -------------------------------------------
    def _off_couplings(self):
        """Any L-operator of the basis with entries off the diagonal
        (hops_basis.py:1223-1227)."""
        return bool(np.any(self.off_diag[self.tb.absL]))

    def _off_tables(self, sc):
        """Destination tables of the basis modes: for every destination state
        d of the basis states (hops_system.py:218-226) M_d[m, i] = L_m[d, s_i]
        for s_i != d (hops_basis.py:1030-1100).  Returns off2[m, i] = sum_d
        |M_d[m, i]|^2 (added over d in ascending order) and the tables of the
        destinations outside the basis, {d: M_d}, for d in sc."""
        tb = self.tb
        st = [int(s) for s in tb.state_list]
        dests = set()
        for s in st:
            dests.update(self.dest[s])
        want = set(sc) & dests
        nm, n = tb.n_mode, tb.n_state
        off2 = np.zeros((nm, n))
        Md = {d: np.zeros((nm, n), complex) for d in sorted(want)}
        for lm in range(nm):
            Lc = self.L_csc[int(tb.absL[tb.mode_lrel[lm]])]
            for i, s in enumerate(st):
                lo, hi = Lc.indptr[s], Lc.indptr[s + 1]
                acc = 0.0
                for d, v in zip(Lc.indices[lo:hi], Lc.data[lo:hi]):
                    d = int(d)
                    if d == s or d not in dests:
                        continue
                    acc += abs(v) ** 2
                    if d in want:
                        Md[d][lm, i] = v
                off2[lm, i] = acc
        return off2, Md

    def _h2(self, z_step):
        """-iH over the full state space plus, with off-diagonal L-operators in
        the basis, their noise and noise-memory operator sum_j (conj(z1_j) -
        i z2_j + sum_{m of L_j} z_mem,m) L_j (hops_basis.py:914-953)."""
        tb = self.tb
        H2 = -1j * self.info.H
        if not self._off_couplings():
            return H2
        z1, z2, zm = z_step
        comp = np.zeros(len(tb.absL), dtype=np.complex128)
        for lm, m in enumerate(tb.modes):
            comp[tb.mode_lrel[lm]] += zm[int(m)]
        coef = np.conj(np.asarray(z1)) - 1j * np.asarray(z2) + comp
        Z = np.zeros_like(H2)
        for j, l in enumerate(tb.absL):
            if self.off_diag[l]:
                Z = Z + coef[j] * self.info.l_ops[int(l)].toarray()
        return H2 + Z

    def _ltc_ops(self):
        """(T_phys, T_hier) of the off-diagonal L-operators with a
        low-temperature coefficient (hops_basis.py:955-998): T_hier = sum
        conj(c_j) <L_j> L_j and T_phys = T_hier + T_hier^dagger - sum c_j
        L_j^2; (None, None) without such operators in the basis."""
        tb = self.tb
        lt = np.asarray(self.info.lt_param, dtype=np.complex128)[tb.absL]
        if not (np.any(lt) and self._off_couplings()):
            return None, None
        N = self.info.nstates
        Th = np.zeros((N, N), complex)
        Lsq = np.zeros((N, N), complex)
        for j, l in enumerate(tb.absL):
            if self.off_diag[l]:
                Ld = self.info.l_ops[int(l)].toarray()
                Th = Th + np.conj(lt[j]) * self._avg[j] * Ld
                Lsq = Lsq + lt[j] * (Ld @ Ld)
        return Th + Th.conj().T - Lsq, Th

    def _markov_up(self):
        tb = self.tb
        F = np.ones((tb.n_mode, tb.n_hier), bool)
        if self.markov is None:
            return F
        mask = self.markov[tb.modes]
        F[mask, 1:] = False
        for lm in np.nonzero(mask)[0]:
            j = self.aux_index.get(((int(tb.modes[lm]), 1),))
            if j is not None:
                F[:, j] = False
        return F

    def _filter_up(self):
        """Allowed flux up along (mode, aux), hops_fluxfilters.py:313-501: the
        Markovian mask; a Triangular filter closes the filtered modes of an
        auxiliary whose depth in them has reached kmax2; a LongEdge filter,
        from total depth kmax2 on, closes every mode of an auxiliary with depth
        in a filtered mode except the own mode of an edge (single-mode)
        auxiliary, and the filtered modes of any other auxiliary.  Filters of
        one kind act on one mask in order, and the kinds multiply."""
        F = self._markov_up()
        if not self.filters:
            return F
        tb = self.tb
        nm, nh = tb.n_mode, tb.n_hier
        K = np.zeros((nh, self.info.n_hmodes), dtype=np.int64)  # aux depth by absolute mode
        for j, a in enumerate(tb.aux_list):
            for (m, v) in a:
                K[j, m] = v
        depth = K.sum(axis=1)
        n_nonzero = (K > 0).sum(axis=1)
        for kind in ("Triangular", "LongEdge"):
            G = np.ones((nm, nh), bool)
            for name, mask, kmax2 in self.filters:
                if name != kind:
                    continue
                rel = mask[tb.modes]                       # filtered modes of the basis
                fdep = K[:, mask].sum(axis=1)              # depth in the filtered modes
                if kind == "Triangular":
                    G[np.ix_(rel, fdep >= kmax2)] = False
                    continue
                deep = depth >= kmax2
                with_f = deep & (fdep > 0)
                G[:, with_f] = False
                for j in np.nonzero(with_f & (n_nonzero == 1))[0]:
                    G[self.rel_mode[int(tb.aux_list[j][0][0])], j] = True
                G[np.ix_(rel, deep & (fdep == 0))] = False
            F &= G
        return F

    def _list_sc(self):
        info, st = self.info, set(int(s) for s in self.tb.state_list)
        out = set()
        H = info.H
        for s in st:
            out |= set(int(x) for x in np.nonzero(H[s])[0])
            for l in info.l_by_state[s]:
                c = info.l_ops[l]
                out |= set(int(x) for x in c.col[c.row == s])
        return sorted(out - st)

    def define_basis(self, tb, phi, dt, derivative, z_step=None):
        """hops_basis.py:117-182: new state list and auxiliary list.  z_step
        (a callable returning (z1, z2, z_mem): the noise of the basis
        L-operators at the step's start and the noise memory by absolute
        mode) is called only when an L-operator of the basis has entries off
        the diagonal."""
        self._bind(tb)
        self._deriv = derivative
        # the error terms of this call, kept for inspection: per auxiliary,
        # per basis state, and per outside state sc
        self.err_aux = self.err_state = None
        self.err_outside = {}
        n, nh = tb.n_state, tb.n_hier
        self._pn = phi / np.linalg.norm(phi[:n])
        Pn = self._pn
        C2 = Pn.reshape(nh, n).T
        absC2 = np.abs(C2) ** 2
        M2, K2, g, w, X2 = self._mode_tables()
        sc = self._list_sc()
        off = self._off_couplings()
        off2, Md = self._off_tables(sc) if off else (0.0, {})
        H2 = self._h2(z_step() if off else None)
        T = self._ltc_ops()
        st = [int(s) for s in tb.state_list]
        if self.delta_a > 0:
            # hier_stable_error (hops_basis.py:560-665)
            D = self._deriv(Pn) / hbar + Pn / dt
            err = np.sum(np.abs(D.reshape(nh, n).T) ** 2, axis=0)
            if sc:
                # flux out of the state basis (basis_functions_adaptive.py:76-163)
                if T[0] is None:
                    Hsc = H2[np.ix_(sc, st)]
                    err = err + np.sum(np.abs(Hsc @ C2 / hbar) ** 2, axis=0)
                else:
                    Hp = (H2 + T[0])[np.ix_(sc, st)]
                    Hh = (H2 + T[1])[np.ix_(sc, st)]
                    Caux = C2.copy()
                    Caux[:, 0] = 0
                    e = np.sum(np.abs(Hh @ Caux / hbar) ** 2, axis=0)
                    e[0] = np.sum(np.abs(Hp @ C2[:, :1] / hbar) ** 2)
                    err = err + e
            KW = np.abs(w).reshape(-1, 1) * (1 + K2)
            Msup = np.abs(M2) ** 2
            if off:
                Msup = Msup + off2
            E_up = np.abs(KW) ** 2 * (Msup @ absC2) / hbar ** 2
            F_up = self._filter_up()
            depth = np.array([_depth(a) for a in tb.aux_list])
            F_up[:, depth == self.maxhier] = False
            err = err + np.sum(F_up * E_up, axis=0)
            gw2 = np.square(np.abs(g / w)).reshape(-1, 1)
            Dsup = np.abs(M2 - X2) ** 2
            if off:
                Dsup = Dsup + off2
            E_dn = gw2 * (Dsup @ absC2) / hbar ** 2
            err = err + np.sum((K2 > 0) * E_dn, axis=0)
            self.err_aux = err
            dsq = self.delta_a ** 2
            idx_st, aux_st = keep_above(err, dsq / 2.0, tb.aux_list)
            unstable = np.setdiff1d(np.arange(nh), idx_st)
            dbound = dsq - np.sum(err[unstable])
            aux_b = self._boundary_aux(E_up, E_dn, idx_st, dbound, depth)
            aux_new = list(set(aux_st) | set(aux_b))
            stable_idx = sorted(self.aux_index[a] for a in aux_st)
        else:
            aux_new, aux_b, stable_idx = list(tb.aux_list), [], list(range(nh))
        if self.delta_s > 0:
            st_stable, st_bound = self._state_basis(Pn, C2, absC2, M2, K2, g, w, X2, sc, dt, stable_idx, aux_b,
                                                    aux_new, off2 if off else None, Md, H2, T)
            states_new = sorted(set(st_stable) | set(st_bound))
        else:
            states_new = st
        return states_new, aux_new

    def _boundary_aux(self, E_up, E_dn, idx_st, dbound, depth):
        """hops_basis.py:667-793."""
        tb = self.tb
        nm, nh = tb.n_mode, tb.n_hier
        F_up = np.ones((nm, nh), bool)
        F_dn = np.ones((nm, nh), bool)
        for j, a in enumerate(tb.aux_list):
            d = dict(a)
            if depth[j] < self.maxhier:
                for lm, m in enumerate(tb.modes):
                    if _step_aux(a, int(m), 1) in self.aux_index:
                        F_up[lm, j] = False
            else:
                F_up[:, j] = False
            for lm, m in enumerate(tb.modes):
                m = int(m)
                if m not in d or _step_aux(a, m, -1) in self.aux_index:
                    F_dn[lm, j] = False
        F_up &= self._filter_up()
        Eu = (E_up * F_up)[:, idx_st]
        Ed = (E_dn * F_dn)[:, idx_st]
        srt = np.sort(np.append(Eu[Eu != 0], Ed[Ed != 0]))
        thr = error_threshold(srt, dbound * self.f_discard * self.f_discard)
        discarded = np.sum(Eu[Eu <= thr]) + np.sum(Ed[Ed <= thr])
        Eu[Eu <= thr] = 0
        Ed[Ed <= thr] = 0
        acc = {}
        for col, j in enumerate(idx_st):
            a = tb.aux_list[j]
            for lm in np.nonzero(Eu[:, col])[0]:
                key = _step_aux(a, int(tb.modes[lm]), 1)
                acc[key] = acc.get(key, 0.0) + Eu[lm, col]
            d = dict(a)
            for lm in np.nonzero(Ed[:, col])[0]:
                m = int(tb.modes[lm])
                if m in d:
                    key = _step_aux(a, m, -1)
                    acc[key] = acc.get(key, 0.0) + Ed[lm, col]
        thr = error_threshold(np.sort(np.array(list(acc.values()))), dbound, offset=discarded)
        return [k for k, v in acc.items() if v > thr]

    def _aux_filters(self, aux_set):
        """(F_down, F_up) over [mode, aux]: F_down marks the auxiliaries k with
        k - e_m in aux_set, F_up those with k + e_m in aux_set
        (hops_fluxfilters.py:292-352)."""
        tb = self.tb
        F_dn = np.zeros((tb.n_mode, tb.n_hier), bool)
        F_up = np.zeros((tb.n_mode, tb.n_hier), bool)
        for a in aux_set:
            for lm, m in enumerate(tb.modes):
                j = self.aux_index.get(_step_aux(a, int(m), 1))
                if j is not None:
                    F_dn[lm, j] = True
            for (m, v) in a:
                j = self.aux_index.get(_step_aux(a, m, -1))
                if j is not None and m in self.rel_mode:
                    F_up[self.rel_mode[m], j] = True
        return F_dn, F_up

    def _state_basis(self, Pn, C2, absC2, M2, K2, g, w, X2, sc, dt, stable_idx, aux_b, aux_new, off2, Md, H2, T):
        """hops_basis.py:294-482 and :795-912.  off2 / Md: the destination
        tables of off-diagonal L-operators (None / {} without them); H2: -iH
        plus their noise operator; T: their low-temperature operators."""
        tb = self.tb
        n, nh, nm = tb.n_state, tb.n_hier, tb.n_mode
        st = [int(s) for s in tb.state_list]
        P_st = np.zeros(nh * n, complex)
        P_st.reshape(nh, n)[stable_idx] = Pn.reshape(nh, n)[stable_idx]
        D = self._deriv(P_st) / hbar + P_st / dt
        err = np.sum(np.abs(D.reshape(nh, n).T[:, stable_idx]) ** 2, axis=1)
        # flux out of each basis state into every other state through H2
        # (basis_functions_adaptive.py:893-995)
        Hc = H2 - np.diag(np.diag(H2))
        C2s = C2[:, stable_idx]
        if T[0] is None:
            V1 = np.sum(np.abs(Hc[:, st]) ** 2, axis=0)
            err = err + V1 * np.sum(np.abs(C2s) ** 2, axis=1) / hbar ** 2
        else:
            Tp, Th = T
            Vp = np.sum(np.abs((Hc + Tp - np.diag(np.diag(Tp)))[:, st]) ** 2, axis=0)
            Vh = np.sum(np.abs((Hc + Th - np.diag(np.diag(Th)))[:, st]) ** 2, axis=0)
            Caux = C2s.copy()
            Caux[:, 0] = 0
            err = err + (Vp * np.abs(C2s[:, 0]) ** 2 + Vh * np.sum(np.abs(Caux) ** 2, axis=1)) / hbar ** 2
        F_dn, F_up = self._aux_filters(aux_b)
        gw2 = np.square(np.abs(g / w)).reshape(-1, 1)
        KW = np.abs(w).reshape(-1, 1) * (1 + K2)
        E_dn = ((np.abs(M2 - X2) ** 2).T @ (F_dn * gw2)) * absC2 / hbar ** 2
        E_up = ((np.abs(M2) ** 2).T @ (np.abs(F_up * KW) ** 2)) * absC2 / hbar ** 2
        if off2 is not None:
            # flux through the destination tables, filtered by the new basis
            # (basis_functions_adaptive.py:268-384, :626-758)
            F_dn_o, F_up_o = self._aux_filters(aux_new)
            if np.sum(off2) != 0:
                E_dn = (off2.T @ (F_dn_o * gw2)) * absC2 / hbar ** 2 + E_dn
                E_up = E_up + (off2.T @ (np.abs(F_up_o * KW) ** 2)) * absC2 / hbar ** 2
        err = err + np.sum(E_dn[:, stable_idx], axis=1)
        err = err + np.sum(E_up[:, stable_idx], axis=1)
        self.err_state = err
        dsq = self.delta_s ** 2
        rel_st, st_stable = keep_above(err, dsq / 2.0, st)
        dbound = dsq - np.sum(err[np.setdiff1d(np.arange(n), rel_st)])
        bound = []
        if len(rel_st) < self.info.nstates and sc:
            # flux into the outside states (basis_functions_adaptive.py:997-1132)
            Cs = C2[np.ix_(rel_st, stable_idx)]
            if T[0] is None:
                E = np.sum(np.abs(H2[np.ix_(sc, st_stable)] @ (Cs / hbar)) ** 2, axis=1)
            else:
                E = (np.abs((H2 + T[0])[np.ix_(sc, st_stable)] @ (Cs[:, :1] / hbar)) ** 2)[:, 0]
                if Cs.shape[1] > 1:
                    E = E + np.sum(np.abs((H2 + T[1])[np.ix_(sc, st_stable)] @ (Cs[:, 1:] / hbar)) ** 2, axis=1)
            E = np.array(E, dtype=float)
            if Md:
                # flux up and down into each outside destination state, out of
                # the stable states and auxiliaries (basis_functions_adaptive.py:386-515, :760-891)
                keep = np.zeros(n, bool)
                keep[rel_st] = True
                C2sq = np.abs(np.where(keep[:, None], C2, 0)[:, stable_idx]) ** 2
                F_dn_o, F_up_o = self._aux_filters(aux_new)
                KWo = np.abs(F_up_o[:, stable_idx] * KW[:, stable_idx]) ** 2
                G = F_dn_o[:, stable_idx] * gw2
                for d, M in Md.items():
                    A = np.abs(M) ** 2
                    f_up = np.sum((A.T @ KWo) * C2sq / hbar ** 2)
                    f_dn = np.sum((A.T @ G) * C2sq / hbar ** 2)
                    E[sc.index(d)] += f_up + f_dn
            self.err_outside = dict(zip(sc, E))
            nz = np.nonzero(E)[0]
            if len(nz):
                _, bound = keep_above(E[nz], dbound, list(np.array(sc)[nz]))
        bound = sorted(set(int(b) for b in bound) - set(st))
        return [int(s) for s in st_stable], bound

class AdaptiveTrajectoryEngine:
    """One adaptive trajectory: device state + basis bookkeeping."""

    def __init__(self, setup, noise1, noise2, precision_mode, device, delta_a, delta_s, update_step,
                 f_discard, early_integrator="INCH_WORM", early_steps=5, inchworm_cap=5, static_basis=None):
        self.setup = setup
        self.info = setup.info
        self.noise1, self.noise2 = noise1, noise2
        self.prec = {"fp64": _lib.HOPS_PREC_FP64, "mixed": _lib.HOPS_PREC_MIXED}[precision_mode]
        self.device = torch.device(device)
        self.delta_a, self.delta_s = float(delta_a), float(delta_s)
        self.update_step = int(update_step) if update_step else 1
        self.f_discard = float(f_discard)
        if early_integrator not in ("INCH_WORM", "STATIC", "STATIC_STATE_INCHWORM_HIERARCHY"):
            raise UnsupportedRequest(early_integrator, "early time integrator clause of the propagate")
        self.early_integrator = early_integrator
        # STATIC_BASIS: [state list, auxiliary list] kept in the basis from the
        # start (trajectory/hops_trajectory.py:303-317)
        self.static_basis = None
        if static_basis is not None:
            self.static_basis = ([int(x) for x in static_basis[0]], [aux_tuple(a) for a in static_basis[1]])
        if early_integrator == "STATIC_STATE_INCHWORM_HIERARCHY" and self.static_basis is None:
            raise UnsupportedRequest("STATIC_STATE_INCHWORM_HIERARCHY without a STATIC_BASIS",
                                     "early time integrator clause of the propagate")
        self.early_steps, self.inchworm_cap = int(early_steps), int(inchworm_cap)
        self.maxhier = setup.hier["MAXHIER"]
        self.markov = None
        filters = []
        for name, prm in setup.hier["STATIC_FILTERS"]:
            if name == "Markovian":
                mk = np.array(prm, bool)
                self.markov = mk if self.markov is None else (self.markov | mk)
            elif name in ("Triangular", "LongEdge"):
                filters.append((name, [list(prm[0]), int(prm[1])]))
            else:
                raise UnsupportedRequest("static filter %s with an adaptive hierarchy" % name, "make_adaptive")
        self.selector = BasisSelector(self.info, self.delta_a, self.delta_s, self.maxhier, self.f_discard,
                                      self.markov, filters)
        self.early_counter = 0
        self.raw_pending = False  # the next step takes Phi as an operator left it
        self.sp = _stream_ptr(self.device)
        # noise of every hierarchy L-operator; a basis takes its active rows
        from .noise import device_noise
        self.interp = noise1.interpolate
        if noise2.model != "ZERO" and noise2.interpolate != self.interp:
            raise UnsupportedRequest("INTERPOLATE set for one noise and not the other", "HopsTrajectory")
        nl = self.info.n_l2
        # complex128 noise: interpolated, or PRE_CALCULATED complex128 used as given
        c128 = noise1.needs_c128() or noise2.needs_c128()
        if c128 and setup.integ["EFFECTIVE_NOISE_INTEGRATION"]:
            raise UnsupportedRequest("EFFECTIVE_NOISE_INTEGRATION with complex128 (interpolated or "
                                     "PRE_CALCULATED complex128) noise", "HopsTrajectory")
        self.z1_all = device_noise([noise1], list(range(nl)), self.device, self.sp, c128=c128)[0]
        self.z2_all = None if noise2.model == "ZERO" else device_noise([noise2], list(range(nl)), self.device,
                                                                       self.sp, c128=c128)[0]
        self.n_t = noise1.n_t
        if self.interp:
            # the grid noise stays; each propagate call reads the cubic noise
            # on its own half-step times (begin_call)
            self._z_grid = (self.z1_all, self.z2_all)
            self.begin_call(0.0, 2.0 * noise1.param["TAU"], 0)
        self.zmem = torch.zeros(self.info.n_hmodes + 1, dtype=torch.complex128, device=self.device)
        # EFFECTIVE_NOISE_INTEGRATION: the RK stages read moving averages of
        # the noise (integrator/integrator_rk.py:115-128); the error bound
        # keeps the point samples (hops_trajectory.py:693-714)
        self.effective = bool(setup.integ["EFFECTIVE_NOISE_INTEGRATION"])
        self.z1_eff = self.z2_eff = None
        self._eff_w = None

    def _rows(self, z, act):
        return None if z is None else z.index_select(0, act).unsqueeze(0).contiguous()

    def _use_effective_noise(self, w):
        if not self.effective or self._eff_w == w:
            return
        from .noise import window_mean
        self.z1_eff = window_mean(self.z1_all, w, self.sp)
        self.z2_eff = None if self.z2_all is None else window_mean(self.z2_all, w, self.sp)
        self._eff_w = w
        self._attach_noise()

    def _attach_noise(self):
        """Point the RK state at the (effective) noise rows of its basis."""
        if not self.effective or self.z1_eff is None:
            return
        act = torch.as_tensor(self.tb.absL_active, dtype=torch.long, device=self.device)
        st = self.st
        st.z1 = self._rows(self.z1_eff, act)
        st.bt.z1_d = st.z1.data_ptr()
        if self.z2_eff is not None:
            st.z2 = self._rows(self.z2_eff, act)
            st.bt.z2_d = st.z2.data_ptr()

    # -- basis objects -------------------------------------------------------
    def _make(self, states, aux):
        tb = BasisTables(self.info, states, aux, self.setup.eom_kind, dense_h=False)
        dop = DeviceOperator(tb, self.prec, self.device)
        act = torch.as_tensor(tb.absL_active, dtype=torch.long, device=self.device)
        z1 = self._rows(self.z1_all, act)
        z2 = self._rows(self.z2_all, act)
        st = BatchState(dop, 1, self.setup.eom_kind, z1, z2, self.n_t, store_slots=0, lanes=1)
        st.zmem[: self.info.n_hmodes].copy_(self.zmem[: self.info.n_hmodes])
        self.tb, self.dop, self.st = tb, dop, st
        # derivative evaluation runs in FP64 on its own operator when the
        # propagation is mixed precision
        self.dop64 = dop if self.prec == _lib.HOPS_PREC_FP64 else DeviceOperator(tb, _lib.HOPS_PREC_FP64, self.device)
        self.st64 = st if self.prec == _lib.HOPS_PREC_FP64 else BatchState(
            self.dop64, 1, self.setup.eom_kind, z1, z2, self.n_t, store_slots=0, lanes=1)
        self._z_point = (z1, z2)
        self._attach_noise()

    def begin_call(self, t0, tau, n_steps):
        """INTERPOLATE: the cubic noise of every L-operator at t0 + j tau/2,
        j = 0..2 n_steps (noise index j for this call), for the current and
        every later basis of the call."""
        if not self.interp:
            return
        from .noise import spline_noise
        n_out = 2 * int(n_steps) + 1
        z1g, z2g = self._z_grid
        p1, p2 = self.noise1.param, self.noise2.param
        self.z1_all = spline_noise(z1g, p1["TAU"], t0, 0.5 * tau, n_out, self.sp, p1["FLAG_REAL"])
        self.z2_all = None if z2g is None else spline_noise(z2g, p2["TAU"], t0, 0.5 * tau, n_out, self.sp,
                                                            p2["FLAG_REAL"])
        self.n_t = n_out
        if getattr(self, "tb", None) is not None:
            act = torch.as_tensor(self.tb.absL_active, dtype=torch.long, device=self.device)
            z1, z2 = self._rows(self.z1_all, act), self._rows(self.z2_all, act)
            self.st.set_noise(z1, z2, n_out)
            if self.st64 is not self.st:
                self.st64.set_noise(z1, z2, n_out)
            self._z_point = (z1, z2)

    def phi_host(self):
        return self.st.phi_host()[0]

    def _sync_zmem_from_state(self):
        self.zmem[: self.info.n_hmodes].copy_(self.st.zmem[: self.info.n_hmodes])

    # -- device derivative ---------------------------------------------------
    def _derivative(self, y_host, tidx):
        st = self.st64
        if st is not self.st:
            st.zmem[: self.info.n_hmodes].copy_(self.st.zmem[: self.info.n_hmodes])
        y = torch.from_numpy(np.ascontiguousarray(y_host, dtype=np.complex128)).to(self.device)
        d = torch.empty_like(y)
        saved = (st.bt.z1_d, st.bt.z2_d)
        z1, z2 = self._z_point  # point samples for the error bound
        st.bt.z1_d = z1.data_ptr()
        st.bt.z2_d = None if z2 is None else z2.data_ptr()
        try:
            check(lib().hops_derivative(ctypes.byref(self.dop64.op), ctypes.byref(st.bt),
                                        ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(d.data_ptr()), int(tidx),
                                        self.sp))
        finally:
            st.bt.z1_d, st.bt.z2_d = saved
        return d.cpu().numpy()

    def _z_step(self, tidx):
        """(z1, z2, z_mem) for the selection's noise operator: the point
        noise of the basis L-operators at noise index tidx (the step's start,
        hops_trajectory.py:693-714) and the noise memory by absolute mode."""
        absL = torch.as_tensor(self.tb.absL, dtype=torch.long, device=self.device)
        z1 = self.z1_all.index_select(0, absL)[:, int(tidx)].cpu().numpy()
        if self.z2_all is None:
            z2 = np.zeros_like(z1)
        else:
            z2 = self.z2_all.index_select(0, absL)[:, int(tidx)].cpu().numpy()
        zm = self.st.zmem[: self.info.n_hmodes].cpu().numpy()
        return z1, z2, zm

    def define_basis(self, phi, dt, tidx):
        return self.selector.define_basis(self.tb, phi, dt, lambda y: self._derivative(y, tidx),
                                          lambda: self._z_step(tidx))

    # -- basis change ----------------------------------------------------------
    def update_basis(self, phi_host_vec, states, aux):
        """hops_basis.py:184-292: new basis, Phi moved on the device and
        rescaled to the old physical norm."""
        old = self.tb
        if set(states) == set(int(s) for s in old.state_list) and set(aux) == set(old.aux_list):
            return False
        self._sync_zmem_from_state()
        old_st = self.st
        n_old = old.n_state
        old_pos = {int(s): i for i, s in enumerate(old.state_list)}
        old_aux = {a: i for i, a in enumerate(old.aux_list)}
        self._make(states, aux)
        tb = self.tb
        new_pos = {int(s): i for i, s in enumerate(tb.state_list)}
        stable_states = sorted(set(old_pos) & set(new_pos))
        src, dst = [], []
        for j, a in enumerate(tb.aux_list):
            i = old_aux.get(a)
            if i is None:
                continue
            for s in stable_states:
                src.append(i * n_old + old_pos[s])
                dst.append(j * tb.n_state + new_pos[s])
        norm_old = np.linalg.norm(phi_host_vec[:n_old])
        norm_new = np.linalg.norm(phi_host_vec[[old_pos[s] for s in stable_states]])
        scale = norm_old / norm_new
        si = torch.as_tensor(np.array(src, dtype=np.int32), device=self.device)
        di = torch.as_tensor(np.array(dst, dtype=np.int32), device=self.device)
        check(lib().hops_permute_c128(ctypes.c_void_p(old_st.phi.data_ptr()), ctypes.c_void_p(self.st.phi.data_ptr()),
                                      ctypes.c_void_p(si.data_ptr()), ctypes.c_void_p(di.data_ptr()), len(src),
                                      float(scale), self.sp))
        return True

    # -- time loop ---------------------------------------------------------------
    def initialize(self, psi_0, t0_index):
        """Initial basis and its first update (trajectory/hops_trajectory.py:279-317)."""
        psi_0 = np.asarray(psi_0, dtype=np.complex128)
        if self.delta_s > 0:
            states = [int(s) for s in np.where(np.abs(psi_0) > 0)[0]]
        else:
            states = list(range(self.info.nstates))
        if self.delta_a > 0:
            aux = [()]
        else:
            from .basis import apply_static_filters, triangular_aux_list
            aux = triangular_aux_list(self.info.n_hmodes, self.maxhier)
            aux = apply_static_filters(aux, self.setup.hier["STATIC_FILTERS"], self.info.n_hmodes)
        self._make(states, aux)
        tb = self.tb
        phi = np.zeros(tb.n_hier * tb.n_state, complex)
        phi[: tb.n_state] = psi_0[tb.state_list]
        self.st.set_phi(phi)
        if self.static_basis is None:
            st, ax = self.define_basis(phi, 1.0, t0_index)
        else:
            st = sorted(set(int(s) for s in tb.state_list) | set(self.static_basis[0]))
            ax = list(set(tb.aux_list) | set(self.static_basis[1]))
        self.update_basis(phi, st, ax)

    def _rk(self, it, tau, ns):
        self._use_effective_noise(ns)
        self.st.raw_first = self.raw_pending
        self.st.propagate(it, 1, tau, ns, 0)

    def _same_basis(self, states, aux):
        return (set(states) == set(int(s) for s in self.tb.state_list)
                and set(aux) == set(self.tb.aux_list))

    def step(self, index_t, it, tau, ns):
        """One adaptive RK4 step from noise index ``it``
        (trajectory/hops_trajectory.py:456-533, :636-714)."""
        early = self.early_counter < self.early_steps
        inchworm = self.early_integrator in ("INCH_WORM", "STATIC_STATE_INCHWORM_HIERARCHY")
        static_states = self.early_integrator == "STATIC_STATE_INCHWORM_HIERARCHY"
        if early and inchworm:
            phi_old = self.phi_host()
            zmem_old = self.st.zmem.clone()
        self._rk(it, tau, ns)
        if early:
            if inchworm:
                phi = self.phi_host()
                s_new, a_new = self.define_basis(phi, tau, it)
                if static_states:
                    s_new = list(self.static_basis[0])
                k = 0
                while not self._same_basis(s_new, a_new):
                    # union basis, then the step again from the state before it
                    s_u = sorted(set(s_new) | set(int(s) for s in self.tb.state_list))
                    a_u = list(set(a_new) | set(self.tb.aux_list))
                    self.st.set_phi(phi_old)
                    self.st.zmem.copy_(zmem_old)
                    self.update_basis(phi_old, s_u, a_u)
                    phi_old = self.phi_host()
                    zmem_old = self.st.zmem.clone()
                    self._rk(it, tau, ns)
                    phi = self.phi_host()
                    s_new, a_new = self.define_basis(phi, tau, it)
                    if static_states:
                        s_new = list(self.static_basis[0])
                    k += 1
                    if k >= self.inchworm_cap:
                        break
                self.update_basis(phi, s_new, a_new)
            self.early_counter += 1
        elif (index_t + 1) % self.update_step == 0:
            phi = self.phi_host()
            s_new, a_new = self.define_basis(phi, tau, it)
            self.update_basis(phi, s_new, a_new)
        self.raw_pending = False
        self._sync_zmem_from_state()

    # -- operators -----------------------------------------------------------
    def apply_operator(self, op):
        """HopsTrajectory._operator (trajectory/hops_trajectory.py:540-577):
        with state adaptivity the states that op populates join the basis
        first; op, restricted to the state basis, acts on every auxiliary on
        the device; then the early-time integrator starts over.  The
        reference's define_basis / update_basis calls after the product keep
        the basis as it is."""
        import scipy.sparse as sps
        sparse = sps.issparse(op)
        if sparse:
            op = sps.csr_matrix(op)
        if self.delta_s > 0:
            cur = [int(s) for s in self.tb.state_list]
            block = op[:, cur]
            rows = block.nonzero()[0] if sparse else np.nonzero(np.asarray(block))[0]
            states = sorted(set(cur) | set(int(r) for r in rows))
            self.update_basis(self.phi_host(), states, list(self.tb.aux_list))
        sl = [int(s) for s in self.tb.state_list]
        sub = op[sl][:, sl]
        sub = sub.toarray() if sparse else np.asarray(sub)
        self.st.apply_state_op(sub)
        self.raw_pending = True
        if self.delta_s > 0:
            self.early_counter = 0
