"""Host-side bookkeeping of the HOPS basis: the derived system dictionary, the
auxiliary multi-indices and their order, and the integer/complex tables the
device uses to build the hierarchy super-operator of a basis.

The tables are what the HIP engine needs; every arithmetic operation on the
hierarchy vector happens on the device (mesohops_amd/csrc/hops_engine.hip).
"""
from __future__ import annotations

import itertools
from collections import Counter

import numpy as np
import scipy.sparse as sp

from .errors import AuxError, UnsupportedRequest


# ---------------------------------------------------------------------------
# System dictionary (basis/system_functions.py:10-171)
# ---------------------------------------------------------------------------
def operator_signature(op):
    """Hashable nonzero pattern of an L-operator; two L-operators with the same
    pattern are the same bath coupling (util/helper_functions.py:8-32)."""
    if sp.issparse(op):
        return tuple(tuple(int(i) for i in x) for x in op.nonzero()) if op.getnnz() else tuple()
    a = np.asarray(op)
    return tuple(tuple(int(i) for i in x) for x in np.nonzero(a)) if len(a) else tuple()


def _touched_states(op):
    r, c = op.nonzero() if sp.issparse(op) else np.nonzero(np.asarray(op))
    return tuple(sorted(set(int(i) for i in r) | set(int(j) for j in c)))


class SystemInfo:
    """Derived system parameters (the reference's system.param dictionary,
    basis/system_functions.py:10-171) in array form."""

    def __init__(self, system_param):
        self.param = system_param
        ham = system_param["HAMILTONIAN"]
        H = ham.toarray() if sp.issparse(ham) else np.asarray(ham)
        self.H = np.array(H, dtype=np.complex128)
        self.nstates = self.H.shape[0]
        gw = system_param["GW_SYSBATH"]
        self.n_hmodes = len(gw)
        self.G = np.array([g for (g, w) in gw], dtype=np.complex128)
        self.W = np.array([w for (g, w) in gw], dtype=np.complex128)
        lh = system_param["L_HIER"]
        sig = [operator_signature(L) for L in lh]
        order = {}
        self.l_index_by_hmode = np.zeros(self.n_hmodes, dtype=np.int64)
        self.l_ops = []
        states_by_l = []
        for i, s in enumerate(sig):
            if s not in order:
                order[s] = len(order)
                c = sp.coo_matrix(lh[i])
                c.eliminate_zeros()
                self.l_ops.append(sp.coo_matrix(c, dtype=np.complex128))
                states_by_l.append(_touched_states(lh[i]))
            self.l_index_by_hmode[i] = order[s]
        self.l_signature = order
        self.n_l2 = len(order)
        self.hmodes_by_state = [[] for _ in range(self.nstates)]
        for hm, L in enumerate(lh):
            for s in _touched_states(L):
                self.hmodes_by_state[s].append(hm)
        self.l_by_state = [[] for _ in range(self.nstates)]
        for il, sts in enumerate(states_by_l):
            for s in sts:
                self.l_by_state[s].append(il)
        lt = np.zeros(self.n_l2, dtype=np.complex128)
        for L, c in zip(system_param.get("L_LT_CORR", []), system_param.get("PARAM_LT_CORR", [])):
            s = operator_signature(L)
            if s in order:
                lt[order[s]] += c
        self.lt_param = lt
        # noise maps: hierarchy L index -> noise modes of that L-operator
        self.noise_modes_by_l = {}
        self.noise_n_l2 = {}  # distinct L-operators of each noise
        for tag in ("1", "2"):
            key = "L_NOISE" + tag
            if key not in system_param:
                continue
            by_l = [[] for _ in range(len(sig))]
            seen = set()
            for i, L in enumerate(system_param[key]):
                s = operator_signature(L)
                seen.add(s)
                if s in order:
                    by_l[order[s]].append(i)
            self.noise_modes_by_l[tag] = by_l
            self.noise_n_l2[tag] = len(seen)
        self.n_noise_modes = len(system_param.get("L_NOISE1", gw))


# ---------------------------------------------------------------------------
# Auxiliary multi-indices (basis/hops_aux.py:13-201,460-480;
# basis/hops_hierarchy.py:340-368)
# ---------------------------------------------------------------------------
def aux_order_key(aux):
    """The reference orders auxiliaries by the length of their ID string
    (their depth) and then by the ID string, whose fixed-width mode fields make
    that the lexicographic order of the sorted mode sequence."""
    seq = tuple(m for (m, v) in aux for _ in range(v))
    return (len(seq), seq)


def triangular_aux_list(n_hmodes, maxhier):
    """All (mode, depth) multi-indices of total depth <= maxhier, in the
    reference's order (basis/hops_hierarchy.py:340-368, :582)."""
    out = []
    for depth in range(maxhier + 1):
        for raw in itertools.combinations_with_replacement(range(n_hmodes), depth):
            c = Counter(raw)
            out.append(tuple((m, c[m]) for m in sorted(c)))
    return out


# a state basis above this size whose H has more than this many nonzeros per
# row on average gets the dense system block on the matrix cores
DENSE_H_MIN_STATES = 64
DENSE_H_MIN_ROW_NNZ = 8


class BasisTables:
    """Relative indexing of one basis (state list x aux list) and the tables of
    hops_op_t (include/hops_mi355x.h).  Follows the relative/absolute index
    conventions of basis/hops_system.py:206-286 and basis/hops_modes.py:111-171."""

    def __init__(self, info: SystemInfo, state_list, aux_list, eom_kind, dense_h=None):
        """dense_h: apply -iH over the state basis as a dense block on the
        matrix cores (True), in the sparse pattern (False), or the first when
        the basis has more than DENSE_H_MIN_STATES states and H more than
        DENSE_H_MIN_ROW_NNZ nonzeros per state row on average (None)."""
        self.info = info
        st = np.array(sorted(int(s) for s in state_list), dtype=np.int64)
        aux = sorted(set(aux_list), key=aux_order_key)
        if len(aux) == 0 or len(aux[0]) != 0:
            raise AuxError("Zero Vector not contained in list_aux!")
        self.state_list = st
        self.aux_list = aux
        n = len(st)
        nh = len(aux)
        self.n_state, self.n_hier = n, nh
        index = {a: i for i, a in enumerate(aux)}
        in_use = sorted({m for a in aux for (m, v) in a})
        st_modes = sorted({m for s in st for m in info.hmodes_by_state[s]})
        modes = np.array(sorted(set(in_use) | set(st_modes)), dtype=np.int64)
        self.modes = modes
        rel_mode = {int(m): i for i, m in enumerate(modes)}
        absL = sorted({int(info.l_index_by_hmode[m]) for m in modes})
        self.absL = np.array(absL, dtype=np.int64)
        relL = {l: i for i, l in enumerate(absL)}
        mode_lrel = np.array([relL[int(info.l_index_by_hmode[m])] for m in modes], dtype=np.int64)
        act_abs = sorted({l for s in st for l in info.l_by_state[s]})
        for l in act_abs:
            if l not in relL:
                raise UnsupportedRequest("an active L-operator without a hierarchy mode", "BasisTables")
        self.absL_active = np.array(act_abs, dtype=np.int64)
        act_lrel = np.array([relL[l] for l in act_abs], dtype=np.int64)
        lact_of_lrel = {int(lr): j for j, lr in enumerate(act_lrel)}
        mode_lact = np.array([lact_of_lrel.get(int(lr), -1) for lr in mode_lrel], dtype=np.int64)

        # reduced L-operators over the state list (basis/hops_system.py:318-354)
        pos = {int(s): i for i, s in enumerate(st)}
        self.L = []
        for l in absL:
            c = info.l_ops[l]
            keep = [(pos[int(i)], pos[int(j)], d) for i, j, d in zip(c.row, c.col, c.data)
                    if int(i) in pos and int(j) in pos]
            if keep:
                r, cc, dd = zip(*keep)
                self.L.append(sp.csr_array(sp.coo_matrix((np.array(dd, dtype=np.complex128), (r, cc)),
                                                         shape=(n, n))))
            else:
                self.L.append(sp.csr_array((n, n), dtype=np.complex128))

        # multi-index CSR
        aux_ptr = np.zeros(nh + 1, dtype=np.int32)
        aux_mode, aux_val = [], []
        for i, a in enumerate(aux):
            aux_ptr[i + 1] = aux_ptr[i] + len(a)
            for (m, v) in a:
                aux_mode.append(m)
                aux_val.append(v)
        # connections (basis/hops_hierarchy.py:470-523)
        up = [[] for _ in range(nh)]
        dn = [[] for _ in range(nh)]
        for i, a in enumerate(aux):
            d = dict(a)
            for m in in_use:
                d2 = dict(d)
                d2[m] = d2.get(m, 0) + 1
                key = tuple((mm, d2[mm]) for mm in sorted(d2))
                j = index.get(key)
                if j is not None:
                    lm = rel_mode[m]
                    up[i].append((j, lm))
                    dn[j].append((i, lm, d2[m]))
        up_ptr = np.zeros(nh + 1, dtype=np.int32)
        dn_ptr = np.zeros(nh + 1, dtype=np.int32)
        for i in range(nh):
            up[i].sort(key=lambda x: x[1])
            dn[i].sort(key=lambda x: x[1])
            up_ptr[i + 1] = up_ptr[i] + len(up[i])
            dn_ptr[i + 1] = dn_ptr[i] + len(dn[i])
        self.up = up
        self.dn = dn
        up_aux = np.array([j for row in up for (j, lm) in row], dtype=np.int32)
        up_mode = np.array([lm for row in up for (j, lm) in row], dtype=np.int32)
        dn_aux = np.array([j for row in dn for (j, lm, v) in row], dtype=np.int32)
        dn_mode = np.array([lm for row in dn for (j, lm, v) in row], dtype=np.int32)
        dn_val = np.array([v for row in dn for (j, lm, v) in row], dtype=np.int32)

        # per relative L: row CSR and column CSR over states
        nl = len(absL)
        lr_ptr = np.zeros(nl * (n + 1), dtype=np.int32)
        lc_ptr = np.zeros(nl * (n + 1), dtype=np.int32)
        lr_col, lr_val, lc_row, lc_val = [], [], [], []
        base_r = base_c = 0
        rownnz = np.zeros((nl, n), dtype=np.int64)
        colnnz = np.zeros((nl, n), dtype=np.int64)
        for j, Lj in enumerate(self.L):
            csr = sp.csr_matrix(Lj)
            csr.sort_indices()
            csc = sp.csc_matrix(Lj)
            csc.sort_indices()
            lr_ptr[j * (n + 1):(j + 1) * (n + 1)] = csr.indptr + base_r
            lc_ptr[j * (n + 1):(j + 1) * (n + 1)] = csc.indptr + base_c
            base_r += csr.nnz
            base_c += csc.nnz
            lr_col.extend(csr.indices.tolist())
            lr_val.extend(csr.data.tolist())
            lc_row.extend(csc.indices.tolist())
            lc_val.extend(csc.data.tolist())
            rownnz[j] = np.diff(csr.indptr)
            colnnz[j] = np.diff(csc.indptr)
        # ELL width = max over rows (k, s) of the K2_kp1 + K2_km1 entries
        nm = len(modes)
        UP = np.zeros((nh, nm), dtype=np.int64)
        DN = np.zeros((nh, nm), dtype=np.int64)
        for i in range(nh):
            for (j, lm) in up[i]:
                UP[i, lm] += 1
            for (j, lm, v) in dn[i]:
                DN[i, lm] += 1
        cnt = UP @ rownnz[mode_lrel] + DN @ colnnz[mode_lrel]
        self.ell_w = int(cnt.max()) if cnt.size else 0

        # active L-operators: COO of L and L^2, LTC coefficients
        la_ptr = [0]
        la_row, la_col, la_val = [], [], []
        lq_ptr = [0]
        lq_row, lq_col, lq_val = [], [], []
        lt = info.lt_param[self.absL_active] if len(act_abs) else np.zeros(0, np.complex128)
        self.has_lt = bool(np.any(lt))
        Lsq = []
        for j, lr in enumerate(act_lrel):
            c = sp.coo_matrix(self.L[lr])
            la_row.extend(c.row.tolist())
            la_col.extend(c.col.tolist())
            la_val.extend(c.data.tolist())
            la_ptr.append(len(la_row))
            q = sp.csr_array(self.L[lr] @ self.L[lr])
            Lsq.append(q)
            qc = sp.coo_matrix(q)
            lq_row.extend(qc.row.tolist())
            lq_col.extend(qc.col.tolist())
            lq_val.extend(qc.data.tolist())
            lq_ptr.append(len(lq_row))
        # relative modes per active L (eom/eom_functions.py:32-62)
        am = [[] for _ in act_abs]
        for i, la in enumerate(mode_lact):
            if la >= 0:
                am[la].append(i)
        am_ptr = np.cumsum([0] + [len(x) for x in am]).astype(np.int32)
        am_mode = np.array([i for x in am for i in x], dtype=np.int32)
        # absolute modes -> active L index (-2 not in basis, -1 inactive)
        abs_lact = np.full(info.n_hmodes, -2, dtype=np.int32)
        for i, m in enumerate(modes):
            abs_lact[m] = mode_lact[i]
        # first-level auxiliaries e_m with an active L (eom/hops_eom.py:178-193)
        first = []
        for i, m in enumerate(modes):
            j = index.get(((int(m), 1),))
            if j is not None and mode_lact[i] >= 0:
                first.append((j, int(mode_lact[i]), i))
        # system-block pattern: union of H, active L and (LTC) L^2; with a
        # dense system block, of the L and L^2 entries only (-iH is applied
        # as a dense matrix, hd)
        Hn = info.H[np.ix_(st, st)]
        if dense_h is None:
            dense_h = n > DENSE_H_MIN_STATES and np.count_nonzero(Hn) > DENSE_H_MIN_ROW_NNZ * n
        self.dense_h = bool(dense_h)
        self.hd = -1j * Hn if self.dense_h else None
        pat = sp.csr_matrix((n, n), dtype=np.int8) if self.dense_h else sp.csr_matrix(np.abs(Hn) > 0, dtype=np.int8)
        for lr in act_lrel:
            pat = pat + sp.csr_matrix(abs(self.L[lr]) > 0, dtype=np.int8)
        if self.has_lt:
            for q in Lsq:
                pat = pat + sp.csr_matrix(abs(q) > 0, dtype=np.int8)
        pat = sp.csr_matrix(pat)
        pat.sort_indices()
        pat.data[:] = 1
        pat_ptr = pat.indptr.astype(np.int32)
        pat_col = pat.indices.astype(np.int32)
        pidx = {}
        for s in range(n):
            for p in range(pat_ptr[s], pat_ptr[s + 1]):
                pidx[(s, int(pat_col[p]))] = p
        pat_h = np.array([0.0 if self.dense_h else -1j * Hn[s, int(pat_col[p])] for s in range(n)
                          for p in range(pat_ptr[s], pat_ptr[s + 1])], dtype=np.complex128)
        contrib = [[] for _ in range(len(pat_col))]
        for j, lr in enumerate(act_lrel):
            c = sp.coo_matrix(self.L[lr])
            for r0, c0, v in zip(c.row, c.col, c.data):
                contrib[pidx[(int(r0), int(c0))]].append((j, 0, v))
            if self.has_lt:
                qc = sp.coo_matrix(Lsq[j])
                for r0, c0, v in zip(qc.row, qc.col, qc.data):
                    contrib[pidx[(int(r0), int(c0))]].append((j, 1, v))
        pc_ptr = np.cumsum([0] + [len(x) for x in contrib]).astype(np.int32)

        cplx = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).view(np.float64)
        i32 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int32))
        self.eom_kind = eom_kind
        self.n_mode = nm
        self.n_lop = nl
        self.n_act = len(act_abs)
        self.n_pat = int(len(pat_col))
        self.n_mode_abs = int(info.n_hmodes)
        self.mode_lact = mode_lact
        self.mode_lrel = mode_lrel
        self.act_lrel = act_lrel
        self.first = first
        self.tables = {
            "aux_ptr": i32(aux_ptr), "aux_mode": i32(aux_mode), "aux_val": i32(aux_val),
            "up_ptr": i32(up_ptr), "up_aux": up_aux, "up_mode": up_mode,
            "dn_ptr": i32(dn_ptr), "dn_aux": dn_aux, "dn_mode": dn_mode, "dn_val": dn_val,
            "lr_ptr": i32(lr_ptr), "lr_col": i32(lr_col), "lr_val": cplx(lr_val),
            "lc_ptr": i32(lc_ptr), "lc_row": i32(lc_row), "lc_val": cplx(lc_val),
            "mode_lrel": i32(mode_lrel), "mode_lact": i32(mode_lact), "mode_abs": i32(modes),
            "mode_g": cplx(info.G[modes]), "mode_w": cplx(info.W[modes]),
            "abs_g": cplx(info.G), "abs_w": cplx(info.W), "abs_lact": i32(abs_lact),
            "act_lrel": i32(act_lrel), "act_zrow": None, "am_ptr": i32(am_ptr), "am_mode": i32(am_mode),
            "la_ptr": i32(la_ptr), "la_row": i32(la_row), "la_col": i32(la_col), "la_val": cplx(la_val),
            "lq_ptr": i32(lq_ptr), "lq_row": i32(lq_row), "lq_col": i32(lq_col), "lq_val": cplx(lq_val),
            "lt": cplx(lt) if self.has_lt else None,
            "first_aux": i32([f[0] for f in first]), "first_lact": i32([f[1] for f in first]),
            "first_mode": i32([f[2] for f in first]),
            "pat_ptr": pat_ptr, "pat_col": pat_col, "pat_h": cplx(pat_h),
            "pc_ptr": pc_ptr, "pc_act": i32([e[0] for x in contrib for e in x]),
            "pc_kind": i32([e[1] for x in contrib for e in x]),
            "pc_val": cplx([e[2] for x in contrib for e in x]),
        }


# ---------------------------------------------------------------------------
# Static hierarchy filters (basis/hierarchy_functions.py:7-140,
# basis/hops_hierarchy.py:238-296)
# ---------------------------------------------------------------------------
def _keep_triangular(aux, flagged, kmax2):
    return sum(v for (m, v) in aux if flagged[m]) <= kmax2


def _keep_longedge(aux, flagged, kdepth):
    depth = sum(v for (m, v) in aux)
    return depth <= kdepth or sum(v for (m, v) in aux if flagged[m]) == 0 or len(aux) <= 1


def _keep_markovian(aux, flagged):
    depth = sum(v for (m, v) in aux)
    return depth == 0 or not (any(flagged[m] for (m, v) in aux) and depth > 1)


def apply_static_filters(aux_list, filters, n_hmodes):
    """Apply [(name, params), ...] in order (Triangular, LongEdge, Markovian)."""
    for name, prm in filters:
        if name == "Triangular":
            flags, kmax2 = list(prm[0]), prm[1]
            if len(flags) != n_hmodes:
                raise UnsupportedRequest("The number of entries in the list of static filter booleans does "
                                         "not match the number of hierarchy modes", "hierarchy_static_filter",
                                         override=True)
            aux_list = [a for a in aux_list if _keep_triangular(a, flags, kmax2)]
        elif name == "LongEdge":
            flags, kdepth = list(prm[0]), prm[1]
            if len(flags) != n_hmodes:
                raise UnsupportedRequest("The number of entries in the list of static filter booleans does "
                                         "not match the number of hierarchy modes", "hierarchy_static_filter",
                                         override=True)
            aux_list = [a for a in aux_list if _keep_longedge(a, flags, kdepth)]
        elif name == "Markovian":
            flags = list(prm)
            if len(flags) != n_hmodes:
                raise UnsupportedRequest("The number of entries in the list of static filter booleans does "
                                         "not match the number of hierarchy modes", "hierarchy_static_filter",
                                         override=True)
            aux_list = [a for a in aux_list if _keep_markovian(a, flags)]
        else:
            raise UnsupportedRequest(name, "hierarchy_static_filter")
    return aux_list
