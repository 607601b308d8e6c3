"""The native runtime's host half (mesohops_amd/csrc/hops_runtime.cpp), without
a GPU: its basis tables equal mesohops_amd/basis.py::BasisTables (which the
oracle pins), and its basis selection equals the oracle's define_basis on the
bases and hierarchy vectors of oracle adaptive trajectories (derivative sums
from the oracle's dsystem_dt)."""
import ctypes
import math

import numpy as np
import pytest

import systems
from oracle import hops_oracle as O

INT_TABLES = ["aux_ptr", "aux_mode", "aux_val", "up_ptr", "up_aux", "up_mode", "dn_ptr", "dn_aux", "dn_mode",
              "dn_val", "lr_ptr", "lr_col", "lc_ptr", "lc_row", "mode_lrel", "mode_lact", "mode_abs", "abs_lact",
              "act_lrel", "am_ptr", "am_mode", "la_ptr", "la_row", "la_col", "lq_ptr", "lq_row", "lq_col",
              "first_aux", "first_lact", "first_mode", "pat_ptr", "pat_col", "pc_ptr", "pc_act", "pc_kind"]
CPLX_TABLES = ["lr_val", "lc_val", "mode_g", "mode_w", "abs_g", "abs_w", "la_val", "lq_val", "pat_h", "pc_val"]


def _markov_chain(nsite):
    import scipy.sparse as sp
    g0, w0 = O.bcf_convert_dl_to_exp(50.0, 50.0, 295.0)
    gw, ll = [], []
    for i in range(nsite):
        L = sp.coo_matrix(([1.0], ([i], [i])), shape=(nsite, nsite))
        gw += [[g0, w0], [-1j * np.imag(g0), 500.0]]
        ll += [L, L]
    H = (np.eye(nsite, k=1) + np.eye(nsite, k=-1)) * 50.0
    return dict(HAMILTONIAN=np.array(H, complex), GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll,
                ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw), [False, True] * nsite


def _bases(sysp, rng):
    from mesohops_amd.basis import SystemInfo, triangular_aux_list
    info = SystemInfo(sysp)
    full = triangular_aux_list(info.n_hmodes, 3)
    out = [(list(range(info.nstates)), full)]
    for _ in range(3):
        k = rng.integers(1, info.nstates + 1)
        st = sorted(rng.choice(info.nstates, size=k, replace=False).tolist())
        aux = [()] + [a for a in full[1:] if rng.random() < 0.5]
        out.append((st, aux))
    return info, out


@pytest.mark.parametrize("name", ["dod", "chain_ltc", "three_diag", "three_general", "lattice_ltc"])
def test_runtime_tables_equal_host_tables(name):
    from mesohops_amd.basis import BasisTables
    from mesohops_amd.ragged import DebugBasis, RtSystem
    sysp = {"dod": systems.dimer_of_dimers(), "chain_ltc": systems.chain(5, ltc=12.0 - 3.0j),
            "three_diag": systems.three_site(False), "three_general": systems.three_site(True),
            "lattice_ltc": systems.lattice(3, 2)}[name]
    rng = np.random.default_rng(5)
    info, bases = _bases(sysp, rng)
    rs = RtSystem(info)
    for states, aux in bases:
        try:
            tb = BasisTables(info, states, aux, 2)
        except Exception:
            continue
        db = DebugBasis(rs, states, aux, 2)
        sizes = db.table("sizes", np.int32)
        assert list(sizes[:8]) == [tb.n_state, tb.n_hier, tb.n_mode, tb.n_lop, tb.n_act, len(tb.first), tb.ell_w,
                                   tb.n_pat], (name, states)
        for t in INT_TABLES:
            np.testing.assert_array_equal(db.table(t, np.int32), np.asarray(tb.tables[t], np.int32).reshape(-1),
                                          err_msg=t)
        for t in CPLX_TABLES:
            np.testing.assert_array_equal(db.table(t, np.complex128), tb.tables[t].view(np.complex128), err_msg=t)
        assert bool(db.table("has_lt", np.int32)[0]) == tb.has_lt
        if tb.has_lt:
            np.testing.assert_array_equal(db.table("lt", np.complex128), tb.tables["lt"].view(np.complex128))
        np.testing.assert_array_equal(db.table("act_zrow", np.int32), tb.absL_active.astype(np.int32))


CASES = {
    "dod_both": (systems.dimer_of_dimers(), None, 1e-3, 1e-3, 3, "NORMALIZED NONLINEAR"),
    "chain_hier_only": (systems.chain(6), None, 1e-3, 0.0, 4, "NORMALIZED NONLINEAR"),
    "chain_states_only": (systems.chain(8), None, 0.0, 2e-3, 2, "NONLINEAR"),
    "markov_chain": (*_markov_chain(7), 2e-3, 2e-3, 4, "NORMALIZED NONLINEAR"),
    "chain12": (systems.chain(12), None, 2e-3, 2e-3, 5, "NORMALIZED NONLINEAR"),
    "lattice4_ltc": (systems.lattice(4, 4), None, 2e-3, 2e-3, 3, "NORMALIZED NONLINEAR"),
    # Triangular / LongEdge static filters on the flux up (hops_fluxfilters.py:372-501)
    "dod_triangular": (systems.dimer_of_dimers(), None, 1e-3, 1e-3, 3, "NORMALIZED NONLINEAR",
                       [("Triangular", [[True, False] * 4, 1])]),
    "chain_longedge": (systems.chain(6), None, 1e-3, 0.0, 4, "NORMALIZED NONLINEAR",
                       [("LongEdge", [[False, True] * 6, 1])]),
    "markov_triangular_longedge": (*_markov_chain(7), 2e-3, 2e-3, 4, "NORMALIZED NONLINEAR",
                                   [("Triangular", [[True, False] * 7, 2]), ("LongEdge", [[True, False] * 7, 3])]),
    # off-diagonal (Peierls) L-operators (hops_basis.py:914-998, :1030-1100)
    "peierls_both": (systems.peierls_chain(), None, 2e-3, 2e-3, 3, "NORMALIZED NONLINEAR"),
    "peierls_ltc_hier_only": (systems.peierls_chain(True), None, 1e-3, 0.0, 3, "NONLINEAR"),
    "peierls_ring_linear": (systems.peierls_ring(8, ltc=2.0 + 0.3j), None, 2e-3, 2e-3, 3, "LINEAR"),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_runtime_selection_matches_oracle(name):
    from mesohops_amd.basis import SystemInfo
    from mesohops_amd.ragged import DebugBasis, RtSystem, traj_param
    sysp, markov, da, ds, K, eom, *rest = CASES[name]
    filters = rest[0] if rest else None
    nst = sysp["HAMILTONIAN"].shape[0]
    noise = {"SEED": 9, "MODEL": "FFT_FILTER", "TLEN": 200.0, "TAU": 1.0}
    tr = O.AdaptiveTrajectory(sysp, {"EQUATION_OF_MOTION": eom}, noise, hierarchy_param={"MAXHIER": K},
                              delta_a=da, delta_s=ds, markov=markov, filters=filters)
    psi0 = np.zeros(nst, complex)
    psi0[nst // 2] = 1.0
    tr.initialize(psi0)
    info = SystemInfo(sysp)
    rs = RtSystem(info)
    prm = traj_param(2, K, da, ds, 1, 0.01, 2.0, 1, 10, 0, markov=markov, filters=filters)
    dt = 2.0
    sizes = set()
    for step in range(14):
        tr.propagate(dt, dt)
        b = tr.basis
        z1, z2, zm = tr.z_step(tr.z_mem)
        ref = tr.ad.define_basis(tr.dsystem_dt, b, tr.phi, dt, [z1, z2, zm])
        n, nh = b.n_state, b.n_hier
        pn = tr.phi / np.linalg.norm(tr.phi[:n])
        D = tr.dsystem_dt(pn, zm, z1, z2)[0] / O.HBAR + pn / dt
        e_aux = np.sum(np.abs(D.reshape(nh, n)) ** 2, axis=1)
        db = DebugBasis(rs, b.state_list, b.aux_list, 2)
        db.set_noise(z1, z2, zm)
        stable = db.select(prm, pn, e_aux, dt)
        e_state = None
        if ds > 0:
            pst = np.zeros_like(pn)
            pst.reshape(nh, n)[stable] = pn.reshape(nh, n)[stable]
            D2 = tr.dsystem_dt(pst, zm, z1, z2)[0] / O.HBAR + pst / dt
            e_state = np.sum(np.abs(D2.reshape(nh, n)[stable]) ** 2, axis=0)
        db.select_states(e_state)
        st_new, aux_new = db.result()
        assert st_new == [int(s) for s in ref[0]], (step, st_new, ref[0])
        assert set(aux_new) == set(ref[1]), step
        sizes.add((n, nh))
    assert len(sizes) > 1


def test_threshold_prefix_sort_matches_full_sort():
    """The selection's error threshold sorts only the prefix of the errors it
    needs (values binned by binary exponent).  It equals the oracle's
    determine_error_thresh on the fully sorted errors: heavy-tailed values over
    30 decades, zeros, ties, max_err from 0 to above the total, max_err equal
    to an exact prefix sum, and offsets below and above max_err."""
    from mesohops_amd._lib import lib
    L = lib()
    rng = np.random.default_rng(11)
    for trial in range(240):
        n = int(rng.choice([0, 1, 2, 7, 100, 5000, 70000]))
        v = 10.0 ** rng.uniform(-30, 0, n)
        if n and trial % 3 == 0:
            v[rng.random(n) < 0.2] = 0.0
        if n and trial % 5 == 0:
            v = np.round(v, 3)
        v = np.ascontiguousarray(v)
        total = float(v.sum())
        choices = [0.0, total * 1e-8, total * 1e-3, total * 0.5, total, total * 2 + 1e-300]
        if n:
            choices.append(float(np.cumsum(np.sort(v))[rng.integers(n)]))
        max_err = float(rng.choice(choices))
        offset = float(rng.choice([0.0, max_err * 0.3, max_err * 1.5]))
        below = ctypes.c_double()
        thr = L.hops_rt_debug_threshold(v.ctypes.data, n, max_err, offset, ctypes.byref(below))
        ref = O.determine_error_thresh(np.sort(v), max_err, offset)
        assert thr == ref, (trial, n, max_err, offset)
        assert below.value == pytest.approx(math.fsum(v[v <= ref]), rel=1e-12, abs=0), trial


def test_threshold_crossing_far_below_total():
    """The shape of the selection's error lists: a few large errors carry
    almost all of the total, and the threshold crossing lies among the small
    ones, max_err being many decades below the total.  The threshold equals
    the oracle's on the fully sorted values, also when max_err is a running
    sum of the sorted values exactly or one ulp either side of one (where
    another summation order could decide differently)."""
    from mesohops_amd._lib import lib
    L = lib()
    rng = np.random.default_rng(5)
    for trial in range(120):
        n_small = int(rng.choice([300, 5000, 40000]))
        small = 10.0 ** rng.uniform(-16, -9, n_small)
        big = 10.0 ** rng.uniform(-3, 1, int(rng.integers(1, 2000)))
        v = np.concatenate([small, big])
        rng.shuffle(v)
        v = np.ascontiguousarray(v)
        cs = np.cumsum(np.sort(v))
        k = int(rng.integers(1, n_small))
        base = float(cs[k])
        max_err = float(rng.choice([base, np.nextafter(base, np.inf), np.nextafter(base, -np.inf),
                                    base * (1 + 1e-9), float(small.sum()) * 0.5]))
        offset = float(rng.choice([0.0, max_err * 0.25]))
        below = ctypes.c_double()
        thr = L.hops_rt_debug_threshold(v.ctypes.data, len(v), max_err, offset, ctypes.byref(below))
        ref = O.determine_error_thresh(np.sort(v), max_err, offset)
        assert thr == ref, (trial, len(v), max_err, offset)


@pytest.mark.parametrize("name", ["peierls_both", "peierls_strong_ltc", "peierls_states_nonlinear",
                                  "peierls_ring_linear"])
def test_runtime_error_terms_match_oracle(name):
    """Off-diagonal L-operators in the native runtime's selection: the error
    per auxiliary, per basis state and per outside state equal the oracle's
    at every step to 1e-10 relative (destination tables, noise operator,
    low-temperature operators, flux into outside destinations), and the
    selected bases are the oracle's."""
    from mesohops_amd.basis import SystemInfo
    from mesohops_amd.ragged import DebugBasis, RtSystem, traj_param
    from test_adaptive_host import TERM_CASES, _oracle_terms
    sysp, da, ds, K, eom, s0 = TERM_CASES[name]
    nst = sysp["HAMILTONIAN"].shape[0]
    noise = {"SEED": 7, "MODEL": "FFT_FILTER", "TLEN": 200.0, "TAU": 1.0}
    tr = O.AdaptiveTrajectory(sysp, {"EQUATION_OF_MOTION": eom}, noise, hierarchy_param={"MAXHIER": K},
                              delta_a=da, delta_s=ds)
    psi0 = np.zeros(nst, complex)
    psi0[s0] = 1.0
    tr.initialize(psi0)
    rs = RtSystem(SystemInfo(sysp))
    prm = traj_param(2, K, da, ds, 1, 0.01, 2.0, 1, 10, 0)
    dt = 2.0
    n_out = 0
    for step in range(15):
        tr.propagate(dt, dt)
        b = tr.basis
        z1, z2, zm = tr.z_step(tr.z_mem)
        ref = tr.ad.define_basis(tr.dsystem_dt, b, tr.phi, dt, [z1, z2, zm])
        n, nh = b.n_state, b.n_hier
        pn = tr.phi / np.linalg.norm(tr.phi[:n])
        D = tr.dsystem_dt(pn, zm, z1, z2)[0] / O.HBAR + pn / dt
        e_aux = np.sum(np.abs(D.reshape(nh, n)) ** 2, axis=1)
        db = DebugBasis(rs, b.state_list, b.aux_list, 2)
        db.set_noise(z1, z2, zm)
        stable = db.select(prm, pn, e_aux, dt)
        e_state = None
        if ds > 0:
            pst = np.zeros_like(pn)
            pst.reshape(nh, n)[stable] = pn.reshape(nh, n)[stable]
            D2 = tr.dsystem_dt(pst, zm, z1, z2)[0] / O.HBAR + pst / dt
            e_state = np.sum(np.abs(D2.reshape(nh, n)[stable]) ** 2, axis=0)
        db.select_states(e_state)
        st_new, aux_new = db.result()
        assert st_new == [int(s) for s in ref[0]], (step, st_new, ref[0])
        assert set(aux_new) == set(ref[1]), step
        eh, es, eo = _oracle_terms(tr.ad, tr.dsystem_dt, b, tr.phi, dt, [z1, z2, zm])
        if eh is not None:
            np.testing.assert_allclose(db.errors("aux"), eh, rtol=1e-10, atol=1e-14 * np.max(eh))
        if es is not None:
            np.testing.assert_allclose(db.errors("state"), es, rtol=1e-10, atol=1e-14 * np.max(es))
        got_out = db.errors("outside") if es is not None else {}
        assert sorted(got_out) == sorted(eo), (step, got_out, eo)
        for d, v in eo.items():
            assert abs(got_out[d] - v) <= 1e-10 * abs(v) + 1e-300, (step, d, got_out[d], v)
        n_out += len(eo)
    assert n_out > 0


def test_zobrist_collisions_fail_loudly():
    """Auxiliary identity in the native runtime is a 128-bit Zobrist key, and
    every key hit is confirmed against the auxiliaries themselves: with two
    (mode, depth) entries given the same key words, a basis holding both is
    refused, and a selection whose lookups meet the collision raises instead
    of returning a basis."""
    from mesohops_amd import _lib
    from mesohops_amd.basis import SystemInfo
    from mesohops_amd.errors import EngineError
    from mesohops_amd.ragged import DebugBasis, RtSystem, traj_param
    sysp = systems.chain(4)
    info = SystemInfo(sysp)
    rs = RtSystem(info)
    L = _lib.lib()
    # distinct auxiliaries are accepted before the tables are changed
    DebugBasis(rs, [0, 1, 2, 3], [(), ((0, 1),), ((2, 1),)], 2)
    for m in (0, 2):
        assert L.hops_rt_debug_set_zobrist(rs.handle, m, 1, 0x1234567890ABCDEF, 0x0FEDCBA987654321) == 0
    with pytest.raises(EngineError, match="Zobrist"):
        DebugBasis(rs, [0, 1, 2, 3], [(), ((0, 1),), ((2, 1),)], 2)
    # the first-level lookup of e_2 hits e_0: even a basis holding only e_0
    # is refused
    with pytest.raises(EngineError, match="Zobrist"):
        DebugBasis(rs, [0, 1, 2, 3], [(), ((0, 1),)], 2)
    # a collision one level deeper that only a selection meets: 2 e_4 gets the
    # key of e_5 + e_6; the basis holds e_4 and e_5 + e_6, the selection grows
    # the hierarchy (tiny delta) and either refuses the selection or returns a
    # list that no basis accepts
    rs2 = RtSystem(info)
    A, B2 = 0x1111222233334444, 0x5555666677778888
    C, D = 0x9999AAAABBBBCCCC, 0xDDDDEEEEFFFF0001
    assert L.hops_rt_debug_set_zobrist(rs2.handle, 5, 1, A, C) == 0
    assert L.hops_rt_debug_set_zobrist(rs2.handle, 6, 1, B2, D) == 0
    assert L.hops_rt_debug_set_zobrist(rs2.handle, 4, 2, A ^ B2, C ^ D) == 0
    aux = [(), ((4, 1),), ((5, 1),), ((6, 1),), ((5, 1), (6, 1))]
    db = DebugBasis(rs2, [0, 1, 2, 3], aux, 2)
    rng = np.random.default_rng(2)
    pn = rng.normal(size=4 * len(aux)) + 1j * rng.normal(size=4 * len(aux))
    pn /= np.linalg.norm(pn[:4])
    prm = traj_param(2, 3, 1e-9, 0.0, 1, 0.0, 1.0, 1, 10, 0)
    with pytest.raises(EngineError, match="Zobrist"):
        db.select(prm, pn, np.ones(len(aux)), 1.0)
        db.select_states(None)
        st_new, aux_new = db.result()
        assert ((4, 2),) in aux_new
        DebugBasis(rs2, st_new, aux_new, 2)
