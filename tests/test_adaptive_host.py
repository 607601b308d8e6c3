"""The product's host basis selection (mesohops_amd/adaptive.py BasisSelector)
equals the oracle's define_basis on the bases and hierarchy vectors of oracle
adaptive trajectories, with the oracle's derivative standing in for the
device derivative (no GPU needed)."""
import numpy as np
import pytest

import systems
from oracle import hops_oracle as O


def _markov_chain(nsite):
    g0, w0 = O.bcf_convert_dl_to_exp(50.0, 50.0, 295.0)
    import scipy.sparse as sp
    gw, ll = [], []
    for i in range(nsite):
        L = sp.coo_matrix(([1.0], ([i], [i])), shape=(nsite, nsite))
        gw += [[g0, w0], [-1j * np.imag(g0), 500.0]]
        ll += [L, L]
    H = (np.eye(nsite, k=1) + np.eye(nsite, k=-1)) * 50.0
    return dict(HAMILTONIAN=np.array(H, complex), GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll,
                ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw), [False, True] * nsite


CASES = {
    "dod_both": (systems.dimer_of_dimers(), None, 1e-3, 1e-3, 3, "NORMALIZED NONLINEAR"),
    "chain_hier_only": (systems.chain(6), None, 1e-3, 0.0, 4, "NORMALIZED NONLINEAR"),
    "chain_states_only": (systems.chain(8), None, 0.0, 2e-3, 2, "NONLINEAR"),
    "markov_chain": (*_markov_chain(7), 2e-3, 2e-3, 4, "NORMALIZED NONLINEAR"),
    # Triangular / LongEdge static filters on the flux up (hops_fluxfilters.py:372-501)
    "dod_triangular": (systems.dimer_of_dimers(), None, 1e-3, 1e-3, 3, "NORMALIZED NONLINEAR",
                       [("Triangular", [[True, False] * 4, 1])]),
    "chain_longedge": (systems.chain(6), None, 1e-3, 0.0, 4, "NORMALIZED NONLINEAR",
                       [("LongEdge", [[False, True] * 6, 1])]),
    "markov_triangular": (*_markov_chain(7), 2e-3, 2e-3, 4, "NORMALIZED NONLINEAR",
                          [("Triangular", [[True, False] * 7, 2]), ("LongEdge", [[True, False] * 7, 3])]),
    # off-diagonal (Peierls) L-operators: destination-state flux terms, the
    # noise operator in the state flux, and with LTC the T operators
    "peierls_both": (systems.peierls_chain(), None, 2e-3, 2e-3, 3, "NORMALIZED NONLINEAR"),
    "peierls_ltc_both": (systems.peierls_chain(True), None, 2e-3, 2e-3, 3, "NORMALIZED NONLINEAR"),
    "peierls_hier_only": (systems.peierls_chain(), None, 1e-3, 0.0, 3, "NONLINEAR"),
    "peierls_states_only": (systems.peierls_chain(True), None, 0.0, 1e-3, 2, "NORMALIZED NONLINEAR"),
    "peierls_ring_linear": (systems.peierls_ring(8, ltc=2.0 + 0.3j), None, 2e-3, 2e-3, 3, "LINEAR"),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_selector_matches_oracle_define_basis(name):
    from mesohops_amd.adaptive import BasisSelector
    from mesohops_amd.basis import BasisTables, SystemInfo
    sysp, markov, da, ds, K, eom, *rest = CASES[name]
    filters = rest[0] if rest else None
    nst = sysp["HAMILTONIAN"].shape[0]
    noise = {"SEED": 7, "MODEL": "FFT_FILTER", "TLEN": 200.0, "TAU": 1.0}
    tr = O.AdaptiveTrajectory(sysp, {"EQUATION_OF_MOTION": eom}, noise, hierarchy_param={"MAXHIER": K},
                              delta_a=da, delta_s=ds, markov=markov, filters=filters)
    psi0 = np.zeros(nst, complex)
    psi0[nst // 2] = 1.0
    tr.initialize(psi0)
    info = SystemInfo(sysp)
    sel = BasisSelector(info, da, ds, K, 0.01, markov, filters)
    n_checked = 0
    sizes = set()
    for step in range(12):
        tr.propagate(2.0, 2.0)
        b = tr.basis
        z_step = tr.z_step(tr.z_mem)
        ref = tr.ad.define_basis(tr.dsystem_dt, b, tr.phi, 2.0, z_step)
        tb = BasisTables(info, b.state_list, b.aux_list, 2)
        assert tb.aux_list == b.aux_list
        got = sel.define_basis(tb, tr.phi, 2.0,
                               lambda y: tr.dsystem_dt(y, z_step[2], z_step[0], z_step[1])[0],
                               lambda: z_step)
        assert got[0] == ref[0], (step, got[0], ref[0])
        assert set(got[1]) == set(ref[1]), step
        sizes.add((b.n_state, b.n_hier))
        n_checked += 1
    assert n_checked == 12
    # the basis moved during the run, so the comparison covered several bases
    assert len(sizes) > 1, sizes


def test_threshold_closed_form():
    """error_threshold: the largest sorted error whose prefix sum stays below
    the bound (basis/basis_functions.py:8-41), on hand-worked cases."""
    from mesohops_amd.adaptive import error_threshold
    e = np.array([1.0, 2.0, 3.0, 4.0])
    assert error_threshold(e, 3.5) == 2.0        # 1+2 = 3 < 3.5 <= 6
    assert error_threshold(e, 0.5) == 0.0        # even the smallest is too large
    assert error_threshold(e, 100.0) == np.inf   # everything can be dropped
    assert error_threshold(e, 3.5, offset=1.0) == 1.0
    assert error_threshold(np.array([]), 1.0) == 0.0


def _strong_ltc_peierls(c):
    p = systems.peierls_chain(True)
    p["PARAM_LT_CORR"] = [c] * len(p["L_LT_CORR"])
    return p


# off-diagonal L-operators, run from a state that reaches the middle of the
# chain only through the Peierls couplings; strong low-temperature terms
TERM_CASES = {
    "peierls_both": (systems.peierls_chain(), 2e-3, 3e-4, 3, "NORMALIZED NONLINEAR", 0),
    "peierls_strong_ltc": (_strong_ltc_peierls(40.0 + 15j), 2e-3, 3e-4, 3, "NORMALIZED NONLINEAR", 0),
    "peierls_states_nonlinear": (_strong_ltc_peierls(40.0 + 15j), 0.0, 2e-4, 2, "NONLINEAR", 0),
    "peierls_ring_linear": (systems.peierls_ring(8, ltc=30.0 + 5j), 2e-3, 3e-4, 3, "LINEAR", 3),
}


def _oracle_terms(ad, dsys, b, phi, dt, z):
    """The oracle's error terms of one define_basis call: per auxiliary, per
    basis state, and per outside state (the steps of AdaptiveState.state_basis)."""
    n, nh, nm = b.n_state, b.n_hier, len(b.modes)
    Pn = phi / np.linalg.norm(phi[:n])
    err_h = aux_new = aux_bound = None
    if ad.adaptive_h:
        err_h, E_ud = ad.hier_stable_error(dsys, b, Pn, dt, z)
        idx_st, aux_st = O.basis_from_list(err_h, ad.delta_a ** 2 / 2.0, b.aux_list)
        dbound = ad.delta_a ** 2 - np.sum(err_h[np.setdiff1d(np.arange(nh), idx_st)])
        aux_bound = ad.boundary_hier(b, E_ud, idx_st, dbound)
        aux_new = list(set(aux_st) | set(aux_bound))
        stable_idx = sorted(b.aux_index[a] for a in aux_st)
    else:
        aux_new, aux_bound, stable_idx = list(b.aux_list), [], list(range(nh))
    if not ad.adaptive_s:
        return err_h, None, {}
    err_s = ad.state_stable_error(dsys, b, Pn, dt, z, stable_idx, aux_bound, aux_new)
    rel_st, st_st = O.basis_from_list(err_s, ad.delta_s ** 2 / 2.0, list(b.state_list))
    sc = ad.list_sc(b)
    avg = ad.expectations(b, Pn[:n])
    T_phys, T_hier = ad.ltc_operators(b, avg)
    M2, M2_off, dests = ad.tables(b)
    sc_dest, flux = [], []
    if len(M2_off):
        M2_sc = []
        for i, d in enumerate(dests):
            if d in sc:
                sc_dest.append(sc.index(d))
                M2_sc.append(M2_off[i])
        F_dn_o, F_up_o = ad.state_filters(b, aux_new)
        K2 = ad.K2_aux_by_mode(b)
        flux = list(O.error_flux_up_by_dest_state(Pn, n, nh, nm, b.w, K2, M2_sc, stable_idx, F_up_o, rel_st)
                    + O.error_flux_down_by_dest_state(Pn, n, nh, nm, b.g, b.w, M2_sc, stable_idx, F_dn_o, rel_st))
    out = {}
    if len(rel_st) < ad.sysd["NSTATES"] and sc:
        C2 = O._C2(Pn, n, nh)[np.ix_(list(rel_st), stable_idx)]
        H2 = ad._H2(b, z)
        if T_phys is None:
            E = np.sum(np.abs(H2[np.ix_(sc, st_st)] @ (C2 / O.HBAR)) ** 2, axis=1)
        else:
            E = (np.abs((H2 + T_phys)[np.ix_(sc, st_st)] @ (C2[:, :1] / O.HBAR)) ** 2)[:, 0]
            if C2.shape[1] > 1:
                E = E + np.sum(np.abs((H2 + T_hier)[np.ix_(sc, st_st)] @ (C2[:, 1:] / O.HBAR)) ** 2, axis=1)
        E = np.array(E, dtype=float)
        for i, f in zip(sc_dest, flux):
            E[i] += f
        out = dict(zip(sc, E))
    return err_h, err_s, out


@pytest.mark.parametrize("name", sorted(TERM_CASES))
def test_selector_error_terms_match_oracle(name):
    """Off-diagonal L-operators: the selector's error per auxiliary, per
    basis state and per outside state equal the oracle's at every step to
    1e-10 relative, so each term (destination tables, noise operator,
    low-temperature operators, flux into outside destinations) is checked,
    not only the selections it moves."""
    from mesohops_amd.adaptive import BasisSelector
    from mesohops_amd.basis import BasisTables, SystemInfo
    sysp, da, ds, K, eom, s0 = TERM_CASES[name]
    nst = sysp["HAMILTONIAN"].shape[0]
    noise = {"SEED": 7, "MODEL": "FFT_FILTER", "TLEN": 200.0, "TAU": 1.0}
    tr = O.AdaptiveTrajectory(sysp, {"EQUATION_OF_MOTION": eom}, noise, hierarchy_param={"MAXHIER": K},
                              delta_a=da, delta_s=ds)
    psi0 = np.zeros(nst, complex)
    psi0[s0] = 1.0
    tr.initialize(psi0)
    info = SystemInfo(sysp)
    sel = BasisSelector(info, da, ds, K, 0.01)
    n_out = 0
    for step in range(15):
        tr.propagate(2.0, 2.0)
        b = tr.basis
        z = tr.z_step(tr.z_mem)
        tb = BasisTables(info, b.state_list, b.aux_list, 2)
        got = sel.define_basis(tb, tr.phi, 2.0, lambda y: tr.dsystem_dt(y, z[2], z[0], z[1])[0], lambda: z)
        ref = tr.ad.define_basis(tr.dsystem_dt, b, tr.phi, 2.0, z)
        assert got[0] == ref[0] and set(got[1]) == set(ref[1]), step
        eh, es, eo = _oracle_terms(tr.ad, tr.dsystem_dt, b, tr.phi, 2.0, z)
        if eh is not None:
            np.testing.assert_allclose(sel.err_aux, eh, rtol=1e-10, atol=1e-14 * np.max(eh))
        if es is not None:
            np.testing.assert_allclose(sel.err_state, es, rtol=1e-10, atol=1e-14 * np.max(es))
        assert sorted(sel.err_outside) == sorted(eo), step
        for d, v in eo.items():
            assert abs(sel.err_outside[d] - v) <= 1e-10 * abs(v) + 1e-300, (step, d)
        n_out += len(eo)
    assert n_out > 0
