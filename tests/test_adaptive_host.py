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
                               lambda y: tr.dsystem_dt(y, z_step[2], z_step[0], z_step[1])[0])
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
