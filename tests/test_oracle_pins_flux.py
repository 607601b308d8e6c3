"""Pins of the oracle's adaptive error bound (the derivative-error terms that
choose the hierarchy and state bases).

- Hand values of tests/test_basis_functions_adaptive.py: deletion flux,
  state flux out of each auxiliary and each state (with and without the
  low-temperature operators), boundary-state flux, and the mode-from-state
  and expectation tables.
- Flux up / flux down for general (Peierls, mixed, mismatched) L-operators:
  the oracle's table-based sums against brute-force loops over
  (mode, auxiliary, destination, source) written from the definition of each
  term, on the inputs of the same reference tests.
- HopsBasis-level answers of tests/test_hops_basis.py: the per-destination
  tables, the noise and low-temperature operators of off-diagonal couplings,
  the analytic state error, the boundary-auxiliary binning (three sets), and
  the stable/boundary auxiliaries and states chosen on the reference's own
  noise fixtures.
CPU only."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import hops_oracle as O

HB = O.HBAR
FX = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_fixtures.npz"))


# --- tests/test_basis_functions_adaptive.py hand values ----------------------
def test_error_deriv_hand_values():
    """:15-46."""
    phi = np.ones(6, complex)

    def d_if(p, z1, z2, z3):
        return [np.array([1, 1, 1, 1, 0, 0], complex)]

    def d_else(p, z1, z2, z3):
        return [np.ones(6, complex)]

    E = O.error_deriv(d_if, phi, [[0, 0], [0, 0], [1, 1, 1, 1]], 2, 3, 2, stable_idx=[0, 1])
    np.testing.assert_allclose(E, np.full((2, 2), abs(1 / HB + 0.5) ** 2))
    E = O.error_deriv(d_else, phi, [[0, 0], [0, 0], [1, 1, 1, 1]], 2, 3, 2)
    np.testing.assert_allclose(E, np.full((2, 3), abs(1 / HB + 0.5) ** 2))


def test_error_sflux_hier_hand_values():
    """:48-90."""
    hs = np.zeros((5, 5))
    hs[4, 3] = hs[3, 4] = 100
    hs[1, 2] = hs[2, 1] = 50
    phi = np.ones(9)
    E = O.error_sflux_hier(phi, [1, 2, 3], [4], 3, 3, sp.csr_array(hs))
    np.testing.assert_allclose(E, [1e4 / HB ** 2] * 3)
    tp = np.zeros((5, 5), complex)
    tp[0, 3] = tp[3, 0] = 20
    tp[4, 3] = tp[3, 4] = -50
    tp[4, 4] = 5e8
    th = np.zeros((5, 5), complex)
    th[0, 2] = th[2, 0] = -10
    th[4, 0] = 10 - 1e8j
    th[0, 4] = 10 + 1e8j
    E = O.error_sflux_hier(phi, [1, 2, 3], [0, 4], 3, 3, sp.csr_array(hs), sp.csr_array(tp), sp.csr_array(th))
    np.testing.assert_allclose(E, np.array([2500 + 400, 10000 + 100, 10000 + 100]) / HB ** 2)


def test_error_sflux_stable_state_hand_values():
    """:931-978."""
    hs = np.zeros((5, 5), complex)
    hs[1, 1] = -50000
    hs[1, 2] = hs[2, 1] = 100
    hs[4, 2] = hs[2, 4] = 100
    hs[0, 4] = hs[4, 0] = 10000
    tp = np.zeros_like(hs)
    tp[0, 0] = 20000
    tp[1, 1] = 30000
    tp[0, 4] = -1000j
    tp[4, 0] = 1000j
    tp[1, 4] = tp[4, 1] = 100
    tp[1, 2] = tp[2, 1] = -50
    th = np.zeros_like(hs)
    th[0, 0] = 20000
    th[1, 1] = 30000
    th[0, 4] = -1000j
    th[4, 0] = 1000j
    th[3, 0] = th[0, 3] = -200
    phi = np.ones(12)
    E = O.error_sflux_stable_state(phi, 3, 4, sp.csr_array(hs), [0, 1, 2], [1, 2, 3])
    np.testing.assert_allclose(E, np.array([30000, 60000, 0]) / HB ** 2)
    E = O.error_sflux_stable_state(phi, 3, 4, sp.csr_array(hs), [0, 1, 2], [1, 2, 3],
                                   sp.csr_array(tp), sp.csr_array(th))
    np.testing.assert_allclose(E, np.array([32500, 52500, 80000]) / HB ** 2)


def _chain10():
    hs = np.zeros((10, 10))
    for i, v in enumerate([10, 20, 30, 40, 50, 60, 70, 80, 90]):
        hs[i, i + 1] = hs[i + 1, i] = v
    hs[3, 8] = hs[8, 3] = 1000
    return hs


def test_error_sflux_boundary_state_hand_values():
    """:980-1127."""
    hs = sp.csr_array(_chain10())
    phi = np.ones(20)
    s0, sc = [3, 4, 5, 6], [0, 1, 2, 7, 9]
    sc_dest, flux = [1, 2, 4], [0.1, 0.2, 0.0]
    idx, E = O.error_sflux_boundary_state(phi, s0, sc, 5, 4, hs, np.arange(10), [0, 1, 2], sc_dest, flux)
    assert len(idx) == 0 and len(E) == 0
    st = [0, 1, 2, 3]
    idx, E = O.error_sflux_boundary_state(phi, s0, sc, 5, 4, hs, st, [0, 1, 2], sc_dest, flux)
    assert list(idx) == [1, 2, 7]
    np.testing.assert_allclose(E, [0.1, 3 * 30 ** 2 / HB ** 2 + 0.2, 3 * 70 ** 2 / HB ** 2])
    idx, E = O.error_sflux_boundary_state(phi, s0, sc, 5, 4, hs, st, [0, 1, 2], [], [])
    assert list(idx) == [2, 7]
    np.testing.assert_allclose(E, [3 * 30 ** 2 / HB ** 2, 3 * 70 ** 2 / HB ** 2])
    tp = np.zeros((10, 10), complex)
    tp[0, 3] = tp[3, 0] = 20
    tp[3, 4] = tp[4, 3] = 5e8
    tp[2, 3] = tp[3, 2] = -5
    th = np.zeros((10, 10), complex)
    th[0, 4] = -10
    th[4, 3] = 10 - 1e8j
    th[3, 4] = 10 + 1e8j
    idx, E = O.error_sflux_boundary_state(phi, s0, sc, 5, 4, hs, st, [0, 1, 2], sc_dest, flux,
                                          sp.csr_array(tp), sp.csr_array(th))
    assert list(idx) == [0, 1, 2, 7]
    np.testing.assert_allclose(E, [(400 + 100 + 100) / HB ** 2, 0.1, (625 + 900 + 900) / HB ** 2 + 0.2,
                                   3 * 70 ** 2 / HB ** 2], rtol=1e-8)


LOP_A = np.array([[1, 0, -1j], [0, 1, 0], [1j, 0, -2]])
LOP_B = np.array([[0, 1, 0], [1, 0, 1], [1, 0, 1]])


def test_mode_from_state_tables_hand_values():
    """:209-307 (destinations covering every state)."""
    diag, off = O.mode_from_state_tables([LOP_A, LOP_B], [0, 2], [0, 1, 2])
    np.testing.assert_allclose(off[0], [[0, -1j], [0, 0]])
    np.testing.assert_allclose(off[1], [[0, 0], [1, 1]])
    np.testing.assert_allclose(off[2], [[1j, 0], [1, 0]])
    np.testing.assert_allclose(diag, [[1, -2], [0, 1]])
    _, off = O.mode_from_state_tables([LOP_A, LOP_B], [0, 2], [0])
    np.testing.assert_allclose(off[0], [[0, -1j], [0, 0]])


def test_expectation_table_hand_values():
    """:659-713."""
    np.testing.assert_allclose(O.expectation_table([1.0, -0.5], 3), [[1.0, 1.0, 1.0], [-0.5, -0.5, -0.5]])


# --- flux up / down against brute force ------------------------------------
PHI = np.array([0.1j, -0.5j, -0.3j, 0.2, -1.5j, 0.4])      # 2 states x 3 auxiliaries
C2 = PHI.reshape(3, 2).T
W4 = np.array([50.0, 500.0, 50.0, 500.0], complex)
G4 = np.full(4, 1000.0 + 1000j)
F_EXCL = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 0], [0, 0, 0]])
PX = np.array([[0, 1], [1, 0]])
PY = np.array([[0, -1j], [1j, 0]])
L3 = np.array([[0, -1j, 1], [1j, 0, 1], [1, 1, 0]])
CASES = [   # (L-operator of each mode, source states, destination states)
    ([PX, PX, PY, PY], [0, 1], [0, 1]),
    ([np.diag([0.3, -0.7])] * 2 + [PY] * 2, [0, 1], [0, 1]),
    ([np.diag([0.3, -0.7, 0.5])] * 2 + [L3] * 2, [0, 1], [0, 1, 2]),
    ([np.diag([0.3, -0.7, 0.5])] * 2 + [L3] * 2, [1, 2], [0, 1, 2]),
    ([np.diag([0.3, -0.7, 0.5])] * 2 + [L3] * 2, [0, 1], [0]),
]


def _tables(Ls, states, dests):
    diag, off = O.mode_from_state_tables(Ls, states, dests)
    # a source state that is no destination receives no diagonal flux
    for i, s in enumerate(states):
        if s not in dests:
            diag[:, i] = 0
    return diag, off


def _bf_up(Ls, w, kvec, C, states, dests, F=None):
    """Brute force of the flux up: per (mode, aux) and per (state, aux), and the
    total into each destination."""
    nm, (n, nk) = len(Ls), C.shape
    F = np.ones((nm, nk)) if F is None else F
    Emk = np.zeros((nm, nk))
    Esk = np.zeros((n, nk))
    by_d = np.zeros(len(dests))
    for m in range(nm):
        for k in range(nk):
            amp = abs(w[m] * (1 + kvec[k][m])) ** 2
            for di, d in enumerate(dests):
                for i, s in enumerate(states):
                    t = amp * abs(Ls[m][d, s]) ** 2 * abs(C[i, k]) ** 2 / HB ** 2
                    Emk[m, k] += t
                    if F[m, k]:
                        Esk[i, k] += t
                        by_d[di] += t
    return Emk, Esk, by_d


def _bf_down(Ls, g, w, lexp, C, states, dests, F=None):
    nm, (n, nk) = len(Ls), C.shape
    F = np.ones((nm, nk)) if F is None else F
    Emk = np.zeros((nm, nk))
    Esk = np.zeros((n, nk))
    by_d = np.zeros(len(dests))
    for m in range(nm):
        gw = abs(g[m] / w[m]) ** 2
        for k in range(nk):
            for i, s in enumerate(states):
                for di, d in enumerate(dests):
                    amp = Ls[m][d, s] - (lexp[m] if d == s else 0.0)
                    t = gw * abs(amp) ** 2 * abs(C[i, k]) ** 2 / HB ** 2
                    Emk[m, k] += t
                    if F[m, k]:
                        Esk[i, k] += t
                        by_d[di] += t
                if s not in dests:     # the <L> part of the diagonal still acts on s
                    t = gw * abs(lexp[m]) ** 2 * abs(C[i, k]) ** 2 / HB ** 2
                    Emk[m, k] += t
                    if F[m, k]:
                        Esk[i, k] += t
    return Emk, Esk, by_d


@pytest.mark.parametrize("F", [None, F_EXCL], ids=["unfiltered", "filtered"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_flux_up_general_against_brute_force(case, F):
    """tests/test_basis_functions_adaptive.py:310-505."""
    Ls, states, dests = CASES[case]
    kvec = [[0, 0, 0, 0], [0, 0, 1, 0], [0, 0, 2, 0]]
    K2 = np.array(kvec).T
    diag, off = _tables(Ls, states, dests)
    Emk, Esk, by_d = _bf_up(Ls, W4, kvec, C2, states, dests, F)
    np.testing.assert_allclose(O.error_flux_up_hier_stable(PHI, 2, 3, 4, W4, K2, diag, off), Emk, rtol=1e-12)
    np.testing.assert_allclose(O.error_flux_up_state_stable(PHI, 2, 3, 4, W4, K2, diag, off, F, F), Esk,
                               rtol=1e-12)
    out = [i for i, d in enumerate(dests) if d not in states]
    got = O.error_flux_up_by_dest_state(PHI, 2, 3, 4, W4, K2, off, None, F, None)
    np.testing.assert_allclose(got[out], by_d[out], rtol=1e-12)
    # only the stable auxiliaries 0 and 2, then only the stable state 1
    sub = [0, 2]
    _, _, by_d = _bf_up(Ls, W4, [kvec[k] for k in sub], C2[:, sub], states, dests,
                        None if F is None else F[:, sub])
    got = O.error_flux_up_by_dest_state(PHI, 2, 3, 4, W4, K2, off, sub, F, None)
    np.testing.assert_allclose(got[out], by_d[out], rtol=1e-12)
    C_st = C2 * np.array([[0], [1]])
    _, _, by_d = _bf_up(Ls, W4, kvec, C_st, states, dests, F)
    got = O.error_flux_up_by_dest_state(PHI, 2, 3, 4, W4, K2, off, None, F, [1])
    np.testing.assert_allclose(got[out], by_d[out], rtol=1e-12)


@pytest.mark.parametrize("F", [None, F_EXCL], ids=["unfiltered", "filtered"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_flux_down_general_against_brute_force(case, F):
    """tests/test_basis_functions_adaptive.py:715-929."""
    Ls, states, dests = CASES[case]
    lexp = [np.conj(C2[:, 0]) @ L[np.ix_(states, states)] @ C2[:, 0] for L in Ls]
    X2 = O.expectation_table(lexp, 2)
    diag, off = _tables(Ls, states, dests)
    Emk, Esk, by_d = _bf_down(Ls, G4, W4, lexp, C2, states, dests, F)
    np.testing.assert_allclose(O.error_flux_down_hier_stable(PHI, 2, 3, 4, G4, W4, diag, off, X2), Emk,
                               rtol=1e-12)
    np.testing.assert_allclose(O.error_flux_down_state_stable(PHI, 2, 3, 4, G4, W4, diag, off, X2, F, F), Esk,
                               rtol=1e-12)
    out = [i for i, d in enumerate(dests) if d not in states]
    got = O.error_flux_down_by_dest_state(PHI, 2, 3, 4, G4, W4, off, None, F, None)
    np.testing.assert_allclose(got[out], by_d[out], rtol=1e-12)
    sub = [0, 2]
    _, _, by_d = _bf_down(Ls, G4, W4, lexp, C2[:, sub], states, dests, None if F is None else F[:, sub])
    got = O.error_flux_down_by_dest_state(PHI, 2, 3, 4, G4, W4, off, sub, F, None)
    np.testing.assert_allclose(got[out], by_d[out], rtol=1e-12)


# --- HopsBasis-level answers of tests/test_hops_basis.py --------------------
def _gw0():
    return O.bcf_convert_dl_to_exp(20.0, 50.0, 140.0)


def _peierls_system(extra_holstein_on_first_peierls=False, lt=None):
    g0, w0 = _gw0()
    n = 4

    def hol(i):
        a = np.zeros((n, n), complex)
        a[i, i] = 1
        return a

    def pei(i):
        a = np.zeros((n, n), complex)
        a[i, i + 1] = 1j
        a[i + 1, i] = -1j
        return a

    base = [hol(0), hol(1), hol(2), hol(3), pei(0), pei(1), pei(2)]
    if extra_holstein_on_first_peierls:
        base[4] = base[4] + hol(0)
    gw, ll = [], []
    for L in base:
        gw += [[g0, w0], [-1j * np.imag(g0), 500.0]]
        ll += [L, L]
    hs = np.zeros((n, n))
    for i in range(n - 1):
        hs[i, i + 1] = hs[i + 1, i] = 10
    p = dict(HAMILTONIAN=hs.astype(np.complex128), GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll,
             ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw)
    if lt is not None:
        p["L_LT_CORR"] = base
        p["PARAM_LT_CORR"] = lt
    return p, base


def _state(p, maxhier=2):
    sysd = O.derive_system(p)
    return sysd, O.AdaptiveState(sysd, "NORMALIZED NONLINEAR", maxhier, 1e-3, 1e-3, 0.01)


def test_tables_by_destination_hand_values():
    """tests/test_hops_basis.py:1351-1640."""
    p, _ = _peierls_system()
    sysd, ad = _state(p)
    b = O.Basis(sysd, [0, 3], [()])
    assert list(b.modes) == [0, 1, 6, 7, 8, 9, 12, 13]
    diag, off, dests = ad.tables(b)
    assert dests == [0, 1, 2, 3]
    z = np.zeros((8, 2), complex)
    want = [z.copy() for _ in range(4)]
    want[1][4:6, 0] = -1j
    want[2][6:8, 1] = 1j
    for got, exp in zip(off, want):
        np.testing.assert_allclose(got, exp)
    d = z.copy()
    d[0:2, 0] = 1
    d[2:4, 1] = 1
    np.testing.assert_allclose(diag, d)
    psi = np.array([1, -1]) / np.sqrt(2)
    X2 = O.expectation_table(ad.expectations(b, psi)[b.relL_by_mode], 2)
    np.testing.assert_allclose(X2, np.vstack([np.full((4, 2), 0.5), np.zeros((4, 2))]))
    # one dense coupling operator on states {0, 2}
    lop = np.zeros((4, 4), complex)
    lop[0, 0] = 2
    lop[0, 1], lop[1, 0] = 1j, -1j
    lop[0, 2] = lop[2, 0] = 1
    lop[0, 3], lop[3, 0] = -1j, 1j
    g0, w0 = _gw0()
    p1 = dict(p, GW_SYSBATH=[[g0, w0]], L_HIER=[lop], L_NOISE1=[lop], PARAM_NOISE1=[[g0, w0]])
    sysd, ad = _state(p1)
    b = O.Basis(sysd, [0, 2], [()])
    diag, off, dests = ad.tables(b)
    assert dests == [0, 1, 2, 3]
    for got, exp in zip(off, [[[0, 1]], [[-1j, 0]], [[1, 0]], [[1j, 0]]]):
        np.testing.assert_allclose(got, exp)
    np.testing.assert_allclose(diag, [[2, 0]])
    X2 = O.expectation_table(ad.expectations(b, np.array([1, 1]) / np.sqrt(2))[b.relL_by_mode], 2)
    np.testing.assert_allclose(X2, [[2, 2]])


def test_noise_operator_hand_values():
    """tests/test_hops_basis.py:1645-1795."""
    p, base = _peierls_system()
    sysd, ad = _state(p)
    b = O.Basis(sysd, [0, 3], [()])
    assert list(b.absL) == [0, 3, 4, 6]
    n1 = 1j * np.arange(7)
    n2 = 2 * np.ones(7)
    zm = np.arange(14).astype(complex)
    Z = ad.noise_operator(b, [n1[b.absL], n2[b.absL], zm])
    want = sum(zm[m] * p["L_HIER"][m] for m in (8, 9, 12, 13))
    want = want + sum((np.conj(n1) - 1j * n2)[l] * base[l] for l in (4, 6))
    np.testing.assert_allclose(Z, want)
    ph = dict(p, GW_SYSBATH=p["GW_SYSBATH"][:8], L_HIER=p["L_HIER"][:8], L_NOISE1=p["L_NOISE1"][:8],
              PARAM_NOISE1=p["PARAM_NOISE1"][:8])
    sysd, ad = _state(ph)
    b = O.Basis(sysd, [0, 3], [()])
    assert ad.noise_operator(b, [n1[b.absL], n2[b.absL], zm[:8]]) is None


def test_ltc_operators_hand_values():
    """tests/test_hops_basis.py:1797-1987."""
    G = [10, 5 + 5j, -10, 5 - 5j, 10j, -5j, 10 + 10j]
    p, base = _peierls_system(extra_holstein_on_first_peierls=True, lt=G)
    sysd, ad = _state(p)
    b = O.Basis(sysd, [0, 3], [(), ((10, 1),)])
    assert list(b.modes) == [0, 1, 6, 7, 8, 9, 10, 12, 13]
    psi_full = np.array([1, 0, 0, -1]) / np.sqrt(2)
    T_phys, T_hier = ad.ltc_operators(b, ad.expectations(b, psi_full[[0, 3]]))
    Tp = np.zeros((4, 4), complex)
    Th = np.zeros((4, 4), complex)
    for l in (0, 3, 4, 5, 6):
        L = base[l]
        if not np.allclose(L, np.diag(np.diag(L))):
            X = np.conj(psi_full) @ L @ psi_full
            Tp += 2 * np.real(G[l]) * X * L - G[l] * (L @ L)
            Th += np.conj(G[l]) * X * L
    np.testing.assert_allclose(T_phys, Tp)
    np.testing.assert_allclose(T_hier, Th)
    ph, _ = _peierls_system(lt=[10000] * 4 + [0, 0, 0])
    ph = dict(ph, GW_SYSBATH=ph["GW_SYSBATH"][:8], L_HIER=ph["L_HIER"][:8], L_NOISE1=ph["L_NOISE1"][:8],
              PARAM_NOISE1=ph["PARAM_NOISE1"][:8], L_LT_CORR=ph["L_LT_CORR"][:4], PARAM_LT_CORR=[10000] * 4)
    sysd, ad = _state(ph)
    b = O.Basis(sysd, [0, 3], [()])
    assert ad.ltc_operators(b, ad.expectations(b, psi_full[[0, 3]])) == (None, None)


def _chain_system(nsite, couplings):
    g0, w0 = _gw0()
    gw, ll = [], []
    for i in range(nsite):
        L = np.zeros((nsite, nsite))
        L[i, i] = 1.0
        gw += [[g0, w0], [-1j * np.imag(g0), 500.0]]
        ll += [sp.coo_matrix(L), L]
    hs = np.zeros((nsite, nsite))
    for i, v in enumerate(couplings):
        hs[i, i + 1] = hs[i + 1, i] = v
    return dict(HAMILTONIAN=hs.astype(np.complex128), GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll,
                ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw)


def test_state_stable_error_analytic():
    """tests/test_hops_basis.py:1227-1349: states {4, 5, 6}, auxiliaries
    {0, e_10 + e_11}, boundary auxiliaries e_10 and e_11."""
    p = _chain_system(10, [40, 10] * 4 + [40])
    sysd, ad = _state(p)
    b = O.Basis(sysd, [4, 5, 6], [(), ((10, 1), (11, 1))])
    dsys = O.make_derivative(b, "NORMALIZED NONLINEAR")
    phi = np.array([0, 1.0, 0, 0, 0.8, 0], complex)
    rng = np.random.default_rng(0)
    z_step = [rng.normal(size=len(b.absL)) + 1j * rng.normal(size=len(b.absL)),
              np.zeros(len(b.absL), complex), np.zeros(20, complex)]
    bound = [((10, 1),), ((11, 1),)]
    got = ad.state_stable_error(dsys, b, phi, 2.0, z_step, [0, 1], bound, list(b.aux_list) + bound)
    deletion = np.sum(O.error_deriv(dsys, phi, z_step, 3, 2, 2.0, [0, 1]), axis=1)
    gw = p["GW_SYSBATH"]
    sflux = np.array([0, (1.0 ** 2 + 0.8 ** 2) * (10 ** 2 + 40 ** 2), 0]) / HB ** 2
    up = np.array([0, abs(gw[10][1]) ** 2 + abs(gw[11][1]) ** 2, 0]) / HB ** 2
    down = np.array([0, 0.8 ** 2 * (1.0 - 1.0 ** 2) ** 2 * (abs(gw[10][0] / gw[10][1]) ** 2
                                                          + abs(gw[11][0] / gw[11][1]) ** 2), 0]) / HB ** 2
    np.testing.assert_allclose(got, deletion + sflux + up + down, rtol=1e-12)
    # the same terms worked out for psi = (0.6, 0.8, 0), where <L_5> = 0.64 and
    # the flux down out of e_10 + e_11 no longer vanishes
    phi = np.array([0.6, 0.8, 0, 0, 0.8, 0], complex)
    got = ad.state_stable_error(dsys, b, phi, 2.0, z_step, [0, 1], bound, list(b.aux_list) + bound)
    deletion = np.sum(O.error_deriv(dsys, phi, z_step, 3, 2, 2.0, [0, 1]), axis=1)
    sflux = np.array([0.36, 0.64 + 0.64, 0]) * 1700 / HB ** 2
    up = np.array([0, 0.64 * (abs(gw[10][1]) ** 2 + abs(gw[11][1]) ** 2), 0]) / HB ** 2
    down = np.array([0, 0.64 * (1 - 0.64) ** 2 * (abs(gw[10][0] / gw[10][1]) ** 2
                                                   + abs(gw[11][0] / gw[11][1]) ** 2), 0]) / HB ** 2
    np.testing.assert_allclose(got, deletion + sflux + up + down, rtol=1e-12)


def _two_site_state(maxhier=6):
    p = _chain_system(2, [40])
    return p, _state(p, maxhier)


def test_boundary_hierarchy_binning_hand_sets():
    """tests/test_hops_basis.py:690-979: which auxiliaries outside the basis
    are kept from given flux-up / flux-down arrays, with and without the
    f_discard pre-trim, including several fluxes into one auxiliary."""
    _, (sysd, ad) = _two_site_state()
    b = O.Basis(sysd, [1], [()])
    assert list(b.modes) == [2, 3]
    up = np.array([[0.1 ** 2], [0.001 ** 2]])
    got = ad.boundary_hier(b, [up, np.zeros((2, 1))], np.array([0]), 0.0015 ** 2)
    assert set(got) == {((2, 1),)}

    A = lambda *t: tuple(t)   # noqa: E731  (an auxiliary as (mode, depth) pairs)
    aux = [(), A((1, 1)), A((2, 1)), A((3, 1)), A((0, 2)), A((1, 2)), A((2, 1), (3, 1)), A((2, 1), (3, 2)),
           A((3, 3)), A((2, 1), (3, 3)), A((3, 4))]
    b = O.Basis(sysd, [0, 1], aux)
    ix = {a: i for i, a in enumerate(aux)}
    up = np.zeros((4, 11))
    dn = np.zeros((4, 11))
    up[0, ix[A((0, 2))]] = 0.00003 ** 2
    up[0, ix[A((2, 1), (3, 1))]] = 0.01 ** 2
    up[1, ix[A((2, 1), (3, 2))]] = 0.00004 ** 2
    up[1, ix[A((0, 2))]] = 0.01 ** 2
    up[2, ix[A((0, 2))]] = 0.01 ** 2
    dn[0, ix[A((0, 2))]] = 0.1 ** 2
    # the Basis sorts its auxiliaries: put the arrays in its order
    perm = [ix[a] for a in b.aux_list]
    got = ad.boundary_hier(b, [up[:, perm], dn[:, perm]], np.arange(11), 0.001 ** 2)
    assert set(got) == {A((0, 1), (2, 1), (3, 1)), A((0, 2), (1, 1)), A((0, 2), (2, 1)), A((0, 1))}

    names = ["", "1", "2", "3", "00", "11", "23", "122", "123", "223", "1123", "1133", "1233", "11123",
             "11223", "11233"]

    def mk(s):
        from collections import Counter
        c = Counter(int(ch) for ch in s)
        return tuple((m, c[m]) for m in sorted(c))

    aux = [mk(s) for s in names]
    ix = {s: i for i, s in enumerate(names)}
    up = np.zeros((4, 16))
    dn = np.zeros((4, 16))
    for m, srcs in {0: ["", "1", "2", "3", "00", "11", "23", "122", "123", "223", "1123", "1133", "1233"],
                    1: ["2", "3", "00", "11", "122", "223", "1133"],
                    2: ["1", "2", "00", "11", "122", "123", "223", "1233"],
                    3: ["1", "3", "00", "11", "23", "122", "223", "1133", "1233"]}.items():
        for s in srcs:
            up[m, ix[s]] = 1.0
    for m, srcs in {0: ["00"], 1: ["122", "1133", "1233", "11223"], 2: ["122", "123", "1123", "1233", "11123"],
                    3: ["123", "223", "1123", "1133", "11123", "11223"]}.items():
        for s in srcs:
            dn[m, ix[s]] = 1.0
    b = O.Basis(sysd, [0, 1], aux)
    perm = [ix[next(s for s in names if mk(s) == a)] for a in b.aux_list]
    up, dn = up[:, perm], dn[:, perm]
    n1, n2, n3 = 26, 5, 3
    cases = [(1.0 * n1 + 2.0 * n2 + 3.0 * n3 + 0.001, ["12", "1223"]),
             (1.0 * n1 + 2.0 * n2 + 0.001, ["12", "1223", "13", "22", "113"]),
             (1.0 * n1 + 0.001, ["12", "1223", "13", "22", "113", "0", "1122", "112", "233", "133"])]
    ad0 = O.AdaptiveState(sysd, "NORMALIZED NONLINEAR", 6, 1e-3, 1e-3, 0.0)
    for bound, want in cases:
        got = ad0.boundary_hier(b, [up, dn], np.arange(16), bound)
        assert set(got) == {mk(s) for s in want}, bound


def test_boundary_fraction_discard():
    """tests/test_hops_basis.py:981-1120 (f_discard 0.99 and 0), and two cases
    that follow from the reference's definitions and separate plausible
    slips: the pre-trim budget is (f_discard * delta)^2, not f_discard *
    delta^2; and flux down along a mode that the auxiliary does not hold
    (hops_fluxfilters.py:263-290 zeroes it) never enters the pre-trim."""
    _, (sysd, _) = _two_site_state(4)
    b = O.Basis(sysd, [1], [(), ((2, 1),), ((3, 1),)])
    up = np.zeros((2, 3))
    up[0, 1] = 0.14 ** 2     # into (2, 2)
    up[1, 1] = 0.1 ** 2      # into (2, 1)(3, 1)
    up[0, 2] = 0.099 ** 2    # into (2, 1)(3, 1)
    up[1, 2] = 0.201 ** 2    # into (3, 2)
    dn = np.zeros((2, 3))
    for f, want in ((0.99, {((2, 2),), ((3, 2),)}), (0.0, {((2, 1), (3, 1)), ((3, 2),)})):
        ad = O.AdaptiveState(sysd, "NORMALIZED NONLINEAR", 4, 0.15, 0.0, f)
        assert set(ad.boundary_hier(b, [up, dn], np.arange(3), 0.15 ** 2)) == want
    # impossible flux down out of the physical wave function: ignored
    ad = O.AdaptiveState(sysd, "NORMALIZED NONLINEAR", 4, 0.15, 0.0, 0.99)
    dn_bad = dn.copy()
    dn_bad[0, 0] = 0.0099
    assert set(ad.boundary_hier(b, [up, dn_bad], np.arange(3), 0.15 ** 2)) == {((2, 2),), ((3, 2),)}
    # f = 0.5, budget 1: two 0.3 components bin into (2, 1)(3, 1), 0.45 goes
    # into (3, 2).  (f delta)^2 = 0.25 trims no component, the bins are 0.6
    # and 0.45, and only the 0.6 bin is needed.  A budget of f delta^2 = 0.5
    # would trim both 0.3 components and keep (3, 2) instead.
    up = np.zeros((2, 3))
    up[1, 1] = 0.3
    up[0, 2] = 0.3
    up[1, 2] = 0.45
    ad = O.AdaptiveState(sysd, "NORMALIZED NONLINEAR", 4, 1.0, 0.0, 0.5)
    assert set(ad.boundary_hier(b, [up, dn], np.arange(3), 1.0)) == {((2, 1), (3, 1))}


def _fixture_noise(key, rows=None):
    z = FX[key]
    return {"SEED": z if rows is None else z[:rows], "MODEL": "FFT_FILTER", "TLEN": 250.0, "TAU": 1.0}


def test_define_hierarchy_basis_on_reference_noise():
    """tests/test_hops_basis.py:586-688: 2 sites, MAXHIER 4, the reference's
    raw noise rows; stable and boundary auxiliaries before and after a flat
    Phi."""
    p = _chain_system(2, [40])
    psi0 = np.array([0.0, 1.0], complex)
    tr = O.AdaptiveTrajectory(p, {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"},
                              _fixture_noise("hops_basis_noise_2site"), hierarchy_param={"MAXHIER": 4},
                              delta_a=1e-3, delta_s=1e-3)
    tr.initialize(psi0)
    z = tr.z_step(tr.z_mem)
    stable, bound = tr.ad.hierarchy_basis(tr.dsystem_dt, tr.basis, tr.phi, 2.0, z)
    assert tuple(stable) == ((), ((2, 1),), ((3, 1),))
    assert bound == []
    phi = np.array([np.sqrt(0.5), np.sqrt(0.5)] * tr.basis.n_hier, complex)
    stable, bound = tr.ad.hierarchy_basis(tr.dsystem_dt, tr.basis, phi, 2.0, z)
    assert set(stable) == {(), ((2, 1),), ((3, 1),)}
    assert set(bound) == {((1, 1), (2, 1)), ((3, 2),), ((0, 1), (3, 1)), ((0, 1),), ((1, 1),), ((2, 2),),
                          ((2, 1), (3, 1)), ((0, 1), (2, 1)), ((1, 1), (3, 1))}


def test_define_state_basis_on_reference_noise():
    """tests/test_hops_basis.py:481-584: 10 sites, MAXHIER 2, the reference's
    raw noise rows; stable and boundary states before and after a flat psi."""
    p = _chain_system(10, [40, 10] * 4 + [40])
    psi0 = np.zeros(10, complex)
    psi0[5] = 1.0
    tr = O.AdaptiveTrajectory(p, {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"},
                              _fixture_noise("hops_basis_noise_10site"), hierarchy_param={"MAXHIER": 2},
                              delta_a=1e-3, delta_s=1e-3)
    tr.initialize(psi0)
    b = tr.basis
    z = tr.z_step(tr.z_mem)
    st, bd = tr.ad.state_basis(tr.dsystem_dt, b, tr.phi, 2.0, z, [0, 1, 2], [], list(b.aux_list))
    assert st == [4, 5, 6] and bd == []
    phi = 0 * tr.phi
    phi[:b.n_state] = 1 / np.sqrt(b.n_state)
    st, bd = tr.ad.state_basis(tr.dsystem_dt, b, phi, 2.0, z, [0, 1, 2], [], list(b.aux_list))
    assert st == [4, 5, 6] and bd == [3, 7]
