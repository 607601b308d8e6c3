"""Pins of the oracle's low-temperature correction (LTC).

The helper terms are checked against the reference's hand-computed matrices
(tests/test_low_temperature_correction.py:277-433).  The derivative is checked
the way tests/integrated_tests/test_LTC_eom.py:71-590 checks it: the
derivative of a system with the correction minus the same system without it,
on the same Phi and noise, must equal the correction alone.  Here that
correction is written out in closed form for projector L-operators
(L^2 = L, so <L^2> = <L> = p, the population it projects on), independently
of the operator algebra the oracle uses.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import hops_oracle as O

L0 = sp.coo_matrix(np.diag([1.0, 0.0]))
L1 = sp.coo_matrix(np.diag([0.0, 1.0]))
LSEC = sp.coo_matrix(np.array([[0, 2j], [-2j, 0]]))


def _d(x):
    return x.toarray() if sp.issparse(x) else np.asarray(x)


def test_lt_corr_hand_values():
    """tests/test_low_temperature_correction.py:277-373."""
    full = np.array([[-5 - 10j, 0], [0, 5 + 10j]])
    not_phys = np.array([[2.5 - 2.5j, 0], [0, -2.5 + 2.5j]])
    c_phys, c_hier = O.lt_corr([10 + 10j, -10 - 10j], [L0, L1], [0.25, 0.25], [L0 @ L0, L1 @ L1])
    np.testing.assert_allclose(_d(c_phys), full - not_phys)
    np.testing.assert_allclose(_d(c_hier), not_phys)
    # a Peierls (off-diagonal, imaginary) operator
    full_sec = np.array([[-40 - 40j, 10j], [-10j, -40 - 40j]])
    not_phys_sec = np.array([[0, 5 + 5j], [-5 - 5j, 0]])
    c_phys, c_hier = O.lt_corr([10 + 10j], [LSEC], [0.25], [LSEC @ LSEC])
    np.testing.assert_allclose(_d(c_phys), full_sec - not_phys_sec)
    np.testing.assert_allclose(_d(c_hier), not_phys_sec)
    # vanishing coefficients, expectations or operators
    c_phys, c_hier = O.lt_corr([0, 0], [L0, L1], [0.25, 0.25], [L0 @ L0, L1 @ L1])
    assert np.allclose(_d(c_phys), 0) and np.allclose(_d(c_hier), 0)
    c_phys, c_hier = O.lt_corr([10, -10], [L0, L1], [0.5, 0.5], [L0 @ L0, L1 @ L1])
    np.testing.assert_allclose(_d(c_phys), -_d(c_hier))
    c_phys, c_hier = O.lt_corr([10, -10], [L0, L1], [0, 0], [L0 @ L0, L1 @ L1])
    assert np.allclose(_d(c_hier), 0)
    z = np.zeros((2, 2))
    c_phys, c_hier = O.lt_corr([10 + 10j, -10 - 10j], [z, z], [0.25, 0.25], [z, z])
    assert np.allclose(c_phys, 0) and np.allclose(c_hier, 0)
    assert O.lt_corr([], [], [], []) == (0, 0)


def test_lt_corr_to_norm_corr_hand_values():
    """tests/test_low_temperature_correction.py:376-403: 9 * 0.25."""
    assert np.isclose(O.lt_corr_to_norm_corr([1 + 10j, -10 - 1j], [0.5, 0.5], [0.75, 0.75]), 9 * 0.25)
    assert np.isclose(O.lt_corr_to_norm_corr([0, 0], [0.5, 0.5], [0.75, 0.75]), 0)
    assert np.isclose(O.lt_corr_to_norm_corr([1 + 10j, -10 - 1j], [0.5, 0.5], [0.5, 0.5]), 0)
    assert O.lt_corr_to_norm_corr([], [], []) == 0


def test_lt_corr_linear_hand_values():
    """tests/test_low_temperature_correction.py:405-433."""
    np.testing.assert_allclose(_d(O.lt_corr_linear([-1, -1j], [L0, L1])), [[1, 0], [0, 1j]])
    assert np.allclose(_d(O.lt_corr_linear([0, 0], [L0, L1])), 0)
    z = np.zeros((2, 2))
    assert np.allclose(O.lt_corr_linear([-1, -1j], [z, z]), 0)
    assert np.allclose(O.lt_corr_linear([], []), 0)
    with pytest.raises(ValueError):
        O.lt_corr_linear([1.0], [L0, L1])


# --- derivative deltas (tests/integrated_tests/test_LTC_eom.py) -------------
def _site_projector(n, states):
    return sp.coo_matrix(np.diag([1.0 if i in states else 0.0 for i in range(n)]))


def _ltc_system(n_site, proj_states, gw_hier, lt_noise, with_ltc):
    """Sites with two hierarchy modes per projector L_n and one extra noise mode
    per projector whose g/w is the correction coefficient."""
    H = np.zeros((n_site, n_site))
    for i in range(n_site - 1):
        H[i, i + 1] = H[i + 1, i] = 1.0
    Ls = [_site_projector(n_site, st) for st in proj_states]
    gw, lh = [], []
    for L in Ls:
        for g, w in gw_hier:
            gw.append([g, w])
            lh.append(L)
    p = dict(HAMILTONIAN=H.astype(np.complex128), GW_SYSBATH=gw, L_HIER=lh,
             L_NOISE1=lh + Ls, ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw + [list(x) for x in lt_noise])
    if with_ltc:
        p["L_LT_CORR"] = Ls
        p["PARAM_LT_CORR"] = [g / w for g, w in lt_noise]
    return p


def _closed_form_delta(states, proj_states, c, phi, n_hier, eom):
    """The correction's share of hbar*dPhi/dt for projector L-operators:
    LINEAR: -sum_n c_n L_n psi on the physical block; NONLINEAR:
    sum_n conj(c_n) p_n L_n phi_k on every block plus sum_n c_n (p_n - 1) L_n psi
    on the physical block; NORMALIZED NONLINEAR also -sum_n Re(c_n)(2 p_n^2 - p_n) phi."""
    n = len(states)
    P = np.asarray(phi).reshape(n_hier, n)
    psi = P[0]
    w = np.abs(psi) ** 2 / np.sum(np.abs(psi) ** 2)
    # operators with no state in the basis are not active and drop out
    act = [i for i, st in enumerate(proj_states) if any(s in st for s in states)]
    mask = [np.array([1.0 if s in proj_states[i] else 0.0 for s in states]) for i in act]
    p = [float(np.sum(w * m)) for m in mask]
    out = np.zeros_like(P, dtype=complex)
    if eom == "LINEAR":
        for k, i in enumerate(act):
            out[0] -= c[i] * mask[k] * psi
        return out.reshape(-1)
    for k, i in enumerate(act):
        out += np.conj(c[i]) * p[k] * mask[k][None, :] * P
        out[0] += c[i] * (p[k] - 1.0) * mask[k] * psi
    if eom == "NORMALIZED NONLINEAR":
        delta = sum(np.real(c[i]) * (2 * p[k] ** 2 - p[k]) for k, i in enumerate(act))
        out -= delta * P
    return out.reshape(-1)


def _delta(sys_on, sys_off, states, aux, eom, phi):
    outs = []
    for p in (sys_on, sys_off):
        sysd = O.derive_system(p)
        b = O.Basis(sysd, states, aux)
        d = O.make_derivative(b, eom)
        nl = len(b.absL)
        z1 = (0.3 + 0.7j) * np.arange(1, nl + 1)
        z2 = np.zeros(nl, complex)
        zm = np.zeros(len(p["L_NOISE1"]), complex)
        outs.append(d(np.asarray(phi, complex), zm, z1, z2)[0])
    return outs[0] - outs[1]


THREE = [[0], [1], [2]]
GW3 = [[10.0, 10.0], [5.0, 5.0]]
LT3 = [(250.0, 1000.0), (250.0, 2000.0), (250.0, 3000.0)]
PHI3 = np.array([0.7071, 0.5773, 0.4082, 0.5, 0, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0.5, 0, 0, 0, 0.5, 0.5, 0, 0, 0.5])


@pytest.mark.parametrize("eom", ["LINEAR", "NONLINEAR", "NORMALIZED NONLINEAR"])
def test_ltc_derivative_delta_full_basis(eom):
    """tests/integrated_tests/test_LTC_eom.py:71-287: 3 sites, MAXHIER 1."""
    on = _ltc_system(3, THREE, GW3, LT3, True)
    off = _ltc_system(3, THREE, GW3, LT3, False)
    aux = O.triangular_hierarchy(6, 1)
    got = _delta(on, off, [0, 1, 2], aux, eom, PHI3)
    c = [g / w for g, w in LT3]
    want = _closed_form_delta([0, 1, 2], THREE, c, PHI3, len(aux), eom)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    assert np.max(np.abs(want)) > 0.01


def test_ltc_derivative_delta_state_sub_basis():
    """tests/integrated_tests/test_LTC_eom.py:290-389: states {0, 2} of the
    3-site system with a subset of the auxiliaries; the operator of site 1 has
    no state in the basis and drops out."""
    on = _ltc_system(3, THREE, GW3, LT3, True)
    off = _ltc_system(3, THREE, GW3, LT3, False)
    full = O.triangular_hierarchy(6, 1)
    aux = [full[0], full[1], full[3], full[4]]
    rng = np.random.default_rng(3)
    phi = rng.normal(size=2 * len(aux)) + 1j * rng.normal(size=2 * len(aux))
    c = [g / w for g, w in LT3]
    got = _delta(on, off, [0, 2], aux, "NORMALIZED NONLINEAR", phi)
    want = _closed_form_delta([0, 2], THREE, c, phi, len(aux), "NORMALIZED NONLINEAR")
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)


def test_ltc_derivative_delta_multiparticle_complex():
    """tests/integrated_tests/test_LTC_eom.py:392-590: L-operators projecting on
    two states each, complex coefficients (250+10j)/w, states {0, 2}."""
    proj = [[0, 1], [0, 2], [2, 3], [1, 3]]
    lt = [(250.0 + 10j, 1000.0), (250.0 + 10j, 2000.0), (250.0 + 10j, 3000.0), (250.0 + 10j, 4000.0)]
    on = _ltc_system(4, proj, GW3, lt, True)
    off = _ltc_system(4, proj, GW3, lt, False)
    full = O.triangular_hierarchy(8, 1)
    aux = [full[i] for i in (0, 1, 3, 4, 5, 6)]
    P = np.zeros((len(aux), 2), complex)
    P[0] = [0.7071, 0.4082]
    for i in range(1, len(aux)):
        P[i] = [0.5 * (i % 2 == 1), 0.5 * (i > 3)]
    c = [g / w for g, w in lt]
    for eom in ("LINEAR", "NONLINEAR", "NORMALIZED NONLINEAR"):
        got = _delta(on, off, [0, 2], aux, eom, P.reshape(-1))
        want = _closed_form_delta([0, 2], proj, c, P.reshape(-1), len(aux), eom)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14, err_msg=eom)
