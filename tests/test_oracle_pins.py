"""Pins of the FP64 oracle: the reference's golden fixtures, an independent
brute-force restatement of the EOM, closed forms and invariants.  CPU only."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import systems
from oracle import hops_oracle as O

FX = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_fixtures.npz"))


def _dense(prefix, shape):
    a = np.zeros(shape, complex)
    idx = FX[prefix + "_idx"]
    a[idx[0], idx[1]] = FX[prefix + "_val"]
    return a


@pytest.fixture(scope="module")
def dod():
    tr = O.Trajectory(systems.dimer_of_dimers(), {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"},
                      {"SEED": 0, "MODEL": "FFT_FILTER", "TLEN": 25000.0, "TAU": 1.0, "RAND_MODEL": "BOX_MULLER"},
                      hierarchy_param={"MAXHIER": 2})
    psi0 = np.zeros(4, complex)
    psi0[2] = 1.0
    tr.initialize(psi0)
    return tr


def test_super_operators_equal_golden(dod):
    """tests/test_dimer_of_dimers.py:163-187 (exact equality)."""
    b = dod.basis
    assert np.array_equal(b.K2_k.toarray(), _dense("K2_k", tuple(FX["K2_k_shape"])))
    assert np.array_equal(b.K2_kp1.toarray(), _dense("K2_kp1", tuple(FX["K2_kp1_shape"])))
    assert np.array_equal(b.K2_km1.toarray(), _dense("K2_km1", tuple(FX["K2_km1_shape"])))
    shp = tuple(FX["Z2_shape"])
    for i in range(shp[0]):
        assert np.array_equal(b.Z2_kp1[i].toarray(), _dense("Z2_%d" % i, shp[1:]))


def test_alpha_equals_golden(dod):
    """tests/test_dimer_of_dimers.py:213-220."""
    a = dod.noise1.alpha(list(range(4)))
    np.testing.assert_allclose(a[:, ::50], FX["alpha_stride50"])


def test_eta_equals_golden(dod):
    """tests/test_dimer_of_dimers.py:223-239 (rtol 1e-10)."""
    a = dod.noise1.alpha(list(range(4)))
    z = O.uncorrelated_noise(0, list(range(4)), len(dod.noise1.t_axis), "BOX_MULLER")
    eta = O.correlated_noise(a, z)[0]
    np.testing.assert_allclose(eta[::5], FX["eta_stride5"], rtol=1e-10)


def test_dimer_of_dimers_trajectory(dod):
    """tests/test_dimer_of_dimers.py:242-260 (index 25, rtol 1e-5) and the
    whole recorded trajectory."""
    import copy
    tr = copy.deepcopy(dod)
    tr.propagate(200.0, 4.0)
    traj = FX["traj_dimer_of_dimers"]
    assert np.allclose(tr.psi_traj[25], traj[25], rtol=1e-5)
    assert np.max(np.abs(np.array(tr.psi_traj) - traj)) < 1e-5


def test_raw_noise_equals_golden():
    """Seed-0 SUM_GAUSSIAN raw noise, TLEN 10, TAU 1 (tests/test_noise_fft.py:86-120)."""
    z = O.uncorrelated_noise(0, [0, 1], 11, "SUM_GAUSSIAN")
    np.testing.assert_allclose(z, FX["pre_calculated_uncorrelated_noise_test"])


def test_raw_index_prefix_invariant():
    """A prefix of the raw-slot assignment does not depend on the axis length
    (tests/test_noise_fft.py:376-400); the slots of one axis are a permutation
    of all 4(n-1) draws."""
    big_a, big_b = O.raw_index(1000)
    for n in range(2, 1000):
        a, b = O.raw_index(n)
        assert list(big_a[:n]) == list(a[:n])
        assert list(big_b[:n]) == list(b[:n])
    for n in (3, 4, 17):
        a, b = O.raw_index(n)
        assert len(a) == len(b) == 2 * (n - 1)
        assert len(set(list(a) + list(b))) == 4 * (n - 1) or n < 3


def test_known_bath_parameters():
    """G and W derived for the system of tests/test_hops_system.py:25-55 (the
    dimer of dimers) equal the golden known_param_g/w (:88-100)."""
    sysd = O.derive_system(systems.dimer_of_dimers())
    assert np.allclose(sysd["G"], FX["known_param_g"])
    assert np.allclose(sysd["W"], FX["known_param_w"])
    lam, gam, T = 20.0, 50.0, 140.0
    g0, w0 = O.bcf_convert_dl_to_exp(lam, gam, T)
    assert np.isclose(g0, 2 * lam * O.KB * T - 1j * lam * gam) and w0 == gam


# ---------------------------------------------------------------------------
# Brute force: dense per-auxiliary loops over k +- e_m (an independent
# formulation of the EOM, after tests/test_hops_eom.py:14-229)
# ---------------------------------------------------------------------------
def brute_force(phi, basis, eom, z1, z_mem):
    n, nh = basis.n_state, basis.n_hier
    P = phi.reshape(nh, n)
    D = np.zeros_like(P)
    modes = basis.modes
    Lm = [basis.L_csr[basis.relL_by_mode[i]].toarray() for i in range(len(modes))]
    g, w = basis.g, basis.w
    kvec = np.zeros((nh, basis.sysd["N_HMODES"]), int)
    for i, a in enumerate(basis.aux_list):
        for (m, v) in a:
            kvec[i, m] = v
    # noise per unique L in relative order, conj; plus memory for nonlinear
    zc = np.conj(z1).astype(complex)
    if eom != "LINEAR":
        for i, m in enumerate(modes):
            zc[basis.relL_by_mode[i]] += z_mem[m]
    Lrel = [basis.L_csr[j].toarray() for j in range(len(basis.L_csr))]
    noise_mat = sum(zc[j] * Lrel[j] for j in range(len(Lrel)))
    psi0 = P[0]
    lexp = [np.conj(psi0) @ L @ psi0 / (np.conj(psi0) @ psi0) for L in Lrel]
    I = np.eye(n)
    norm_corr = 0.0
    if eom == "NORMALIZED NONLINEAR":
        norm_corr = np.real(np.sum(np.array(lexp) * zc))
        for i, m in enumerate(modes):
            e = np.zeros(basis.sysd["N_HMODES"], int)
            e[m] = 1
            hit = np.where((kvec == e).all(1))[0]
            if len(hit):
                pm = P[hit[0]]
                r = g[i] / w[i]
                norm_corr -= np.real(r * np.conj(psi0) @ Lm[i] @ pm)
                norm_corr += np.real(lexp[basis.relL_by_mode[i]] * r * np.conj(psi0) @ pm)
    for k in range(nh):
        D[k] += (-1j * basis.H - (kvec[k] @ basis.sysd["W"]) * I - norm_corr * I + noise_mat) @ P[k]
        for i, m in enumerate(modes):
            e = np.zeros(basis.sysd["N_HMODES"], int)
            e[m] = 1
            dn = np.where((kvec == kvec[k] - e).all(1))[0]
            if len(dn):
                D[k] += kvec[k][m] * w[i] * Lm[i] @ P[dn[0]]
            up = np.where((kvec == kvec[k] + e).all(1))[0]
            if len(up):
                op = Lm[i] if eom == "LINEAR" else Lm[i] - lexp[basis.relL_by_mode[i]] * I
                D[k] -= (g[i] / w[i]) * op @ P[up[0]]
    return D.reshape(-1)


@pytest.mark.parametrize("general", [False, True])
@pytest.mark.parametrize("eom", ["LINEAR", "NONLINEAR", "NORMALIZED NONLINEAR"])
def test_derivative_equals_brute_force(general, eom):
    sysp = systems.three_site(general)
    sysd = O.derive_system(sysp)
    basis = O.Basis(sysd, np.arange(3), O.triangular_hierarchy(6, 3))
    f = O.make_derivative(basis, eom)
    rng = np.random.default_rng(11)
    phi = rng.normal(size=basis.n_state * basis.n_hier) + 1j * rng.normal(size=basis.n_state * basis.n_hier)
    z1 = rng.normal(size=len(basis.absL)) + 1j * rng.normal(size=len(basis.absL))
    zmem = (rng.normal(size=6) + 1j * rng.normal(size=6)) * (0 if eom == "LINEAR" else 1)
    d, _ = f(phi, zmem, z1, np.zeros(len(basis.absL)))
    bf = brute_force(phi, basis, eom, z1, zmem)
    assert np.max(np.abs(d - bf)) < 1e-10 * np.max(np.abs(bf))


def test_zmem_derivative_closed_form():
    """dz_m = <L_m> conj(g_m) - conj(w_m) z_m (eom/eom_functions.py:65-160)."""
    sysp = systems.three_site(False)
    sysd = O.derive_system(sysp)
    basis = O.Basis(sysd, np.arange(3), O.triangular_hierarchy(6, 2))
    f = O.make_derivative(basis, "NONLINEAR")
    rng = np.random.default_rng(2)
    phi = rng.normal(size=basis.n_state * basis.n_hier) + 0j
    zmem = rng.normal(size=6) + 1j * rng.normal(size=6)
    _, dz = f(phi, zmem, np.zeros(3), np.zeros(3))
    psi = phi[:3]
    for m in range(6):
        L = basis.L_csr[basis.relL_by_mode[m]].toarray()
        lexp = np.conj(psi) @ L @ psi / (np.conj(psi) @ psi)
        assert np.isclose(dz[m], lexp * np.conj(sysd["G"][m]) - np.conj(sysd["W"][m]) * zmem[m])


def test_rk4_exact_for_uncoupled_decay():
    """No coupling, no noise, depth 0: psi(t) = exp(-i E t / hbar) psi(0); RK4
    error after 50 steps of h = 1 fs is O(h^5) per step."""
    H = np.diag([100.0, -50.0]).astype(complex)
    L = [sp.coo_matrix(([1.0], ([i], [i])), shape=(2, 2)) for i in range(2)]
    gw = [[10.0, 10.0], [10.0, 10.0]]
    sysp = dict(HAMILTONIAN=H, GW_SYSBATH=gw, L_HIER=L, L_NOISE1=L, ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw)
    tr = O.Trajectory(sysp, {"EQUATION_OF_MOTION": "LINEAR"},
                      {"SEED": 0, "MODEL": "ZERO", "TLEN": 100.0, "TAU": 0.5}, hierarchy_param={"MAXHIER": 0})
    psi0 = np.array([1.0, 1.0], complex) / np.sqrt(2)
    tr.initialize(psi0)
    tr.propagate(50.0, 1.0)
    exact = np.exp(-1j * np.diag(H) * 50.0 / O.HBAR) * psi0
    assert np.max(np.abs(tr.psi - exact)) < 1e-9


def test_noise_correlation_reconstructs_alpha():
    """Time-averaged autocorrelation of the correlated noise, averaged over an
    ensemble, reproduces alpha(t) on its first lags within 2 %
    (tests/integrated_tests/test_rand_models.py:13-28,117-170)."""
    g1, w1, g2, w2 = 10000.0 - 1000.0j, 100.0, 1000.0j, 1000.0
    L = [np.diag([1.0, 0.0]), np.diag([1.0, 0.0]), np.diag([0.0, 1.0]), np.diag([0.0, 1.0])]
    gw = [[g1, w1], [g2, w2], [g1, w1], [g2, w2]]
    sysp = dict(HAMILTONIAN=np.array([[0, 10.0], [10.0, 0]]), GW_SYSBATH=gw, L_HIER=L, L_NOISE1=L,
                ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw)
    sysd = O.derive_system(sysp)
    n_traj = 400
    t_axis = np.arange(0.0, 2500.0, 1.0)
    ct = np.zeros(len(t_axis), complex)
    for s in range(n_traj):
        nz = O.Noise({"SEED": 1000 + s, "MODEL": "FFT_FILTER", "TLEN": 2500.0, "TAU": 1.0,
                      "RAND_MODEL": "SUM_GAUSSIAN"}, sysd, 1)
        x = nz.get_noise(t_axis, [0])[0].astype(complex)
        r = np.correlate(x, x, mode="full")
        ct += r[r.size // 2:]
    ct /= n_traj * len(t_axis)
    cf = g1 * np.exp(-t_axis * w1 / O.HBAR) + g2 * np.exp(-t_axis * w2 / O.HBAR)
    assert np.allclose(ct[:10], cf[:10], rtol=2e-2)


# ---------------------------------------------------------------------------
# Adaptive basis pins (tests/test_dimer_of_dimers.py:263-305,
# tests/test_adap_basis.py:110-361)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("da,ds", [(1e-15, 0), (0, 1e-100), (1e-15, 1e-100)])
def test_adaptive_delta_to_zero_equals_full(da, ds):
    """With vanishing thresholds the adaptive trajectory equals the plain one."""
    sysp = systems.dimer_of_dimers()
    noise = dict(SEED=0, MODEL="FFT_FILTER", TLEN=400.0, TAU=1.0, RAND_MODEL="BOX_MULLER")
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    psi0 = np.zeros(4, complex)
    psi0[2] = 1
    ref = O.Trajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 2})
    ref.initialize(psi0)
    ref.propagate(200.0, 4.0)
    tr = O.AdaptiveTrajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 2}, delta_a=da, delta_s=ds)
    tr.initialize(psi0)
    tr.propagate(200.0, 4.0)
    np.testing.assert_allclose(np.array(tr.psi_traj)[33], np.array(ref.psi_traj)[33])
    assert np.max(np.abs(np.array(tr.psi_traj) - np.array(ref.psi_traj))) < 1e-12


@pytest.mark.parametrize("kind", ["Triangular", "LongEdge", "both"])
def test_adaptive_static_filters_reach_the_filtered_hierarchy(kind):
    """With vanishing thresholds the adaptive hierarchy grows into exactly the
    statically filtered hierarchy: every legal auxiliary and no other, where
    legality is the reference's membership rule, applied here by brute force
    to the triangular hierarchy (basis/hierarchy_functions.py:7-95: Triangular
    keeps depth <= kmax2 in the filtered modes; LongEdge keeps total depth <=
    k, no depth in the filtered modes, or a single mode)."""
    import itertools
    from collections import Counter
    sysp = systems.dimer_of_dimers()
    noise = dict(SEED=0, MODEL="FFT_FILTER", TLEN=200.0, TAU=1.0)
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    mask = [True, False] * 4
    mask2 = [False, True] * 4
    filters = []
    if kind in ("Triangular", "both"):
        filters.append(("Triangular", [mask, 1]))
    if kind in ("LongEdge", "both"):
        filters.append(("LongEdge", [mask2 if kind == "both" else mask, 1]))
    maxhier = 3
    tr = O.AdaptiveTrajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": maxhier}, delta_a=1e-14, delta_s=0,
                              filters=filters)
    psi0 = np.zeros(4, complex)
    psi0[1] = 1
    tr.initialize(psi0)
    tr.propagate(40.0, 2.0)
    legal = set()
    for depth in range(maxhier + 1):
        for raw in itertools.combinations_with_replacement(range(8), depth):
            a = tuple(sorted(Counter(raw).items()))
            ok = True
            for name, (mk, k2) in filters:
                fd = sum(v for m, v in a if mk[m])
                if name == "Triangular":
                    ok &= fd <= k2
                else:
                    ok &= sum(v for _, v in a) <= k2 or fd == 0 or len(a) <= 1
            if ok:
                legal.add(a)
    got = set(tr.basis.aux_list)
    print(kind, "legal", len(legal), "adaptive", len(got))
    assert got == legal
    assert len(legal) < 165  # the filter removed auxiliaries


def _embed(b_small, b_full, phi):
    out = np.zeros(b_full.n_hier * b_full.n_state, complex)
    pos = {int(s): i for i, s in enumerate(b_full.state_list)}
    for i, a in enumerate(b_small.aux_list):
        j = b_full.aux_index[a]
        for si, s in enumerate(b_small.state_list):
            out[j * b_full.n_state + pos[int(s)]] = phi[i * b_small.n_state + si]
    return out


@pytest.mark.parametrize("mode", ["hier", "state"])
def test_adaptive_derivative_error_below_delta(mode):
    """|D_full(phi_adaptive) - D_adaptive(phi_adaptive)| <= delta at every step,
    for the chains of tests/test_adap_basis.py:110-361 (Markovian filter on
    the fast mode of every site, seed 100100)."""
    if mode == "hier":
        nsite, V, lam, maxhier, t_max = 3, 40.0, 40.0, 4, 200.0
        da, ds = 5e-4, 0
    else:
        nsite, V, lam, maxhier, t_max = 10, 20.0, 60.0, 2, 100.0
        da, ds = 0, 5e-4
    g0, w0 = O.bcf_convert_dl_to_exp(lam, 50.0, 140.0)
    gw, ll = [], []
    for i in range(nsite):
        L = sp.coo_matrix(([1.0], ([i], [i])), shape=(nsite, nsite))
        gw += [[g0, w0], [-1j * np.imag(g0), 500.0]]
        ll += [L, L]
    markov = [False, True] * nsite
    H = (np.eye(nsite, k=1) + np.eye(nsite, k=-1)) * V
    sysp = dict(HAMILTONIAN=np.array(H, complex), GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll, ALPHA_NOISE1=O.bcf_exp,
                PARAM_NOISE1=gw)
    noise = dict(SEED=100100, MODEL="FFT_FILTER", TLEN=2100.0, TAU=1.0)
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    psi0 = np.zeros(nsite, complex)
    psi0[int(np.round(nsite / 2))] = 1.0
    aux_full = [a for a in O.triangular_hierarchy(2 * nsite, maxhier)
                if sum(v for _, v in a) <= 1 or not any(markov[m] for m, _ in a)]
    full = O.Basis(O.derive_system(sysp), np.arange(nsite), aux_full)
    f_full = O.make_derivative(full, "NORMALIZED NONLINEAR")
    tr = O.AdaptiveTrajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": maxhier}, delta_a=da, delta_s=ds,
                              markov=markov)
    tr.initialize(psi0)
    errs = []
    for t in np.arange(0, t_max, 4.0):
        tr.propagate(4.0, 4.0)
        b = tr.basis
        lops = list(b.absL)
        z1 = tr.noise1.get_noise([t], lops)[:, 0]
        d_ad = tr.dsystem_dt(tr.phi, tr.z_mem, z1, np.zeros(len(lops)))[0] / O.HBAR
        z1f = tr.noise1.get_noise([t], list(full.absL))[:, 0]
        d_full = f_full(_embed(b, full, tr.phi), tr.z_mem, z1f, np.zeros(len(full.absL)))[0] / O.HBAR
        errs.append(np.linalg.norm(d_full - _embed(b, full, d_ad)))
    assert b.n_hier < full.n_hier or b.n_state < full.n_state
    assert max(errs) <= max(da, ds), max(errs)


def test_effective_noise_matches_reference_hand_values():
    """tests/test_integrator_rk.py:45-73: PRE_CALCULATED noise arange(-10,
    10.5, 0.5) (and its negative), TAU 0.25, step 1.5 at t = 5: the effective
    noise is the mean of samples [0:3], [3:6], [6:9] of the window from t."""
    sysp = dict(HAMILTONIAN=np.array([[0, 10.0], [10.0, 0]]), GW_SYSBATH=[[10.0, 10.0], [5.0, 5.0]],
                L_HIER=[np.diag([1.0, 0.0]), np.diag([0.0, 1.0])], L_NOISE1=[np.diag([1.0, 0.0]), np.diag([0.0, 1.0])],
                ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=[[10.0, 10.0], [5.0, 5.0]])
    seed = np.array([np.arange(-10, 10.5, 0.5), -1 * np.arange(-10, 10.5, 0.5)])
    nz = O.Noise({"SEED": seed, "MODEL": "PRE_CALCULATED", "TLEN": 10.0, "TAU": 0.25}, O.derive_system(sysp), 1)
    got = O.effective_noise(nz, 5.0, 1.5, [0, 1])
    known = seed[:, 5 * 4:8 * 4]
    want = np.array([np.mean(known[:, :3], axis=1), np.mean(known[:, 3:6], axis=1),
                     np.mean(known[:, 6:9], axis=1)]).T
    assert np.allclose(got, want)
    # the point samples of the plain integrator are columns 0, 3, 6 of the same window
    assert np.allclose(nz.get_noise([5.0, 5.75, 6.5], [0, 1]), known[:, [0, 3, 6]])


# ---------------------------------------------------------------------------
# _operator and STATIC_BASIS (trajectory/hops_trajectory.py:303-317, 540-577)
# ---------------------------------------------------------------------------
def _blockwise(op_sub, phi, n):
    out = phi.copy()
    for k in range(len(phi) // n):
        out[k * n:(k + 1) * n] = op_sub @ phi[k * n:(k + 1) * n]
    return out


@pytest.mark.parametrize("sparse", [False, True])
def test_operator_acts_on_every_auxiliary(sparse):
    """The oracle's _operator equals op applied aux block by aux block (a
    reshape in the wrong order would mix auxiliaries), for a plain and for a
    state-adaptive trajectory (populated states join the basis, with zero
    amplitude before the product)."""
    sysp = systems.chain(6)
    noise = dict(SEED=3, MODEL="FFT_FILTER", TLEN=100.0, TAU=0.5)
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    psi0 = np.zeros(6, complex)
    psi0[2] = 1
    rng = np.random.default_rng(4)
    op = rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6))
    op[np.abs(op) < 1.0] = 0
    op[5, :] = 0
    op[5, 2] = 0.5  # populates state 5 from state 2
    if sparse:
        op = sp.csr_matrix(op)
    dense = op.toarray() if sparse else op
    tr = O.Trajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 2})
    tr.initialize(psi0)
    tr.propagate(10.0, 1.0)
    before = tr.phi.copy()
    tr.operator(op)
    np.testing.assert_allclose(tr.phi, _blockwise(dense, before, 6), rtol=1e-13, atol=1e-15)
    ad = O.AdaptiveTrajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 2}, delta_a=1e-3, delta_s=1e-3)
    ad.initialize(psi0)
    ad.propagate(10.0, 1.0)
    old_states = [int(s) for s in ad.basis.state_list]
    assert 5 not in old_states
    full_before = {(a, int(s)): ad.phi[i * ad.basis.n_state + j]
                   for i, a in enumerate(ad.basis.aux_list) for j, s in enumerate(ad.basis.state_list)}
    ad.early_counter = 7
    ad.operator(op)
    st = [int(s) for s in ad.basis.state_list]
    assert 5 in st and set(old_states) <= set(st)
    n = len(st)
    vec = np.array([full_before.get((a, s), 0) for a in ad.basis.aux_list for s in st])
    np.testing.assert_allclose(ad.phi, _blockwise(dense[np.ix_(st, st)], vec, n), rtol=1e-13, atol=1e-15)
    assert ad.early_counter == 0


def test_static_basis_is_kept_from_the_start():
    """STATIC_BASIS joins the initial basis without a define_basis pass, and
    STATIC_STATE_INCHWORM_HIERARCHY holds the state basis at the static states
    through the early steps."""
    sysp = systems.chain(6)
    noise = dict(SEED=3, MODEL="FFT_FILTER", TLEN=100.0, TAU=0.5)
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    psi0 = np.zeros(6, complex)
    psi0[2] = 1
    static = [[1, 2, 3], [((4, 1),), ((5, 1),)]]
    tr = O.AdaptiveTrajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 3}, delta_a=1e-3, delta_s=1e-3,
                              early_integrator="STATIC_STATE_INCHWORM_HIERARCHY", static_basis=static)
    tr.initialize(psi0)
    assert [int(s) for s in tr.basis.state_list] == [1, 2, 3]
    assert set(tr.basis.aux_list) == {(), ((4, 1),), ((5, 1),)}
    tr.propagate(5.0, 1.0)
    assert all(s == [1, 2, 3] for s in tr.state_traj[1:6])


def test_interpolated_noise_reproduces_cubics():
    """INTERPOLATE (noise/hops_noise.py:242-243, 379-385): the oracle's noise
    lookup goes through scipy's cubic interp1d along the time axis; a
    not-a-knot cubic spline reproduces a cubic polynomial exactly, so noise
    rows that are cubics in t are returned exactly at off-grid times, and the
    grid values are returned at grid times."""
    sysp = systems.dimer_of_dimers()
    sysd = O.derive_system(sysp)
    t = O.noise_time_axis(20.0, 0.5)
    n_l2 = len(set(O.l_signature(L) for L in sysd["L_NOISE1"]))
    rows = np.array([(0.3 + 0.1j * l) * t ** 3 - (1.0 + l) * t + 2.0j for l in range(n_l2)])
    nz = O.Noise({"SEED": rows, "MODEL": "PRE_CALCULATED", "TLEN": 20.0, "TAU": 0.5, "INTERPOLATE": True}, sysd, 1)
    tq = [0.1, 3.75, 7.3, 19.9]
    got = nz.get_noise(tq, list(range(n_l2)))
    want = np.array([(0.3 + 0.1j * l) * np.array(tq) ** 3 - (1.0 + l) * np.array(tq) + 2.0j for l in range(n_l2)])
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(nz.get_noise([1.5, 2.0], [0]), rows[[0]][:, [3, 4]], rtol=1e-12, atol=1e-12)


def _peierls_chain(ltc):
    """The 5-site Peierls chain of systems.peierls_chain: state 2 is reached
    only through the off-diagonal L-operators of bonds 1-2 and 2-3."""
    return systems.peierls_chain(ltc)


@pytest.mark.parametrize("ltc", [False, True], ids=["plain", "ltc"])
def test_adaptive_offdiagonal_delta_to_zero_equals_full(ltc):
    """Off-diagonal L-operators: with vanishing thresholds the adaptive
    trajectory (its state basis grown only through the Peierls fluxes and
    noise) equals the plain one."""
    sysp = _peierls_chain(ltc)
    noise = dict(SEED=4, MODEL="FFT_FILTER", TLEN=200.0, TAU=1.0)
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    psi0 = np.zeros(5, complex)
    psi0[0] = 1
    ref = O.Trajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 2})
    ref.initialize(psi0)
    ref.propagate(60.0, 2.0)
    tr = O.AdaptiveTrajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 2}, delta_a=1e-15, delta_s=1e-100)
    tr.initialize(psi0)
    tr.propagate(60.0, 2.0)
    assert 2 in tr.state_traj[-1] and 2 not in tr.state_traj[0]
    assert np.max(np.abs(np.array(tr.psi_traj) - np.array(ref.psi_traj))) < 1e-12


@pytest.mark.parametrize("mode", ["hier", "state"])
@pytest.mark.parametrize("ltc", [False, True], ids=["plain", "ltc"])
def test_adaptive_offdiagonal_derivative_error_below_delta(mode, ltc):
    """|D_full(phi_adaptive) - D_adaptive(phi_adaptive)| <= delta at every step
    for off-diagonal L-operators (the bound tests/test_adap_basis.py:110-361
    checks for site projectors): a missing or underestimated flux term of the
    Peierls couplings lets the truncation error exceed delta."""
    sysp = _peierls_chain(ltc)
    da, ds = (3e-3, 0) if mode == "hier" else (0, 3e-3)
    maxhier = 3
    noise = dict(SEED=7, MODEL="FFT_FILTER", TLEN=200.0, TAU=1.0)
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    psi0 = np.zeros(5, complex)
    psi0[0] = 1.0
    full = O.Basis(O.derive_system(sysp), np.arange(5), O.triangular_hierarchy(7, maxhier))
    f_full = O.make_derivative(full, "NORMALIZED NONLINEAR")
    tr = O.AdaptiveTrajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": maxhier}, delta_a=da, delta_s=ds)
    tr.initialize(psi0)
    errs = []
    for t in np.arange(0, 80.0, 2.0):
        tr.propagate(2.0, 2.0)
        b = tr.basis
        # the basis was chosen with the noise at the step's start
        z1 = tr.noise1.get_noise([t], list(b.absL))[:, 0]
        d_ad = tr.dsystem_dt(tr.phi, tr.z_mem, z1, np.zeros(len(b.absL)))[0] / O.HBAR
        z1f = tr.noise1.get_noise([t], list(full.absL))[:, 0]
        d_full = f_full(_embed(b, full, tr.phi), tr.z_mem, z1f, np.zeros(len(full.absL)))[0] / O.HBAR
        errs.append(np.linalg.norm(d_full - _embed(b, full, d_ad)))
    assert b.n_hier < full.n_hier or b.n_state < full.n_state
    assert max(errs) <= max(da, ds), max(errs)
