"""Adaptive HOPS on the device against the FP64 oracle.

hops_derivative (the derivative the adaptive error bound needs) equals the
oracle's dsystem_dt to 1e-10 relative; whole adaptive trajectories reproduce
the oracle's basis sequence (state_list and aux_list at every step) exactly
and its psi(t) to 1e-10.  Noise arrays are shared via PRE_CALCULATED.
"""
import ctypes

import numpy as np
import pytest

import systems
from oracle import hops_oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _noise(sysp, seed, tlen, tau):
    nz = O.Noise({"SEED": seed, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau}, O.derive_system(sysp), 1)
    return nz.noise


@pytest.mark.parametrize("eom", ["NORMALIZED NONLINEAR", "NONLINEAR", "LINEAR"])
@pytest.mark.parametrize("sys_name", ["dod", "chain_ltc", "three_general"])
def test_device_derivative_matches_oracle(eom, sys_name):
    _need_gpu()
    from mesohops_amd import _lib
    from mesohops_amd.basis import BasisTables, SystemInfo
    from mesohops_amd.engine import BatchState, DeviceOperator, _stream_ptr
    from mesohops_amd.trajectory import _EOM_KIND
    sysp = {"dod": systems.dimer_of_dimers(), "chain_ltc": systems.chain(5, ltc=20.0 - 3.0j),
            "three_general": systems.three_site(True)}[sys_name]
    info = SystemInfo(sysp)
    sysd = O.derive_system(sysp)
    rng = np.random.default_rng(3)
    # a ragged aux set; for the dimer of dimers also a state sub-basis whose
    # aux list keeps modes of the dropped state (L-operators without an
    # active state).  The LTC chain keeps every state: with an inactive
    # L-operator the reference's LINEAR LTC term has mismatched shapes
    # (eom/hops_eom.py:527-530 multiplies the active-L coefficients with all
    # basis L^2) and raises.
    full = O.triangular_hierarchy(sysd["N_HMODES"], 2)
    aux = [full[0]] + [a for i, a in enumerate(full[1:]) if rng.random() < 0.7]
    states = list(range(info.nstates - 1)) if sys_name == "dod" else list(range(info.nstates))
    tb = BasisTables(info, states, aux, _EOM_KIND[eom])
    b = O.Basis(sysd, states, tb.aux_list)
    dsys = O.make_derivative(b, eom)
    dop = DeviceOperator(tb, _lib.HOPS_PREC_FP64)
    n_t = 8
    z1h = (rng.normal(size=(tb.n_act, n_t)) + 1j * rng.normal(size=(tb.n_act, n_t))).astype(np.complex64)
    z2h = (rng.normal(size=(tb.n_act, n_t)) + 1j * rng.normal(size=(tb.n_act, n_t))).astype(np.complex64)
    z1 = torch.from_numpy(z1h[None]).cuda()
    z2 = torch.from_numpy(z2h[None]).cuda()
    st = BatchState(dop, 1, _EOM_KIND[eom], z1, z2, n_t, lanes=1)
    zmem = (rng.normal(size=info.n_hmodes) + 1j * rng.normal(size=info.n_hmodes)) * 30.0
    st.zmem[: info.n_hmodes].copy_(torch.from_numpy(zmem))
    phi = rng.normal(size=tb.n_hier * tb.n_state) + 1j * rng.normal(size=tb.n_hier * tb.n_state)
    y = torch.from_numpy(phi).cuda()
    d = torch.empty_like(y)
    tidx = 5
    _lib.check(_lib.lib().hops_derivative(ctypes.byref(dop.op), ctypes.byref(st.bt), ctypes.c_void_p(y.data_ptr()),
                                          ctypes.c_void_p(d.data_ptr()), tidx, _stream_ptr(dop.device)))
    got = d.cpu().numpy()
    # oracle noise is indexed by the basis L-operators (relative order of
    # absL) and, like the reference's get_noise output, is complex64
    act = {int(l): j for j, l in enumerate(tb.absL_active)}
    zo1 = np.array([z1h[act[int(l)], tidx] if int(l) in act else 0 for l in b.absL], dtype=np.complex64)
    zo2 = np.array([z2h[act[int(l)], tidx] if int(l) in act else 0 for l in b.absL], dtype=np.complex64)
    ref = dsys(phi, zmem, zo1, zo2)[0]
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print("derivative", sys_name, eom, "rel err %.3e" % err)
    assert err < FP64_TOL


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


ADAPTIVE_CASES = {
    "dod_both": (systems.dimer_of_dimers(), None, 1e-3, 1e-3, 3, "NORMALIZED NONLINEAR", 80.0, 2.0),
    "chain_hier_only": (systems.chain(6), None, 1e-3, 0.0, 4, "NORMALIZED NONLINEAR", 60.0, 2.0),
    "chain_states_only": (systems.chain(8), None, 0.0, 2e-3, 2, "NONLINEAR", 60.0, 2.0),
    "markov_chain": (*_markov_chain(7), 2e-3, 2e-3, 4, "NORMALIZED NONLINEAR", 60.0, 2.0),
    "chain10_update3": (systems.chain(10), None, 3e-3, 3e-3, 3, "NORMALIZED NONLINEAR", 60.0, 1.0),
    # 2D lattice with a low-temperature correction and noise modes beyond the
    # hierarchy modes (the Matsubara modes treated by the correction)
    "lattice4_ltc": (systems.lattice(4, 4), None, 2e-3, 2e-3, 3, "NORMALIZED NONLINEAR", 40.0, 1.0),
    # Triangular / LongEdge static filters on the flux up (hops_fluxfilters.py:372-501)
    "dod_triangular": (systems.dimer_of_dimers(), None, 1e-3, 1e-3, 3, "NORMALIZED NONLINEAR", 60.0, 2.0,
                       [["Triangular", [[True, False] * 4, 1]]]),
    "chain_longedge": (systems.chain(6), None, 1e-3, 0.0, 4, "NORMALIZED NONLINEAR", 60.0, 2.0,
                       [["LongEdge", [[False, True] * 6, 1]]]),
    "markov_triangular_longedge": (*_markov_chain(7), 2e-3, 2e-3, 4, "NORMALIZED NONLINEAR", 60.0, 2.0,
                                   [["Triangular", [[True, False] * 7, 2]], ["LongEdge", [[True, False] * 7, 3]]]),
}


@pytest.mark.parametrize("name", sorted(ADAPTIVE_CASES))
def test_adaptive_trajectory_matches_oracle(name):
    _need_gpu()
    from mesohops_amd import HopsTrajectory
    sysp, markov, da, ds, K, eom, t_max, dt, *rest = ADAPTIVE_CASES[name]
    filters = rest[0] if rest else []
    update = 3 if name.endswith("update3") else 1
    nst = sysp["HAMILTONIAN"].shape[0]
    tlen, tau = t_max + 20.0, 0.5
    z = _noise(sysp, 11, tlen, tau)
    noise = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    psi0 = np.zeros(nst, complex)
    psi0[nst // 2] = 1.0
    ref = O.AdaptiveTrajectory(sysp, {"EQUATION_OF_MOTION": eom}, noise, hierarchy_param={"MAXHIER": K},
                               delta_a=da, delta_s=ds, update_step=update, markov=markov,
                               filters=[(f[0], f[1]) for f in filters])
    ref.initialize(psi0)
    ref.propagate(t_max, dt)
    hp = {"MAXHIER": K}
    if markov is not None:
        hp["STATIC_FILTERS"] = [["Markovian", markov]]
    if filters:
        hp["STATIC_FILTERS"] = hp.get("STATIC_FILTERS", []) + [[f[0], [list(f[1][0]), f[1][1]]] for f in filters]
    tr = HopsTrajectory(sysp, {"EQUATION_OF_MOTION": eom}, dict(noise), hierarchy_param=hp)
    tr.make_adaptive(da, ds, update_step=update)
    tr.initialize(psi0)
    tr.propagate(t_max, dt)
    got_states = tr.storage["state_list"]
    got_aux = tr.storage["list_nhier"]
    assert len(got_states) == len(ref.state_traj)
    for i, (s_got, s_ref) in enumerate(zip(got_states, ref.state_traj)):
        assert list(s_got) == list(s_ref), (i, s_got, s_ref)
        assert got_aux[i] == len(ref.aux_traj[i]), i
    assert [tuple(a) for a in tr.auxiliary_list] == list(ref.basis.aux_list)
    psi = tr.storage["psi_traj"]
    err = np.max(np.abs(psi - np.array(ref.psi_traj)))
    sizes = sorted(set(zip(tr.storage["list_nstate"], tr.storage["list_nhier"])))
    print("adaptive %s: max |dpsi| = %.3e, (n_state, n_hier) from %s to %s" % (name, err, sizes[0], sizes[-1]))
    assert err < FP64_TOL
    assert np.max(np.abs(tr.phi - ref.phi)) < FP64_TOL
    assert len(sizes) > 1   # the basis moved during the run


def _strong_ltc_peierls(c):
    p = systems.peierls_chain(True)
    p["PARAM_LT_CORR"] = [c] * len(p["L_LT_CORR"])
    return p


# off-diagonal (Peierls) L-operators; the run starts on a state from which the
# middle of the chain is reached only through the Peierls couplings
OFFDIAG_CASES = {
    "peierls_both": (systems.peierls_chain(), 2e-3, 3e-4, 3, "NORMALIZED NONLINEAR", 0, 50.0, 1.0),
    "peierls_strong_ltc": (_strong_ltc_peierls(40.0 + 15j), 2e-3, 3e-4, 3, "NORMALIZED NONLINEAR", 0, 50.0, 1.0),
    "peierls_hier_only": (systems.peierls_chain(), 1e-3, 0.0, 4, "NONLINEAR", 0, 40.0, 2.0),
    "peierls_states_nonlinear": (_strong_ltc_peierls(40.0 + 15j), 0.0, 2e-4, 2, "NONLINEAR", 0, 40.0, 1.0),
    "peierls_ring_linear": (systems.peierls_ring(8, ltc=30.0 + 5j), 2e-3, 3e-4, 3, "LINEAR", 3, 40.0, 1.0),
}


@pytest.mark.parametrize("name", sorted(OFFDIAG_CASES))
def test_adaptive_offdiagonal_matches_oracle(name):
    """Adaptive HopsTrajectory with off-diagonal L-operators (the reference's
    Peierls couplings, hops_basis.py:914-998, :1030-1100): basis sequence
    equal to the oracle's at every step and psi(t) to 1e-10."""
    _need_gpu()
    from mesohops_amd import HopsTrajectory
    sysp, da, ds, K, eom, s0, t_max, dt = OFFDIAG_CASES[name]
    nst = sysp["HAMILTONIAN"].shape[0]
    tlen, tau = t_max + 20.0, 0.5
    z = _noise(sysp, 23, tlen, tau)
    noise = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    psi0 = np.zeros(nst, complex)
    psi0[s0] = 1.0
    print("case", name, flush=True)
    ref = O.AdaptiveTrajectory(sysp, {"EQUATION_OF_MOTION": eom}, noise, hierarchy_param={"MAXHIER": K},
                               delta_a=da, delta_s=ds)
    ref.initialize(psi0)
    ref.propagate(t_max, dt)
    tr = HopsTrajectory(sysp, {"EQUATION_OF_MOTION": eom}, dict(noise), hierarchy_param={"MAXHIER": K})
    tr.make_adaptive(da, ds)
    tr.initialize(psi0)
    tr.propagate(t_max, dt)
    assert len(tr.storage["state_list"]) == len(ref.state_traj)
    for i, (s_got, s_ref) in enumerate(zip(tr.storage["state_list"], ref.state_traj)):
        assert list(s_got) == list(s_ref), (i, s_got, s_ref)
        assert tr.storage["list_nhier"][i] == len(ref.aux_traj[i]), i
    assert set(tr.auxiliary_list) == set(ref.basis.aux_list)
    err = np.max(np.abs(tr.storage["psi_traj"] - np.array(ref.psi_traj)))
    sizes = sorted(set(zip(tr.storage["list_nstate"], tr.storage["list_nhier"])))
    print("off-diagonal %s: max |dpsi| = %.3e, (n_state, n_hier) from %s to %s" % (name, err, sizes[0], sizes[-1]))
    assert err < FP64_TOL
    assert len(sizes) > 1


@pytest.mark.parametrize("early", ["INCH_WORM", "STATIC_STATE_INCHWORM_HIERARCHY"])
def test_static_basis_matches_oracle(early):
    """STATIC_BASIS and the STATIC_STATE_INCHWORM_HIERARCHY early integrator
    (trajectory/hops_trajectory.py:303-317, :469-509): basis sequence equal to
    the oracle's at every step and psi(t) to 1e-10."""
    _need_gpu()
    from mesohops_amd import HopsTrajectory
    sysp = systems.chain(6)
    tlen, tau = 60.0, 0.5
    z = _noise(sysp, 5, tlen, tau)
    noise = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    psi0 = np.zeros(6, complex)
    psi0[2] = 1
    static = [[1, 2, 3], [((4, 1),), ((5, 1),), ((4, 1), (5, 1))]]
    ref = O.AdaptiveTrajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 3}, delta_a=1e-3, delta_s=1e-3,
                               early_integrator=early, static_basis=static)
    ref.initialize(psi0)
    ref.propagate(30.0, 1.0)
    tr = HopsTrajectory(sysp, eom, dict(noise), hierarchy_param={"MAXHIER": 3},
                        integration_param={"STATIC_BASIS": static, "EARLY_ADAPTIVE_INTEGRATOR": early})
    tr.make_adaptive(1e-3, 1e-3)
    tr.initialize(psi0)
    tr.propagate(30.0, 1.0)
    for i, (s_got, s_ref) in enumerate(zip(tr.storage["state_list"], ref.state_traj)):
        assert list(s_got) == list(s_ref), (i, s_got, s_ref)
        assert tr.storage["list_nhier"][i] == len(ref.aux_traj[i]), i
    assert set(tr.auxiliary_list) == set(ref.basis.aux_list)
    err = np.max(np.abs(tr.storage["psi_traj"] - np.array(ref.psi_traj)))
    print("static basis", early, "max |dpsi| = %.3e" % err)
    assert err < FP64_TOL


@pytest.mark.parametrize("adaptive", [False, True])
def test_operator_matches_oracle(adaptive):
    """HopsTrajectory._operator between two propagate calls
    (trajectory/hops_trajectory.py:540-577): with state adaptivity the
    populated states join the basis and the early integrator restarts; psi(t)
    and the basis sequence follow the oracle to 1e-10."""
    _need_gpu()
    import scipy.sparse as sp
    from mesohops_amd import HopsTrajectory
    sysp = systems.chain(6)
    tlen, tau = 60.0, 0.5
    z = _noise(sysp, 8, tlen, tau)
    noise = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    psi0 = np.zeros(6, complex)
    psi0[2] = 1
    op = sp.csr_matrix(([1.0, 1.0, 0.3j], ([2, 5, 0], [5, 2, 2])), shape=(6, 6)) + sp.eye(6) * 0.5
    if adaptive:
        ref = O.AdaptiveTrajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 2}, delta_a=1e-3, delta_s=1e-3)
    else:
        ref = O.Trajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 2})
    ref.initialize(psi0)
    ref.propagate(10.0, 1.0)
    ref.operator(op)
    ref.propagate(10.0, 1.0)
    tr = HopsTrajectory(sysp, eom, dict(noise), hierarchy_param={"MAXHIER": 2})
    if adaptive:
        tr.make_adaptive(1e-3, 1e-3)
    tr.initialize(psi0)
    tr.propagate(10.0, 1.0)
    tr._operator(op)
    if adaptive:
        assert 5 in tr.state_list and 0 in tr.state_list
    tr.propagate(10.0, 1.0)
    psi = np.array(tr.storage["psi_traj"])
    err = np.max(np.abs(psi - np.array(ref.psi_traj)))
    print("operator adaptive=%s max |dpsi| = %.3e" % (adaptive, err))
    assert err < FP64_TOL
    if adaptive:
        for i, (s_got, s_ref) in enumerate(zip(tr.storage["state_list"], ref.state_traj)):
            assert list(s_got) == list(s_ref), (i, s_got, s_ref)
    assert np.max(np.abs(tr.phi.reshape(-1) - ref.phi)) < FP64_TOL


def test_adaptive_interpolated_noise_matches_oracle():
    """INTERPOLATE with an adaptive basis and stage times between noise
    points (TAU 1, step 1.5 fs): each propagate call reads the cubic noise on
    its half-step times; basis sequence and psi(t) follow the oracle (scipy's
    interp1d, as the reference) to 1e-10, across two propagate calls."""
    _need_gpu()
    from mesohops_amd import HopsTrajectory
    sysp = systems.chain(6)
    sysd = O.derive_system(sysp)
    tlen, tau = 60.0, 1.0
    nz = O.Noise({"SEED": 12, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau}, sysd, 1)
    nz.prepare()
    z = O.correlated_noise(np.complex64(nz.alpha(list(range(nz.n_l2)))), nz.z_unc)
    noise = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau, "INTERPOLATE": True}
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    psi0 = np.zeros(6, complex)
    psi0[2] = 1
    ref = O.AdaptiveTrajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 3}, delta_a=1e-3, delta_s=1e-3)
    ref.initialize(psi0)
    ref.propagate(21.0, 1.5)
    ref.propagate(15.0, 1.5)
    tr = HopsTrajectory(sysp, eom, dict(noise), hierarchy_param={"MAXHIER": 3})
    tr.make_adaptive(1e-3, 1e-3)
    tr.initialize(psi0)
    tr.propagate(21.0, 1.5)
    tr.propagate(15.0, 1.5)
    for i, (s_got, s_ref) in enumerate(zip(tr.storage["state_list"], ref.state_traj)):
        assert list(s_got) == list(s_ref), (i, s_got, s_ref)
    err = np.max(np.abs(np.array(tr.storage["psi_traj"]) - np.array(ref.psi_traj)))
    print("adaptive interpolated noise max |dpsi| = %.3e" % err)
    assert len(tr.storage["psi_traj"]) == len(ref.psi_traj) == 25
    assert err < FP64_TOL
