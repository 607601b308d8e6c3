"""Ragged batches on the device (native runtime + slot kernels) against the
FP64 oracle: every trajectory of a batch keeps its own adaptive basis; each
must reproduce the oracle's basis sequence exactly and its psi(t) to 1e-10
(FP64), be bitwise the same alone and in a mix, and keep the population
error <= 1e-4 in mixed precision."""
import numpy as np
import pytest

import systems
from oracle import hops_oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
ARENA = 1 << 30


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _noise(sysp, seed, tlen, tau):
    return O.Noise({"SEED": seed, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau}, O.derive_system(sysp), 1).noise


def _oracle(sysp, eom, z, tlen, tau, K, da, ds, t_max, dt, psi0, update=1, markov=None, filters=None):
    noise = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    tr = O.AdaptiveTrajectory(sysp, {"EQUATION_OF_MOTION": eom}, noise, hierarchy_param={"MAXHIER": K},
                              delta_a=da, delta_s=ds, update_step=update, markov=markov,
                              filters=None if filters is None else [(f[0], f[1]) for f in filters])
    tr.initialize(psi0)
    tr.propagate(t_max, dt)
    return tr


CASES = {
    "dod": (systems.dimer_of_dimers(), "NORMALIZED NONLINEAR", 3, 1e-3, 1e-3, 60.0, 2.0, 1),
    "chain8_states": (systems.chain(8), "NONLINEAR", 2, 0.0, 2e-3, 40.0, 2.0, 1),
    "chain10_update2": (systems.chain(10), "NORMALIZED NONLINEAR", 4, 2e-3, 2e-3, 40.0, 1.0, 2),
    "lattice4_ltc": (systems.lattice(4, 4), "NORMALIZED NONLINEAR", 3, 2e-3, 2e-3, 40.0, 1.0, 1),
    # Triangular and LongEdge static filters on the flux up (hops_fluxfilters.py:372-501)
    "dod_triangular_longedge": (systems.dimer_of_dimers(), "NORMALIZED NONLINEAR", 3, 1e-3, 1e-3, 60.0, 2.0, 1,
                                [["Triangular", [[True, False] * 4, 1]], ["LongEdge", [[False, True] * 4, 2]]]),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_adaptive_batch_matches_oracle(name):
    _need_gpu()
    from mesohops_amd.ragged import AdaptiveBatch
    sysp, eom, K, da, ds, t_max, dt, update, *rest = CASES[name]
    filters = rest[0] if rest else None
    nst = sysp["HAMILTONIAN"].shape[0]
    tlen, tau = t_max + 20.0, 0.5
    seeds = [3, 17, 29, 41]
    zs = [_noise(sysp, s, tlen, tau) for s in seeds]
    psi0 = np.zeros(nst, complex)
    psi0[nst // 2] = 1.0
    nps = [{"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau} for z in zs]
    hp = {"MAXHIER": K}
    if filters:
        hp["STATIC_FILTERS"] = [[f[0], [list(f[1][0]), f[1][1]]] for f in filters]
    ab = AdaptiveBatch(sysp, {"EQUATION_OF_MOTION": eom}, nps[0], None, hp, noise_params=nps,
                       precision_mode="fp64", arena_bytes=ARENA)
    ab.make_adaptive(da, ds, update_step=update)
    ab.initialize(psi0)
    ab.propagate(t_max, dt)
    for b, z in enumerate(zs):
        print("case", name, "trajectory", b, flush=True)
        ref = _oracle(sysp, eom, z, tlen, tau, K, da, ds, t_max, dt, psi0, update, filters=filters)
        psi, states, nhier = ab.results(b)
        assert len(states) == len(ref.state_traj)
        for i, (s_got, s_ref) in enumerate(zip(states, ref.state_traj)):
            assert list(s_got) == list(s_ref), (b, i, s_got, s_ref)
            assert nhier[i] == len(ref.aux_traj[i]), (b, i)
        assert set(ab.auxiliary_list(b)) == set(ref.basis.aux_list)
        err = np.max(np.abs(psi - np.array(ref.psi_traj)))
        print("ragged %s traj %d: max |dpsi| = %.3e, n_hier %d..%d" % (name, b, err, min(nhier), max(nhier)))
        assert err < FP64_TOL
        assert np.max(np.abs(ab.phi(b) - ref.phi)) < FP64_TOL
        np.testing.assert_allclose(ab.z_mem(b), ref.z_mem, rtol=0, atol=1e-9 * max(1, np.max(np.abs(ref.z_mem))))


def _strong_ltc_peierls(c):
    p = systems.peierls_chain(True)
    p["PARAM_LT_CORR"] = [c] * len(p["L_LT_CORR"])
    return p


# off-diagonal (Peierls) L-operators in the native runtime; each run starts on
# a state from which the middle of the chain is reached only through the
# Peierls couplings
OFFDIAG_CASES = {
    "peierls_both": (systems.peierls_chain(), "NORMALIZED NONLINEAR", 3, 2e-3, 3e-4, 50.0, 1.0, 0),
    "peierls_strong_ltc": (_strong_ltc_peierls(40.0 + 15j), "NORMALIZED NONLINEAR", 3, 2e-3, 3e-4, 50.0, 1.0, 0),
    "peierls_hier_only": (systems.peierls_chain(), "NONLINEAR", 4, 1e-3, 0.0, 40.0, 2.0, 0),
    "peierls_ring_linear": (systems.peierls_ring(8, ltc=30.0 + 5j), "LINEAR", 3, 2e-3, 3e-4, 40.0, 1.0, 3),
}


@pytest.mark.parametrize("name", sorted(OFFDIAG_CASES))
def test_adaptive_batch_offdiagonal_matches_oracle(name):
    """AdaptiveBatch with off-diagonal L-operators (hops_basis.py:914-998,
    :1030-1100): every trajectory's basis sequence equals the oracle's and
    its psi(t) agrees to 1e-10 in FP64."""
    _need_gpu()
    from mesohops_amd.ragged import AdaptiveBatch
    sysp, eom, K, da, ds, t_max, dt, s0 = OFFDIAG_CASES[name]
    nst = sysp["HAMILTONIAN"].shape[0]
    tlen, tau = t_max + 20.0, 0.5
    seeds = [5, 19, 31]
    zs = [_noise(sysp, s, tlen, tau) for s in seeds]
    psi0 = np.zeros(nst, complex)
    psi0[s0] = 1.0
    nps = [{"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau} for z in zs]
    print("case", name, flush=True)
    ab = AdaptiveBatch(sysp, {"EQUATION_OF_MOTION": eom}, nps[0], None, {"MAXHIER": K}, noise_params=nps,
                       precision_mode="fp64", arena_bytes=ARENA)
    ab.make_adaptive(da, ds)
    ab.initialize(psi0)
    ab.propagate(t_max, dt)
    moved = False
    for b, z in enumerate(zs):
        ref = _oracle(sysp, eom, z, tlen, tau, K, da, ds, t_max, dt, psi0)
        psi, states, nhier = ab.results(b)
        assert len(states) == len(ref.state_traj)
        for i, (s_got, s_ref) in enumerate(zip(states, ref.state_traj)):
            assert list(s_got) == list(s_ref), (b, i, s_got, s_ref)
            assert nhier[i] == len(ref.aux_traj[i]), (b, i)
        assert set(ab.auxiliary_list(b)) == set(ref.basis.aux_list)
        err = np.max(np.abs(psi - np.array(ref.psi_traj)))
        print("ragged off-diagonal %s traj %d: max |dpsi| = %.3e, n_state %d..%d, n_hier %d..%d"
              % (name, b, err, min(len(x) for x in states), max(len(x) for x in states), min(nhier), max(nhier)))
        assert err < FP64_TOL
        moved |= len(set(nhier)) > 1 or len(set(len(x) for x in states)) > 1
    assert moved


def _c128_noise(sysp, seed, tlen, tau):
    """The filter's complex128 values (not rounded to complex64)."""
    nz = O.Noise({"SEED": seed, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau}, O.derive_system(sysp), 1)
    nz.prepare()
    return O.correlated_noise(np.complex64(nz.alpha(list(range(nz.n_l2)))), nz.z_unc)


@pytest.mark.parametrize("name,sub", [("chain6", 1), ("peierls", 1), ("chain6", 3)])
def test_adaptive_batch_interpolated_noise_matches_oracle(name, sub):
    """INTERPOLATE in AdaptiveBatch (noise/hops_noise.py:242-243, 379-385):
    stage times between noise points (TAU 1, step 1.5 fs), two propagate
    calls, each reading the cubic noise of its own half-step grid in
    complex128; with off-diagonal L-operators the selection's noise operator
    reads the same complex128 values.  Basis sequence and psi(t) of every
    trajectory follow the oracle (scipy's interp1d, as the reference) to
    1e-10."""
    _need_gpu()
    from mesohops_amd.ragged import AdaptiveBatch
    sysp = systems.chain(6) if name == "chain6" else systems.peierls_chain()
    s0 = 2 if name == "chain6" else 0
    nst = sysp["HAMILTONIAN"].shape[0]
    tlen, tau = 60.0, 1.0
    zs = [_c128_noise(sysp, seed, tlen, tau) for seed in (12, 40, 77)]
    nps = [{"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau, "INTERPOLATE": True} for z in zs]
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    psi0 = np.zeros(nst, complex)
    psi0[s0] = 1
    da, ds = (1e-3, 1e-3) if name == "chain6" else (2e-3, 3e-4)
    print("case", name, "sub-batches", sub, flush=True)
    ab = AdaptiveBatch(sysp, eom, nps[0], None, {"MAXHIER": 3}, noise_params=nps, precision_mode="fp64",
                       arena_bytes=ARENA, sub_batches=sub)
    ab.make_adaptive(da, ds)
    ab.initialize(psi0)
    ab.propagate(21.0, 1.5)
    ab.propagate(15.0, 1.5)
    for b, n in enumerate(nps):
        ref = O.AdaptiveTrajectory(sysp, eom, n, hierarchy_param={"MAXHIER": 3}, delta_a=da, delta_s=ds)
        ref.initialize(psi0)
        ref.propagate(21.0, 1.5)
        ref.propagate(15.0, 1.5)
        psi, states, nhier = ab.results(b)
        assert len(states) == len(ref.state_traj) == 25
        for i, (s_got, s_ref) in enumerate(zip(states, ref.state_traj)):
            assert list(s_got) == list(s_ref), (b, i, s_got, s_ref)
            assert nhier[i] == len(ref.aux_traj[i]), (b, i)
        err = np.max(np.abs(psi - np.array(ref.psi_traj)))
        print("ragged interpolated %s traj %d: max |dpsi| = %.3e" % (name, b, err))
        assert err < FP64_TOL


def test_adaptive_batch_complex128_precalculated_matches_oracle():
    """A PRE_CALCULATED complex128 noise array whose values complex64 would
    round is used as given in AdaptiveBatch, as the reference uses it
    (noise/hops_noise.py:262-263): psi(t) and bases follow the oracle to
    1e-10, and differ from the run on the same noise rounded to complex64."""
    _need_gpu()
    from mesohops_amd.ragged import AdaptiveBatch
    sysp = systems.chain(6)
    tlen, tau = 50.0, 0.5
    zs = [_c128_noise(sysp, seed, tlen, tau) for seed in (5, 6)]
    assert not np.array_equal(zs[0].astype(np.complex64).astype(np.complex128), zs[0])
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    psi0 = np.zeros(6, complex)
    psi0[2] = 1

    def run(arrs):
        nps = [{"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau} for z in arrs]
        ab = AdaptiveBatch(sysp, eom, nps[0], None, {"MAXHIER": 3}, noise_params=nps, precision_mode="fp64",
                           arena_bytes=ARENA)
        ab.make_adaptive(1e-3, 1e-3)
        ab.initialize(psi0)
        ab.propagate(30.0, 1.0)
        return ab

    ab = run(zs)
    ab32 = run([z.astype(np.complex64) for z in zs])
    for b, z in enumerate(zs):
        ref = _oracle(systems.chain(6), "NORMALIZED NONLINEAR", z, tlen, tau, 3, 1e-3, 1e-3, 30.0, 1.0, psi0)
        psi, states, _ = ab.results(b)
        for i, (s_got, s_ref) in enumerate(zip(states, ref.state_traj)):
            assert list(s_got) == list(s_ref), (b, i)
        err = np.max(np.abs(psi - np.array(ref.psi_traj)))
        d32 = np.max(np.abs(psi - ab32.results(b)[0]))
        print("ragged complex128 noise traj %d: max |dpsi| = %.3e; against the complex64-rounded noise %.3e"
              % (b, err, d32))
        assert err < FP64_TOL
        assert d32 > 0


@pytest.mark.parametrize("early", ["INCH_WORM", "STATIC_STATE_INCHWORM_HIERARCHY"])
def test_adaptive_batch_static_basis_matches_oracle(early):
    """STATIC_BASIS in AdaptiveBatch (trajectory/hops_trajectory.py:299-317,
    :478-491): the first basis joins the static states and auxiliaries, and
    with STATIC_STATE_INCHWORM_HIERARCHY the early steps keep the static
    states; every trajectory's basis sequence and psi(t) follow the oracle
    (1e-10)."""
    _need_gpu()
    from mesohops_amd.ragged import AdaptiveBatch
    sysp = systems.chain(6)
    tlen, tau = 60.0, 0.5
    zs = [_noise(sysp, s, tlen, tau) for s in (5, 9, 13)]
    nps = [{"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau} for z in zs]
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    psi0 = np.zeros(6, complex)
    psi0[2] = 1
    static = [[1, 2, 3], [((4, 1),), ((5, 1),), ((4, 1), (5, 1))]]
    ab = AdaptiveBatch(sysp, eom, nps[0], None, {"MAXHIER": 3}, noise_params=nps, precision_mode="fp64",
                       arena_bytes=ARENA,
                       integration_param={"STATIC_BASIS": static, "EARLY_ADAPTIVE_INTEGRATOR": early})
    ab.make_adaptive(1e-3, 1e-3)
    ab.initialize(psi0)
    ab.propagate(30.0, 1.0)
    for b, z in enumerate(zs):
        noise = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
        ref = O.AdaptiveTrajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 3}, delta_a=1e-3, delta_s=1e-3,
                                   early_integrator=early, static_basis=static)
        ref.initialize(psi0)
        ref.propagate(30.0, 1.0)
        psi, states, nhier = ab.results(b)
        assert len(states) == len(ref.state_traj)
        for i, (s_got, s_ref) in enumerate(zip(states, ref.state_traj)):
            assert list(s_got) == list(s_ref), (b, i, s_got, s_ref)
            assert nhier[i] == len(ref.aux_traj[i]), (b, i)
        err = np.max(np.abs(psi - np.array(ref.psi_traj)))
        print("ragged static basis %s traj %d: max |dpsi| = %.3e" % (early, b, err))
        assert err < FP64_TOL


@pytest.mark.parametrize("sub", [1, 2])
def test_adaptive_batch_effective_noise_matches_oracle(sub):
    """EFFECTIVE_NOISE_INTEGRATION in AdaptiveBatch (integrator/
    integrator_rk.py:115-128): the RK stages read window means of the noise
    while the selection keeps the point samples (hops_trajectory.py:693-714);
    every trajectory's basis sequence and psi(t) follow the oracle (1e-10),
    in one runtime batch and split over two."""
    _need_gpu()
    from mesohops_amd.ragged import AdaptiveBatch
    sysp = systems.chain(6)
    tau_n = 0.25
    zs = [_noise(sysp, s, 60.0, tau_n) for s in (4, 8, 15)]
    nps = [{"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": 60.0, "TAU": tau_n} for z in zs]
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    psi0 = np.zeros(6, complex)
    psi0[3] = 1.0
    ab = AdaptiveBatch(sysp, eom, nps[0], None, {"MAXHIER": 3}, noise_params=nps, precision_mode="fp64",
                       arena_bytes=ARENA, integration_param={"EFFECTIVE_NOISE_INTEGRATION": True}, sub_batches=sub)
    ab.make_adaptive(2e-3, 2e-3)
    ab.initialize(psi0)
    ab.propagate(30.0, 1.5)
    for b, n in enumerate(nps):
        ref = O.AdaptiveTrajectory(sysp, eom, n, hierarchy_param={"MAXHIER": 3}, delta_a=2e-3, delta_s=2e-3,
                                   effective_noise_integration=True)
        ref.initialize(psi0)
        ref.propagate(30.0, 1.5)
        psi, states, nhier = ab.results(b)
        assert [list(x) for x in states] == [list(x) for x in ref.state_traj], b
        err = np.max(np.abs(psi - np.array(ref.psi_traj)))
        print("ragged effective noise traj %d: max |dpsi| = %.3e" % (b, err))
        assert err < FP64_TOL


def test_alone_equals_in_mix():
    """A trajectory's result does not depend on which others share its batch."""
    _need_gpu()
    from mesohops_amd.ragged import AdaptiveBatch
    sysp = systems.chain(10)
    tlen, tau = 80.0, 0.5
    psi0 = np.zeros(10, complex)
    psi0[5] = 1.0
    zs = [_noise(sysp, s, tlen, tau) for s in (5, 6, 7)]
    nps = [{"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau} for z in zs]

    def run(params, sub=None):
        ab = AdaptiveBatch(sysp, {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}, params[0], None, {"MAXHIER": 4},
                           noise_params=params, arena_bytes=ARENA, sub_batches=sub)
        ab.make_adaptive(2e-3, 2e-3)
        ab.initialize(psi0)
        ab.propagate(30.0, 1.0)
        ab.propagate(20.0, 1.0)
        return ab

    mix = run(nps)
    alone = run([nps[1]])
    p_mix, s_mix, h_mix = mix.results(1)
    p_alone, s_alone, h_alone = alone.results(0)
    assert h_mix == h_alone
    assert np.array_equal(p_mix, p_alone)
    # the same batch split over concurrently stepped runtime batches
    # (RaggedGroup): two (2 + 1 trajectories) and three (one each)
    for sub in (2, 3):
        split = run(nps, sub)
        assert split.K == sub
        for b in range(3):
            p_s, s_s, h_s = split.results(b)
            p_m, s_m, h_m = mix.results(b)
            assert h_s == h_m, (sub, b)
            assert all(np.array_equal(x, y) for x, y in zip(s_s, s_m)), (sub, b)
            assert np.array_equal(p_s, p_m), (sub, b)
        split.rt.close()


def test_mixed_precision_population_error():
    _need_gpu()
    from mesohops_amd.ragged import AdaptiveBatch
    sysp = systems.chain(10)
    tlen, tau = 80.0, 0.5
    psi0 = np.zeros(10, complex)
    psi0[5] = 1.0
    zs = [_noise(sysp, s, tlen, tau) for s in (1, 2)]
    nps = [{"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau} for z in zs]
    ab = AdaptiveBatch(sysp, {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}, nps[0], None, {"MAXHIER": 4},
                       noise_params=nps, precision_mode="mixed", arena_bytes=ARENA)
    ab.make_adaptive(2e-3, 2e-3)
    ab.initialize(psi0)
    ab.propagate(50.0, 1.0)
    for b, z in enumerate(zs):
        ref = _oracle(sysp, "NORMALIZED NONLINEAR", z, tlen, tau, 4, 2e-3, 2e-3, 50.0, 1.0, psi0)
        psi, _, _ = ab.results(b)
        err = np.max(np.abs(np.abs(psi) ** 2 - np.abs(np.array(ref.psi_traj)) ** 2))
        print("mixed ragged traj %d: max population error %.3e" % (b, err))
        assert err < 1e-4


def test_ragged_batch_of_different_systems():
    """Slots with different systems and hierarchy depths in one batch."""
    _need_gpu()
    from mesohops_amd.basis import SystemInfo
    from mesohops_amd.ragged import RaggedBatch, RtSystem, traj_param
    tlen, tau = 60.0, 0.5
    specs = [(systems.chain(6), 3, 2e-3), (systems.dimer_of_dimers(), 2, 1e-3), (systems.chain(9, seed_h=4), 4, 3e-3)]
    rb = RaggedBatch("fp64", ARENA, max_slots=8)
    keep = []
    for i, (sysp, K, d) in enumerate(specs):
        info = SystemInfo(sysp)
        rs = RtSystem(info)
        z = _noise(sysp, 100 + i, tlen, tau)
        zt = torch.from_numpy(np.ascontiguousarray(z)).cuda()
        psi0 = np.zeros(info.nstates, complex)
        psi0[info.nstates // 2] = 1.0
        prm = traj_param(2, K, d, d, 1, 0.01, 2.0, 2, z.shape[1], 0)
        slot = rb.add(rs, prm, psi0, zt)
        keep.append((slot, sysp, K, d, z, psi0, info, rs, zt, prm))
    rb.step(20)
    for slot, sysp, K, d, z, psi0, info, *_ in keep:
        ref = _oracle(sysp, "NORMALIZED NONLINEAR", z, tlen, tau, K, d, d, 40.0, 2.0, psi0)
        psi, states, nhier = rb.records(slot, info.nstates)
        assert [list(s) for s in states] == [list(s) for s in ref.state_traj]
        err = np.max(np.abs(psi - np.array(ref.psi_traj)))
        print("mixed-system slot %d: max |dpsi| = %.3e" % (slot, err))
        assert err < FP64_TOL
