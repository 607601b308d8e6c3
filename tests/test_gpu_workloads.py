"""The benchmark workloads at their full problem sizes on the device, each run
for a few steps so the oracle can follow:

- batched LINEAR HOPS, 10-site chain, 2 modes/site, depth 4 (10,626
  auxiliaries x 10 states), 256 trajectories sharing one operator, dt 2 fs:
  complex64 stages against an FP64 run of the same batch (population error
  <= 1e-4), three trajectories alone against the batch (bitwise), and two
  trajectories against the FP64 oracle (1e-10);
- adaptive NORMALIZED NONLINEAR HOPS, 100-site chain, depth 10,
  delta = 1e-3, psi0 on site 50, dt 1 fs: two trajectories against the oracle
  (basis sequence equal at every step, psi(t) to 1e-10), and 256 mixed-
  precision trajectories in one ragged batch (normalised, two of them bitwise
  the same alone, four within 1e-4 of FP64);
- the serving stream: six requests drawn from the serving distribution
  (chain 2-200 sites, depth 2-10, delta in [1e-4, 1e-2]), each cut to 20 fs,
  bitwise the same alone and in the mix;
- adaptive 30 x 30 lattice with the low-temperature correction at 77 K,
  depth 8, delta = 1e-3, psi0 in the middle: two trajectories against the
  oracle, and 64 mixed-precision trajectories within 1e-4 of FP64 on two.
"""
import numpy as np
import pytest

from oracle import hops_oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
POP_TOL = 1e-4


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _oracle_noise(sysp, seed, tlen, tau):
    z = O.Noise({"SEED": seed, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau}, O.derive_system(sysp), 1).noise
    return {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}


def _pop_err(a, b):
    return float(np.max(np.abs(np.abs(np.asarray(a)) ** 2 - np.abs(np.asarray(b)) ** 2)))


# --- 10-site chain, depth 4, one trajectory over all 1000 steps --------------
def test_chain10_depth4_linear_full_length_against_oracle():
    """One trajectory of the batched LINEAR workload over its full 1000 steps
    (2000 fs): FP64 on the device against the FP64 oracle, and complex64
    stages against both.  LINEAR HOPS does not normalise, so the populations
    grow by orders of magnitude; the test prints the absolute population
    differences next to the largest population, which shows how far an
    absolute population bound is met by two FP64 codes at that scale."""
    _need_gpu()
    from mesohops_amd import HopsBatch
    from mesohops_amd.workloads import chain_system
    sysp = chain_system(10)
    eom = {"EQUATION_OF_MOTION": "LINEAR"}
    psi0 = np.zeros(10, complex)
    psi0[0] = 1.0
    tlen, tau, steps, dt = 2048.0, 1.0, 1000, 2.0
    noise = _oracle_noise(sysp, 0, tlen, tau)
    print("chain10 depth 4: one trajectory, %d steps, oracle" % steps, flush=True)
    ref = O.Trajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 4})
    ref.initialize(psi0)
    ref.propagate(steps * dt, dt)
    ref_psi = np.array(ref.psi_traj)
    out = {}
    for prec in ("fp64", "mixed"):
        print("chain10 depth 4: one trajectory, %d steps, %s" % (steps, prec), flush=True)
        b = HopsBatch(sysp, eom, noise, None, {"MAXHIER": 4}, noise_params=[noise], precision_mode=prec)
        b.initialize(psi0)
        b.propagate(steps * dt, dt)
        out[prec] = np.array(b.psi_traj[0])
        del b
    pmax = float(np.max(np.abs(ref_psi) ** 2))
    rel64 = float(np.max(np.abs(out["fp64"] - ref_psi)) / np.max(np.abs(ref_psi)))
    abs64 = _pop_err(out["fp64"], ref_psi)
    absmx = _pop_err(out["mixed"], ref_psi)
    absmx64 = _pop_err(out["mixed"], out["fp64"])
    print("largest population %.4e; FP64 device vs oracle: psi rel %.3e, population abs %.3e (rel %.3e); "
          "complex64 stages vs oracle: population abs %.3e (rel %.3e), vs FP64 device abs %.3e"
          % (pmax, rel64, abs64, abs64 / pmax, absmx, absmx / pmax, absmx64))
    assert len(ref_psi) == steps + 1
    assert rel64 < FP64_TOL
    assert absmx / pmax <= POP_TOL


# --- 10-site chain, depth 4, 256 trajectories, LINEAR ----------------------
def test_chain10_depth4_linear_batch256():
    _need_gpu()
    from mesohops_amd import HopsBatch
    from mesohops_amd.workloads import chain_system
    sysp = chain_system(10)
    eom = {"EQUATION_OF_MOTION": "LINEAR"}
    hier = {"MAXHIER": 4}
    psi0 = np.zeros(10, complex)
    psi0[0] = 1.0
    tlen, tau, steps, dt = 64.0, 1.0, 20, 2.0
    nps = [{"SEED": s, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau} for s in range(256)]
    runs = {}
    for prec in ("mixed", "fp64"):
        print("chain10 depth 4: 256 trajectories, %d steps, %s" % (steps, prec), flush=True)
        b = HopsBatch(sysp, eom, nps[0], None, hier, noise_params=nps, precision_mode=prec)
        b.initialize(psi0)
        assert b.tables.n_hier == 10626 and b.tables.n_state == 10
        b.propagate(steps * dt, dt)
        runs[prec] = np.array(b.psi_traj)
        del b
        torch.cuda.empty_cache()
    err = _pop_err(runs["mixed"], runs["fp64"])
    print("population error, complex64 stages vs FP64 over 256 trajectories: %.3e" % err)
    assert err <= POP_TOL
    for k in (0, 97, 255):
        print("chain10 depth 4: trajectory %d alone" % k, flush=True)
        one = HopsBatch(sysp, eom, nps[k], None, hier, noise_params=[nps[k]], precision_mode="mixed")
        one.initialize(psi0)
        one.propagate(steps * dt, dt)
        assert np.array_equal(np.array(one.psi_traj[0]), runs["mixed"][k]), k
    # two trajectories against the FP64 oracle on identical noise
    print("chain10 depth 4: 2 trajectories against the oracle, 5 steps", flush=True)
    ons = [_oracle_noise(sysp, s, tlen, tau) for s in (0, 1)]
    b = HopsBatch(sysp, eom, ons[0], None, hier, noise_params=ons, precision_mode="fp64")
    b.initialize(psi0)
    b.propagate(5 * dt, dt)
    for k in range(2):
        ref = O.Trajectory(sysp, eom, ons[k], hierarchy_param=hier)
        ref.initialize(psi0)
        ref.propagate(5 * dt, dt)
        err = np.max(np.abs(np.array(b.psi_traj[k]) - np.array(ref.psi_traj)))
        print("trajectory %d: max |psi - psi_oracle| = %.3e" % (k, err))
        assert err < FP64_TOL


# --- 100-site chain, depth 10, adaptive --------------------------------------
def _chain100():
    from mesohops_amd.workloads import chain_system
    sysp = chain_system(100)
    psi0 = np.zeros(100, complex)
    psi0[50] = 1.0
    return sysp, psi0


def test_chain100_depth10_adaptive_against_oracle():
    _need_gpu()
    from mesohops_amd.ragged import AdaptiveBatch
    sysp, psi0 = _chain100()
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    tlen, tau, t_max = 64.0, 0.5, 30.0
    ons = [_oracle_noise(sysp, s, tlen, tau) for s in (0, 1)]
    print("chain100 depth 10: 2 trajectories, %g fs, FP64" % t_max, flush=True)
    ab = AdaptiveBatch(sysp, eom, ons[0], None, {"MAXHIER": 10}, noise_params=ons, precision_mode="fp64")
    ab.make_adaptive(1e-3, 1e-3)
    ab.initialize(psi0, tau=1.0)
    ab.propagate(t_max, 1.0)
    for k in range(2):
        ref = O.AdaptiveTrajectory(sysp, eom, ons[k], hierarchy_param={"MAXHIER": 10}, delta_a=1e-3, delta_s=1e-3)
        ref.initialize(psi0)
        ref.propagate(t_max, 1.0)
        psi, states, nhier = ab.results(k)
        assert len(states) == len(ref.state_traj)
        for i, (s_got, s_ref) in enumerate(zip(states, ref.state_traj)):
            assert list(s_got) == list(s_ref), (k, i)
            assert nhier[i] == len(ref.aux_traj[i]), (k, i)
        assert set(ab.auxiliary_list(k)) == set(ref.basis.aux_list)
        err = np.max(np.abs(psi - np.array(ref.psi_traj)))
        print("trajectory %d: max |psi - psi_oracle| = %.3e, n_hier up to %d" % (k, err, max(nhier)))
        assert err < FP64_TOL
        assert max(nhier) > 100
    ab.rt.close()


def test_chain100_depth10_adaptive_batch256():
    _need_gpu()
    from mesohops_amd.ragged import AdaptiveBatch
    sysp, psi0 = _chain100()
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    noise = {"MODEL": "FFT_FILTER", "TLEN": 64.0, "TAU": 0.5}
    steps = 20

    def run(seeds, prec):
        print("chain100 depth 10: %d trajectories, %d steps, %s" % (len(seeds), steps, prec), flush=True)
        ab = AdaptiveBatch(sysp, eom, dict(noise, SEED=0), None, {"MAXHIER": 10}, seeds=list(seeds),
                           precision_mode=prec)
        ab.make_adaptive(1e-3, 1e-3)
        ab.initialize(psi0, tau=1.0)
        ab.propagate(float(steps), 1.0)
        out = [ab.results(k)[0] for k in range(len(seeds))]
        ab.rt.close()
        return out

    batch = run(range(256), "mixed")
    norms = np.array([np.linalg.norm(p, axis=1) for p in batch])
    assert np.all(np.isfinite(norms))
    assert np.max(np.abs(norms - 1.0)) < 1e-6
    for k in (0, 201):
        alone = run([k], "mixed")[0]
        assert np.array_equal(alone, batch[k]), k
    ref = run(range(4), "fp64")
    err = max(_pop_err(batch[k], ref[k]) for k in range(4))
    print("population error, mixed vs FP64 on 4 trajectories: %.3e" % err)
    assert err <= POP_TOL


# --- serving stream ------------------------------------------------------------
def test_serving_stream_requests_alone_equal_in_mix():
    _need_gpu()
    from mesohops_amd.serving import ServingLoop
    from mesohops_amd.workloads import c4_requests
    reqs = c4_requests(6)
    for r in reqs:
        r.t_max = 20.0
    print("serving: requests", [(r.rid, r.nsite, r.maxhier, round(r.delta, 5)) for r in reqs], flush=True)
    loop = ServingLoop(precision_mode="mixed", arena_bytes=4 << 30, max_slots=8, record=True)
    mix, stats = loop.run(reqs)
    loop.close()
    assert stats["completed"] == len(reqs)
    for r in reqs:
        r.arrival = 0.0
        one = ServingLoop(precision_mode="mixed", arena_bytes=4 << 30, max_slots=2, record=True)
        alone, _ = one.run([r])
        one.close()
        a, m = alone[r.rid], mix[r.rid]
        assert a["nhier"] == m["nhier"], r.rid
        assert all(np.array_equal(x, y) for x, y in zip(a["states"], m["states"]))
        assert np.array_equal(a["psi"], m["psi"]), r.rid
        print("request %d (%d sites, depth %d): alone == in mix, n_hier up to %d"
              % (r.rid, r.nsite, r.maxhier, max(a["nhier"])))


# --- 30 x 30 lattice, depth 8, low-temperature correction ---------------------
def _lattice():
    from mesohops_amd.workloads import lattice_system
    sysp = lattice_system(30, 30)
    psi0 = np.zeros(900, complex)
    psi0[15 * 30 + 15] = 1.0
    return sysp, psi0


def test_lattice30_depth8_ltc_against_oracle():
    _need_gpu()
    from mesohops_amd.ragged import AdaptiveBatch
    sysp, psi0 = _lattice()
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    tlen, tau, t_max = 64.0, 0.5, 12.0
    ons = [_oracle_noise(sysp, s, tlen, tau) for s in (0, 1)]
    print("lattice 30x30 depth 8: 2 trajectories, %g fs, FP64" % t_max, flush=True)
    ab = AdaptiveBatch(sysp, eom, ons[0], None, {"MAXHIER": 8}, noise_params=ons, precision_mode="fp64")
    ab.make_adaptive(1e-3, 1e-3)
    ab.initialize(psi0, tau=1.0)
    ab.propagate(t_max, 1.0)
    for k in range(2):
        ref = O.AdaptiveTrajectory(sysp, eom, ons[k], hierarchy_param={"MAXHIER": 8}, delta_a=1e-3, delta_s=1e-3)
        ref.initialize(psi0)
        ref.propagate(t_max, 1.0)
        psi, states, nhier = ab.results(k)
        for i, (s_got, s_ref) in enumerate(zip(states, ref.state_traj)):
            assert list(s_got) == list(s_ref), (k, i)
            assert nhier[i] == len(ref.aux_traj[i]), (k, i)
        err = np.max(np.abs(psi - np.array(ref.psi_traj)))
        print("trajectory %d: max |psi - psi_oracle| = %.3e, n_state up to %d, n_hier up to %d"
              % (k, err, max(len(s) for s in states), max(nhier)))
        assert err < FP64_TOL
    ab.rt.close()


def test_lattice30_depth8_ltc_batch64():
    _need_gpu()
    from mesohops_amd.ragged import AdaptiveBatch
    sysp, psi0 = _lattice()
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    noise = {"MODEL": "FFT_FILTER", "TLEN": 64.0, "TAU": 0.5}
    out = {}
    for prec, nb in (("mixed", 64), ("fp64", 2)):
        print("lattice 30x30 depth 8: %d trajectories, 12 steps, %s" % (nb, prec), flush=True)
        ab = AdaptiveBatch(sysp, eom, dict(noise, SEED=0), None, {"MAXHIER": 8}, seeds=list(range(nb)),
                           precision_mode=prec)
        ab.make_adaptive(1e-3, 1e-3)
        ab.initialize(psi0, tau=1.0)
        ab.propagate(12.0, 1.0)
        out[prec] = [ab.results(k)[0] for k in range(nb)]
        ab.rt.close()
    assert all(np.all(np.isfinite(p)) for p in out["mixed"])
    err = max(_pop_err(out["mixed"][k], out["fp64"][k]) for k in range(2))
    print("population error, mixed vs FP64 on 2 trajectories: %.3e" % err)
    assert err <= POP_TOL
