"""save_checkpoint / load_checkpoint (trajectory/hops_trajectory.py:853-1034):
a trajectory checkpointed at t1 and reloaded continues exactly as the
uninterrupted one (bitwise: tests/test_checkpoint.py:458-491 uses atol
1e-100), non-adaptive and adaptive, with the storage carried over."""
import numpy as np
import pytest

import systems

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("adaptive", [False, True])
def test_checkpoint_resume_is_exact(adaptive, tmp_path):
    _need_gpu()
    from mesohops_amd import HopsTrajectory
    sysp = systems.dimer_of_dimers() if not adaptive else systems.chain(8)
    n = sysp["HAMILTONIAN"].shape[0]
    psi0 = np.zeros(n, complex)
    psi0[n // 2] = 1.0
    noise = {"SEED": 42, "MODEL": "FFT_FILTER", "TLEN": 200.0, "TAU": 0.5}

    def fresh():
        tr = HopsTrajectory(sysp, {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}, dict(noise),
                            hierarchy_param={"MAXHIER": 3})
        if adaptive:
            tr.make_adaptive(2e-3, 2e-3)
        tr.initialize(psi0)
        return tr

    a = fresh()
    a.propagate(40.0, 2.0)
    path = str(tmp_path / "ck.npz")
    a.save_checkpoint(path)
    a.propagate(40.0, 2.0)
    b = HopsTrajectory.load_checkpoint(path)
    assert b.t == 40.0
    b.propagate(40.0, 2.0)
    pa, pb = np.array(a.storage["psi_traj"]), np.array(b.storage["psi_traj"])
    assert pa.shape == pb.shape == (41, n)
    assert np.array_equal(pa, pb)
    assert np.array_equal(a.phi, b.phi)
    assert np.array_equal(a.z_mem.toarray(), b.z_mem.toarray())
    assert list(a.state_list) == list(b.state_list)
    assert set(a.auxiliary_list) == set(b.auxiliary_list)
    if adaptive:
        assert a.storage["list_nhier"] == b.storage["list_nhier"]


def test_resume_a_checkpoint_in_the_reference_layout(tmp_path):
    """A checkpoint laid out as the reference's save_checkpoint writes it
    (trajectory/hops_trajectory.py:884-936), taken from the FP64 oracle of an
    adaptive trajectory after its early steps, resumes on the engine exactly
    as the oracle continues: the same basis sequence and psi(t) to 1e-10."""
    _need_gpu()
    import scipy.sparse as sp
    from mesohops_amd import HopsTrajectory
    from mesohops_amd.trajectory import write_checkpoint
    from oracle import hops_oracle as O
    sysp = systems.chain(8)
    n = 8
    psi0 = np.zeros(n, complex)
    psi0[4] = 1.0
    tlen, tau = 120.0, 0.5
    z = O.Noise({"SEED": 5, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau}, O.derive_system(sysp), 1).noise
    noise = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    ref = O.AdaptiveTrajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 3}, delta_a=2e-3, delta_s=2e-3)
    ref.initialize(psi0)
    ref.propagate(20.0, 2.0)
    b = ref.basis
    n_noise = len(sysp["L_NOISE1"])
    zm = np.zeros((n_noise, 1), dtype=np.complex128)
    zm[: len(ref.z_mem), 0] = ref.z_mem
    params = {"system_param": {k: sysp[k] for k in ("HAMILTONIAN", "GW_SYSBATH", "L_HIER", "L_NOISE1",
                                                    "ALPHA_NOISE1", "PARAM_NOISE1")},
              "hierarchy_param": {"MAXHIER": 3, "TERMINATOR": False, "STATIC_FILTERS": []},
              "eom_param": {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR", "TIME_DEPENDENCE": False,
                            "ADAPTIVE_H": True, "ADAPTIVE_S": True, "DELTA_A": 2e-3, "DELTA_S": 2e-3,
                            "UPDATE_STEP": 1, "F_DISCARD": 0.01},
              "noise1_param": dict(noise, ADAPTIVE=True),
              "noise2_param": {"SEED": None, "MODEL": "ZERO", "TLEN": tlen, "TAU": tau},
              "integration_param": {"INTEGRATOR": "RUNGE_KUTTA", "EARLY_ADAPTIVE_INTEGRATOR": "INCH_WORM",
                                    "EARLY_INTEGRATOR_STEPS": 5, "INCHWORM_CAP": 5, "STATIC_BASIS": None,
                                    "EFFECTIVE_NOISE_INTEGRATION": False},
              "storage_param": {}}
    storage = {k: [] for k in ("psi_traj", "t_axis", "aux_list", "state_list", "list_nhier", "list_nstate",
                               "list_aux_norm")}
    path = str(tmp_path / "from_reference.npz")
    write_checkpoint(path, ref.phi, sp.coo_array(zm).tocsr(), ref.t, b.state_list, b.aux_list, ref.early_counter,
                     storage, {"INITIALIZATION_TIME": 0, "LIST_PROPAGATION_TIME": []}, params)
    tr = HopsTrajectory.load_checkpoint(path)
    assert tr.t == 20.0
    tr.propagate(20.0, 2.0)
    ref.propagate(20.0, 2.0)
    got_states = tr.storage["state_list"]
    assert [list(s) for s in got_states] == [list(s) for s in ref.state_traj[-len(got_states):]]
    err = np.max(np.abs(np.array(tr.storage["psi_traj"]) - np.array(ref.psi_traj[-len(got_states):])))
    print("resumed from a reference-layout checkpoint: max |psi - psi_oracle| = %.3e over %d steps"
          % (err, len(got_states)))
    assert err < 1e-10
    assert np.max(np.abs(tr.phi - ref.phi)) < 1e-10
