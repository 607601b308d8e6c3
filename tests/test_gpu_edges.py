"""Edge cases of the device path: the smallest systems, the ZERO noise model,
batch sizes that do not fill the lane interleave, and the error paths of a
trajectory (hops_trajectory.py:253-336, :401-538, :579-614).

FP64 results are compared with the oracle to 1e-10; layouts with each other
bit for bit.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import systems
from oracle import hops_oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-10


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _one_site():
    g, w = O.bcf_convert_dl_to_exp(35.0, 50.0, 295.0)
    L = [sp.coo_matrix(([1.0], ([0], [0])), shape=(1, 1))]
    return {"HAMILTONIAN": np.array([[25.0]], complex), "GW_SYSBATH": [[g, w]], "L_HIER": L, "L_NOISE1": L,
            "ALPHA_NOISE1": O.bcf_exp, "PARAM_NOISE1": [[g, w]]}


def _run_pair(sysp, eom, maxhier, noise, psi0, t_max=20.0, dt=1.0):
    from mesohops_amd import HopsTrajectory
    ref = O.Trajectory(sysp, {"EQUATION_OF_MOTION": eom}, noise, hierarchy_param={"MAXHIER": maxhier})
    ref.initialize(psi0)
    ref.propagate(t_max, dt)
    dev = HopsTrajectory(sysp, {"EQUATION_OF_MOTION": eom}, dict(noise), hierarchy_param={"MAXHIER": maxhier})
    dev.initialize(psi0)
    dev.propagate(t_max, dt)
    return ref, dev


@pytest.mark.parametrize("eom", ["LINEAR", "NONLINEAR", "NORMALIZED NONLINEAR"])
@pytest.mark.parametrize("case", ["one_state", "depth_zero", "zero_noise"])
def test_smallest_systems_match_oracle(case, eom):
    """One state with one mode; MAXHIER 0 (the physical wave function alone);
    the ZERO noise model."""
    _need_gpu()
    tlen, tau = 30.0, 0.5
    if case == "one_state":
        sysp, maxhier, psi0 = _one_site(), 3, np.array([1.0], complex)
    else:
        sysp, maxhier = systems.chain(3, k_mats=1), (0 if case == "depth_zero" else 2)
        psi0 = np.array([0.0, 1.0, 0.0], complex)
    if case == "zero_noise":
        noise = {"SEED": None, "MODEL": "ZERO", "TLEN": tlen, "TAU": tau}
    else:
        z = O.Noise({"SEED": 2, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau}, O.derive_system(sysp), 1).noise
        noise = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    ref, dev = _run_pair(sysp, eom, maxhier, noise, psi0)
    want = np.array(ref.psi_traj)
    got = np.array(dev.storage["psi_traj"])
    assert got.shape == want.shape
    err = np.max(np.abs(got - want)) / max(1.0, np.max(np.abs(want)))
    print(case, eom, "max |dpsi| =", err)
    assert err < TOL
    assert np.max(np.abs(dev.phi - ref.phi)) < TOL * max(1.0, np.max(np.abs(ref.phi)))


def test_odd_batch_with_wide_lanes_equals_alone():
    """An odd batch with 4 lanes interleaved and 2 lanes per thread requested:
    the kernel falls back to one lane per thread for the batch (no vector
    straddles the last trajectory), and every trajectory equals its run alone."""
    _need_gpu()
    from mesohops_amd import HopsBatch, _lib
    sysp = systems.chain(4, k_mats=1)
    tlen, tau = 40.0, 0.5
    B = 35
    zs = [O.Noise({"SEED": s, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau}, O.derive_system(sysp), 1).noise
          for s in range(B)]
    nps = [{"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau} for z in zs]
    psi0 = np.zeros(4, complex)
    psi0[2] = 1
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    for lds in (0, 1):
        _lib.set_option("vec_lanes", 2)
        _lib.set_option("lds_stage", lds)
        try:
            b = HopsBatch(sysp, eom, nps[0], None, {"MAXHIER": 3}, noise_params=nps, precision_mode="mixed",
                          lanes=4)
            b.initialize(psi0)
            b.propagate(20.0, 1.0)
        finally:
            _lib.set_option("vec_lanes", 1)
            _lib.set_option("lds_stage", 0)
        for i in (0, 17, B - 1):
            one = HopsBatch(sysp, eom, nps[i], None, {"MAXHIER": 3}, noise_params=[nps[i]], precision_mode="mixed",
                            lanes=1)
            one.initialize(psi0)
            one.propagate(20.0, 1.0)
            assert np.array_equal(np.array(b.psi_traj[i]), np.array(one.psi_traj[0])), (lds, i)


def test_trajectory_error_paths():
    """Re-initialisation, propagation before initialisation, past the noise
    length, and a step that is not a whole number of noise half-steps raise the
    reference's exceptions."""
    _need_gpu()
    from mesohops_amd import HopsTrajectory, LockedException, TrajectoryError, UnsupportedRequest
    sysp = systems.chain(3, k_mats=0)
    noise = {"SEED": 0, "MODEL": "FFT_FILTER", "TLEN": 20.0, "TAU": 1.0}
    tr = HopsTrajectory(sysp, {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}, dict(noise),
                        hierarchy_param={"MAXHIER": 2})
    with pytest.raises(TrajectoryError):
        tr.propagate(4.0, 2.0)
    psi0 = np.array([1.0, 0.0, 0.0], complex)
    tr.initialize(psi0)
    with pytest.raises(LockedException):
        tr.initialize(psi0)
    with pytest.raises(TrajectoryError):
        tr.propagate(40.0, 2.0)
    with pytest.raises((TrajectoryError, UnsupportedRequest)):
        tr.propagate(4.0, 1.0)
    tr.propagate(4.0, 2.0)  # still usable after the refused calls
    assert abs(tr.t - 4.0) < 1e-12
    with pytest.raises(TrajectoryError):
        tr.make_adaptive(1e-3, 1e-3)
