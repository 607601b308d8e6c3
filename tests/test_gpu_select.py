"""Adaptive basis selection on the device (hops_select.hip) against the same
selection on the host threads (hops_runtime.cpp select_hier /
select_states): two AdaptiveBatch runs of the same trajectories on the same
noise, one with device_select on and one off, give the same state and
auxiliary lists at every step and the same psi(t), bit for bit.  The host
selection is itself held to the oracle's define_basis by test_gpu_ragged.py
and test_runtime_host.py."""
import numpy as np
import pytest

import systems
from oracle import hops_oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ARENA = 1 << 30


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _noise(sysp, seed, tlen, tau):
    return O.Noise({"SEED": seed, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau}, O.derive_system(sysp), 1).noise


def _markov(sysp, every):
    n_modes = len(O.derive_system(sysp)["G"])
    return [[bool(m % every == 0) for m in range(n_modes)]]


# system, EOM, depth, delta_a, delta_s, t_max, dt, update_step, precision, extra hierarchy param
CASES = {
    "chain12": (systems.chain(12), "NORMALIZED NONLINEAR", 4, 1e-3, 1e-3, 30.0, 1.0, 1, "fp64", None),
    "chain10_hier_only": (systems.chain(10), "NORMALIZED NONLINEAR", 5, 5e-4, 0.0, 30.0, 1.0, 1, "fp64", None),
    "chain10_linear_update2": (systems.chain(10), "LINEAR", 4, 2e-3, 2e-3, 30.0, 1.0, 2, "fp64", None),
    "dod_markov": (systems.dimer_of_dimers(), "NORMALIZED NONLINEAR", 3, 1e-3, 1e-3, 40.0, 2.0, 1, "fp64",
                   "markov"),
    "lattice4_ltc": (systems.lattice(4, 4), "NORMALIZED NONLINEAR", 3, 2e-3, 2e-3, 30.0, 1.0, 1, "fp64", None),
    "chain16_mixed": (systems.chain(16), "NORMALIZED NONLINEAR", 6, 1e-3, 1e-3, 30.0, 1.0, 1, "mixed", None),
}


def _run(name, device_select):
    from mesohops_amd.ragged import AdaptiveBatch
    sysp, eom, K, da, ds, t_max, dt, update, prec, extra = CASES[name]
    nst = sysp["HAMILTONIAN"].shape[0]
    tlen, tau = t_max + 20.0, 0.5
    zs = [_noise(sysp, s, tlen, tau) for s in (3, 17, 29, 41, 53)]
    nps = [{"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau} for z in zs]
    hp = {"MAXHIER": K}
    if extra == "markov":
        hp["STATIC_FILTERS"] = [["Markovian", _markov(sysp, 3)[0]]]
    psi0 = np.zeros(nst, complex)
    psi0[nst // 2] = 1.0
    ab = AdaptiveBatch(sysp, {"EQUATION_OF_MOTION": eom}, nps[0], None, hp, noise_params=nps,
                       precision_mode=prec, arena_bytes=ARENA, device_select=device_select)
    ab.make_adaptive(da, ds, update_step=update)
    ab.initialize(psi0)
    ab.propagate(t_max, dt)
    return ab, len(zs)


@pytest.mark.skip(reason="the device selection made an illegal memory access on the GPU in its first run "
                         "(case chain10_hier_only, first propagate); no cause is found: its kernels run bounds-clean "
                         "and bitwise equal to the host selection in the CPU emulation of "
                         "tests/test_select_emulation.py, which does not cover the rocPRIM segmented sorts. "
                         "The path stays off (device_select=0) until a cause is shown.")
@pytest.mark.parametrize("name", sorted(CASES))
def test_device_selection_equals_host(name):
    _need_gpu()
    print("case", name, "device selection", flush=True)
    dev, B = _run(name, True)
    print("case", name, "host selection", flush=True)
    host, _ = _run(name, False)
    t = dev.rt.timers()
    print("device run timers:", t)
    for b in range(B):
        p_d, s_d, h_d = dev.results(b)
        p_h, s_h, h_h = host.results(b)
        assert h_d == h_h, (b, h_d, h_h)
        assert [list(x) for x in s_d] == [list(x) for x in s_h], b
        assert sorted(dev.auxiliary_list(b)) == sorted(host.auxiliary_list(b)), b
        assert np.array_equal(p_d, p_h), (b, np.max(np.abs(p_d - p_h)))
        assert np.array_equal(dev.phi(b), host.phi(b)), b
        print("case %s traj %d: bases and psi bitwise equal over %d records, n_hier %d..%d"
              % (name, b, len(h_d), min(h_d), max(h_d)))
