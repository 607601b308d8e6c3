"""Parity of the HIP engine with the FP64 oracle, on the device.

Tolerances: FP64 runs must match the oracle's psi(t) to 1e-10 (absolute, psi
is normalised or O(1)); mixed precision (complex64 stages, complex128 carry)
must keep the population error max_t,n | |psi_n|^2 - |psi_n^ref|^2 | <= 1e-4.
The noise is the same array on both sides (PRE_CALCULATED ingestion,
noise/hops_noise.py:250-300) unless a test says otherwise.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import systems
from oracle import hops_oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
POP_TOL = 1e-4
RESIDENT_DEFAULT = 1024  # the library's default "resident_rows"


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.fixture(autouse=True)
def _launched_forms():
    """Small bases run in the resident form by default; this module tests the
    launched stage kernels (the resident tests switch it on themselves)."""
    if not torch.cuda.is_available():
        yield
        return
    from mesohops_amd import _lib
    _lib.set_option("resident_rows", 0)
    try:
        yield
    finally:
        _lib.set_option("resident_rows", RESIDENT_DEFAULT)


def oracle_run(sysp, eom, noise_param, maxhier, t_max, dt, noise2_param=None):
    tr = O.Trajectory(sysp, {"EQUATION_OF_MOTION": eom}, noise_param, noise2_param,
                      hierarchy_param={"MAXHIER": maxhier})
    psi0 = np.zeros(tr.sysd["NSTATES"], complex)
    psi0[tr.sysd["NSTATES"] // 2] = 1.0
    tr.initialize(psi0)
    tr.propagate(t_max, dt)
    return tr, psi0


def engine_run(sysp, eom, noise_arrays, tlen, tau_noise, maxhier, t_max, dt, psi0, prec="fp64",
               noise2_arrays=None, lanes=None, tile_order=False):
    from mesohops_amd import HopsBatch
    nps = [{"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau_noise} for z in noise_arrays]
    n2 = None
    if noise2_arrays is not None:
        n2 = {"SEED": noise2_arrays[0], "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau_noise,
              "FLAG_REAL": False}
    b = HopsBatch(sysp, {"EQUATION_OF_MOTION": eom}, nps[0], n2, {"MAXHIER": maxhier},
                  noise_params=nps, precision_mode=prec, lanes=lanes)
    b.initialize(psi0)
    if tile_order:
        b.state.use_tile_order(True)
    b.propagate(t_max, dt)
    return b


def _oracle_noise(sysp, seed, tlen, tau, rand="SUM_GAUSSIAN"):
    sysd = O.derive_system(sysp)
    nz = O.Noise({"SEED": seed, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau, "RAND_MODEL": rand}, sysd, 1)
    return nz.noise


@pytest.mark.parametrize("eom", ["NORMALIZED NONLINEAR", "NONLINEAR", "LINEAR"])
def test_fp64_dimer_of_dimers(eom):
    _need_gpu()
    sysp = systems.dimer_of_dimers()
    tlen, tau = 400.0, 1.0
    z = _oracle_noise(sysp, 0, tlen, tau, "BOX_MULLER")
    nparam = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    tr, psi0 = oracle_run(sysp, eom, nparam, 2, 200.0, 4.0)
    b = engine_run(sysp, eom, [z], tlen, tau, 2, 200.0, 4.0, psi0)
    ref = np.array(tr.psi_traj)
    got = np.array(b.psi_traj[0])
    assert got.shape == ref.shape
    err = np.max(np.abs(got - ref))
    print("fp64 dimer-of-dimers", eom, "max |dpsi| =", err)
    assert err < FP64_TOL
    # full hierarchy vector too
    assert np.max(np.abs(b.phi[0] - tr.phi)) < FP64_TOL


@pytest.mark.parametrize("general", [False, True])
@pytest.mark.parametrize("eom", ["NORMALIZED NONLINEAR", "NONLINEAR", "LINEAR"])
def test_fp64_three_site_general_L(general, eom):
    """Diagonal and generalised (complex off-diagonal) L-operators, depth 3
    (the systems of tests/test_hops_eom.py:231-316)."""
    _need_gpu()
    sysp = systems.three_site(general)
    tlen, tau = 60.0, 0.5
    z = _oracle_noise(sysp, 3, tlen, tau)
    nparam = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    tr, psi0 = oracle_run(sysp, eom, nparam, 3, 40.0, 1.0)
    b = engine_run(sysp, eom, [z], tlen, tau, 3, 40.0, 1.0, psi0)
    err = np.max(np.abs(np.array(b.psi_traj[0]) - np.array(tr.psi_traj)))
    print("fp64 3-site general=%s %s max |dpsi| = %.3e" % (general, eom, err))
    assert err < FP64_TOL
    if eom != "LINEAR":
        zm = b.state.zmem_host()[0]
        assert np.max(np.abs(zm[: len(tr.z_mem)] - tr.z_mem)) < 1e-8 * max(1.0, np.max(np.abs(tr.z_mem)))


@pytest.mark.parametrize("eom", ["NORMALIZED NONLINEAR", "LINEAR"])
def test_fp64_low_temperature_correction(eom):
    _need_gpu()
    sysp = systems.chain(4, k_mats=0, ltc=37.5 - 12.0j)
    tlen, tau = 120.0, 0.5
    z = _oracle_noise(sysp, 5, tlen, tau)
    nparam = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    tr, psi0 = oracle_run(sysp, eom, nparam, 3, 100.0, 1.0)
    b = engine_run(sysp, eom, [z], tlen, tau, 3, 100.0, 1.0, psi0)
    err = np.max(np.abs(np.array(b.psi_traj[0]) - np.array(tr.psi_traj)))
    print("fp64 LTC", eom, "max |dpsi| =", err)
    assert err < FP64_TOL


def test_fp64_linear_dimer_c1():
    """2-site LINEAR dimer, depth 4, dt 4 fs to 1000 fs (TLEN 1024, TAU 2)."""
    _need_gpu()
    sysp = systems.dimer_c1()
    tlen, tau = 1024.0, 2.0
    z = _oracle_noise(sysp, 0, tlen, tau)
    nparam = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    tr = O.Trajectory(sysp, {"EQUATION_OF_MOTION": "LINEAR"}, nparam, hierarchy_param={"MAXHIER": 4})
    psi0 = np.array([1.0, 0.0], complex)
    tr.initialize(psi0)
    tr.propagate(1000.0, 4.0)
    b = engine_run(sysp, "LINEAR", [z], tlen, tau, 4, 1000.0, 4.0, psi0)
    err = np.max(np.abs(np.array(b.psi_traj[0]) - np.array(tr.psi_traj)))
    print("fp64 C1 dimer max |dpsi| =", err)
    assert err < FP64_TOL * max(1.0, np.max(np.abs(tr.psi_traj)))


def test_super_operator_matches_fixture_exactly():
    """Device-built K2_kp1 + K2_km1 and K2_k equal the reference's golden
    matrices for the dimer of dimers bit for bit (test_dimer_of_dimers.py:163-187)."""
    _need_gpu()
    import os
    from mesohops_amd import _lib
    from mesohops_amd.basis import BasisTables, SystemInfo, triangular_aux_list
    from mesohops_amd.engine import DeviceOperator
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_fixtures.npz"))
    sysp = systems.dimer_of_dimers()
    info = SystemInfo(sysp)
    tb = BasisTables(info, np.arange(4), triangular_aux_list(info.n_hmodes, 2), _lib.HOPS_EOM_NORMALIZED_NONLINEAR)
    dop = DeviceOperator(tb, _lib.HOPS_PREC_FP64)
    r, c, v = dop.ell_dense()
    N = tb.n_state * tb.n_hier
    got = sp.coo_matrix((v, (r, c)), shape=(N, N)).toarray()
    want = np.zeros((N, N), complex)
    for name in ("K2_kp1", "K2_km1"):
        idx, val = fx[name + "_idx"], fx[name + "_val"]
        want[idx[0], idx[1]] += val
    assert np.array_equal(got, want)
    k0 = dop.k0.cpu().numpy()
    idx, val = fx["K2_k_idx"], fx["K2_k_val"]
    want_k = np.zeros(tb.n_hier, complex)
    want_k[idx[0]] = val
    assert np.array_equal(k0, want_k)


def test_mixed_precision_population_error():
    """complex64 stages with a complex128 carry stay within 1e-4 in population."""
    _need_gpu()
    sysp = systems.chain(6, k_mats=1)
    tlen, tau = 520.0, 1.0
    z = _oracle_noise(sysp, 1, tlen, tau)
    for eom in ("LINEAR", "NORMALIZED NONLINEAR"):
        nparam = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
        tr, psi0 = oracle_run(sysp, eom, nparam, 3, 500.0, 2.0)
        b = engine_run(sysp, eom, [z], tlen, tau, 3, 500.0, 2.0, psi0, prec="mixed")
        ref = np.abs(np.array(tr.psi_traj)) ** 2
        got = np.abs(np.array(b.psi_traj[0])) ** 2
        err = np.max(np.abs(got - ref))
        print("mixed", eom, "max population error =", err)
        assert err < POP_TOL


def test_batch_equals_alone():
    """Each trajectory of a batch gives the same result as when run alone."""
    _need_gpu()
    sysp = systems.chain(5, k_mats=1)
    tlen, tau = 130.0, 1.0
    zs = [_oracle_noise(sysp, s, tlen, tau) for s in range(9)]
    psi0 = np.zeros(5, complex)
    psi0[2] = 1
    for prec in ("fp64", "mixed"):
        bb = engine_run(sysp, "NORMALIZED NONLINEAR", zs, tlen, tau, 3, 100.0, 2.0, psi0, prec=prec)
        for s in (0, 4, 8):
            one = engine_run(sysp, "NORMALIZED NONLINEAR", [zs[s]], tlen, tau, 3, 100.0, 2.0, psi0, prec=prec)
            d = np.max(np.abs(np.array(bb.psi_traj[s]) - np.array(one.psi_traj[0])))
            print(prec, "batch vs alone traj", s, d)
            assert d == 0.0


@pytest.mark.parametrize("tlen,tau", [(1024.0, 2.0), (1000.0, 1.0), (60.0, 0.5), (2048.0, 0.5), (1536.0, 0.5)])
def test_device_noise_filter(tlen, tau):
    """Device circulant-embedding filter vs the oracle (power-of-two and
    Bluestein lengths, LDS-resident rows and, for TLEN 1536 at TAU 0.5 with a
    16384-point Bluestein transform, global scratch rows); both start from
    the same NumPy raw draws.  The
    reference transforms the complex64 raw noise in single precision, which
    bounds the agreement at ~1e-7 of the noise scale; the bar is 1e-6."""
    _need_gpu()
    from mesohops_amd.basis import SystemInfo
    from mesohops_amd.engine import _stream_ptr
    from mesohops_amd.noise import HopsNoise, device_noise
    sysp = systems.dimer_of_dimers()
    info = SystemInfo(sysp)
    for rand in ("SUM_GAUSSIAN", "BOX_MULLER"):
        p = {"SEED": 7, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau, "RAND_MODEL": rand}
        nz = HopsNoise(p, info, sysp, 1)
        dev = torch.device("cuda:0")
        got = device_noise([nz], list(range(info.n_l2)), dev, _stream_ptr(dev)).cpu().numpy()[0]
        want = _oracle_noise(sysp, 7, tlen, tau, rand)
        scale = np.max(np.abs(want))
        err = np.max(np.abs(got - want)) / scale
        print("noise", tlen, tau, rand, "max rel err", err)
        assert err < 1e-6


@pytest.mark.parametrize("sysname", ["chain", "three_general"])
def test_stationary_kernel_equals_streaming(sysname):
    """The operator-stationary stage kernel (batches of >= 8) and the streaming
    kernel give bit-identical trajectories."""
    _need_gpu()
    from mesohops_amd import _lib
    sysp = systems.chain(5, k_mats=1) if sysname == "chain" else systems.three_site(True)
    tlen, tau = 70.0, 0.5
    zs = [_oracle_noise(sysp, s, tlen, tau) for s in range(12)]
    psi0 = np.zeros(len(sysp["HAMILTONIAN"]), complex)
    psi0[1] = 1
    for prec in ("fp64", "mixed"):
        for eom in ("LINEAR", "NORMALIZED NONLINEAR"):
            _lib.set_option("stationary_stage", 1)
            a = engine_run(sysp, eom, zs, tlen, tau, 3, 50.0, 1.0, psi0, prec=prec)
            _lib.set_option("stationary_stage", 0)
            b = engine_run(sysp, eom, zs, tlen, tau, 3, 50.0, 1.0, psi0, prec=prec)
            _lib.set_option("stationary_stage", 1)
            d = max(np.max(np.abs(np.array(a.psi_traj[i]) - np.array(b.psi_traj[i]))) for i in range(12))
            print(sysname, prec, eom, "stationary vs streaming", d)
            assert d == 0.0


def test_stationary_kernel_options_bit_identical():
    """The nontemporal hint on the streamed arrays, the trajectory-group split
    and the locality tile order of the stationary kernel change caching and
    work distribution only: trajectories are bit-identical to the default
    launch."""
    _need_gpu()
    from mesohops_amd import _lib
    sysp = systems.chain(4, k_mats=1)
    tlen, tau = 50.0, 0.5
    B = 70
    zs = [_oracle_noise(sysp, s, tlen, tau) for s in range(B)]
    psi0 = np.zeros(4, complex)
    psi0[0] = 1
    for prec in ("mixed", "fp64"):
        ref = engine_run(sysp, "NORMALIZED NONLINEAR", zs, tlen, tau, 3, 30.0, 1.0, psi0, prec=prec, lanes=2)
        for nt, q in ((1, 0), (2, 0), (3, 0), (4, 0), (0, 2), (1, 5)):
            _lib.set_option("nt_stream", nt)
            _lib.set_option("groups_q", q)
            try:
                got = engine_run(sysp, "NORMALIZED NONLINEAR", zs, tlen, tau, 3, 30.0, 1.0, psi0, prec=prec, lanes=2)
            finally:
                _lib.set_option("nt_stream", 0)
                _lib.set_option("groups_q", 0)
            d = max(np.max(np.abs(np.array(got.psi_traj[i]) - np.array(ref.psi_traj[i]))) for i in range(B))
            print(prec, "nt", nt, "q", q, "max diff", d)
            assert d == 0.0
        got = engine_run(sysp, "NORMALIZED NONLINEAR", zs, tlen, tau, 3, 30.0, 1.0, psi0, prec=prec, lanes=2,
                         tile_order=True)
        d = max(np.max(np.abs(np.array(got.psi_traj[i]) - np.array(ref.psi_traj[i]))) for i in range(B))
        print(prec, "locality tile order max diff", d)
        assert d == 0.0


@pytest.mark.parametrize("eom", ["LINEAR", "NORMALIZED NONLINEAR"])
def test_vector_lanes_bit_identical(eom):
    """Two or four trajectory lanes per thread of the stationary kernel change
    only the width of its memory accesses, and its LDS-staged form only where
    the pattern values and coefficients are read from: every lane performs the
    same operations, so trajectories are bit-identical to one lane per thread,
    at every interleave and hint level, in both precisions."""
    _need_gpu()
    from mesohops_amd import _lib
    sysp = systems.chain(4, k_mats=1)
    tlen, tau = 50.0, 0.5
    B = 72
    zs = [_oracle_noise(sysp, s, tlen, tau) for s in range(B)]
    psi0 = np.zeros(4, complex)
    psi0[0] = 1
    for prec in ("mixed", "fp64"):
        ref = engine_run(sysp, eom, zs, tlen, tau, 3, 30.0, 1.0, psi0, prec=prec, lanes=4)
        for lanes, vec, nt, lds in ((4, 2, 0, 0), (4, 4, 0, 0), (8, 2, 2, 0), (8, 4, 2, 0), (8, 4, 3, 0),
                                    (2, 4, 0, 0), (4, 1, 0, 1), (4, 2, 2, 1), (8, 2, 2, 1), (8, 4, 3, 1),
                                    (1, 1, 0, 1)):
            _lib.set_option("vec_lanes", vec)
            _lib.set_option("nt_stream", nt)
            _lib.set_option("lds_stage", lds)
            try:
                got = engine_run(sysp, eom, zs, tlen, tau, 3, 30.0, 1.0, psi0, prec=prec, lanes=lanes)
            finally:
                _lib.set_option("vec_lanes", 1)
                _lib.set_option("nt_stream", 0)
                _lib.set_option("lds_stage", 0)
            d = max(np.max(np.abs(np.array(got.psi_traj[i]) - np.array(ref.psi_traj[i]))) for i in range(B))
            print(prec, "lanes", lanes, "vec", vec, "nt", nt, "lds", lds, "max diff", d)
            assert d == 0.0


def test_fused_linear_prologue_bit_identical():
    """The four LINEAR stage prologues of a step in one launch (the default)
    give the same trajectories as one prologue launch per stage, bit for bit:
    with the low-temperature correction and a second noise, in both
    precisions, with the streaming (lanes 1, B < 8 groups) and stationary
    kernels."""
    _need_gpu()
    from mesohops_amd import _lib
    sysp = systems.chain(4, k_mats=1, ltc=37.5 - 12.0j)
    sysp.update(L_NOISE2=sysp["L_NOISE1"], ALPHA_NOISE2=O.bcf_exp, PARAM_NOISE2=sysp["PARAM_NOISE1"])
    tlen, tau = 50.0, 0.5
    B = 40
    zs = [_oracle_noise(sysp, s, tlen, tau) for s in range(B)]
    z2 = [_oracle_noise(sysp, 100 + s, tlen, tau) for s in range(B)]
    psi0 = np.zeros(4, complex)
    psi0[1] = 1
    for prec in ("mixed", "fp64"):
        for lanes in (1, 4):
            _lib.set_option("fuse_prologue", 0)
            try:
                ref = engine_run(sysp, "LINEAR", zs, tlen, tau, 3, 30.0, 1.0, psi0, prec=prec, noise2_arrays=z2,
                                 lanes=lanes)
            finally:
                _lib.set_option("fuse_prologue", 1)
            got = engine_run(sysp, "LINEAR", zs, tlen, tau, 3, 30.0, 1.0, psi0, prec=prec, noise2_arrays=z2,
                             lanes=lanes)
            d = max(np.max(np.abs(np.array(got.psi_traj[i]) - np.array(ref.psi_traj[i]))) for i in range(B))
            print(prec, "lanes", lanes, "fused vs per-stage prologue max diff", d)
            assert d == 0.0
            assert np.array_equal(got.phi, ref.phi)


def _batch_two_calls(sysp, eom, zs, tlen, tau, psi0, prec, lanes, z2=None):
    """Two propagate calls (17 and 13 steps of 1 fs) of one batch."""
    from mesohops_amd import HopsBatch
    nps = [{"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau} for z in zs]
    n2 = None
    if z2 is not None:
        n2 = {"SEED": z2[0], "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau, "FLAG_REAL": False}
    b = HopsBatch(sysp, {"EQUATION_OF_MOTION": eom}, nps[0], n2, {"MAXHIER": 3}, noise_params=nps,
                  precision_mode=prec, lanes=lanes)
    b.initialize(psi0)
    b.propagate(17.0, 1.0)
    b.propagate(13.0, 1.0)
    return b


@pytest.mark.parametrize("eom", ["LINEAR", "NORMALIZED NONLINEAR"])
def test_graph_replay_bit_identical(eom):
    """RK4 steps replayed from a captured HIP graph give the same trajectories
    and hierarchy as eager launches, bit for bit: graphs of 1, 3 and 4 steps
    (a call replays its graph only when it has at least 4 replays; remainder
    steps run eagerly), two propagate calls in a row (17 and 13 steps), both
    precisions, streaming and stationary stage kernels, the LINEAR fused
    prologue and the per-stage prologues."""
    _need_gpu()
    from mesohops_amd import _lib
    sysp = systems.chain(4, k_mats=1, ltc=37.5 - 12.0j)
    tlen, tau = 80.0, 0.5
    B = 40
    zs = [_oracle_noise(sysp, s, tlen, tau) for s in range(B)]
    psi0 = np.zeros(4, complex)
    psi0[1] = 1
    for prec in ("mixed", "fp64"):
        for lanes in (1, 4):
            _lib.set_option("graph_steps", 0)
            try:
                ref = _batch_two_calls(sysp, eom, zs, tlen, tau, psi0, prec, lanes)
                for K in (1, 3, 4):
                    _lib.set_option("graph_steps", K)
                    got = _batch_two_calls(sysp, eom, zs, tlen, tau, psi0, prec, lanes)
                    d = max(np.max(np.abs(np.array(got.psi_traj[i]) - np.array(ref.psi_traj[i])))
                            for i in range(B))
                    print(eom, prec, "lanes", lanes, "graph_steps", K, "max diff vs eager", d, flush=True)
                    assert len(got.psi_traj[0]) == len(ref.psi_traj[0]) == 31
                    assert d == 0.0
                    assert np.array_equal(got.phi, ref.phi)
            finally:
                _lib.set_option("graph_steps", 16)


@pytest.mark.parametrize("eom", ["LINEAR", "NONLINEAR", "NORMALIZED NONLINEAR"])
def test_resident_steps_bit_identical(eom):
    """The resident form (one workgroup per trajectory running every step of
    a call) gives the same trajectories and hierarchy as the launched stage
    kernels, bit for bit: two propagate calls, both precisions, one and four
    interleaved lanes, the low-temperature correction and a second noise."""
    _need_gpu()
    from mesohops_amd import _lib
    sysp = systems.chain(4, k_mats=1, ltc=37.5 - 12.0j)
    sysp.update(L_NOISE2=sysp["L_NOISE1"], ALPHA_NOISE2=O.bcf_exp, PARAM_NOISE2=sysp["PARAM_NOISE1"])
    tlen, tau = 80.0, 0.5
    B = 40
    zs = [_oracle_noise(sysp, s, tlen, tau) for s in range(B)]
    z2 = [_oracle_noise(sysp, 100 + s, tlen, tau) for s in range(B)]
    psi0 = np.zeros(4, complex)
    psi0[1] = 1
    for prec in ("mixed", "fp64"):
        for lanes in (1, 4):
            _lib.set_option("resident_rows", 0)
            ref = _batch_two_calls(sysp, eom, zs, tlen, tau, psi0, prec, lanes, z2=z2)
            _lib.set_option("resident_rows", 1 << 20)
            try:
                got = _batch_two_calls(sysp, eom, zs, tlen, tau, psi0, prec, lanes, z2=z2)
            finally:
                _lib.set_option("resident_rows", 0)
            d = max(np.max(np.abs(np.array(got.psi_traj[i]) - np.array(ref.psi_traj[i]))) for i in range(B))
            print(eom, prec, "lanes", lanes, "resident vs launched max diff", d, flush=True)
            assert d == 0.0
            assert np.array_equal(got.phi, ref.phi)


def test_resident_steps_match_oracle_c1():
    """The C1 dimer (2 sites, depth 4, dt 4 fs to 1000 fs) in the resident
    form against the FP64 oracle."""
    _need_gpu()
    from mesohops_amd import _lib
    sysp = systems.dimer_c1()
    tlen, tau = 1024.0, 2.0
    z = _oracle_noise(sysp, 0, tlen, tau)
    nparam = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    tr = O.Trajectory(sysp, {"EQUATION_OF_MOTION": "LINEAR"}, nparam, hierarchy_param={"MAXHIER": 4})
    psi0 = np.array([1.0, 0.0], complex)
    tr.initialize(psi0)
    tr.propagate(1000.0, 4.0)
    _lib.set_option("resident_rows", 1 << 20)
    try:
        b = engine_run(sysp, "LINEAR", [z], tlen, tau, 4, 1000.0, 4.0, psi0)
    finally:
        _lib.set_option("resident_rows", 0)
    err = np.max(np.abs(np.array(b.psi_traj[0]) - np.array(tr.psi_traj)))
    print("resident C1 dimer max |dpsi| =", err)
    assert err < FP64_TOL * max(1.0, np.max(np.abs(tr.psi_traj)))


@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_interleaved_lanes_equal_plain_layout(lanes):
    """Interleaving T trajectories per row changes only the memory layout:
    trajectories are bit-identical to the one-per-row layout (both kernels)."""
    _need_gpu()
    from mesohops_amd import _lib
    sysp = systems.chain(4, k_mats=1)
    tlen, tau = 60.0, 0.5
    B = 70
    zs = [_oracle_noise(sysp, s, tlen, tau) for s in range(B)]
    psi0 = np.zeros(4, complex)
    psi0[0] = 1
    for eom in ("LINEAR", "NORMALIZED NONLINEAR"):
        ref = engine_run(sysp, eom, zs, tlen, tau, 3, 40.0, 1.0, psi0, prec="mixed", lanes=1)
        for stationary in (1, 0):
            _lib.set_option("stationary_stage", stationary)
            got = engine_run(sysp, eom, zs, tlen, tau, 3, 40.0, 1.0, psi0, prec="mixed", lanes=lanes)
            _lib.set_option("stationary_stage", 1)
            d = max(np.max(np.abs(np.array(got.psi_traj[i]) - np.array(ref.psi_traj[i]))) for i in range(B))
            dphi = np.max(np.abs(got.phi - ref.phi))
            print("lanes", lanes, eom, "stationary", stationary, d, dphi)
            assert d == 0.0 and dphi == 0.0


def test_chunked_propagation_equals_whole_batch():
    """Propagating cache-sized chunks of the batch one after another (all
    steps of a chunk before the next) gives bit-identical trajectories."""
    _need_gpu()
    from mesohops_amd import _lib
    sysp = systems.chain(4, k_mats=1)
    tlen, tau = 60.0, 0.5
    B = 70
    zs = [_oracle_noise(sysp, s, tlen, tau) for s in range(B)]
    psi0 = np.zeros(4, complex)
    psi0[0] = 1
    for eom in ("LINEAR", "NORMALIZED NONLINEAR"):
        ref = engine_run(sysp, eom, zs, tlen, tau, 3, 40.0, 1.0, psi0, prec="mixed", lanes=4)
        _lib.set_option("chunk_traj", 32)
        try:
            got = engine_run(sysp, eom, zs, tlen, tau, 3, 40.0, 1.0, psi0, prec="mixed", lanes=4)
        finally:
            _lib.set_option("chunk_traj", 0)
        d = max(np.max(np.abs(np.array(got.psi_traj[i]) - np.array(ref.psi_traj[i]))) for i in range(B))
        dphi = np.max(np.abs(got.phi - ref.phi))
        dz = np.max(np.abs(got.state.zmem_host() - ref.state.zmem_host()))
        print("chunked", eom, d, dphi, dz)
        assert d == 0.0 and dphi == 0.0 and dz == 0.0


@pytest.mark.parametrize("flag_real", [False, True])
@pytest.mark.parametrize("eom", ["NORMALIZED NONLINEAR", "NONLINEAR", "LINEAR"])
def test_fp64_second_noise(eom, flag_real):
    """A nonzero noise 2 (PRE_CALCULATED, complex or FLAG_REAL): the
    coefficient conj(z1) - i z2 of LINEAR HOPS is formed in complex64, as the
    reference forms it from its complex64 noise arrays."""
    _need_gpu()
    from mesohops_amd import HopsBatch
    sysp = systems.three_site(False)
    sysp.update(L_NOISE2=sysp["L_NOISE1"], ALPHA_NOISE2=O.bcf_exp, PARAM_NOISE2=sysp["PARAM_NOISE1"])
    tlen, tau = 60.0, 0.5
    z1 = _oracle_noise(sysp, 3, tlen, tau)
    z2 = _oracle_noise(sysp, 4, tlen, tau)
    n1 = {"SEED": z1, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    n2 = {"SEED": z2, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau, "FLAG_REAL": flag_real}
    tr, psi0 = oracle_run(sysp, eom, n1, 3, 40.0, 1.0, noise2_param=n2)
    b = HopsBatch(sysp, {"EQUATION_OF_MOTION": eom}, n1, dict(n2), {"MAXHIER": 3}, noise_params=[n1],
                  precision_mode="fp64")
    b.initialize(psi0)
    b.propagate(40.0, 1.0)
    ref = np.array(tr.psi_traj)
    err = np.max(np.abs(np.array(b.psi_traj[0]) - ref)) / max(1.0, np.max(np.abs(ref)))
    print("fp64 noise2 %s flag_real=%s max |dpsi| = %.3e" % (eom, flag_real, err))
    assert err < FP64_TOL


@pytest.mark.parametrize("ratio", [2, 6, 10])
@pytest.mark.parametrize("eom", ["NORMALIZED NONLINEAR", "LINEAR"])
def test_fp64_effective_noise_integration(eom, ratio):
    """EFFECTIVE_NOISE_INTEGRATION (integrator/integrator_rk.py:115-128): the
    device moving average reproduces NumPy's complex64 means (sequential below
    4 samples, blocked pairwise from 4), so psi(t) matches the oracle to 1e-10;
    ratio = tau / TAU gives windows of 1, 3 and 5 samples."""
    _need_gpu()
    from mesohops_amd import HopsTrajectory
    sysp = systems.three_site(False)
    tau_n = 0.25
    dt = ratio * tau_n
    tlen = 40.0
    z = _oracle_noise(sysp, 9, tlen, tau_n)
    nparam = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau_n}
    ref = O.Trajectory(sysp, {"EQUATION_OF_MOTION": eom}, nparam, hierarchy_param={"MAXHIER": 3},
                       effective_noise_integration=True)
    psi0 = np.array([0, 1, 0], complex)
    ref.initialize(psi0)
    t_max = dt * int(20.0 / dt)
    ref.propagate(t_max, dt)
    tr = HopsTrajectory(sysp, {"EQUATION_OF_MOTION": eom}, dict(nparam), hierarchy_param={"MAXHIER": 3},
                        integration_param={"EFFECTIVE_NOISE_INTEGRATION": True})
    tr.initialize(psi0)
    tr.propagate(t_max, dt)
    got = np.array(tr.storage["psi_traj"])
    err = np.max(np.abs(got - np.array(ref.psi_traj))) / max(1.0, np.max(np.abs(ref.psi_traj)))
    print("effective noise", eom, "ratio", ratio, "max |dpsi| = %.3e" % err)
    assert err < FP64_TOL


def test_adaptive_effective_noise_integration():
    """Adaptive trajectory with EFFECTIVE_NOISE_INTEGRATION: RK stages on the
    averaged noise, the error bound on point samples, as the reference."""
    _need_gpu()
    from mesohops_amd import HopsTrajectory
    sysp = systems.chain(6)
    tau_n = 0.25
    z = _oracle_noise(sysp, 4, 60.0, tau_n)
    nparam = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": 60.0, "TAU": tau_n}
    psi0 = np.zeros(6, complex)
    psi0[3] = 1.0
    ref = O.AdaptiveTrajectory(sysp, {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}, nparam,
                               hierarchy_param={"MAXHIER": 3}, delta_a=2e-3, delta_s=2e-3,
                               effective_noise_integration=True)
    ref.initialize(psi0)
    ref.propagate(30.0, 1.5)
    tr = HopsTrajectory(sysp, {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}, dict(nparam),
                        hierarchy_param={"MAXHIER": 3}, integration_param={"EFFECTIVE_NOISE_INTEGRATION": True})
    tr.make_adaptive(2e-3, 2e-3)
    tr.initialize(psi0)
    tr.propagate(30.0, 1.5)
    assert [list(s) for s in tr.storage["state_list"]] == [list(s) for s in ref.state_traj]
    err = np.max(np.abs(tr.storage["psi_traj"] - np.array(ref.psi_traj)))
    print("adaptive effective noise max |dpsi| = %.3e" % err)
    assert err < FP64_TOL


def test_noise_spline_matches_scipy():
    """hops_noise_spline against scipy.interpolate.interp1d(kind="cubic") (the
    reference's interpolation, noise/hops_noise.py:242-243) on random complex
    rows: 4, 5 and 6 points (the smallest not-a-knot systems) and long rows,
    evaluated on and between grid points from an offset."""
    _need_gpu()
    from scipy.interpolate import interp1d
    from mesohops_amd.engine import _stream_ptr
    from mesohops_amd.noise import spline_noise
    rng = np.random.default_rng(2)
    dev = torch.device("cuda:0")
    for n, h, t0, dt in ((4, 0.5, 0.0, 0.25), (5, 1.0, 0.3, 0.35), (6, 0.5, 0.0, 0.5), (301, 0.5, 1.25, 0.75),
                         (2049, 0.5, 0.0, 0.375)):
        y = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
        tmax = (n - 1) * h
        n_out = int(np.floor((tmax - t0) / dt + 1e-9)) + 1
        got = spline_noise(torch.from_numpy(y).to(dev), h, t0, dt, n_out, _stream_ptr(dev)).cpu().numpy()
        tq = t0 + np.arange(n_out) * dt
        want = interp1d(np.arange(n) * h, y, kind="cubic", axis=1)(tq)
        err = np.max(np.abs(got - want)) / np.max(np.abs(y))
        print("spline n=%d h=%g t0=%g dt=%g: max rel err %.2e" % (n, h, t0, dt, err))
        assert err < 1e-12


@pytest.mark.parametrize("eom", ["NORMALIZED NONLINEAR", "LINEAR"])
def test_fp64_interpolated_noise(eom):
    """INTERPOLATE = True with a half step between noise points (TAU 1, step
    1.5 fs: stage times t + 0.75 fs): the equations take the complex128 cubic
    interpolation of the noise, as the reference does
    (noise/hops_noise.py:242-243, 379-385); noise 1 and a FLAG_REAL noise 2
    share the oracle's complex128 correlated rows (PRE_CALCULATED)."""
    _need_gpu()
    from mesohops_amd import HopsTrajectory
    sysp = systems.three_site(False)
    sysp.update(L_NOISE2=sysp["L_NOISE1"], ALPHA_NOISE2=O.bcf_exp, PARAM_NOISE2=sysp["PARAM_NOISE1"])
    sysd = O.derive_system(sysp)
    tlen, tau = 60.0, 1.0
    zs = []
    for seed in (3, 4):
        nz = O.Noise({"SEED": seed, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau}, sysd, 1)
        nz.prepare()
        a = np.complex64(nz.alpha(list(range(nz.n_l2))))
        zs.append(O.correlated_noise(a, nz.z_unc))  # complex128, before the reference's cast
    n1 = {"SEED": zs[0], "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau, "INTERPOLATE": True}
    n2 = {"SEED": zs[1], "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau, "INTERPOLATE": True,
          "FLAG_REAL": True}
    ref, psi0 = oracle_run(sysp, eom, n1, 3, 45.0, 1.5, noise2_param=n2)
    tr = HopsTrajectory(sysp, {"EQUATION_OF_MOTION": eom}, dict(n1), dict(n2), hierarchy_param={"MAXHIER": 3})
    tr.initialize(psi0)
    tr.propagate(45.0, 1.5)
    got = np.array(tr.storage["psi_traj"])
    want = np.array(ref.psi_traj)
    err = np.max(np.abs(got - want)) / max(1.0, np.max(np.abs(want)))
    print("interpolated noise %s: max |dpsi| = %.3e over %d steps" % (eom, err, len(want) - 1))
    assert len(got) == len(want) == 31
    assert err < FP64_TOL


def test_device_noise_filter_complex128():
    """hops_noise_filter_c128: the filter's complex128 values (on which the
    reference builds its interpolation) against the oracle's, at the bar of
    the complex64 filter test; rounding them to complex64 gives the
    complex64 filter's output exactly."""
    _need_gpu()
    from mesohops_amd.basis import SystemInfo
    from mesohops_amd.engine import _stream_ptr
    from mesohops_amd.noise import HopsNoise, device_noise
    sysp = systems.dimer_of_dimers()
    info = SystemInfo(sysp)
    sysd = O.derive_system(sysp)
    p = {"SEED": 7, "MODEL": "FFT_FILTER", "TLEN": 1000.0, "TAU": 1.0}
    nz = HopsNoise(p, info, sysp, 1)
    dev = torch.device("cuda:0")
    rows = list(range(info.n_l2))
    got = device_noise([nz], rows, dev, _stream_ptr(dev), c128=True).cpu().numpy()[0]
    g64 = device_noise([nz], rows, dev, _stream_ptr(dev)).cpu().numpy()[0]
    on = O.Noise(p, sysd, 1)
    on.prepare()
    want = O.correlated_noise(np.complex64(on.alpha(rows)), on.z_unc)
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print("complex128 filter max rel err %.2e" % err)
    assert got.dtype == np.complex128 and err < 1e-6
    assert np.array_equal(np.complex64(got), g64)


@pytest.mark.parametrize("adaptive", [False, True])
def test_fp64_precalculated_complex128_noise(adaptive):
    """A PRE_CALCULATED complex128 array that complex64 cannot hold is used as
    given (noise/hops_noise.py:255-263), not rounded: psi(t) to 1e-10 against
    the oracle on the same array, plain and adaptive."""
    _need_gpu()
    from mesohops_amd import HopsTrajectory
    sysp = systems.chain(5)
    tlen, tau = 50.0, 0.5
    z = np.complex128(_oracle_noise(sysp, 6, tlen, tau))
    z = z * (1.0 + 1e-9 * np.arange(z.size).reshape(z.shape))  # below complex64 resolution
    noise = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    psi0 = np.zeros(5, complex)
    psi0[2] = 1
    if adaptive:
        ref = O.AdaptiveTrajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 3}, delta_a=1e-3, delta_s=1e-3)
    else:
        ref = O.Trajectory(sysp, eom, noise, hierarchy_param={"MAXHIER": 3})
    ref.initialize(psi0)
    ref.propagate(30.0, 1.0)
    tr = HopsTrajectory(sysp, eom, dict(noise), hierarchy_param={"MAXHIER": 3})
    if adaptive:
        tr.make_adaptive(1e-3, 1e-3)
    tr.initialize(psi0)
    tr.propagate(30.0, 1.0)
    err = np.max(np.abs(np.array(tr.storage["psi_traj"]) - np.array(ref.psi_traj)))
    print("complex128 PRE_CALCULATED adaptive=%s max |dpsi| = %.3e" % (adaptive, err))
    assert err < FP64_TOL


def test_two_host_threads_two_streams_bitwise_equal_serial():
    """Two host threads propagate two batches at once, each on its own stream
    (graph-replayed steps, LINEAR fused prologue and NORMALIZED NONLINEAR):
    each result equals the same batch run alone, bit for bit."""
    _need_gpu()
    import threading
    from mesohops_amd import HopsBatch
    sysp = systems.chain(5, k_mats=1)
    tlen, tau = 160.0, 0.5
    psi0 = np.zeros(5, complex)
    psi0[2] = 1

    def make(eom, seeds):
        zs = [_oracle_noise(sysp, s, tlen, tau) for s in seeds]
        nps = [{"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau} for z in zs]
        b = HopsBatch(sysp, {"EQUATION_OF_MOTION": eom}, nps[0], None, {"MAXHIER": 3}, noise_params=nps,
                      precision_mode="mixed", lanes=4)
        b.initialize(psi0)
        return b

    jobs = [("LINEAR", list(range(16))), ("NORMALIZED NONLINEAR", list(range(16, 40)))]
    serial = []
    for eom, seeds in jobs:
        b = make(eom, seeds)
        b.propagate(140.0, 2.0)   # 70 steps: 4 replays of a 16-step graph
        serial.append(np.array(b.psi_traj))
    batches = [make(eom, seeds) for eom, seeds in jobs]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    errs = []

    def work(i):
        try:
            with torch.cuda.stream(streams[i]):
                batches[i].propagate(140.0, 2.0)
            streams[i].synchronize()
        except Exception as e:  # reported below
            errs.append(e)

    th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    for i in range(2):
        assert np.array_equal(np.array(batches[i].psi_traj), serial[i]), jobs[i][0]
    print("two threads: both batches equal their serial runs bit for bit")
