"""Dense system block on the matrix cores (k_dense_h): a fully coupled
aggregate (point-dipole couplings between every pair of sites, so H has no
zero off the diagonal) takes the dense path, whose (-iH) Y is a complex GEMM
on the 16x16x4 MFMA (f64 for complex128 stages, f32 for complex64).

- FP64 psi(t) against the oracle (the reference's (-1j*H) @ Phi,
  eom/hops_eom.py:389-391, 526) to 1e-10, for the three equations of motion;
- the dense path against the sparse pattern path (the same system with the
  dense threshold raised) to 1e-12 in FP64, and mixed precision against FP64
  within the population bar;
- a batch against its trajectories alone; the derivative entry point;
- every tile edge: state counts that are and are not multiples of 16, column
  counts that are not multiples of 16, one and eight trajectory lanes;
- the LDS-staged kernel (k_dense_lds) for these sizes, and the plain kernel
  (k_dense_h) for a basis too large for its LDS block.
"""
import ctypes

import numpy as np
import pytest

import systems
from oracle import hops_oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
POP_TOL = 1e-4


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _noise(sysp, seed, tlen, tau):
    sysd = O.derive_system(sysp)
    return O.Noise({"SEED": seed, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau}, sysd, 1).noise


def _run(sysp, eom, zs, tlen, tau, K, t_max, dt, psi0, prec="fp64", lanes=None, dense=True):
    from mesohops_amd import HopsBatch, basis
    nps = [{"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau} for z in zs]
    old = basis.DENSE_H_MIN_STATES
    if not dense:
        basis.DENSE_H_MIN_STATES = 1 << 30
    try:
        b = HopsBatch(sysp, {"EQUATION_OF_MOTION": eom}, nps[0], None, {"MAXHIER": K}, noise_params=nps,
                      precision_mode=prec, lanes=lanes)
        b.initialize(psi0)
    finally:
        basis.DENSE_H_MIN_STATES = old
    assert b.dop.tables.dense_h == dense
    b.propagate(t_max, dt)
    return b


def _psi0(n):
    psi0 = np.zeros(n, complex)
    psi0[n // 2] = 1.0
    return psi0


@pytest.mark.parametrize("eom", ["NORMALIZED NONLINEAR", "NONLINEAR", "LINEAR"])
def test_dense_fp64_matches_oracle(eom):
    _need_gpu()
    n = 72
    sysp = systems.aggregate(n)
    tlen, tau, t_max, dt = 60.0, 1.0, 20.0, 2.0
    z = _noise(sysp, 3, tlen, tau)
    print("case dense oracle", eom, flush=True)
    b = _run(sysp, eom, [z], tlen, tau, 1, t_max, dt, _psi0(n))
    tr = O.Trajectory(sysp, {"EQUATION_OF_MOTION": eom}, {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen,
                                                          "TAU": tau}, hierarchy_param={"MAXHIER": 1})
    tr.initialize(_psi0(n))
    tr.propagate(t_max, dt)
    err = np.max(np.abs(np.array(b.psi_traj[0]) - np.array(tr.psi_traj)))
    print("dense %s, %d sites depth 1: max |dpsi| vs oracle = %.3e" % (eom, n, err))
    assert err < FP64_TOL
    assert np.max(np.abs(b.phi[0] - tr.phi)) < FP64_TOL


@pytest.mark.parametrize("n,K,lanes", [(80, 1, 1), (67, 2, 8), (96, 1, 8)])
def test_dense_equals_sparse_and_mixed(n, K, lanes):
    _need_gpu()
    sysp = systems.aggregate(n, seed_h=n)
    tlen, tau, t_max, dt = 40.0, 1.0, 12.0, 2.0
    B = 8 if lanes == 8 else 2
    zs = [_noise(sysp, 100 + i, tlen, tau) for i in range(B)]
    eom = "NORMALIZED NONLINEAR"
    print("case dense vs sparse", n, K, lanes, flush=True)
    d = _run(sysp, eom, zs, tlen, tau, K, t_max, dt, _psi0(n), lanes=lanes)
    s = _run(sysp, eom, zs, tlen, tau, K, t_max, dt, _psi0(n), lanes=lanes, dense=False)
    m = _run(sysp, eom, zs, tlen, tau, K, t_max, dt, _psi0(n), prec="mixed", lanes=lanes)
    for i in range(B):
        a, c = np.array(d.psi_traj[i]), np.array(s.psi_traj[i])
        err = np.max(np.abs(a - c))
        pop = np.max(np.abs(np.abs(np.array(m.psi_traj[i])) ** 2 - np.abs(a) ** 2))
        print("n=%d K=%d lanes=%d traj %d: dense vs sparse %.3e, mixed population error %.3e" % (n, K, lanes, i,
                                                                                              err, pop))
        assert err < 1e-12
        assert pop < POP_TOL
    # a trajectory alone gives its result in the batch
    alone = _run(sysp, eom, [zs[1]], tlen, tau, K, t_max, dt, _psi0(n))
    assert np.max(np.abs(np.array(alone.psi_traj[0]) - np.array(d.psi_traj[1]))) < 1e-13


def test_dense_graph_replay_equals_eager():
    """Enough steps that the propagation replays captured graphs (the dense
    kernel inside them) against the same steps launched eagerly."""
    _need_gpu()
    from mesohops_amd import _lib
    n = 70
    sysp = systems.aggregate(n, seed_h=5)
    tlen, tau, t_max, dt = 400.0, 1.0, 160.0, 2.0
    z = _noise(sysp, 9, tlen, tau)
    eom = "NORMALIZED NONLINEAR"
    g = _run(sysp, eom, [z], tlen, tau, 1, t_max, dt, _psi0(n))
    _lib.set_option("graph_steps", 0)
    try:
        e = _run(sysp, eom, [z], tlen, tau, 1, t_max, dt, _psi0(n))
    finally:
        _lib.set_option("graph_steps", 16)
    assert np.array_equal(np.array(g.psi_traj[0]), np.array(e.psi_traj[0]))


def test_dense_derivative_equals_sparse():
    """hops_derivative with the dense block against the pattern form on a
    random hierarchy vector."""
    _need_gpu()
    from mesohops_amd import _lib
    from mesohops_amd.basis import BasisTables, SystemInfo, triangular_aux_list
    from mesohops_amd.engine import BatchState, DeviceOperator, _stream_ptr
    n = 75
    sysp = systems.aggregate(n, seed_h=2)
    info = SystemInfo(sysp)
    aux = triangular_aux_list(info.n_hmodes, 1)
    rng = np.random.default_rng(0)
    nr = n * len(aux)
    y = rng.normal(size=nr) + 1j * rng.normal(size=nr)  # the same vector for both forms
    outs = []
    for dense in (True, False):
        tb = BasisTables(info, np.arange(n), aux, 2, dense_h=dense)
        dop = DeviceOperator(tb, _lib.HOPS_PREC_FP64)
        z = torch.zeros((1, tb.n_act, 64), dtype=torch.complex64, device="cuda")
        st = BatchState(dop, 1, 2, z, None, 64, lanes=1)
        yd = torch.from_numpy(y).cuda()
        dd = torch.zeros_like(yd)
        _lib.check(_lib.lib().hops_derivative(ctypes.byref(dop.op), ctypes.byref(st.bt),
                                              ctypes.c_void_p(yd.data_ptr()), ctypes.c_void_p(dd.data_ptr()), 3,
                                              _stream_ptr(dop.device)))
        torch.cuda.synchronize()
        outs.append(dd.cpu().numpy())
    err = np.max(np.abs(outs[0] - outs[1])) / np.max(np.abs(outs[1]))
    print("derivative dense vs sparse: relative %.3e" % err)
    assert err < 1e-13


def test_dense_plain_form_for_large_blocks_equals_sparse():
    """A state basis whose staged Y block would not fit in 64 KiB of LDS (250
    states in FP64: 16 columns x 258 rows x 16 bytes) takes the plain dense
    kernel (k_dense_h): the same psi(t) as the sparse pattern path to 1e-12,
    over two trajectories in one group of two lanes."""
    _need_gpu()
    n = 250
    sysp = systems.aggregate(n, seed_h=11)
    tlen, tau, t_max, dt = 30.0, 1.0, 8.0, 2.0
    zs = [_noise(sysp, 200 + i, tlen, tau) for i in range(2)]
    eom = "NORMALIZED NONLINEAR"
    print("case dense plain form, %d states" % n, flush=True)
    d = _run(sysp, eom, zs, tlen, tau, 1, t_max, dt, _psi0(n), lanes=2)
    assert d.dop.tables.dense_h
    s = _run(sysp, eom, zs, tlen, tau, 1, t_max, dt, _psi0(n), lanes=2, dense=False)
    for i in range(2):
        err = np.max(np.abs(np.array(d.psi_traj[i]) - np.array(s.psi_traj[i])))
        print("plain dense form, traj %d: dense vs sparse %.3e" % (i, err))
        assert err < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_dense_stationary_forms_equal_streaming(prec):
    """With eight trajectory groups of four lanes the stage runs in the
    operator-stationary forms, which add the dense block's (-iH) Y in
    registers: plain and LDS-staged, one, two and four lanes per thread (four
    is served with two, the grid following), hint levels 0, 2 and 3 — every
    form bit for bit the streaming kernel's result."""
    _need_gpu()
    from mesohops_amd import _lib
    n = 70
    sysp = systems.aggregate(n, seed_h=8)
    tlen, tau, t_max, dt = 40.0, 1.0, 10.0, 2.0
    B = 32
    zs = [_noise(sysp, 300 + i, tlen, tau) for i in range(B)]
    eom = "NORMALIZED NONLINEAR"
    opts = {"stationary_stage": 1, "lds_stage": 0, "vec_lanes": 1, "nt_stream": 0}
    try:
        _lib.set_option("stationary_stage", 0)
        ref = _run(sysp, eom, zs, tlen, tau, 1, t_max, dt, _psi0(n), prec=prec, lanes=4)
        _lib.set_option("stationary_stage", 1)
        for lds, vec, nt in ((0, 1, 0), (0, 2, 2), (1, 1, 0), (1, 2, 2), (0, 4, 2), (1, 4, 3)):
            _lib.set_option("lds_stage", lds)
            _lib.set_option("vec_lanes", vec)
            _lib.set_option("nt_stream", nt)
            print("case dense stationary %s lds=%d vec=%d nt=%d" % (prec, lds, vec, nt), flush=True)
            got = _run(sysp, eom, zs, tlen, tau, 1, t_max, dt, _psi0(n), prec=prec, lanes=4)
            for i in range(B):
                assert np.array_equal(np.array(got.psi_traj[i]), np.array(ref.psi_traj[i])), (lds, vec, nt, i)
    finally:
        for k, v in opts.items():
            _lib.set_option(k, v)
