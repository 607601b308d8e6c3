"""Dyadic HOPS on the device against the FP64 oracle (1e-10).

The noise is the same array on both sides: PRE_CALCULATED oracle noise for
DyadicTrajectory, and for DyadicSpectra (which filters its own FFT noise from
the seed) the engine's filtered noise, read back and given to the oracle.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import hops_oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-10


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _chromophores(n_chrom=2, ltc=False):
    """Chromophores coupled in a chain with one Drude-Lorentz mode each on
    their site projector (ground state 0), optionally a low-temperature
    correction."""
    n = n_chrom + 1
    H = np.zeros((n, n), complex)
    for i in range(n_chrom):
        H[i + 1, i + 1] = [40.0, -60.0, 10.0][i % 3]
        if i + 2 < n:
            H[i + 1, i + 2] = H[i + 2, i + 1] = 50.0
    L = [sp.coo_matrix(([1.0], ([i + 1], [i + 1])), shape=(n, n)) for i in range(n_chrom)]
    g, w = O.bcf_convert_dl_to_exp(35.0, 50.0, 295.0)
    gw = [[g, w]] * n_chrom
    sysp = {"HAMILTONIAN": H, "GW_SYSBATH": gw, "L_HIER": L, "L_NOISE1": L, "ALPHA_NOISE1": O.bcf_exp,
            "PARAM_NOISE1": gw}
    if ltc:
        sysp["L_LT_CORR"] = L
        sysp["PARAM_LT_CORR"] = [25.0 - 4.0j] * n_chrom
    return sysp


def _copy_system(sysp):
    return {k: (list(v) if isinstance(v, list) else v) for k, v in sysp.items()}


@pytest.mark.parametrize("adaptive", [False, True])
def test_dyadic_trajectory_matches_oracle(adaptive):
    """initialize, ket raise (sparse), bra raise (dense), propagate, bra lower,
    propagate: psi(t), the response normalisation factors and the response
    function equal the oracle's."""
    _need_gpu()
    from mesohops_amd.dyadic import DyadicTrajectory
    n_chrom, n = 2, 3
    base = _chromophores(n_chrom, ltc=not adaptive)
    tlen, tau = 60.0, 0.5
    dsys = O.dyadic_system(base)
    z = O.Noise({"SEED": 4, "MODEL": "FFT_FILTER", "TLEN": tlen, "TAU": tau}, O.derive_system(dsys), 1).noise
    noise = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": tlen, "TAU": tau}
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    if adaptive:
        otr = O.AdaptiveTrajectory(dsys, eom, noise, hierarchy_param={"MAXHIER": 3}, delta_a=1e-3, delta_s=1e-3)
    else:
        otr = O.Trajectory(dsys, eom, noise, hierarchy_param={"MAXHIER": 3})
    ref = O.Dyadic(otr)
    dev = DyadicTrajectory(_copy_system(base), eom, dict(noise), hierarchy_param={"MAXHIER": 3})
    if adaptive:
        dev.make_adaptive(1e-3, 1e-3)
    g = np.zeros(n)
    g[0] = 1.0
    up = np.zeros((n, n))
    up[1:, 0] = [0.6, -0.3]
    up2 = np.zeros((n, n))
    up2[1:, 0] = [0.2, 0.5]
    down = np.zeros((n, n))
    down[0, 1:] = [0.4, 0.7]
    for d in (ref, dev):
        d.initialize(g, g.copy())
    ref.side_operator(sp.coo_matrix(up), "ket")
    dev._dyad_operator(sp.coo_matrix(up), "ket")
    ref.side_operator(up2, "bra")
    dev._dyad_operator(up2, "bra")
    ref.tr.propagate(8.0, 1.0)
    dev.propagate(8.0, 1.0)
    ref.side_operator(down, "bra")
    dev._dyad_operator(down, "bra")
    ref.tr.propagate(8.0, 1.0)
    dev.propagate(8.0, 1.0)
    assert np.allclose(dev.list_response_norm_sq, ref.norm_sq, rtol=1e-12, atol=0)
    got = dev.storage["psi_traj_sparse"].toarray() if adaptive else np.array(dev.storage["psi_traj"])
    want = ref.full_traj()
    assert got.shape == want.shape
    err = np.max(np.abs(got - want))
    print("dyadic adaptive=%s max |dpsi| = %.3e" % (adaptive, err))
    assert err < TOL
    F = O.signal_operator(np.array([[0.5, 0.2, 0.1], [0.1, -0.4, 0.3]]), np.array([0.0, 1.0, 0.0]), n_chrom)
    r_dev = dev._response_function_comp(sp.coo_matrix(F), 8)
    r_ref = ref.response(F, 8)
    assert np.max(np.abs(r_dev - r_ref)) < TOL * max(1.0, np.max(np.abs(r_ref)))


@pytest.mark.parametrize("kind,adaptive", [("ABSORPTION", False), ("FLUORESCENCE", False), ("FLUORESCENCE", True)])
def test_dyadic_spectra_matches_oracle(kind, adaptive):
    """DyadicSpectra's response function equals the oracle's pathway on the
    engine's own filtered noise."""
    _need_gpu()
    from mesohops_amd import dyadic as D
    from mesohops_amd.engine import _stream_ptr
    from mesohops_amd.noise import device_noise
    mu = np.array([[0.5, 0.2, 0.1], [0.1, -0.4, 0.3]])
    H = np.zeros((3, 3), complex)
    H[1, 1], H[2, 2] = 40.0, -60.0
    H[1, 2] = H[2, 1] = 50.0
    modes = list(O.bcf_convert_dl_to_exp(35.0, 50.0, 295.0))
    E = np.array([0.0, 0.0, 1.0])
    if kind == "ABSORPTION":
        sd = D.prepare_spectroscopy_input_dict(kind, {"t_1": 12.0}, {"E_1": E}, {"list_ket_sites": [1, 2]})
    else:
        sd = D.prepare_spectroscopy_input_dict(kind, {"t_2": 6.0, "t_3": 8.0}, {"E_1": E, "E_sig": E},
                                               {"list_ket_sites": [1, 2], "list_bra_sites": [1, 2]})
    cd = D.prepare_chromophore_input_dict(mu, H, {"list_modes": modes})
    delta = 1e-3 if adaptive else 0
    vd = D.prepare_convergence_parameter_dict(1.0, 3, delta_a=delta, delta_s=delta)
    spec = D.DyadicSpectra(sd, cd, vd, seed=7)
    R = spec.calculate_spectrum()

    dev = torch.device("cuda:0")
    z = device_noise([spec.noise1], list(range(spec.noise1.n_l2)), dev, _stream_ptr(dev))[0].cpu().numpy()
    sysp = {"HAMILTONIAN": H, "GW_SYSBATH": cd["gw_sysbath_hier"], "L_HIER": cd["lop_list_hier"],
            "L_NOISE1": cd["lop_list_noise"], "ALPHA_NOISE1": O.bcf_exp, "PARAM_NOISE1": cd["gw_sysbath_noise"],
            "L_LT_CORR": cd["lop_list_ltc"], "PARAM_LT_CORR": cd["ltc_param"]}
    noise = {"SEED": z, "MODEL": "PRE_CALCULATED", "TLEN": spec.noise_param["TLEN"], "TAU": spec.noise_param["TAU"]}
    eom = {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR"}
    dsys = O.dyadic_system(sysp)
    if adaptive:
        tr = O.AdaptiveTrajectory(dsys, eom, noise, hierarchy_param={"MAXHIER": 3}, delta_a=delta, delta_s=delta)
    else:
        tr = O.Trajectory(dsys, eom, noise, hierarchy_param={"MAXHIER": 3})
    fields = {"E_1": sd["E_1"], "E_2": sd.get("E_2"), "E_3": sd.get("E_3"), "E_sig": sd["E_sig"]}
    want = O.dyadic_spectrum(O.Dyadic(tr), kind, mu, fields, sd["list_ket_sites"], sd.get("list_bra_sites"),
                             sd["t_1"], sd["t_2"], sd["t_3"], 1.0)
    assert R.shape == want.shape
    err = np.max(np.abs(R - want)) / np.max(np.abs(want))
    print(kind, "adaptive", adaptive, "relative |dR| =", err)
    assert err < TOL
    assert spec.initialized
