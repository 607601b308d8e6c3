"""Checkpoint files in the reference's .npz layout, without a GPU: a file laid
out the way the reference's save_checkpoint writes it
(trajectory/hops_trajectory.py:884-936) is read back key by key, and a file
written here has the keys and value types the reference's load_checkpoint
(:939-1034) reads."""
import numpy as np
import pytest
import scipy.sparse as sp

from mesohops_amd.trajectory import read_checkpoint, write_checkpoint

KEYS = {"phi", "z_mem", "t", "state_list", "aux_list", "early_counter", "storage_data", "storage_meta", "params"}


def _params():
    return {"system_param": {"HAMILTONIAN": np.array([[0, 50], [50, 0]], dtype=np.complex128)},
            "hierarchy_param": {"MAXHIER": 3, "STATIC_FILTERS": []},
            "eom_param": {"EQUATION_OF_MOTION": "NORMALIZED NONLINEAR", "DELTA_A": 1e-3, "DELTA_S": 1e-3},
            "noise1_param": {"SEED": 7, "MODEL": "FFT_FILTER", "TLEN": 100.0, "TAU": 0.5},
            "noise2_param": {"SEED": None, "MODEL": "ZERO", "TLEN": 100.0, "TAU": 0.5},
            "integration_param": {"INTEGRATOR": "RUNGE_KUTTA"},
            "storage_param": {"psi_traj": True, "t_axis": True}}


def _reference_layout(path, aux_arrays, n_noise=6):
    """The file as the reference writes it: its aux list is
    np.array([aux.array_aux_vec ...], dtype=object) (which NumPy stacks into
    one [n_aux, k, 2] array when every auxiliary has the same length), its
    noise memory a (len(L_NOISE1), 1) sparse array."""
    zm = np.zeros((n_noise, 1), dtype=np.complex128)
    zm[1, 0] = 0.25 - 1j
    zm[4, 0] = -3.0 + 0.5j
    ck = {"phi": np.arange(2 * len(aux_arrays), dtype=np.complex128) * (1 + 0.5j),
          "z_mem": sp.coo_array(zm).tocsr(),
          "t": 12.5,
          "state_list": np.array([0, 1], dtype=int),
          "aux_list": np.array(aux_arrays, dtype=object),
          "early_counter": 3,
          "storage_data": {"psi_traj": [np.ones(2, complex)], "t_axis": [0.0]},
          "storage_meta": {"INITIALIZATION_TIME": 0.1},
          "params": _params()}
    np.savez_compressed(path, **ck)
    return ck


@pytest.mark.parametrize("case", ["ragged", "stacked_zero_only", "stacked_equal"])
def test_reads_the_reference_layout(tmp_path, case):
    if case == "ragged":
        aux = [np.zeros((0, 2), int), np.array([[0, 1]]), np.array([[1, 2]]), np.array([[0, 1], [3, 1]])]
        want = [(), ((0, 1),), ((1, 2),), ((0, 1), (3, 1))]
    elif case == "stacked_zero_only":    # a fresh adaptive basis: the zero auxiliary alone
        aux = [np.zeros((0, 2), int)]
        want = [()]
    else:                                # every auxiliary of one length: NumPy stacks them to 3-D
        aux = [np.array([[0, 1], [2, 1]]), np.array([[1, 1], [3, 2]])]
        want = [((0, 1), (2, 1)), ((1, 1), (3, 2))]
    path = tmp_path / "ref.npz"
    ref = _reference_layout(path, aux)
    if case != "ragged":
        assert np.load(path, allow_pickle=True)["aux_list"].ndim == 3
    ck = read_checkpoint(path)
    assert ck["aux_list"] == want
    assert ck["phi"].dtype == np.complex128 and np.array_equal(ck["phi"], ref["phi"])
    assert np.array_equal(ck["z_mem"], ref["z_mem"].toarray()[:, 0])
    assert ck["t"] == 12.5 and isinstance(ck["t"], float)
    assert ck["state_list"] == [0, 1]
    assert ck["early_counter"] == 3
    assert ck["storage_meta"] == ref["storage_meta"]
    assert ck["storage_data"]["t_axis"] == [0.0]
    assert ck["params"]["noise1_param"]["SEED"] == 7
    assert set(ck["params"]) == set(ref["params"])


def test_writes_what_the_reference_reads(tmp_path):
    """Every key, read the way the reference's load_checkpoint reads it."""
    zm = sp.csr_array(np.array([[0.0], [1 - 2j], [0.0]], dtype=np.complex128))
    aux = [(), ((0, 1),), ((0, 1), (2, 3))]
    phi = (np.arange(9) + 1j).astype(np.complex128)
    path = tmp_path / "ours.npz"
    write_checkpoint(path, phi, zm, 4.0, [1, 2, 5], aux, 2, {"psi_traj": [phi[:3]]}, {"X": 1}, _params())
    data = np.load(path, allow_pickle=True)
    assert set(data.files) == KEYS
    params = data["params"].item()
    assert set(params) == {"system_param", "hierarchy_param", "eom_param", "noise1_param", "noise2_param",
                           "integration_param", "storage_param"}
    st = data["state_list"]
    assert st.dtype.kind == "i" and list(st) == [1, 2, 5]
    assert np.array_equal(data["phi"][:st.size], phi[:3]) and data["phi"].dtype == np.complex128
    got = [tuple(map(tuple, np.asarray(a).tolist())) for a in data["aux_list"]]
    assert got == [(), ((0, 1),), ((0, 1), (2, 3))]
    assert all(np.asarray(a).shape[-1] == 2 and np.asarray(a).dtype.kind == "i" for a in data["aux_list"])
    z = data["z_mem"].item()
    assert sp.issparse(z) and z.shape == (3, 1) and np.array_equal(z.toarray(), zm.toarray())
    assert float(data["t"]) == 4.0 and int(data["early_counter"]) == 2
    assert data["storage_data"].item()["psi_traj"][0].shape == (3,)
    assert data["storage_meta"].item() == {"X": 1}
    # and back through the reader
    ck = read_checkpoint(path)
    assert ck["aux_list"] == aux and ck["state_list"] == [1, 2, 5] and np.array_equal(ck["phi"], phi)
