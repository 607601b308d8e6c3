"""Host-side logic and the C ABI, without a GPU: the library loads and exports
every entry point of include/hops_mi355x.h, the host tables reproduce the
oracle's super-operator, and the parameter validation raises the reference's
exceptions."""
import ctypes
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

import systems
from oracle import hops_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    out = set()
    for h in ("hops_mi355x.h", "hops_runtime.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        out |= set(re.findall(r"\b(hops_[a-z_0-9]+)\s*\(", txt))
    return sorted(out)


def test_library_exports_every_declared_symbol():
    from mesohops_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = _declared_symbols()
    assert len(syms) >= 8
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) <= set(_lib.EXPORTED)
    assert _lib.lib().hops_version() == _lib.ABI_VERSION


def test_engine_options_validate_without_a_device():
    """hops_set_option takes every documented option within its range and
    rejects unknown names and out-of-range values with a message (it sets
    library state only, so it runs without a GPU)."""
    from mesohops_amd import _lib
    from mesohops_amd.errors import EngineError
    L = _lib.lib()
    good = {"stationary_stage": 1, "chunk_traj": 0, "nt_stream": 2, "groups_q": 0, "vec_lanes": 2,
            "lds_stage": 0, "fuse_prologue": 1, "graph_steps": 16, "resident_rows": 1024}
    for name, v in good.items():
        assert L.hops_set_option(name.encode(), v) == 0, name
    for name, v in (("nt_stream", 5), ("vec_lanes", 3), ("graph_steps", -1), ("graph_steps", 5000),
                    ("resident_rows", -2), ("chunk_traj", -1), ("no_such_option", 1)):
        assert L.hops_set_option(name.encode(), v) != 0, name
        assert L.hops_last_error().decode(), name
    with pytest.raises(EngineError):
        _lib.set_option("graph_steps", -1)
    # the library defaults again
    for name, v in (("vec_lanes", 1), ("nt_stream", 0), ("graph_steps", 16), ("resident_rows", 1024)):
        assert L.hops_set_option(name.encode(), v) == 0


def test_struct_layout_matches_header():
    """ctypes mirrors of hops_op_t / hops_batch_t have the header's field order."""
    from mesohops_amd._lib import HopsBatch, HopsOp
    txt = open(os.path.join(ROOT, "include", "hops_mi355x.h")).read()
    body = txt[txt.index("typedef struct hops_op {"):txt.index("} hops_op_t;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\*?\s*(\w+)\s*[,;]", body.split("{", 1)[1])
    assert [f[0] for f in HopsOp._fields_] == names
    body = txt[txt.index("typedef struct hops_batch {"):txt.index("} hops_batch_t;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\*?\s*(\w+)\s*[,;]", body.split("{", 1)[1])
    assert [f[0] for f in HopsBatch._fields_] == names


def _host_k2(tb):
    t = tb.tables
    n, nh = tb.n_state, tb.n_hier
    lr_val = t["lr_val"].view(np.complex128)
    lc_val = t["lc_val"].view(np.complex128)
    g = t["mode_g"].view(np.complex128)
    w = t["mode_w"].view(np.complex128)
    rows, cols, vals = [], [], []
    for k in range(nh):
        for s in range(n):
            r = k * n + s
            for u in range(t["up_ptr"][k], t["up_ptr"][k + 1]):
                m, k2 = t["up_mode"][u], t["up_aux"][u]
                j = t["mode_lrel"][m]
                for q in range(t["lr_ptr"][j * (n + 1) + s], t["lr_ptr"][j * (n + 1) + s + 1]):
                    rows.append(r)
                    cols.append(k2 * n + t["lr_col"][q])
                    vals.append(-(g[m] / w[m]) * lr_val[q])
            for u in range(t["dn_ptr"][k], t["dn_ptr"][k + 1]):
                m, k2, kv = t["dn_mode"][u], t["dn_aux"][u], t["dn_val"][u]
                j = t["mode_lrel"][m]
                for q in range(t["lc_ptr"][j * (n + 1) + s], t["lc_ptr"][j * (n + 1) + s + 1]):
                    rows.append(r)
                    cols.append(k2 * n + t["lc_row"][q])
                    vals.append((kv * w[m]) * np.conj(lc_val[q]))
    return sp.coo_matrix((vals, (rows, cols)), shape=(n * nh, n * nh)).toarray()


@pytest.mark.parametrize("name,K", [("dod", 2), ("three_general", 3), ("three_diag", 3), ("chain_ltc", 3)])
def test_host_tables_reproduce_oracle_super_operator(name, K):
    from mesohops_amd.basis import BasisTables, SystemInfo, triangular_aux_list
    sysp = {"dod": systems.dimer_of_dimers(), "three_general": systems.three_site(True),
            "three_diag": systems.three_site(False), "chain_ltc": systems.chain(4, ltc=3.0)}[name]
    info = SystemInfo(sysp)
    tb = BasisTables(info, np.arange(info.nstates), triangular_aux_list(info.n_hmodes, K), 2)
    sysd = O.derive_system(sysp)
    b = O.Basis(sysd, np.arange(sysd["NSTATES"]), O.triangular_hierarchy(sysd["N_HMODES"], K))
    assert [tuple(a) for a in tb.aux_list] == b.aux_list
    assert np.max(np.abs(_host_k2(tb) - b.K2_stable.toarray())) <= 1e-15 * np.max(np.abs(b.K2_stable.toarray()))
    assert tb.first == b.first_level
    # ELL width is the densest row of K2_kp1 + K2_km1
    assert tb.ell_w == int(np.max(np.diff(sp.csr_matrix(b.K2_stable).indptr)))


def test_hierarchy_size_closed_form():
    """A triangular hierarchy of M modes and depth K has C(M+K, K) members."""
    from math import comb
    from mesohops_amd.basis import triangular_aux_list
    for M, K in [(2, 4), (6, 3), (20, 2)]:
        assert len(triangular_aux_list(M, K)) == comb(M + K, K)


def test_static_filters():
    """Triangular / LongEdge / Markovian filters (basis/hierarchy_functions.py)."""
    from mesohops_amd.basis import apply_static_filters, triangular_aux_list
    aux = triangular_aux_list(3, 3)
    mark = apply_static_filters(aux, [("Markovian", [False, True, False])], 3)
    assert all(sum(v for _, v in a) <= 1 or all(m != 1 for m, _ in a) for a in mark)
    tri = apply_static_filters(aux, [("Triangular", ([True, True, False], 1))], 3)
    assert all(sum(v for m, v in a if m < 2) <= 1 for a in tri)
    le = apply_static_filters(aux, [("LongEdge", ([True, False, False], 1))], 3)
    for a in le:
        d = sum(v for _, v in a)
        assert d <= 1 or all(m != 0 for m, _ in a) or len(a) <= 1
    from mesohops_amd import UnsupportedRequest
    with pytest.raises(UnsupportedRequest):
        apply_static_filters(aux, [("Domain", [])], 3)


def test_tile_map_is_xcd_aware():
    from mesohops_amd.engine import N_XCD, tile_map
    t = tile_map(16, 256 * 5)
    live = t[t >= 0]
    assert sorted(live.tolist()) == sorted(b * 65536 + k for b in range(16) for k in range(5))
    for i, v in enumerate(t):
        if v >= 0:
            assert (v >> 16) % N_XCD == i % N_XCD
    small = tile_map(2, 300)
    assert small.tolist() == [0, 1, 65536, 65537]
    # 4 interleaved lanes: a tile is 64 rows of one group
    t4 = tile_map(8, 130, 4)
    assert sorted(t4[t4 >= 0].tolist()) == sorted(g * 65536 + k for g in range(8) for k in range(3))


def test_pool_interleave_roundtrip():
    """[B, n_rows] <-> interleaved pool order is a bijection with element
    (b, r) at ((b >> s) n_rows + r) T + (b & (T-1))."""
    from mesohops_amd.engine import BatchState
    st = BatchState.__new__(BatchState)

    class D:
        n_rows = 7
    st.dop, st.B, st.lanes = D(), 10, 4
    st.G = 3
    a = (np.arange(70) + 1j * np.arange(70)).reshape(10, 7)
    flat = st._to_pool(a)
    for b in range(10):
        for r in range(7):
            assert flat[((b >> 2) * 7 + r) * 4 + (b & 3)] == a[b, r]
    assert np.array_equal(st._from_pool(flat), a)


def test_parameter_validation():
    from mesohops_amd import TrajectoryError, UnsupportedRequest
    from mesohops_amd.params import (EOM_DICT_DEFAULT, EOM_DICT_TYPES, NOISE_DICT_DEFAULT,
                                     NOISE_TYPE_DEFAULT, initialize_dict)
    with pytest.raises(UnsupportedRequest):
        initialize_dict({"BOGUS": 1}, EOM_DICT_DEFAULT, EOM_DICT_TYPES, "HopsEOM")
    with pytest.raises(TypeError):
        initialize_dict({"TLEN": 10}, NOISE_DICT_DEFAULT, NOISE_TYPE_DEFAULT, "HopsNoise")
    p = initialize_dict({"TLEN": 10.0}, NOISE_DICT_DEFAULT, NOISE_TYPE_DEFAULT, "HopsNoise")
    assert p["TAU"] == 1.0 and p["MODEL"] == "FFT_FILTER"
    from mesohops_amd.trajectory import _Setup
    with pytest.raises(UnsupportedRequest):
        _Setup(systems.dimer_c1(), {"EQUATION_OF_MOTION": "BOGUS"}, {"TLEN": 10.0, "TAU": 1.0}, None, None, None)
    with pytest.raises(UnsupportedRequest):
        _Setup(systems.dimer_c1(), None, {"TLEN": 10.0, "TAU": 1.0}, None, None, {"INTEGRATOR": "EULER"})


def test_tau_check():
    """(tau/2)/TAU must be an integer and tau >= TAU (hops_trajectory.py:579-614)."""
    from mesohops_amd.basis import SystemInfo
    from mesohops_amd.noise import HopsNoise
    from mesohops_amd.trajectory import _check_tau
    sysp = systems.dimer_c1()
    info = SystemInfo(sysp)
    n1 = HopsNoise({"TLEN": 100.0, "TAU": 1.0}, info, sysp, 1)
    n2 = HopsNoise({"TLEN": 100.0, "TAU": 1.0, "MODEL": "ZERO"}, info, sysp, 2)
    assert _check_tau(2.0, n1, n2)
    assert not _check_tau(0.5, n1, n2)
    # the reference compares a signed difference, so a half step of 1.5 TAU
    # passes this check and fails later at the noise lookup
    assert _check_tau(3.0, n1, n2)


def test_host_raw_noise_matches_golden():
    """The product's host raw draws equal the reference's seed-0 SUM_GAUSSIAN
    fixture (tests/test_noise_fft.py:86-120)."""
    from mesohops_amd.basis import SystemInfo
    from mesohops_amd.noise import HopsNoise
    fx = np.load(os.path.join(ROOT, "tests", "golden", "ref_fixtures.npz"))
    L = [np.diag([1.0, 0.0]), np.diag([0.0, 1.0])]
    gw = [[10.0, 10.0], [5.0, 5.0]]
    sysp = dict(HAMILTONIAN=np.array([[0, 10.0], [10.0, 0]]), GW_SYSBATH=gw, L_HIER=L, L_NOISE1=L,
                ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw)
    nz = HopsNoise({"SEED": 0, "TLEN": 10.0, "TAU": 1.0}, SystemInfo(sysp), sysp, 1)
    np.testing.assert_allclose(nz.raw(), fx["pre_calculated_uncorrelated_noise_test"])
    np.testing.assert_array_equal(nz.raw(), O.uncorrelated_noise(0, [0, 1], 11, "SUM_GAUSSIAN"))


def test_ragged_group_steps_every_batch_and_raises_after_all_finish():
    """RaggedGroup's host logic, without a device: listed (batch, slot) pairs
    go to their batch, batches without listed slots are not stepped, every
    batch runs on its own host thread, and a batch's error is raised only
    after the other batches have finished their steps."""
    import threading
    import time
    from mesohops_amd import ragged
    from mesohops_amd.errors import ArenaFull

    class Part:
        def __init__(self, k, fail=False):
            self.k, self.fail, self.calls, self.done, self.thread = k, fail, [], False, None

        def step(self, n, slots=None):
            self.thread = threading.current_thread().name
            self.calls.append((n, slots))
            if self.fail:
                raise ArenaFull("batch %d full" % self.k)
            time.sleep(0.05)
            self.done = True

    g = ragged.RaggedGroup.__new__(ragged.RaggedGroup)
    g.parts = [Part(0), Part(1, fail=True), Part(2)]
    g.device = "cpu"
    g._lock = threading.Lock()
    g._profiling = False
    g._prof = ([0.0] * 4, [0] * 4)
    orig = ragged.torch.cuda.synchronize
    ragged.torch.cuda.synchronize = lambda *a, **k: None
    try:
        with pytest.raises(ArenaFull):
            g.step(1, slots=[(0, 3), (1, 0), (0, 5)])
        assert g.parts[0].calls == [(1, [3, 5])] and g.parts[0].done
        assert g.parts[1].calls == [(1, [0])]
        assert g.parts[2].calls == []  # no listed slot: not stepped
        g.parts[1].fail = False
        g.step(2)
        assert [p.calls[-1] for p in g.parts] == [(2, None)] * 3
        assert [p.thread for p in g.parts] == ["hops-batch-%d" % k for k in range(3)]
    finally:
        ragged.torch.cuda.synchronize = orig


def _fake_serving(K, max_slots, capacity):
    """A ServingLoop whose runtime batches are host fakes: a slot holds
    ``est`` bytes at admission and grows by ``grow`` bytes a step; a step
    that cannot place a slot's growth marks it out of room and raises
    ArenaFull after stepping the others, as the runtime does."""
    import threading
    import types
    from concurrent.futures import ThreadPoolExecutor
    from mesohops_amd import serving
    from mesohops_amd.errors import ArenaFull

    class Part:
        def __init__(self):
            self.slots, self.next, self.peak, self.lock = {}, 0, 0, threading.Lock()
            self.threads = set()

        def free_bytes(self):
            return capacity - sum(s["bytes"] for s in self.slots.values())

        def estimate(self, est, n_state, n_hier):
            return est

        def add(self, r):
            self.next += 1
            self.slots[self.next] = {"bytes": r.est, "grow": r.grow, "oom": False}
            self.peak = max(self.peak, len(self.slots))
            return self.next

        def step(self, n, slots=None):
            self.threads.add(threading.current_thread().name)
            oom = False
            for s in slots:
                sl = self.slots[s]
                if self.free_bytes() < sl["grow"]:
                    sl["oom"] = oom = True
                else:
                    sl["bytes"] += sl["grow"]
            if oom:
                raise ArenaFull("full")

        def info(self, s):
            sl = self.slots[s]
            return {"oom": sl["oom"], "bytes": sl["bytes"], "n_state": 1, "n_hier": 1}

        def remove(self, s):
            del self.slots[s]

    sl = serving.ServingLoop.__new__(serving.ServingLoop)
    sl.K, sl.n_threads, sl.max_slots, sl.record, sl.reserve = K, 1, max_slots, False, (12, 4000)
    sl.slot_cap = serving.ServingLoop.split_slots(max_slots, K)
    sl.device = "cpu"
    sl.rt = types.SimpleNamespace(parts=[Part() for _ in range(K)])
    sl._prep = ThreadPoolExecutor(max_workers=2)
    sl._prepare = lambda r: (None, None, r.est, None, None)
    sl._admit = lambda r, prep, k: ((k, sl.rt.parts[k].add(r)), None)
    return sl


def _requests(n, est=10, grow=0, steps=5, gap=0.0):
    import types
    return [types.SimpleNamespace(rid=i, arrival=i * gap, n_steps=steps, nsite=4, maxhier=2, delta=1e-3, t_max=1.0,
                                  est=est, grow=grow) for i in range(n)]


def test_serving_batches_run_independently_and_finish_every_request(monkeypatch):
    """ServingLoop.run's host logic, without a device: every request is
    admitted, stepped exactly n_steps times and retired; no batch holds more
    than its share of slots; each batch is stepped only by its own thread;
    a request that outgrows the arena with others is preempted and served
    again from the start."""
    import contextlib
    from mesohops_amd import serving
    monkeypatch.setattr(serving.torch.cuda, "device", lambda d: contextlib.nullcontext())
    assert serving.ServingLoop.split_slots(7, 3) == [3, 2, 2]
    assert serving.ServingLoop.split_slots(512, 16) == [32] * 16
    sl = _fake_serving(K=3, max_slots=7, capacity=100)
    reqs = _requests(20, est=10, steps=4, gap=0.001)
    done, st = sl.run(reqs)
    assert sorted(done) == list(range(20)) and all(d["steps"] == 4 for d in done.values())
    assert st["completed"] == 20 and st["traj_steps"] == 80 and st["preempted"] == 0
    assert [p.peak <= c for p, c in zip(sl.rt.parts, [3, 2, 2])] == [True] * 3
    assert not any(p.slots for p in sl.rt.parts)
    assert all(p.threads <= {"hops-serve-%d" % k} for k, p in enumerate(sl.rt.parts))
    sl._prep.shutdown()
    # bases that grow: two slots of a batch cannot both reach 68 bytes
    sl = _fake_serving(K=2, max_slots=4, capacity=100)
    done, st = sl.run(_requests(6, est=20, grow=8, steps=6))
    assert sorted(done) == list(range(6)) and all(d["steps"] == 6 for d in done.values())
    assert st["preempted"] > 0 and st["traj_steps"] >= 36
    sl._prep.shutdown()
    # stopped by the time limit: unfinished requests leave their batches
    sl = _fake_serving(K=2, max_slots=4, capacity=100)
    done, st = sl.run(_requests(4, est=10, steps=10 ** 9), time_limit=0.2)
    assert not done and st["admitted_unfinished"] == 4
    assert not any(p.slots for p in sl.rt.parts)
    sl._prep.shutdown()


def test_serving_refuses_requests_no_batch_can_hold(monkeypatch):
    import contextlib
    from mesohops_amd import serving
    from mesohops_amd.errors import UnsupportedRequest
    monkeypatch.setattr(serving.torch.cuda, "device", lambda d: contextlib.nullcontext())
    sl = _fake_serving(K=2, max_slots=4, capacity=100)
    with pytest.raises(UnsupportedRequest, match="needs more device memory"):
        sl.run(_requests(3, est=60))  # twice the footprint exceeds an empty batch
    sl._prep.shutdown()
    # alone in its batch and still outgrowing it: one restart, then refused
    sl = _fake_serving(K=1, max_slots=1, capacity=100)
    with pytest.raises(UnsupportedRequest, match="while running alone"):
        sl.run(_requests(1, est=20, grow=30, steps=6))
    sl._prep.shutdown()
