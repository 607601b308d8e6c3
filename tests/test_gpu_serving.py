"""Serving loop (continuous batching) on the device: every request's result is
the same alone and in the mix — bitwise, since a slot's kernels and basis
bookkeeping touch only that slot — while requests of different sizes arrive,
join and leave the batch at different times."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _requests():
    from mesohops_amd.workloads import Request
    return [Request(rid=11, arrival=0.0, nsite=6, maxhier=3, delta=2e-3, t_max=40.0),
            Request(rid=12, arrival=0.02, nsite=9, maxhier=4, delta=5e-3, t_max=25.0),
            Request(rid=13, arrival=0.04, nsite=4, maxhier=2, delta=1e-3, t_max=60.0),
            Request(rid=14, arrival=0.05, nsite=12, maxhier=3, delta=8e-3, t_max=15.0)]


@pytest.mark.parametrize("prec,sub", [("fp64", 1), ("mixed", 1), ("mixed", 3), ("mixed", 16)])
def test_alone_equals_in_mix(prec, sub):
    """sub: runtime batches, each served by its own host thread (3: the four
    requests spread over three batches, each with its own stream and worker
    threads; 16: clamped to the 9 slots, one slot and one worker thread per
    batch)."""
    _need_gpu()
    from mesohops_amd.serving import ServingLoop
    reqs = _requests()
    loop = ServingLoop(precision_mode=prec, arena_bytes=1 << 30, max_slots=9, record=True, sub_batches=sub)
    assert loop.K == min(sub, 9) and loop.slot_cap == ServingLoop.split_slots(9, loop.K)
    mix, stats = loop.run(reqs)
    loop.close()
    print(prec, "sub-batches", sub, "mix stats:", stats)
    assert stats["completed"] == len(reqs)
    assert stats["p50_latency_s"] is not None and stats["p99_latency_s"] >= stats["p50_latency_s"]
    # the requests overlapped on the device
    spans = sorted((mix[r.rid]["admitted"], mix[r.rid]["finished"]) for r in reqs)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    for r in reqs:
        r.arrival = 0.0
        one = ServingLoop(precision_mode=prec, arena_bytes=1 << 30, max_slots=2, record=True)
        alone, _ = one.run([r])
        one.close()
        a, m = alone[r.rid], mix[r.rid]
        assert a["nhier"] == m["nhier"], r.rid
        assert all(np.array_equal(x, y) for x, y in zip(a["states"], m["states"]))
        assert np.array_equal(a["psi"], m["psi"]), r.rid
        assert a["psi"].shape[0] == r.n_steps + 1
        print("request", r.rid, "alone == in mix over", a["psi"].shape[0], "records; n_hier", min(a["nhier"]),
              "..", max(a["nhier"]))


def _big_requests():
    from mesohops_amd.workloads import Request
    return [Request(rid=21 + i, arrival=0.005 * i, nsite=46 + 3 * i, maxhier=7, delta=6e-4, t_max=50.0)
            for i in range(8)]


def test_small_arena_preempts_and_requeues():
    """An arena that holds any one request but not all of them at once: the
    requests whose grown bases do not fit are preempted and run again later;
    every request completes with the result it has alone (bitwise)."""
    _need_gpu()
    from mesohops_amd.serving import ServingLoop
    big = 1 << 30
    alone = {}
    overhead, peak = 0, 0
    for r in _big_requests():
        one = ServingLoop(precision_mode="fp64", arena_bytes=big, max_slots=2, record=True)
        res, _ = one.run([r])
        overhead = max(overhead, big - one.rt.free_bytes())   # descriptors and staging, after the run
        one.close()
        alone[r.rid] = res[r.rid]
        peak = max(peak, res[r.rid]["peak_bytes"])
    total = sum(a["peak_bytes"] for a in alone.values())
    # a growing slot holds its old and new blocks during a basis update and
    # leaves holes below them in a first-fit arena (about 3x its peak block);
    # the staging area grows with the rows of every slot stepped together
    arena = overhead + int(3.5 * peak) + int(0.1 * total)
    print("arena %d B: overhead %d, largest request %d, all requests %d" % (arena, overhead, peak, total))
    assert overhead + total > arena    # they cannot all run at their peaks together
    loop = ServingLoop(precision_mode="fp64", arena_bytes=arena, max_slots=8, record=True, reserve_states=2,
                       reserve_aux=4)
    mix, stats = loop.run(_big_requests())
    loop.close()
    print("small arena stats:", stats)
    assert stats["completed"] == len(alone) and stats["preempted"] >= 1
    for rid, a in alone.items():
        m = mix[rid]
        assert a["nhier"] == m["nhier"], rid
        assert np.array_equal(a["psi"], m["psi"]), rid


def test_request_larger_than_arena_raises():
    """A request that cannot fit in the whole arena is refused, not waited on."""
    _need_gpu()
    from mesohops_amd.errors import UnsupportedRequest
    from mesohops_amd.serving import ServingLoop
    from mesohops_amd.workloads import Request
    loop = ServingLoop(precision_mode="fp64", arena_bytes=8 << 20, max_slots=2, record=False)
    try:
        with pytest.raises(UnsupportedRequest):
            loop.run([Request(rid=31, arrival=0.0, nsite=200, maxhier=10, delta=1e-4, t_max=100.0)],
                     time_limit=120.0)
    finally:
        loop.close()


def test_long_noise_filtered_on_concurrent_streams_equals_sequential():
    """Noise of long requests (t_max 1.1-1.7 ps: 5,120- to 7,168-point
    transforms done as 16,384-point Bluestein transforms, longer than the
    LDS holds, so the filter works in device scratch rows) filtered
    from four host threads on four streams at once, as the serving loop's
    batch threads do at admission, equals the same noise filtered one
    request at a time — bitwise."""
    _need_gpu()
    import ctypes
    import threading
    from mesohops_amd.basis import SystemInfo
    from mesohops_amd.engine import _stream_ptr
    from mesohops_amd.noise import HopsNoise, device_noise
    from mesohops_amd.workloads import Request, chain_system
    dev = torch.device("cuda:0")
    reqs = [Request(rid=40 + i, arrival=0.0, nsite=n, maxhier=2, delta=1e-3, t_max=t)
            for i, (n, t) in enumerate([(24, 1500.0), (7, 1100.0), (40, 1700.0), (15, 1500.0)])]
    preps = []
    for r in reqs:
        sysp = chain_system(r.nsite)
        info = SystemInfo(sysp)
        nz = HopsNoise({"SEED": r.rid, "MODEL": "FFT_FILTER", "TLEN": r.tlen, "TAU": r.tau_noise}, info, sysp, 1)
        n = 2 * (nz.n_t - 1)
        assert n & (n - 1) and 2 * n - 1 > 8192  # Bluestein, over the LDS length
        preps.append((nz, list(range(info.n_l2)), nz.raw()))
    seq = [device_noise([nz], rows, dev, _stream_ptr(dev), raws=[raw])[0].cpu() for nz, rows, raw in preps]
    for rnd in range(2):
        out = [None] * len(preps)
        go = threading.Barrier(len(preps))

        def run(i):
            s = torch.cuda.Stream(device=dev)
            nz, rows, raw = preps[i]
            go.wait()
            with torch.cuda.stream(s):
                z = device_noise([nz], rows, dev, ctypes.c_void_p(s.cuda_stream), raws=[raw])[0]
                s.synchronize()
            out[i] = z.cpu()

        th = [threading.Thread(target=run, args=(i,)) for i in range(len(preps))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for i, (a, b) in enumerate(zip(seq, out)):
            assert b is not None and torch.equal(a, b), (rnd, i)
