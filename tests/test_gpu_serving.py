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


@pytest.mark.parametrize("prec,sub", [("fp64", 1), ("mixed", 1), ("mixed", 3)])
def test_alone_equals_in_mix(prec, sub):
    """sub: runtime batches stepped concurrently (3: the four requests spread
    over three batches, each with its own stream and worker threads)."""
    _need_gpu()
    from mesohops_amd.serving import ServingLoop
    reqs = _requests()
    loop = ServingLoop(precision_mode=prec, arena_bytes=1 << 30, max_slots=9, record=True, sub_batches=sub)
    assert loop.K == sub
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
