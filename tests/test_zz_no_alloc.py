"""hops_propagate_noalloc (the entry point that allocates no device memory)
against hops_propagate: two calls in a row through the C ABI, the second on a
smaller batch, each bit-identical to hops_propagate on the same input; and the
error for a batch whose scratch exceeds the library's array."""
import numpy as np
import pytest

import systems
from oracle import hops_oracle as O


def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")


def _batch(nsite, depth, B, prec, noise_model="FFT_FILTER"):
    from mesohops_amd import HopsBatch
    sysp = systems.chain(nsite)
    nps = [{"SEED": 100 + i, "MODEL": noise_model, "TLEN": 100.0, "TAU": 1.0} for i in range(B)]
    b = HopsBatch(sysp, {"EQUATION_OF_MOTION": "LINEAR"}, nps[0], None, {"MAXHIER": depth},
                  noise_params=nps, precision_mode=prec)
    psi0 = np.zeros(nsite, complex)
    psi0[nsite // 2] = 1.0
    b.initialize(psi0)
    return b


@pytest.mark.gpu
def test_noalloc_equals_propagate_on_two_calls():
    _need_gpu()
    import torch
    # (sites, depth, trajectories, steps, precision): both above the resident
    # size, so the fused LINEAR prologue runs and uses the scratch; the
    # second call is the smaller one
    for nsite, depth, B, steps, prec in [(6, 3, 32, 40, "mixed"), (4, 4, 8, 20, "fp64")]:
        print("noalloc case: %d sites, depth %d, %d trajectories, %d steps, %s" % (nsite, depth, B, steps, prec),
              flush=True)
        out = []
        for noalloc in (False, True):
            b = _batch(nsite, depth, B, prec)
            assert b.tables.n_state * b.tables.n_hier > 1024
            b.state.propagate(0, steps, 2.0, 1, noalloc=noalloc)
            torch.cuda.synchronize()
            out.append(b.state.phi_host())
        assert np.array_equal(out[0], out[1])
        assert np.max(np.abs(out[0])) > 0


@pytest.mark.gpu
def test_noalloc_rejects_batches_beyond_its_scratch():
    _need_gpu()
    import torch
    from mesohops_amd.errors import EngineError
    print("noalloc limit case: 60 sites, depth 1, 1024 trajectories", flush=True)
    b = _batch(60, 1, 1024, "mixed", noise_model="ZERO")
    before = b.state.phi_host()
    with pytest.raises(EngineError, match="scratch"):
        b.state.propagate(0, 2, 2.0, 1, noalloc=True)
    torch.cuda.synchronize()
    assert np.array_equal(b.state.phi_host(), before)   # nothing ran
    b.state.propagate(0, 2, 2.0, 1)                      # the allocating form takes it
    torch.cuda.synchronize()
