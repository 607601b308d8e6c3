"""Launch-shape autotuner for the RK4 stage kernels.

For a given batch it times a few RK4 steps for each candidate layout (number
of trajectories interleaved per row) and stage kernel (operator-stationary or
streaming) and keeps the fastest.  The batch is rewound to its initial state
afterwards, so tuning does not change results (every candidate computes the
same numbers; only the memory layout and the launch shape differ).
"""
import sys
import time

import numpy as np
import torch

from . import _lib

# (lanes, stationary stage kernel, trajectories per cache-resident chunk (0 =
# whole batch), nontemporal hint level (0 none, 1 accumulator and carry, 2 also
# the stage outputs, 3 also the row's own aux block, 4 also the ELL
# neighbours), trajectory-group split of the stationary kernel (0 =
# automatic), locality tile order, trajectory lanes per thread of the
# stationary kernel (1, 2 or 4), LDS-staged form of the stationary kernel)
CANDIDATES = [(1, True, 0, 0, 0, 0, 1, 0), (4, True, 0, 0, 0, 0, 1, 0), (8, True, 0, 0, 0, 0, 1, 0),
              (16, True, 0, 0, 0, 0, 1, 0), (4, False, 0, 0, 0, 0, 1, 0), (4, True, 64, 0, 0, 0, 1, 0),
              (4, True, 0, 2, 0, 0, 1, 0), (8, True, 0, 2, 0, 0, 1, 0), (8, True, 0, 3, 0, 0, 1, 0),
              (16, True, 0, 2, 0, 0, 1, 0), (8, True, 0, 2, 0, 1, 1, 0),
              (4, True, 0, 2, 0, 0, 2, 0), (8, True, 0, 0, 0, 0, 2, 0), (8, True, 0, 2, 0, 0, 2, 0),
              (8, True, 0, 3, 0, 0, 2, 0), (8, True, 0, 2, 0, 0, 4, 0), (16, True, 0, 2, 0, 0, 2, 0),
              (4, True, 0, 2, 0, 0, 1, 1), (4, True, 0, 2, 0, 0, 2, 1), (8, True, 0, 2, 0, 0, 1, 1),
              (8, True, 0, 0, 0, 0, 2, 1), (8, True, 0, 2, 0, 0, 2, 1), (8, True, 0, 3, 0, 0, 2, 1),
              (16, True, 0, 2, 0, 0, 2, 1), (16, True, 0, 3, 0, 0, 2, 1), (16, True, 0, 2, 0, 0, 4, 1)]


def apply(choice, batch=None):
    """Set the library options of a candidate (the layout is set by the
    batch; the tile order is a property of its device state)."""
    lanes, stationary, chunk, nt, q, order, vec, lds = choice
    _lib.set_option("stationary_stage", 1 if stationary else 0)
    _lib.set_option("lds_stage", lds)
    _lib.set_option("chunk_traj", chunk)
    _lib.set_option("nt_stream", nt)
    _lib.set_option("groups_q", q)
    _lib.set_option("vec_lanes", vec)
    if batch is not None:
        batch.state.use_tile_order(bool(order))


def tune(batch, psi_0, tau, steps=12, candidates=None, verbose=False):
    """Returns ((lanes, stationary, chunk, nt, q, order, vec, lds), {candidate: seconds per step})."""
    candidates = candidates or CANDIDATES
    nz = batch.noise1[0]
    ns = int(round(tau * 0.5 / nz.param["TAU"]))
    timings = {}
    for cand in candidates:
        lanes, stationary, chunk, nt, q, order, vec, lds = cand
        if lanes > 1 and batch.B < 8 * lanes and stationary:
            continue
        if chunk and (chunk >= batch.B or chunk < 8 * lanes):
            continue
        if vec > 1 and (vec > lanes or batch.B % vec != 0 or not stationary):
            continue
        batch.relayout(lanes)
        batch.rewind(psi_0)
        apply(cand, batch)
        st = batch.state
        st.bt.psi_store_d = None
        st.propagate(0, 1, tau, ns, 0)
        torch.cuda.synchronize(batch.device)
        t0 = time.perf_counter()
        st.propagate(0, steps, tau, ns, 0)
        torch.cuda.synchronize(batch.device)
        timings[cand] = (time.perf_counter() - t0) / steps
        if verbose:
            print("autotune lanes=%d stationary=%d chunk=%d nt=%d q=%d order=%d vec=%d lds=%d: %.3f ms/step"
                  % (lanes, stationary, chunk, nt, q, order, vec, lds, timings[cand] * 1e3), file=sys.stderr)
    best = min(timings, key=timings.get)
    batch.relayout(best[0])
    apply(best, batch)
    batch.rewind(psi_0)
    return best, timings
