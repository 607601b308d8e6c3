"""Device-side runtime: operator upload, batch buffers and the propagate call.

PyTorch is used only to allocate HBM buffers and to name the current stream;
all arithmetic runs in libhops_mi355x.so.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import HopsBatch, HopsOp, check, lib
from .basis import BasisTables
from .errors import EngineError
from .params import hbar

TILE_ROWS = 256
N_XCD = 8


def _require_gpu():
    if not torch.cuda.is_available():
        raise EngineError("no ROCm GPU visible: the HOPS engine runs only on the device")


def _dev(arr, device, dtype=None):
    """Upload a host array (at least one element, so the pointer is valid)."""
    a = np.ascontiguousarray(arr)
    if a.size == 0:
        a = np.zeros(1, dtype=a.dtype)
    t = torch.from_numpy(a.copy())
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device)


def _stream_ptr(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class DeviceOperator:
    """hops_op_t of one basis with its tables resident in HBM; the hierarchy
    super-operator (ELL K2_kp1 + K2_km1) and K0 are generated on the device."""

    def __init__(self, tables: BasisTables, precision: int, device="cuda:0"):
        _require_gpu()
        self.device = torch.device(device)
        self.tables = tables
        self.precision = precision
        self._keep = {}
        op = HopsOp()
        for name in ("n_state", "n_hier", "n_mode", "n_lop", "n_act", "ell_w", "n_pat", "n_mode_abs"):
            setattr(op, name, int(getattr(tables, name)))
        op.n_first = len(tables.first)
        pp = tables.tables["pat_ptr"]
        op.pat_w = int(np.max(np.diff(pp))) if len(pp) > 1 else 0
        for name, arr in tables.tables.items():
            if arr is None:
                setattr(op, name + "_d", None)
                continue
            t = _dev(arr, self.device)
            self._keep[name] = t
            setattr(op, name + "_d", t.data_ptr())
        nr = tables.n_state * tables.n_hier
        w = max(tables.ell_w, 1)
        self.ell_col = torch.empty(w * nr, dtype=torch.int32, device=self.device)
        vdt = torch.complex128 if precision == _lib.HOPS_PREC_FP64 else torch.complex64
        self.ell_val = torch.empty(w * nr, dtype=vdt, device=self.device)
        self.k0 = torch.empty(tables.n_hier, dtype=torch.complex128, device=self.device)
        op.ell_col_d = self.ell_col.data_ptr()
        op.ell_val_d = self.ell_val.data_ptr()
        op.k0_d = self.k0.data_ptr()
        self.hd = None
        if getattr(tables, "hd", None) is not None:
            # dense -iH: real and imaginary planes, zero-padded to a multiple of 16
            n = tables.n_state
            hn = ((n + 15) // 16) * 16
            planes = np.zeros((2, hn, hn), dtype=np.float64 if precision == _lib.HOPS_PREC_FP64 else np.float32)
            planes[0, :n, :n] = tables.hd.real
            planes[1, :n, :n] = tables.hd.imag
            self.hd = torch.from_numpy(planes).to(self.device)
            op.hd_d = self.hd.data_ptr()
            op.hd_n = hn
        self.op = op
        with torch.cuda.device(self.device):
            check(lib().hops_build_super(ctypes.byref(op), precision, _stream_ptr(self.device)))

    @property
    def n_rows(self):
        return self.tables.n_state * self.tables.n_hier

    def ell_dense(self):
        """Download K2_kp1 + K2_km1 as (row, col, value) triplets (for tests)."""
        nr = self.n_rows
        w = self.tables.ell_w
        col = self.ell_col.cpu().numpy()[: w * nr].reshape(w, nr)
        val = self.ell_val.cpu().numpy()[: w * nr].reshape(w, nr)
        rows = np.broadcast_to(np.arange(nr), (w, nr))
        m = col >= 0
        return rows[m], col[m], val[m]


def tile_map(G, n_rows, lanes=1):
    """Workgroup -> (group, tile) list; a tile is 256 / lanes rows of one group
    of ``lanes`` interleaved trajectories.  With at least 8 groups each group's
    tiles are pinned to one XCD (workgroup i runs on XCD i % 8), so its
    neighbour gathers hit that XCD's L2; fewer groups spread over all XCDs."""
    rows = TILE_ROWS // lanes
    tpt = (n_rows + rows - 1) // rows
    if tpt >= 65536 or G >= 32768:
        raise EngineError("batch too large for the tile encoding")
    if G < N_XCD:
        return np.array([g * 65536 + t for g in range(G) for t in range(tpt)], dtype=np.int32)
    queues = [[g * 65536 + t for g in range(x, G, N_XCD) for t in range(tpt)] for x in range(N_XCD)]
    depth = max(len(q) for q in queues)
    out = np.full(depth * N_XCD, -1, dtype=np.int32)
    for x, q in enumerate(queues):
        out[x: x + N_XCD * len(q): N_XCD] = q
    return out


def locality_tile_order(tables, lanes):
    """Visiting order of the operator-stationary tiles (512 / lanes rows):
    a tile whose rows gather from deeper auxiliaries (K2_kp1) is placed just
    before the median tile of those neighbours, so the gathered rows are
    still in L2 when their own tile runs, and their gathers back up the
    hierarchy find the parent rows there.  Other tiles keep their place."""
    n, nh = tables.n_state, tables.n_hier
    nr = n * nh
    rpt = 512 // lanes
    ntile = (nr + rpt - 1) // rpt
    up_ptr = np.asarray(tables.tables["up_ptr"], dtype=np.int64)
    up_aux = np.asarray(tables.tables["up_aux"], dtype=np.int64)
    child_tile = (up_aux * n + n // 2) // rpt
    keys = []
    for t in range(ntile):
        k0, k1 = (t * rpt) // n, min(nh - 1, ((t + 1) * rpt - 1) // n)
        ch = child_tile[up_ptr[k0]:up_ptr[k1 + 1]]
        ch = ch[ch > t]
        keys.append((float(np.median(ch)), 0, t) if ch.size else (float(t), 1, t))
    return np.array([k[2] for k in sorted(keys)], dtype=np.int32)


def default_lanes(B):
    """Trajectories interleaved per row: 4 once there are enough trajectories
    to give every XCD at least one group of 4, else 1."""
    return 4 if B >= 32 else 1


class BatchState:
    """Device buffers of B trajectories that share one DeviceOperator."""

    def __init__(self, dop: DeviceOperator, B: int, eom_kind: int, noise1, noise2, n_t: int,
                 store_slots: int = 0, lanes: int = None):
        self.dop = dop
        self.device = dop.device
        t = dop.tables
        self.B = B
        self.raw_first = False
        self.eom_kind = eom_kind
        nr = dop.n_rows
        prec = dop.precision
        lanes = default_lanes(B) if lanes is None else int(lanes)
        if lanes not in (1, 2, 4, 8, 16):
            raise EngineError("lanes must be a power of two <= 16")
        self.lanes = lanes
        self.G = (B + lanes - 1) // lanes
        P = self.G * lanes * nr
        cR = torch.complex128 if prec == _lib.HOPS_PREC_FP64 else torch.complex64
        dev = self.device
        self.phi = torch.zeros(P, dtype=torch.complex128, device=dev)
        self.y0 = torch.empty(P if prec == _lib.HOPS_PREC_MIXED else 1, dtype=torch.complex64, device=dev)
        self.ya = torch.empty(P, dtype=cR, device=dev)
        self.yb = torch.empty(P, dtype=cR, device=dev)
        self.acc = torch.empty(P, dtype=cR, device=dev)
        # (-iH) Y of each stage input when the operator has a dense system block
        self.hy = torch.empty(P, dtype=cR, device=dev) if dop.op.hd_d else None
        if noise1.shape != (B, t.n_act, n_t):
            raise EngineError("noise1 shape %s != %s" % (tuple(noise1.shape), (B, t.n_act, n_t)))
        self.z1 = noise1.contiguous()
        self.z2 = None if noise2 is None else noise2.contiguous()
        if self.z2 is not None and self.z2.dtype != self.z1.dtype:
            raise EngineError("noise 1 and noise 2 must share one dtype")
        if self.z1.dtype not in (torch.complex64, torch.complex128):
            raise EngineError("noise must be complex64 or complex128")
        self.zmem = torch.zeros(B * t.n_mode_abs + 1, dtype=torch.complex128, device=dev)
        self.ws_stride = 4 * t.n_mode_abs + 2
        self.coef_stride = 4 * t.n_pat + 2 * t.n_mode + 4
        self.ws = torch.zeros(B * self.ws_stride, dtype=torch.float64, device=dev)
        self.coef = torch.zeros(B * self.coef_stride, dtype=torch.float64, device=dev)
        self.coefT_stride = ((B + 63) // 64) * 64
        rowsT = 2 * t.n_pat + t.n_mode + 1
        self.coefT = torch.zeros(rowsT * self.coefT_stride, dtype=cR, device=dev)
        self.sigT = torch.zeros(self.coefT_stride, dtype=torch.float64, device=dev)
        self.store_slots = store_slots
        self.psi_store = (torch.zeros(B * store_slots * t.n_state, dtype=torch.complex128, device=dev)
                          if store_slots > 0 else None)
        self.tiles = torch.from_numpy(tile_map(self.G, nr, lanes)).to(dev)
        bt = HopsBatch()
        bt.B = B
        bt.precision = prec
        bt.eom = eom_kind
        bt.n_t = n_t
        bt.coef_stride = self.coef_stride
        bt.ws_stride = self.ws_stride
        bt.store_stride = store_slots
        bt.lane_shift = int(lanes).bit_length() - 1
        bt.phi_d = self.phi.data_ptr()
        bt.y0_d = self.y0.data_ptr()
        bt.ya_d = self.ya.data_ptr()
        bt.yb_d = self.yb.data_ptr()
        bt.acc_d = self.acc.data_ptr()
        bt.hy_d = self.hy.data_ptr() if self.hy is not None else None
        bt.z1_d = self.z1.data_ptr()
        bt.z2_d = None if self.z2 is None else self.z2.data_ptr()
        bt.zmem_d = self.zmem.data_ptr()
        bt.ws_d = self.ws.data_ptr()
        bt.coef_d = self.coef.data_ptr()
        bt.psi_store_d = None if self.psi_store is None else self.psi_store.data_ptr()
        bt.tile_d = self.tiles.data_ptr()
        bt.n_tiles = int(self.tiles.numel())
        bt.tile_rows = TILE_ROWS
        bt.coefT_d = self.coefT.data_ptr()
        bt.sigT_d = self.sigT.data_ptr()
        bt.coefT_stride = self.coefT_stride
        bt.noise_c128 = 1 if self.z1.dtype == torch.complex128 else 0
        self.bt = bt

    def use_tile_order(self, on):
        """Visit the stationary kernel's tiles in locality order (same results)."""
        if on:
            order = locality_tile_order(self.dop.tables, self.lanes)
            self.tile_order = torch.from_numpy(order).to(self.device)
            self.bt.tile_order_d = self.tile_order.data_ptr()
        else:
            self.tile_order = None
            self.bt.tile_order_d = None

    def set_noise(self, z1, z2, n_t):
        """Point the batch at other noise arrays [B, n_act, n_t] (the
        interpolated noise of one propagate call)."""
        self.z1 = z1.contiguous()
        self.z2 = None if z2 is None else z2.contiguous()
        self.bt.z1_d = self.z1.data_ptr()
        self.bt.z2_d = None if self.z2 is None else self.z2.data_ptr()
        self.bt.n_t = int(n_t)
        self.bt.noise_c128 = 1 if self.z1.dtype == torch.complex128 else 0

    def _to_pool(self, a):
        """[B, n_rows] -> interleaved pool order [G, n_rows, T]."""
        nr, T = self.dop.n_rows, self.lanes
        full = np.zeros((self.G * T, nr), dtype=a.dtype)
        full[: self.B] = a
        return np.ascontiguousarray(full.reshape(self.G, T, nr).transpose(0, 2, 1)).reshape(-1)

    def _from_pool(self, flat):
        nr, T = self.dop.n_rows, self.lanes
        return flat.reshape(self.G, nr, T).transpose(0, 2, 1).reshape(self.G * T, nr)[: self.B]

    def apply_state_op(self, op):
        """phi[b, k, :] <- op @ phi[b, k, :] on the device for every trajectory
        and auxiliary (op: dense [n_state, n_state], the operator restricted to
        the state basis)."""
        n = self.dop.tables.n_state
        opd = torch.from_numpy(np.ascontiguousarray(op, dtype=np.complex128).reshape(n, n)).to(self.device)
        out = torch.zeros_like(self.phi)
        check(lib().hops_apply_state_op(ctypes.c_void_p(self.phi.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                        ctypes.c_void_p(opd.data_ptr()), n, self.dop.tables.n_hier, self.B,
                                        int(self.lanes).bit_length() - 1, _stream_ptr(self.device)))
        self.phi.copy_(out)  # same buffer: the batch descriptor keeps its pointer
        torch.cuda.synchronize(self.device)
        self.raw_first = True

    def set_phi(self, phi_host):
        """phi_host: complex [B, n_rows] (or [n_rows], broadcast)."""
        nr = self.dop.n_rows
        a = np.asarray(phi_host, dtype=np.complex128)
        a = np.broadcast_to(a.reshape(-1, nr), (self.B, nr))
        self.phi.copy_(torch.from_numpy(self._to_pool(a)).to(self.device))

    def phi_host(self):
        return np.ascontiguousarray(self._from_pool(self.phi.cpu().numpy()))

    def zmem_host(self):
        return self.zmem.cpu().numpy()[: self.B * self.dop.tables.n_mode_abs].reshape(self.B, -1)

    def psi_store_host(self, n_slots):
        n = self.dop.tables.n_state
        a = self.psi_store.cpu().numpy().reshape(self.B, self.store_slots, n)
        return a[:, :n_slots]

    def propagate(self, it0: int, n_steps: int, tau: float, noise_step: int, store0: int = 0,
                  noalloc: bool = False):
        """n_steps RK4 steps of size tau from noise index it0.  noalloc: through
        hops_propagate_noalloc, which allocates no device memory."""
        self.bt.dt_over_hbar = float(tau / hbar)
        self.bt.noise_step = int(noise_step)
        if self.bt.psi_store_d and store0 + n_steps > self.store_slots:
            raise EngineError("psi store too small (%d + %d > %d)" % (store0, n_steps, self.store_slots))
        # raw_first: the first step takes Phi as an operator left it (no
        # normalisation before the step), as the reference steps it
        self.bt.sigma_raw = 1 if self.raw_first else 0
        try:
            with torch.cuda.device(self.device):
                fn = lib().hops_propagate_noalloc if noalloc else lib().hops_propagate
                check(fn(ctypes.byref(self.dop.op), ctypes.byref(self.bt), int(it0), int(n_steps), int(store0),
                         _stream_ptr(self.device)))
        finally:
            self.bt.sigma_raw = 0
            if n_steps > 0:
                self.raw_first = False
