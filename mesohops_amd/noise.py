"""Noise of a HOPS trajectory (noise/hops_noise.py, noise/prepare_functions.py).

Raw white noise is drawn on the host with NumPy, because the reference defines
its values by NumPy's PCG64 streams (``PCG64(seed).jumped(lop)``) and
``normal``/``random`` draws; the correlation filter (circulant embedding and
the two FFTs per row) runs on the device in noise_fft.hip.
"""
from __future__ import annotations

import ctypes
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from ._lib import check, lib
from .basis import SystemInfo, operator_signature
from .errors import UnsupportedRequest
from .params import NOISE_DICT_DEFAULT, NOISE_TYPE_DEFAULT, initialize_dict, precision


def noise_time_axis(tlen, tau):
    """T_AXIS of the noise (noise/hops_noise.py:161-163)."""
    return np.arange(int(np.ceil(tlen / tau)) + 1) * tau


def raw_slot_index(ntaus):
    """Slots of the real/imaginary (or radius/phase) draws so that a prefix of
    the noise does not depend on the axis length (noise/hops_noise.py:592-628):
    slot pairs (0,1), (2,3), ... walk up from the start and the last pairs walk
    back down from the end."""
    m1 = ntaus - 1
    a = np.concatenate([[0], np.arange(2, 4 * (ntaus - 2), 4), [4 * m1 - 2],
                        np.arange(4 * (ntaus - 2), 3, -4)]).astype(np.int64)
    b = np.concatenate([[1], np.arange(3, 4 * (ntaus - 1) - 1, 4), [4 * m1 - 1],
                        np.arange(4 * (m1 - 1) + 1, 4, -4)]).astype(np.int64)
    return a, b


class HopsNoise:
    """Parameters and host-side raw draws of one noise trajectory."""

    def __init__(self, noise_param, info: SystemInfo, system_param, noise_type=1):
        p = dict(noise_param)
        tag = "1"
        if noise_type == 1 or p.get("MODEL") == "ZERO":
            if noise_type == 1 and p.get("FLAG_REAL", False):
                p["FLAG_REAL"] = False
        else:
            tag = "2"
            if "FLAG_REAL" not in p:
                p["FLAG_REAL"] = True
        self.param = initialize_dict(p, NOISE_DICT_DEFAULT, NOISE_TYPE_DEFAULT, "HopsNoise")
        if self.param["MODEL"] not in ("FFT_FILTER", "ZERO", "PRE_CALCULATED"):
            raise UnsupportedRequest("Noise.param[MODEL] {}".format(self.param["MODEL"]), "Noise.prepare_noise")
        # NOISE_WINDOW only bounds the slice of the noise the reference keeps in
        # a working copy (noise/hops_noise.py:385-433); the values are those of
        # the full noise, which stays resident in HBM here, so it changes nothing
        self.t_axis = noise_time_axis(self.param["TLEN"], self.param["TAU"])
        self.n_t = len(self.t_axis)
        key = "L_NOISE" + tag
        if tag in info.noise_n_l2 and key in system_param and info.param.get(key) is system_param[key]:
            self.n_l2 = info.noise_n_l2[tag]  # counted once for the system
        else:
            lops = system_param.get(key, system_param["L_HIER"]) if tag == "1" else system_param[key]
            self.n_l2 = len(set(operator_signature(L) for L in lops))
        self.corr = system_param.get("ALPHA_NOISE" + tag)
        self.corr_param = system_param.get("PARAM_NOISE" + tag)
        self.modes_by_l = info.noise_modes_by_l.get(tag, [])

    @property
    def model(self):
        return self.param["MODEL"]

    @property
    def interpolate(self):
        return bool(self.param["INTERPOLATE"]) and self.model != "ZERO"

    def needs_c128(self):
        """The reference uses this noise in complex128: a cubic interpolation
        of it (noise/hops_noise.py:242-243, 259-261), or a PRE_CALCULATED
        array given in a wider type than complex64 (:262-263, used as given)
        whose values complex64 would round."""
        if self.interpolate:
            return True
        if self.model != "PRE_CALCULATED":
            return False
        arr = np.asarray(self._precalc_array())
        if arr.dtype == np.complex64:
            return False
        with np.errstate(over="ignore"):
            return not np.array_equal(np.complex64(arr), arr)  # values complex64 cannot hold

    def alpha(self):
        """Correlation function of each L-operator, complex64 [n_l2, n_t]
        (noise/hops_noise.py:167-192, cast at :236)."""
        a = np.zeros((self.n_l2, self.n_t), dtype=np.complex128)
        for l in range(self.n_l2):
            for i in self.modes_by_l[l]:
                a[l] += self.corr(self.t_axis, *self.corr_param[i])
        return np.complex64(a)

    def raw(self):
        """Uncorrelated complex64 noise [n_l2, 2(n_t-1)] (noise/hops_noise.py:453-590)."""
        seed = self.param["SEED"]
        nt = self.n_t
        if isinstance(seed, str):
            if not os.path.isfile(seed):
                raise UnsupportedRequest("Noise.param[SEED] {} is not the address of a valid file ".format(seed),
                                         "Noise._prepare_rand", True)
            seed = np.load(seed, allow_pickle=False)
        if isinstance(seed, np.ndarray):
            if seed.shape != (self.n_l2, 2 * (nt - 1)):
                raise UnsupportedRequest("Noise.param[SEED] is an array of the wrong length",
                                         "Noise._prepare_rand", True)
            return np.complex64(seed)
        draws = np.zeros((self.n_l2, 4 * (nt - 1)))
        base = np.random.PCG64(seed=seed)  # jumped() leaves base unchanged
        for l in range(self.n_l2):
            rng = np.random.Generator(base.jumped(l))
            if self.param["RAND_MODEL"] == "BOX_MULLER":
                draws[l] = rng.random(size=4 * (nt - 1))
            elif self.param["RAND_MODEL"] == "SUM_GAUSSIAN":
                draws[l] = rng.normal(loc=0, size=4 * (nt - 1))
            else:
                raise UnsupportedRequest("Noise.param[RAND_MODEL] {}".format(self.param["RAND_MODEL"]),
                                         "Noise._prepare_rand")
        ia, ib = raw_slot_index(nt)
        if self.param["RAND_MODEL"] == "BOX_MULLER":
            return np.complex64(np.sqrt(-2.0 * np.log(draws[:, ia])) * np.exp(2.0j * np.pi * draws[:, ib]))
        out = np.empty((self.n_l2, len(ia)), dtype=np.complex64)
        out.real = draws[:, ia]  # each part rounded to float32, as complex64(a + 1j b) rounds it
        out.imag = draws[:, ib]
        return out

    def _precalc_array(self):
        seed = self.param["SEED"]
        if isinstance(seed, str):
            if not os.path.isfile(seed) or not seed.endswith(".npy"):
                raise UnsupportedRequest("Noise.param[SEED] {} is not the address of a valid file".format(seed),
                                         "Noise.prepare_noise", True)
            seed = np.load(seed, allow_pickle=False)
        arr = np.asarray(seed)
        if arr.shape != (self.n_l2, self.n_t):
            raise UnsupportedRequest("Noise.param[SEED] is an array of the wrong length",
                                     "Noise.prepare_noise", True)
        return arr

    def precalculated(self, c128=False):
        arr = self._precalc_array()
        return np.complex128(arr) if c128 else np.complex64(arr)

    def time_index(self, t):
        """Index of t on the noise grid; off-grid times are errors
        (noise/hops_noise.py:430-447)."""
        hit = np.abs(self.t_axis - t) < precision
        if np.sum(hit) != 1:
            raise UnsupportedRequest("Off axis t-samples when INTERPOLATE = False", "NoiseModel.get_noise()")
        return int(np.where(hit)[0][0])


def device_noise(noises, rows, device, stream_ptr, raws=None, c128=False):
    """Correlated noise of B trajectories for the hierarchy L-operators
    ``rows`` (absolute indices), complex64 [B, len(rows), n_t] in HBM, or
    complex128 with ``c128`` (the filter's values before the complex64 cast,
    PRE_CALCULATED arrays as given).  All noises must share the system (same
    correlation functions).  raws: the host raw draws of each noise
    (HopsNoise.raw()), when already made."""
    B = len(noises)
    nz = noises[0]
    nt = nz.n_t
    rows = [int(r) for r in rows]
    out = torch.zeros((B, len(rows), nt), dtype=torch.complex128 if c128 else torch.complex64, device=device)
    if nz.model == "ZERO":
        return out
    if nz.model == "PRE_CALCULATED":
        host = np.stack([n.precalculated(c128)[rows] for n in noises])
        out.copy_(torch.from_numpy(host).to(device))
    else:
        alpha = nz.alpha()
        N = 2 * (nt - 1)
        # phase and modulus rounded as NumPy rounds them for complex64 input
        ang_d = torch.from_numpy(np.ascontiguousarray(np.angle(alpha))).to(device)
        mag_d = torch.from_numpy(np.ascontiguousarray(np.abs(alpha))).to(device)
        s_w = torch.empty((nz.n_l2, N), dtype=torch.float64, device=device)
        nu = torch.empty(nz.n_l2, dtype=torch.float64, device=device)
        check(lib().hops_noise_embedding(ctypes.c_void_p(ang_d.data_ptr()), ctypes.c_void_p(mag_d.data_ptr()),
                                         nz.n_l2, nt,
                                         ctypes.c_void_p(s_w.data_ptr()), ctypes.c_void_p(nu.data_ptr()),
                                         stream_ptr))
        # raw draws per trajectory on worker threads (NumPy's generators and
        # array arithmetic run without the GIL), filtered chunk by chunk so
        # that about 1 GiB of raw draws is held on the host at a time
        nr = len(rows)
        chunk = max(1, min(B, (1 << 30) // max(1, nr * N * 8)))
        row_lop = torch.from_numpy(np.tile(np.array(rows, dtype=np.int32), chunk)).to(device)
        workers = max(1, min(16, os.cpu_count() or 1, B))
        with ThreadPoolExecutor(max_workers=workers) as pool:
            for b0 in range(0, B, chunk):
                b1 = min(B, b0 + chunk)
                if raws is None:
                    part = list(pool.map(lambda n: n.raw()[rows], noises[b0:b1]))
                else:
                    part = [r[rows] for r in raws[b0:b1]]
                z_d = torch.from_numpy(np.ascontiguousarray(np.stack(part).reshape((b1 - b0) * nr, N))).to(device)
                filt = lib().hops_noise_filter_c128 if c128 else lib().hops_noise_filter
                check(filt(ctypes.c_void_p(z_d.data_ptr()), (b1 - b0) * nr, nt,
                           ctypes.c_void_p(s_w.data_ptr()), ctypes.c_void_p(nu.data_ptr()),
                           ctypes.c_void_p(row_lop.data_ptr()), ctypes.c_void_p(out[b0:b1].data_ptr()), stream_ptr))
                # the filter has read z_d before it is freed: wait for the
                # filter's stream alone when torch queues on it too
                cur = torch.cuda.current_stream(device)
                if cur.cuda_stream == (stream_ptr.value or 0):
                    cur.synchronize()
                else:
                    torch.cuda.synchronize(device)
                del z_d
    if nz.param["FLAG_REAL"]:
        torch.view_as_real(out)[..., 1] = 0
    return out


def spline_noise(z, tau_noise, t0, dt, n_out, stream_ptr, flag_real=False):
    """Cubic interpolation (not-a-knot ends, as scipy's interp1d(kind="cubic")
    that the reference builds, noise/hops_noise.py:242-243) of complex128
    device noise z [..., n_t] on the grid j * tau_noise, evaluated at
    t0 + j * dt for j < n_out: complex128 [..., n_out]."""
    n_t = z.shape[-1]
    R = z.numel() // n_t
    m = max(n_t - 4, 1)
    cp = np.empty(max(n_t, 1))
    c = 0.25
    for k in range(m):  # Thomas pivots of the constant (1, 4, 1) system
        c = 0.25 if k == 0 else 1.0 / (4.0 - c)
        cp[k] = c
    cp_d = torch.from_numpy(cp).to(z.device)
    work = torch.empty_like(z)
    out = torch.empty(tuple(z.shape[:-1]) + (int(n_out),), dtype=torch.complex128, device=z.device)
    # the spline of each row on a shifted grid: evaluation at t0 + j dt
    check(lib().hops_noise_spline(ctypes.c_void_p(z.data_ptr()), R, n_t, float(tau_noise), float(t0), float(dt),
                                  int(n_out), ctypes.c_void_p(cp_d.data_ptr()), ctypes.c_void_p(work.data_ptr()),
                                  ctypes.c_void_p(out.data_ptr()), stream_ptr))
    torch.cuda.synchronize(z.device)
    if flag_real:
        torch.view_as_real(out)[..., 1] = 0
    return out


def window_mean(z, w, stream_ptr):
    """Moving average over w samples of device noise z (complex64 [..., n_t]):
    the noise of EFFECTIVE_NOISE_INTEGRATION (integrator/integrator_rk.py:115-128)."""
    out = torch.empty_like(z)
    n_t = z.shape[-1]
    R = z.numel() // n_t if n_t else 0
    check(lib().hops_noise_window_mean(ctypes.c_void_p(z.data_ptr()), ctypes.c_void_p(out.data_ptr()), R, n_t,
                                       int(w), stream_ptr))
    return out
