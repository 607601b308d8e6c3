"""Parameter dictionaries of the benchmark workloads (synthetic, seeded).

chain_system: nearest-neighbour chain, E_n ~ N(0, sigma) from
np.random.default_rng(seed_h), Drude-Lorentz modes per site from
bcf_convert_dl_to_exp (util/bath_corr_functions.py:70-112), L_n = |n><n|,
optional low-temperature correction on every site.
aggregate_system: every pair of sites coupled (point dipoles), a dense H.
lattice_system: nx x ny nearest-neighbour lattice with the low-temperature
correction: per site the DL correlation function with k_hier + k_ltc Matsubara
terms, the first modes in the hierarchy, sum g/w of the last k_ltc in the
correction, every mode in the noise (trajectory/dyadic_spectra.py:808-840).
c4_requests: the mixed-size serving stream — Poisson arrivals, chain length,
depth, delta, t_max drawn per request from one seeded generator; the noise
seed of a request is its id.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from .params import bcf_convert_dl_to_exp, bcf_exp


def chain_system(nsite, V=50.0, sigma=100.0, seed_h=0, lam=35.0, gam=50.0, temp=295.0, k_mats=1, ltc=None):
    rng = np.random.default_rng(seed_h)
    H = np.diag(rng.normal(0.0, sigma, nsite)).astype(np.complex128)
    for i in range(nsite - 1):
        H[i, i + 1] = H[i + 1, i] = V
    modes = bcf_convert_dl_to_exp(lam, gam, temp, k_matsubara=k_mats)
    gw, ll = [], []
    for i in range(nsite):
        L = sp.coo_matrix(([1.0], ([i], [i])), shape=(nsite, nsite))
        for j in range(0, len(modes), 2):
            gw.append([modes[j], modes[j + 1]])
            ll.append(L)
    p = dict(HAMILTONIAN=H, GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll, ALPHA_NOISE1=bcf_exp, PARAM_NOISE1=gw)
    if ltc is not None:
        p["L_LT_CORR"] = [sp.coo_matrix(([1.0], ([i], [i])), shape=(nsite, nsite)) for i in range(nsite)]
        p["PARAM_LT_CORR"] = [ltc] * nsite
    return p


def aggregate_system(nsite, V=100.0, sigma=100.0, seed_h=0, lam=35.0, gam=50.0, temp=295.0, k_mats=0):
    """A fully coupled aggregate: sites on a line with point-dipole couplings
    J_ij = V / |i - j|^3 between every pair, so H has no zero entry (the
    dense system block), E_n ~ N(0, sigma); Drude-Lorentz modes per site as
    chain_system."""
    rng = np.random.default_rng(seed_h)
    H = np.diag(rng.normal(0.0, sigma, nsite)).astype(np.complex128)
    for i in range(nsite):
        for j in range(i + 1, nsite):
            H[i, j] = H[j, i] = V / (j - i) ** 3
    modes = bcf_convert_dl_to_exp(lam, gam, temp, k_matsubara=k_mats)
    gw, ll = [], []
    for i in range(nsite):
        L = sp.coo_matrix(([1.0], ([i], [i])), shape=(nsite, nsite))
        for j in range(0, len(modes), 2):
            gw.append([modes[j], modes[j + 1]])
            ll.append(L)
    return dict(HAMILTONIAN=H, GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll, ALPHA_NOISE1=bcf_exp, PARAM_NOISE1=gw)


def lattice_system(nx, ny, V=40.0, sigma=150.0, seed_h=0, lam=35.0, gam=50.0, temp=77.0, k_hier=1, k_ltc=3):
    n = nx * ny
    rng = np.random.default_rng(seed_h)
    H = np.diag(rng.normal(0.0, sigma, n)).astype(np.complex128)
    for x in range(nx):
        for y in range(ny):
            i = x * ny + y
            if x + 1 < nx:
                H[i, i + ny] = H[i + ny, i] = V
            if y + 1 < ny:
                H[i, i + 1] = H[i + 1, i] = V
    modes = bcf_convert_dl_to_exp(lam, gam, temp, k_matsubara=k_hier + k_ltc)
    pairs = [(modes[j], modes[j + 1]) for j in range(0, len(modes), 2)]
    gw, ll, gwn, lln, ltc = [], [], [], [], []
    for i in range(n):
        L = sp.coo_matrix(([1.0], ([i], [i])), shape=(n, n))
        c = 0.0
        for j, (g, w) in enumerate(pairs):
            gwn.append([g, w])
            lln.append(L)
            if len(pairs) - j > k_ltc:
                gw.append([g, w])
                ll.append(L)
            else:
                c += g / w
        ltc.append(c)
    return dict(HAMILTONIAN=H, GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=lln, ALPHA_NOISE1=bcf_exp, PARAM_NOISE1=gwn,
                L_LT_CORR=[sp.coo_matrix(([1.0], ([i], [i])), shape=(n, n)) for i in range(n)], PARAM_LT_CORR=ltc)


@dataclass
class Request:
    """One adaptive trajectory request of the serving stream."""
    rid: int
    arrival: float      # seconds after the stream starts
    nsite: int
    maxhier: int
    delta: float
    t_max: float        # fs
    dt: float = 1.0
    tau_noise: float = 0.5

    @property
    def tlen(self):
        return float(256 * math.ceil((self.t_max + 1.0) / 256.0))

    @property
    def n_steps(self):
        return int(math.ceil(self.t_max / self.dt))


def c4_requests(n, rate=200.0, seed=0, nsite=(2, 200), depth=(2, 10), delta=(1e-4, 1e-2), t_max=(200.0, 2000.0)):
    """n requests with Poisson arrivals at ``rate`` per second; chain length
    and depth uniform integers, delta and t_max uniform reals."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate)
        out.append(Request(rid=rid, arrival=t, nsite=int(rng.integers(nsite[0], nsite[1] + 1)),
                           maxhier=int(rng.integers(depth[0], depth[1] + 1)),
                           delta=float(rng.uniform(*delta)), t_max=float(rng.uniform(*t_max))))
    return out
