"""Parameter dictionaries of the test systems (shared by CPU and GPU tests)."""
import numpy as np
import scipy.sparse as sp

from oracle import hops_oracle as O


def dimer_of_dimers(altered=False):
    """The system of the reference's tests/test_dimer_of_dimers.py:19-86:
    4 sites, 2 modes per site (one L given as sparse, one as dense)."""
    g0, w0 = O.bcf_convert_dl_to_exp(20.0, 50.0, 140.0)
    scale = np.sqrt(2) if altered else 1.0
    gdiv = 2.0 if altered else 1.0
    lop = np.zeros([4, 4, 4])
    gw, ll = [], []
    for i in range(4):
        lop[i, i, i] = scale
        gw.append([g0 / gdiv, w0])
        ll.append(sp.coo_matrix(lop[i]))
        gw.append([-1j * np.imag(g0 / gdiv), 500.0])
        ll.append(lop[i])
    hs = np.zeros([4, 4])
    hs[0, 1] = hs[1, 0] = 40
    hs[1, 2] = hs[2, 1] = 10
    hs[2, 3] = hs[3, 2] = 40
    return dict(HAMILTONIAN=np.array(hs, dtype=np.complex128), GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll,
                ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw)


def chain(nsite, V=50.0, sigma=100.0, seed_h=0, lam=35.0, gam=50.0, temp=295.0, k_mats=1,
          ltc=None):
    """Nearest-neighbour chain with E ~ N(0, sigma) and DL modes per site."""
    rng = np.random.default_rng(seed_h)
    H = np.diag(rng.normal(0.0, sigma, nsite)).astype(np.complex128)
    for i in range(nsite - 1):
        H[i, i + 1] = H[i + 1, i] = V
    modes = O.bcf_convert_dl_to_exp(lam, gam, temp, k_matsubara=k_mats)
    gw, ll = [], []
    for i in range(nsite):
        L = sp.coo_matrix(([1.0], ([i], [i])), shape=(nsite, nsite))
        for j in range(0, len(modes), 2):
            gw.append([modes[j], modes[j + 1]])
            ll.append(L)
    p = dict(HAMILTONIAN=H, GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll, ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw)
    if ltc is not None:
        p["L_LT_CORR"] = [sp.coo_matrix(([1.0], ([i], [i])), shape=(nsite, nsite)) for i in range(nsite)]
        p["PARAM_LT_CORR"] = [ltc] * nsite
    return p


def three_site(general=False):
    """The 3-site, 6-mode systems of tests/test_hops_eom.py:231-316 (diagonal
    L-operators, or generalised ones with complex off-diagonal entries)."""
    lop = np.zeros([3, 3, 3], dtype=np.complex128)
    if general:
        lop[0, 0, 1] = 1j
        lop[0, 1, 0] = -1j
        lop[1, 1, 2] = 1
        lop[1, 2, 1] = 1
        lop[2, 0, 0] = 0.5
        lop[2, 1, 1] = 1
        lop[2, 2, 2] = -1
    else:
        lop[0, 0, 0] = 1.0
        lop[0, 1, 1] = -1.0
        lop[1, 1, 1] = 1.0
        lop[1, 2, 2] = -1.0
        lop[2, 2, 2] = 1.0
        lop[2, 0, 0] = -1.0
    H = np.array([[1.0, 14.0, 8.0], [14.0, 0, 10.0], [8.0, 10.0, -1.0]])
    by_mode = [0, 0, 1, 1, 2, 2]
    gw = [[12.0, 18.0], [7.0, 3.0], [11.0, 9.0], [6.0, 4.0], [10.0, 10.0], [5.0, 5.0]]
    ll = [lop[i] for i in by_mode]
    return dict(HAMILTONIAN=H, GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll, ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw)


def dimer_c1():
    """2-site LINEAR dimer: E = 0, V = 50, one DL mode per site (50, 50, 295 K)."""
    g, w = O.bcf_convert_dl_to_exp(50.0, 50.0, 295.0)
    H = np.array([[0.0, 50.0], [50.0, 0.0]], dtype=np.complex128)
    ll = [sp.coo_matrix(([1.0], ([i], [i])), shape=(2, 2)) for i in range(2)]
    gw = [[g, w], [g, w]]
    return dict(HAMILTONIAN=H, GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll, ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw)


def lattice(nx, ny, V=40.0, sigma=150.0, seed_h=0, lam=35.0, gam=50.0, temp=77.0, k_hier=1, k_ltc=3):
    """nx x ny nearest-neighbour lattice, E ~ N(0, sigma), L_n = |n><n|.
    Per site the Drude-Lorentz correlation function with k_hier + k_ltc
    Matsubara terms (util/bath_corr_functions.py:70-112): the high-temperature
    mode and the first k_hier Matsubara modes go into the hierarchy, the last
    k_ltc into the low-temperature correction as sum g/w, and every mode into
    the noise (the recipe of trajectory/dyadic_spectra.py:808-840)."""
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
    modes = O.bcf_convert_dl_to_exp(lam, gam, temp, k_matsubara=k_hier + k_ltc)
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
    return dict(HAMILTONIAN=H, GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=lln, ALPHA_NOISE1=O.bcf_exp,
                PARAM_NOISE1=gwn, L_LT_CORR=[sp.coo_matrix(([1.0], ([i], [i])), shape=(n, n)) for i in range(n)],
                PARAM_LT_CORR=ltc)


def aggregate(nsite, V=100.0, sigma=100.0, seed_h=0, lam=35.0, gam=50.0, temp=295.0, k_mats=0):
    """A fully coupled aggregate: sites on a line with point-dipole couplings
    J_ij = V / |i - j|^3 between every pair (a dense H), E ~ N(0, sigma), one
    high-temperature DL mode per site (plus k_mats Matsubara modes)."""
    rng = np.random.default_rng(seed_h)
    H = np.diag(rng.normal(0.0, sigma, nsite)).astype(np.complex128)
    for i in range(nsite):
        for j in range(i + 1, nsite):
            H[i, j] = H[j, i] = V / (j - i) ** 3
    modes = O.bcf_convert_dl_to_exp(lam, gam, temp, k_matsubara=k_mats)
    gw, ll = [], []
    for i in range(nsite):
        L = sp.coo_matrix(([1.0], ([i], [i])), shape=(nsite, nsite))
        for j in range(0, len(modes), 2):
            gw.append([modes[j], modes[j + 1]])
            ll.append(L)
    return dict(HAMILTONIAN=H, GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll, ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw)


def peierls_chain(ltc=False):
    """5 sites where H couples only 0-1 and 3-4: state 2 is reached only
    through the off-diagonal (Peierls) L-operators of bonds 1-2 and 2-3, next
    to a site projector (Holstein) on every site; ltc adds a low-temperature
    correction on every L-operator."""
    n = 5
    H = np.diag([0.0, 100.0, -50.0, 80.0, 20.0]).astype(complex)
    H[0, 1] = H[1, 0] = 40.0
    H[3, 4] = H[4, 3] = 40.0
    g0, w0 = O.bcf_convert_dl_to_exp(40.0, 50.0, 295.0)
    gw, ll = [], []
    for i in range(n):
        ll.append(sp.coo_matrix(([1.0], ([i], [i])), shape=(n, n)))
        gw.append([g0, w0])
    for i in (1, 2):
        L = np.zeros((n, n), complex)
        L[i, i + 1] = L[i + 1, i] = 1.0
        ll.append(L)
        gw.append([g0 / 4, w0])
    p = dict(HAMILTONIAN=H, GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll, ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw)
    if ltc:
        p["L_LT_CORR"] = ll
        p["PARAM_LT_CORR"] = [3.0 + 0.5j] * n + [1.0, 1.0]
    return p


def peierls_ring(nsite, V=50.0, sigma=80.0, seed_h=0, lam=40.0, gam=50.0, temp=295.0, ltc=None):
    """A chain of nsite sites with nearest-neighbour H couplings on every
    other bond only, a site projector on every site and an off-diagonal
    (Peierls) L-operator |n><n+1| + |n+1><n| on every bond: the bonds
    without an H coupling are crossed only through the Peierls fluxes and
    noise."""
    rng = np.random.default_rng(seed_h)
    n = nsite
    H = np.diag(rng.normal(0.0, sigma, n)).astype(np.complex128)
    for i in range(0, n - 1, 2):
        H[i, i + 1] = H[i + 1, i] = V
    g0, w0 = O.bcf_convert_dl_to_exp(lam, gam, temp)
    gw, ll = [], []
    for i in range(n):
        ll.append(sp.coo_matrix(([1.0], ([i], [i])), shape=(n, n)))
        gw.append([g0, w0])
    for i in range(n - 1):
        L = sp.coo_matrix(([1.0, 1.0], ([i, i + 1], [i + 1, i])), shape=(n, n))
        ll.append(L)
        gw.append([g0 / 4, w0])
    p = dict(HAMILTONIAN=H, GW_SYSBATH=gw, L_HIER=ll, L_NOISE1=ll, ALPHA_NOISE1=O.bcf_exp, PARAM_NOISE1=gw)
    if ltc is not None:
        p["L_LT_CORR"] = ll
        p["PARAM_LT_CORR"] = [ltc] * len(ll)
    return p
