"""Parameter dictionaries, physical constants and bath correlation helpers.

Defaults and type tables mirror the reference so that the same user dicts are
accepted unchanged:
  EOM_DICT_*          eom/hops_eom.py:20-41
  HIERARCHY_DICT_*    basis/hops_hierarchy.py:16-22
  NOISE_DICT_*        noise/hops_noise.py:20-44
  INTEGRATION_DICT_*  trajectory/hops_trajectory.py:34-50
  initialize_dict     util/dynamic_dict.py:46-78
"""
import copy

import numpy as np

from .errors import UnsupportedRequest

# util/physical_constants.py:21-28
h = 33357.179  # cm^-1 fs
hbar = h / (2 * np.pi)  # cm^-1 fs / rad
kB = 0.69503  # cm^-1 / K
c = 299.79  # nm / fs
precision = 1e-8

EOM_DICT_DEFAULT = {
    "TIME_DEPENDENCE": False,
    "EQUATION_OF_MOTION": "NORMALIZED NONLINEAR",
    "ADAPTIVE_H": False,
    "ADAPTIVE_S": False,
    "DELTA_A": 0,
    "DELTA_S": 0,
    "UPDATE_STEP": None,
    "F_DISCARD": 0,
}
EOM_DICT_TYPES = {
    "TIME_DEPENDENCE": [bool],
    "EQUATION_OF_MOTION": [str],
    "ADAPTIVE": [bool],
    "ADAPTIVE_H": [bool],
    "ADAPTIVE_S": [bool],
    "DELTA_A": [float, int],
    "DELTA_S": [float, int],
    "UPDATE_STEP": [float, bool, None, int],
    "F_DISCARD": [float],
}

HIERARCHY_DICT_DEFAULT = {"MAXHIER": 3, "TERMINATOR": False, "STATIC_FILTERS": []}
HIERARCHY_DICT_TYPES = {"MAXHIER": [int], "TERMINATOR": [bool, str], "STATIC_FILTERS": [list]}

NOISE_DICT_DEFAULT = {
    "SEED": None,
    "MODEL": "FFT_FILTER",
    "TLEN": 1000.0,
    "TAU": 1.0,
    "INTERPOLATE": False,
    "RAND_MODEL": "SUM_GAUSSIAN",
    "STORE_RAW_NOISE": False,
    "NOISE_WINDOW": None,
    "ADAPTIVE": False,
    "FLAG_REAL": False,
}
NOISE_TYPE_DEFAULT = {
    "SEED": [int, type(None), str, np.ndarray],
    "MODEL": [str],
    "TLEN": [float],
    "TAU": [float],
    "INTERPOLATE": [bool],
    "RAND_MODEL": [str],
    "STORE_RAW_NOISE": [bool],
    "NOISE_WINDOW": [type(None), float, int],
    "ADAPTIVE": [bool],
    "FLAG_REAL": [bool],
}

INTEGRATION_DICT_DEFAULT = {
    "INTEGRATOR": "RUNGE_KUTTA",
    "EARLY_ADAPTIVE_INTEGRATOR": "INCH_WORM",
    "EARLY_INTEGRATOR_STEPS": 5,
    "INCHWORM_CAP": 5,
    "STATIC_BASIS": None,
    "EFFECTIVE_NOISE_INTEGRATION": False,
}
INTEGRATION_DICT_TYPES = {
    "INTEGRATOR": [str],
    "EARLY_ADAPTIVE_INTEGRATOR": [str],
    "EARLY_INTEGRATOR_STEPS": [int],
    "INCHWORM_CAP": [int],
    "STATIC_BASIS": [type(None), list, np.ndarray],
    "EFFECTIVE_NOISE_INTEGRATION": [bool],
}


def initialize_dict(param, defaults, types, name):
    """Validate a user dict against defaults and allowed types, filling the
    missing keys with deep-copied defaults (util/dynamic_dict.py:46-78):
    an unknown key raises UnsupportedRequest, a wrong type raises TypeError."""
    out = copy.deepcopy(param)
    for key in out:
        if key not in defaults:
            raise UnsupportedRequest(key, name)
    for key, dflt in defaults.items():
        if key in out:
            allowed = types[key]
            if type(out[key]) not in allowed:
                raise TypeError("{}-->{} is of type {} not the supported types {}".format(
                    key, out[key], type(out[key]), allowed))
        else:
            out[key] = copy.deepcopy(dflt)
    return out


def bcf_exp(t_axis, g, w):
    """Exponential correlation function g exp(-w t / hbar) (trajectory/exp_noise.py:5-25)."""
    return g * np.exp(-w * t_axis / hbar)


def bcf_convert_dl_to_exp(lambda_dl, gamma_dl, temp, k_matsubara=0):
    """Drude-Lorentz spectral density as [g0, w0, g1, w1, ...]: the
    high-temperature mode corrected by k_matsubara Matsubara terms, each
    g_k = 2i J(i nu_k)/beta, w_k = nu_k = 2 pi k / beta
    (util/bath_corr_functions.py:70-112)."""
    beta = 1.0 / (kB * temp)
    g0 = 2 * lambda_dl / beta - 1j * lambda_dl * gamma_dl
    out = []
    for k in range(1, k_matsubara + 1):
        nu = k * 2 * np.pi / beta
        jw = 2 * lambda_dl * gamma_dl * (1j * nu) / ((1j * nu) ** 2 + gamma_dl ** 2)
        out += [2j * jw / beta, nu]
        g0 += 2 * lambda_dl * (1 / beta) * (2 * gamma_dl ** 2) / (gamma_dl ** 2 - nu ** 2)
    return [g0, gamma_dl] + out


def ishizaki_decomposition_bcf_dl(lambda_dl, gamma_dl, temp, k_matsubara, epsilon=None):
    """Drude-Lorentz-like correlation function as exponential modes
    [g0, w0, g0', w0', g1, w1, ...] after Ishizaki, J. Phys. Soc. Jpn. 89,
    015001 (2020): a complex-conjugate pair of underdamped-like modes at
    w = 2 gamma +- i epsilon (epsilon defaults to gamma / 10, and must be much
    smaller than gamma), then k_matsubara Matsubara modes g_k = 2i J(i nu_k) /
    beta, w_k = nu_k, each also moving the real part of the first pair's
    prefactor (util/bath_corr_functions.py:115-182)."""
    eps = gamma_dl / 10 if epsilon is None else epsilon
    big = 2 * gamma_dl
    w_up, w_dn = big + 1j * eps, big - 1j * eps
    beta = 1.0 / (kB * temp)
    nu1 = 2 * np.pi / beta

    def spectral(w):
        return 4 * lambda_dl * big ** 3 * w / ((w ** 2 + big ** 2) ** 2)

    g_re = 1j * lambda_dl * w_dn / (eps * beta)
    g_im = 1j * lambda_dl * w_up * w_dn / (2 * eps)
    mats = []
    for k in range(1, int(k_matsubara) + 1):
        nu = k * nu1
        mats += [(2j / beta) * spectral(1j * nu), nu]
        e_tilde = (2 * w_up ** 2) / ((w_up ** 2 - nu ** 2) * beta)
        g_re += 2 * np.real((1j * lambda_dl * w_dn / eps) * e_tilde)
    return [g_re - 1j * g_im, w_up, np.conj(g_re) - 1j * np.conj(g_im), np.conj(w_up)] + mats
