"""Extract the reference's golden test data into tests/golden/ref_fixtures.npz.

Reads the .npy fixtures that MesoHOPS ships with its tests (data only, loaded
with allow_pickle=False) from the reference checkout and stores them in a
compact form: the super-operators as COO triplets, eta on every 5th time
point.  Run from the repository root:  python tests/golden/make_fixtures.py
"""
import os

import numpy as np

REF = "/root/reference/tests"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_fixtures.npz")


def load(rel):
    return np.load(os.path.join(REF, rel), allow_pickle=False)


def coo(a):
    r, c = np.nonzero(a)
    return np.stack([r, c]).astype(np.int32), a[r, c]


out = {}
for name in ("K2_k", "K2_kp1", "K2_km1"):
    idx, val = coo(load("dimer_of_dimers/%s.npy" % name))
    out[name + "_idx"], out[name + "_val"], out[name + "_shape"] = idx, val, np.array(load("dimer_of_dimers/%s.npy" % name).shape)
z = load("dimer_of_dimers/Z2_kp1.npy")
for i in range(z.shape[0]):
    idx, val = coo(z[i])
    out["Z2_%d_idx" % i], out["Z2_%d_val" % i] = idx, val
out["Z2_shape"] = np.array(z.shape)
alpha = load("dimer_of_dimers/alpha.npy")
out["alpha_stride50"] = alpha[:, ::50]
eta = load("dimer_of_dimers/eta.npy")
out["eta_stride5"] = eta[::5]
out["traj_dimer_of_dimers"] = load("dimer_of_dimers/traj_dimer_of_dimers.npy")
out["pre_calculated_uncorrelated_noise_test"] = load("pre_calculated_uncorrelated_noise_test.npy")
out["pre_calculated_correlated_noise_test"] = load("pre_calculated_correlated_noise_test.npy")
out["known_param_g"] = load("known_param_g.npy")
out["known_param_w"] = load("known_param_w.npy")
# raw (uncorrelated) noise rows that the reference's basis tests give as SEED
out["hops_basis_noise_2site"] = load("hops_basis_noise_2site.npy")
out["hops_basis_noise_10site"] = load("hops_basis_noise_10site.npy")
np.savez_compressed(OUT, **out)
print(OUT, os.path.getsize(OUT))
