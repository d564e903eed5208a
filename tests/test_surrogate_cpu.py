"""CPU checks of the surrogate's host-side helpers (eu3d-gpu_amd/surrogate.py):
element bookkeeping, the conservation projector, mixtures, features."""
import importlib
import os

import numpy as np

from eu3d_load import eu3d

sur = importlib.import_module("eu3d-gpu_amd.surrogate")
NAMES = sur.species_names(os.path.join(eu3d.INPUT, "H2_JA919.txt"))
# molar masses of the species (g/mol), from the atomic masses
MW = np.array([sum(sur.ATOMIC_MASS[e] * k for e, k in sur.parse_formula(n).items()) for n in NAMES])


def test_species_and_formulae():
    assert NAMES == ["H", "O", "H2O", "OH", "O2", "H2", "H2O2", "HO2", "N2"]
    assert sur.parse_formula("H2O2") == {"H": 2, "O": 2}
    assert sur.parse_formula("HO2") == {"H": 1, "O": 2}


def test_element_matrix_and_projector():
    A, elems = sur.element_matrix(NAMES)
    assert elems == ["H", "N", "O"]
    assert np.allclose(A.sum(0), 1.0, atol=1e-15)          # each species is all elements
    P = sur.projector(A)
    assert np.allclose(A @ P, 0.0, atol=1e-14)              # projected changes conserve every element
    assert np.allclose(P @ P, P, atol=1e-14) and np.allclose(P, P.T, atol=1e-15)
    # a reaction's mass change, e.g. 2 H2 + O2 -> 2 H2O, is already conservative
    d = np.zeros(len(NAMES))
    i = {n: k for k, n in enumerate(NAMES)}
    d[i["H2"]], d[i["O2"]], d[i["H2O"]] = -2 * MW[i["H2"]], -MW[i["O2"]], 2 * MW[i["H2O"]]
    assert np.allclose(P @ d, d, atol=1e-12)
    assert np.allclose(np.ones(len(NAMES)) @ P, 0.0, atol=1e-14)   # mass is conserved too


def test_mixture_and_products_share_elements():
    A, _ = sur.element_matrix(NAMES)
    for phi in (0.5, 1.0, 2.0):
        u, b = sur.mixture(phi, NAMES, MW), sur.burnt(phi, NAMES, MW)
        assert np.allclose(u.sum(1), 1) and np.allclose(b.sum(1), 1) and np.all(b >= 0)
        assert np.allclose(u @ A.T, b @ A.T, atol=1e-15)
    i = {n: k for k, n in enumerate(NAMES)}
    b = sur.burnt(1.0, NAMES, MW)[0]
    assert b[i["H2"]] == 0 and b[i["O2"]] == 0                # stoichiometric: water and nitrogen only


def test_sample_states_and_features():
    T, p, Y = sur.sample_states(1000, NAMES, MW, 3, T_range=(1200.0, 1300.0), p_atm=(1.0, 2.0))
    assert np.all((T >= 1200) & (T <= 1300)) and np.all((p >= 101325.0) & (p <= 2 * 101325.0))
    assert np.allclose(Y.sum(1), 1.0) and np.all(Y >= 0)
    Y2 = Y.copy()
    Y2[0, 0] = -1e-9                                          # FP64 results can be slightly negative
    f = sur.features(T, p, Y2, 3e-8)
    assert f.shape == (1000, len(NAMES) + 3) and f.dtype == np.float32
    assert np.all(np.isfinite(f))
    assert f[0, 2] == np.float32(np.log(1e-12))
    assert np.allclose(f[:, -1], np.log(3e-8))
