"""The oracle's nuclear attraction (oracle/giao_oracle.c or_nuclear_matrix,
restating nuclear.cpp:24-312) pinned three ways: against GIAOINT's own
nuclear.cpp compiled into oracle/_ref (water cc-pVDZ in a field, random
shells through (g, g)); against the closed form of an s-s element at zero
field, -Z (2 pi / p) K_AB N_a N_b F_0(p |P - C|^2) with
F_0(x) = sqrt(pi / x) erf(sqrt(x)) / 2; and against Hermiticity of the
matrix, V = V^H, which a wrong phase sign or index breaks."""

import math

import numpy as np
import pytest

from giaoint_amd import workloads
from oracle.pyoracle import make_basis


def _water(port):
    sh, B, O = workloads.water_ccpvdz()
    basis = make_basis(sh.shell_list(), port)
    centers = np.array([s[0] for s in sh.shell_list()])
    _, idx = np.unique(centers.round(12), axis=0, return_index=True)
    nuc = centers[np.sort(idx)]
    charges = np.array([8.0 if abs(c[1]) < 1e-9 else 1.0 for c in nuc])
    return basis, charges, nuc, port.field6(B, O)


def test_nuclear_matches_reference_build(port, reference):
    basis, z, c, f = _water(port)
    got = port.nuclear_matrix(basis, z, c, f)
    want = reference.nuclear_matrix(basis, z, c, f)
    err = np.max(np.abs(got - want))
    print(f"water cc-pVDZ B=(0,0,0.1): max |oracle - reference build| = {err:.2e}")
    assert err <= 1e-13


def test_nuclear_random_shells_match_reference_build(port, reference, rng):
    from tests.conftest import random_shells

    shells = random_shells(rng, [0, 1, 2, 3, 4, 2], nprim=(1, 3), ncontr=(1, 3))
    basis = make_basis(shells, port)
    z = np.array([3.0, 1.0])
    c = rng.uniform(-1.5, 1.5, (2, 3))
    f = port.field6((0.2, -0.3, 0.4), (0.1, 0.0, -0.2))
    got = port.nuclear_matrix(basis, z, c, f)
    want = reference.nuclear_matrix(basis, z, c, f)
    assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))


def test_nuclear_hermitian_in_field(port):
    basis, z, c, f = _water(port)
    V = port.nuclear_matrix(basis, z, c, f)
    assert np.max(np.abs(V - V.conj().T)) <= 1e-13
    assert np.max(np.abs(V.imag)) > 1e-6  # the field makes it complex


def test_nuclear_ss_closed_form(port):
    a, b = 0.8, 1.7
    A, B, C = np.array([0.0, 0.1, 0.2]), np.array([0.5, -0.3, 0.9]), np.array([-0.4, 0.2, 0.3])
    basis = make_basis([(A, 0, np.array([a]), np.array([[1.0]])), (B, 0, np.array([b]), np.array([[1.0]]))],
                       port, normalize=False)
    V = port.nuclear_matrix(basis, [2.0], C[None, :], port.field6())
    p = a + b
    P = (a * A + b * B) / p
    K = math.exp(-a * b / p * np.sum((A - B) ** 2))
    x = p * np.sum((P - C) ** 2)
    F0 = 0.5 * math.sqrt(math.pi / x) * math.erf(math.sqrt(x))
    na, nb = (2 * a / math.pi) ** 0.75, (2 * b / math.pi) ** 0.75
    want = -2.0 * (2 * math.pi / p) * K * na * nb * F0
    assert abs(V[0, 1] - want) <= 1e-14 * abs(want)
    assert abs(V[0, 1].imag) == 0.0
