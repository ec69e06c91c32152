"""The C++ host layer (include/giao_hip.hpp) on the GPU against the oracle:
compute_eri, evaluate_eri_shell_quartets (input order, statistics, Schwarz
screening), both compute_eri_tensor overloads, compute_boys, primitive_eri,
detail::cholesky_diagonal/columns, and the exception type and message of
each error path (eri.cpp:454-462,644-688; boys.cpp:245-256;
cholesky.cpp:203-216)."""

import itertools

import numpy as np
import pytest

from oracle.pyoracle import make_basis
from tests import cpp_layer
from tests.conftest import random_shells

pytestmark = pytest.mark.gpu

ATOL = 1e-11
B, O = (0.1, -0.3, 0.25), (0.2, 0.1, -0.1)


def test_cpp_layer_against_oracle(tmp_path, port):
    rng = np.random.default_rng(5)
    shells = random_shells(rng, [1, 0, 2, 1])
    with open(tmp_path / "shells.txt", "w") as fh:
        for c, l, e, k in shells:
            fh.write(" ".join(map(repr, [*map(float, c), l, len(e), k.shape[0], *map(float, e),
                                        *map(float, k.reshape(-1))])) + "\n")
    exe = cpp_layer.build(tmp_path)
    r = cpp_layer.run(exe, "gpu", tmp_path)
    b = make_basis(shells, port)
    f = port.field6(B, O)
    n = b.nao
    c128 = lambda name: np.fromfile(tmp_path / name, dtype=np.complex128)  # noqa: E731
    want, _ = port.eri_batch(b, [(0, 1, 2, 3)], f, 4)
    assert np.max(np.abs(c128("block.bin") - want)) <= ATOL
    allq = list(itertools.product(range(4), repeat=4))
    want, _ = port.eri_batch(b, allq, f, 4)
    assert np.max(np.abs(c128("packed.bin") - want)) <= ATOL
    assert r["stats"] == f"{len(allq)} {len(allq)} 0"
    Q = np.fromfile(tmp_path / "bounds.bin").reshape(4, 4)
    np.testing.assert_allclose(Q, port.schwarz_bounds(b, f), rtol=1e-12)
    thr, req, comp, scr, nvals = r["screened"].split()
    kept = [q for q in allq if Q[q[0], q[1]] * Q[q[2], q[3]] >= float(thr)]
    assert (int(req), int(comp), int(scr)) == (len(allq), len(kept), len(allq) - len(kept))
    assert int(nvals) == sum(int(np.prod([b.ao_counts[i] for i in q])) for q in kept)
    ref = port.eri_tensor(b, f, 4).reshape(-1)
    assert np.max(np.abs(c128("tensor.bin") - ref)) <= ATOL
    assert np.max(np.abs(c128("tensor2.bin") - ref)) <= ATOL
    assert r["tensor_stats"] == f"{(4**4 + 3 * 4**2) // 4} {(4**4 + 3 * 4**2) // 4}"
    boys_ref, info = port.boys(20 - 35j, 8)
    assert np.all(np.abs(c128("boys.bin") - boys_ref) <= 3e-14 + 8e-13 * np.abs(boys_ref))
    assert int(r["boys_region"]) == info["region"]
    prims = [{"exponent": 0.8, "center": (0.0, -0.2, 0.0), "angular": (0, 0, 1)},
             {"exponent": 0.9, "center": (0.1, -0.2, 0.3), "angular": (1, 0, 1)},
             {"exponent": 1.0, "center": (0.2, -0.2, 0.6), "angular": (0, 0, 1)},
             {"exponent": 1.1, "center": (0.3, -0.2, 0.9), "angular": (1, 0, 1)}]
    p = complex(*map(float, r["primitive"].split()))
    assert abs(p - port.primitive_eri(prims, f)) <= ATOL
    pairs = [(a, c) for a in range(4) for c in range(a, 4)]
    d = np.fromfile(tmp_path / "diag.bin")
    assert np.max(np.abs(d - port.cholesky_diagonal(b, pairs, f, 4).reshape(-1))) <= ATOL
    col = c128("column.bin")
    assert np.max(np.abs(col - port.cholesky_columns(b, [(1, 0)], f, 4).reshape(-1))) <= ATOL
    # 1000 warm compute_eri calls on changing quartets: no device allocation
    assert int(r["warm_allocations"]) == 0
    q = tuple(int(x) for x in r["warm_last"].split())
    want, _ = port.eri_batch(b, [q], f, 4)
    assert np.max(np.abs(c128("warm_last.bin") - want)) <= ATOL
    print(f"compute_eri warm: {float(r['warm_us_per_call']):.1f} us per call", flush=True)
    # centre derivatives (derivatives.cpp:952-959,1307-1415)
    want, _ = port.eri_derivative_blocks(b, [(0, 1, 2, 3)], f, 4)
    scale = float(np.max(np.abs(want)))
    assert np.max(np.abs(c128("deriv_block.bin") - want)) <= 1e-11 * scale
    want, _ = port.eri_derivative_blocks(b, [(0, 1, 2, 3), (3, 3, 1, 0), (2, 2, 2, 2)], f, 4)
    assert np.max(np.abs(c128("deriv_packed.bin") - want)) <= 1e-11 * float(np.max(np.abs(want)))
    D = np.fromfile(tmp_path / "deriv_bounds.bin").reshape(4, 4)
    np.testing.assert_allclose(D, port.eri_derivative_pair_bounds(b, f), rtol=1e-12)
    qb = max(D[0, 1] * Q[2, 3], Q[0, 1] * D[2, 3])
    assert float(r["deriv_quartet_bound"]) == np.nextafter(qb * (1 + 128 * np.finfo(float).eps), np.inf)
    assert r["deriv_bad_offsets"] == "length_error derivative batch offsets do not match block sizes"
    assert r["deriv_bad_start"] == "length_error derivative batch offsets have an incorrect size or start"
    assert r["deriv_bad_index"] == "out_of_range derivative shell quartet index is out of range"
    assert r["bad_output_size"] == "length_error ERI output has an incorrect size"
    assert r["bad_index"] == "out_of_range ERI shell-quartet index is out of range"
    assert r["null_consumer"] == "invalid_argument ERI quartet consumer must not be null"
    assert r["screening_without_bounds"] == "invalid_argument positive ERI screening requires Schwarz bounds"
    assert r["bounds_other_field"] == "invalid_argument ERI Schwarz bounds do not match the basis"
    assert r["zero_threads"] == "invalid_argument ERI thread count must be positive"
    assert r["boys_outside_domain"].startswith("BoysNumericalError")
    assert r["boys_order"] == "invalid_argument complex Boys supports orders through 32"
    assert r["real_columns_in_field"] == "invalid_argument real Cholesky columns require zero phase field"
    assert n == sum(b.ao_counts)
