"""Pin the oracle against the reference's own outputs.

- ref_data/density{0.5,1}.dat (the reference's golden stdout, %.10f) must be
  reproduced byte for byte by oracle lattice -> makepair -> force_pair.
- The .npz fixtures (tests/golden/make_golden.py, written by the reference's
  cpu_ref code compiled in place) pin positions, the CSR of makepair/sortpair
  (sha256) and 17-digit momenta of every reference CPU force routine.
"""
import hashlib

import numpy as np
import pytest

import oracle as O
from helpers import REF_DATA_05, REF_DATA_1, load_golden, normwise


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _ref_pipeline(density):
    q = O.lattice_reference(density)
    csr = O.makepair(q)
    p = np.zeros_like(q)
    O.force_aos(csr, q, p, O.REF_DT, steps=100, method="pair")
    return q, csr, p


def test_ref_data_density05_byte_exact():
    _, csr, p = _ref_pipeline(0.5)
    assert csr.n == 62500 and csr.n_entries == 2268138
    assert O.format_reference_output(p) == open(REF_DATA_05).read()


@pytest.mark.slow
def test_ref_data_density1_byte_exact():
    _, csr, p = _ref_pipeline(1.0)
    assert csr.n == 119164 and csr.n_entries == 7839886
    assert O.format_reference_output(p) == open(REF_DATA_1).read()


@pytest.mark.parametrize("method", ["pair", "sorted"])
def test_force_sorted_matches_ref_data(method):
    q = O.lattice_reference(0.5)
    csr = O.makepair(q)
    p = np.zeros_like(q)
    O.force_aos(csr, q, p, O.REF_DT, steps=100, method=method)
    got = np.loadtxt(O.format_reference_output(p).splitlines())
    ref = np.loadtxt(REF_DATA_05)
    assert np.abs(got - ref).max() <= 0.51e-10


def test_fixture_L50_positions_and_list():
    g = load_golden("ref_aos_L50_rho0.5")
    q = O.lattice_reference(0.5, 50.0)
    assert q.shape[0] == int(g["n"])
    assert sha(q[:, :3]) == str(g["q_sha256"])
    np.testing.assert_array_equal(q[g["rows"], :3], g["q_rows"])
    csr = O.makepair(q)
    assert csr.n_entries == int(g["npairs"])
    assert sha(csr.number_of_partners) == str(g["nop_sha256"])
    assert sha(csr.pointer) == str(g["pointer_sha256"])
    assert sha(csr.sorted_list) == str(g["sorted_list_sha256"])


@pytest.mark.parametrize("kind", ["pair", "sorted", "intrin", "next", "mat_transpose",
                                  "sorted_z_intrin"])
def test_fixture_L50_momenta(kind):
    """Every reference AoS routine (incl. AVX2 ones) agrees with the oracle's
    force_pair / force_sorted to 1e-13 normwise (they differ by FMA
    contraction and summation order only)."""
    g = load_golden("ref_aos_L50_rho0.5")
    q = O.lattice_reference(0.5, 50.0)
    csr = O.makepair(q)
    p = np.zeros_like(q)
    O.force_aos(csr, q, p, O.REF_DT, steps=100, method="pair" if kind == "pair" else "sorted")
    rows = g["rows"]
    scale = float(g[f"p_{kind}_absmax"])
    assert np.abs(p[rows, :3] - g[f"p_{kind}_rows"]).max() / scale < 1e-13
    np.testing.assert_allclose(p[:, :3].sum(axis=0), g[f"p_{kind}_sum"], atol=1e-11)
    np.testing.assert_allclose((p[:, :3] ** 2).sum(), float(g[f"p_{kind}_sumsq"]), rtol=1e-12)


def test_fixture_L12_full_arrays():
    g = load_golden("ref_aos_L12_rho0.5")
    q = O.lattice_reference(0.5, 12.0)
    np.testing.assert_array_equal(q[:, :3], g["q_all"])
    csr = O.makepair(q)
    np.testing.assert_array_equal(csr.number_of_partners, g["number_of_partners"])
    np.testing.assert_array_equal(csr.pointer, g["pointer"])
    np.testing.assert_array_equal(csr.sorted_list, g["sorted_list"])
    for kind in ("pair", "sorted"):
        p = np.zeros_like(q)
        O.force_aos(csr, q, p, O.REF_DT, steps=100, method=kind)
        assert normwise(p[:, :3], g[f"p_{kind}_all"]) < 1e-13


@pytest.mark.parametrize("kind", ["pair", "sorted", "intrin"])
def test_fixture_soa(kind):
    g = load_golden("ref_soa_L50_rho0.5")
    q = O.lattice_reference(0.5, 50.0)
    assert sha(q[:, :3]) == str(g["q_sha256"])
    csr = O.makepair(q)
    assert sha(csr.sorted_list) == str(g["sorted_list_sha256"])
    q3 = np.ascontiguousarray(q[:, :3].T)
    p3 = np.zeros_like(q3)
    O.force_soa(csr, q3, p3, O.REF_DT, steps=100, method="pair" if kind == "pair" else "sorted")
    rows = g["rows"]
    scale = float(g[f"p_{kind}_absmax"])
    assert np.abs(p3.T[rows] - g[f"p_{kind}_rows"]).max() / scale < 1e-13


@pytest.mark.slow
def test_fixture_open_rho1():
    g = load_golden("ref_aos_open_rho1")
    q = O.lattice_reference(1.0, 50.0)
    assert sha(q[:, :3]) == str(g["q_sha256"])
    csr = O.makepair(q)
    assert csr.n_entries == int(g["npairs"]) == 7839886
    assert sha(csr.sorted_list) == str(g["sorted_list_sha256"])
    assert int(g["nop_min"]) == csr.number_of_partners.min()
    p = np.zeros_like(q)
    O.force_aos(csr, q, p, O.REF_DT, steps=100, method="sorted")
    rows = g["rows"]
    assert np.abs(p[rows, :3] - g["p_sorted_rows"]).max() / float(g["p_sorted_absmax"]) < 1e-13
