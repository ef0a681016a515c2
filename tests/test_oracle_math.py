"""Pin the oracle's force and list functions with what the mathematics fixes:
two-body closed forms, F = -dU/dr of the Lennard-Jones potential by finite
differences, Newton's third law, half-list/full-list agreement, an all-pairs
brute force without any list, symmetry of the perfect lattice, and the
periodic generator's invariants."""
import numpy as np
import pytest

import oracle as O
from helpers import normwise


def two_body(r_vec, dt=1.0, box=0.0, cl2=O.CL2, origin=5.0):
    q = np.zeros((2, 4))
    q[0, :3] = [origin, origin, origin]
    q[1, :3] = q[0, :3] + np.asarray(r_vec)
    if box > 0:
        q[:, :3] %= box
    csr = O.CSR(np.array([1, 0], np.int32), np.array([0, 1], np.int32),
                np.array([1], np.int32), "half")
    p = np.zeros_like(q)
    O.force_aos(csr, q, p, dt, cl2, box, method="sorted")
    return p


def lj_potential(r):
    return 4.0 * (r ** -12 - r ** -6)


def test_unit_distance_known_answer():
    # df = (24 - 48)/1 * dt = -24 dt; p_i += df * (q_j - q_i)
    p = two_body([1.0, 0.0, 0.0], dt=0.5)
    np.testing.assert_array_equal(p[0, :3], [-12.0, 0.0, 0.0])
    np.testing.assert_array_equal(p[1, :3], [12.0, 0.0, 0.0])


def test_potential_minimum_has_zero_force():
    r0 = 2.0 ** (1.0 / 6.0)
    p = two_body([r0, 0.0, 0.0])
    assert np.abs(p).max() < 1e-13


def test_cutoff_is_inclusive():
    p_in = two_body([3.0, 0.0, 0.0], origin=0.0)  # r2 == 9.0 exactly: kept (r2 > CL2 skips)
    assert p_in[0, 0] != 0.0
    expected = (24.0 * 3.0 ** 6 - 48.0) / 3.0 ** 14
    assert p_in[0, 0] == pytest.approx(expected * 3.0, rel=1e-15)
    p_out = two_body([np.nextafter(3.0, 4.0), 0.0, 0.0], origin=0.0)
    assert np.abs(p_out).max() == 0.0


@pytest.mark.parametrize("r", [0.9, 0.97, 1.0, 1.12, 1.3, 1.7, 2.2, 2.9])
def test_force_is_minus_potential_gradient(r):
    """df/dt * |d| must equal -dU/dr (projected on the pair direction),
    here by a central finite difference of U(r) = 4(r^-12 - r^-6)."""
    direction = np.array([0.48, -0.6, 0.64])
    direction /= np.linalg.norm(direction)
    p = two_body(r * direction)
    h = 1e-6 * r
    dUdr = (lj_potential(r + h) - lj_potential(r - h)) / (2 * h)
    # force on particle 1 (at q0 + r*dir) is -dU/dr along +dir
    f1 = p[1, :3]
    np.testing.assert_allclose(f1, -dUdr * direction, rtol=2e-8, atol=1e-12)


def test_periodic_minimum_image_pair():
    box = 10.0
    p_wrap = two_body([box - 1.1, 0.0, 0.0], box=box)  # nearest image at -1.1
    p_direct = two_body([-1.1, 0.0, 0.0])
    np.testing.assert_allclose(p_wrap[:, :3], p_direct[:, :3], rtol=1e-13)


@pytest.fixture(scope="module")
def small_pbc():
    q, box = O.lattice_pbc(5, 0.5, seed=7)  # N = 500, box = 10 > 2 * 3.3
    half = O.makepair(q, box, "half")
    full = O.makepair(q, box, "full")
    return q, box, half, full


def brute_force_all_pairs(q, box, cl2=O.CL2):
    """F_i = sum_j (24 r^-8 - 48 r^-14) d_ij over every j != i within the
    cutoff, no list at all (numpy, written from U'(r)/r)."""
    x = q[:, :3]
    d = x[None, :, :] - x[:, None, :]
    d -= box * np.rint(d / box)
    r2 = (d * d).sum(-1)
    np.fill_diagonal(r2, np.inf)
    inv_r2 = 1.0 / r2
    coef = 24.0 * inv_r2 ** 4 - 48.0 * inv_r2 ** 7
    coef[r2 > cl2] = 0.0
    return (coef[:, :, None] * d).sum(axis=1)


def test_list_force_equals_all_pairs_brute_force(small_pbc):
    q, box, half, full = small_pbc
    F = brute_force_all_pairs(q, box)
    for csr, method in ((half, "sorted"), (half, "pair"), (full, "full")):
        p = np.zeros_like(q)
        O.force_aos(csr, q, p, 1.0, O.CL2, box, method=method)
        assert normwise(p[:, :3], F) < 1e-13


def test_newton_third_law_and_half_full_agreement(small_pbc):
    q, box, half, full = small_pbc
    ph = np.zeros_like(q)
    pf = np.zeros_like(q)
    O.force_aos(half, q, ph, 0.01, O.CL2, box, steps=3, method="sorted")
    O.force_aos(full, q, pf, 0.01, O.CL2, box, steps=3, method="full")
    scale = np.abs(ph).max()
    assert np.abs(ph[:, :3].sum(axis=0)).max() < 1e-13 * scale * q.shape[0]
    assert normwise(pf, ph) < 1e-13


def test_soa_equals_aos(small_pbc):
    q, box, half, full = small_pbc
    q3 = np.ascontiguousarray(q[:, :3].T)
    for csr, method in ((half, "sorted"), (half, "pair"), (full, "full")):
        p = np.zeros_like(q)
        p3 = np.zeros_like(q3)
        O.force_aos(csr, q, p, 0.01, O.CL2, box, steps=2, method=method)
        O.force_soa(csr, q3, p3, 0.01, O.CL2, box, steps=2, method=method)
        np.testing.assert_array_equal(p[:, :3], p3.T)


def test_perfect_lattice_has_no_net_force():
    q, box = O.lattice_pbc(6, 1.0, seed=1, jitter=(0.0, 0.0))
    full = O.makepair(q, box, "full")
    p = np.zeros_like(q)
    O.force_aos(full, q, p, 1.0, O.CL2, box, method="full")
    # every particle sees a centro-symmetric shell: forces cancel to rounding
    assert np.abs(p).max() < 1e-12


def test_list_is_symmetric_and_matches_predicate(small_pbc):
    q, box, half, full = small_pbc
    i_h, j_h = half.pairs()
    assert np.all(j_h > i_h)
    i_f, j_f = full.pairs()
    fwd = set(zip(i_f.tolist(), j_f.tolist()))
    assert all((j, i) in fwd for i, j in fwd)
    assert set(zip(i_h.tolist(), j_h.tolist())) == {(i, j) for i, j in fwd if i < j}
    # predicate r2 < SL2 on the minimum image, checked independently in numpy
    x = q[:, :3]
    d = x[:, None, :] - x[None, :, :]
    d -= box * np.rint(d / box)
    r2 = (d * d).sum(-1)
    iu, ju = np.nonzero(np.triu(r2 < O.SL2, k=1))
    assert set(zip(iu.tolist(), ju.tolist())) == set(zip(i_h.tolist(), j_h.tolist()))


def test_sortpair_roundtrip(small_pbc):
    q, box, half, _ = small_pbc
    i, j = half.pairs()
    rng = np.random.default_rng(3)
    order = rng.permutation(i.size)
    # the reference's sortpair keeps insertion order within each row
    nop = np.bincount(i, minlength=half.n).astype(np.int32)
    ptr = np.zeros(half.n, np.int32)
    lst = np.zeros(i.size, np.int32)
    O.lib().orc_sortpair(half.n, nop, np.ascontiguousarray(i), np.ascontiguousarray(j),
                         i.size, ptr, lst)
    np.testing.assert_array_equal(ptr, half.pointer)
    np.testing.assert_array_equal(lst, half.sorted_list)
    ii, jj = np.ascontiguousarray(i[order]), np.ascontiguousarray(j[order])
    O.lib().orc_sortpair(half.n, nop, ii, jj, i.size, ptr, lst)
    for r in range(half.n):
        a, b = ptr[r], ptr[r] + nop[r]
        assert sorted(lst[a:b].tolist()) == half.sorted_list[a:b].tolist()


def test_random_shfl_permutes_rows_and_keeps_forces(small_pbc):
    q, box, _, full = small_pbc
    sh = O.random_shfl(full, 10)
    assert not np.array_equal(sh.sorted_list, full.sorted_list)
    for r in range(full.n):
        a, b = full.pointer[r], full.pointer[r] + full.number_of_partners[r]
        assert sorted(sh.sorted_list[a:b].tolist()) == full.sorted_list[a:b].tolist()
    p0 = np.zeros_like(q)
    p1 = np.zeros_like(q)
    O.force_aos(full, q, p0, 1.0, O.CL2, box, method="full")
    O.force_aos(sh, q, p1, 1.0, O.CL2, box, method="full")
    assert normwise(p1, p0) < 1e-14


def test_transposed_layout(small_pbc):
    _, _, _, full = small_pbc
    w = int(full.number_of_partners.max())
    ell = O.make_transposed(full, w + 3)
    n = full.n
    for i in range(0, n, 37):
        row = [ell[i + k * n] for k in range(full.number_of_partners[i])]
        a = full.pointer[i]
        assert row == full.sorted_list[a:a + full.number_of_partners[i]].tolist()
        assert all(ell[i + k * n] == 0 for k in range(full.number_of_partners[i], w + 3))


@pytest.mark.parametrize("k,rho", [(4, 1.0), (7, 0.5), (12, 1.0)])
def test_periodic_lattice_invariants(k, rho):
    q, box = O.lattice_pbc(k, rho, seed=11)
    n = 4 * k ** 3
    assert q.shape == (n, 4)
    assert box == k * (1.0 / (rho * 0.25) ** (1.0 / 3.0))
    assert np.all(q[:, :3] >= 0.0) and np.all(q[:, :3] < box)
    q0, _ = O.lattice_pbc(k, rho, seed=11, jitter=(0.0, 0.0))
    d = q[:, :3] - q0[:, :3]
    d -= box * np.rint(d / box)
    assert np.all(np.abs(d) <= 0.05 + 1e-12)
    assert np.abs(d.mean()) < 0.01
    assert abs(n / box ** 3 - rho) < 1e-12
