"""The product's host-side setup (libljgpu.so, no device needed) against the
oracle: lattices bit for bit, the cell-list Verlet build against the brute
force O(N^2) makepair (exact pair set, exact CSR), validation and the list
utilities."""
import numpy as np
import pytest

import lj_gpu_amd as G
import oracle as O


def same_csr(a, b):
    np.testing.assert_array_equal(a.number_of_partners, b.number_of_partners)
    np.testing.assert_array_equal(a.pointer, b.pointer)
    np.testing.assert_array_equal(a.sorted_list, b.sorted_list)


@pytest.mark.parametrize("density,box", [(0.5, 50.0), (1.0, 50.0), (0.5, 12.0), (0.8, 20.0)])
def test_lattice_reference_matches_oracle(density, box):
    np.testing.assert_array_equal(G.lattice_reference(density, box),
                                  O.lattice_reference(density, box))


@pytest.mark.parametrize("k,rho,seed", [(12, 1.0, 2), (6, 0.5, 1000), (31, 1.0, 2), (9, 0.73, 5)])
def test_lattice_periodic_matches_oracle(k, rho, seed):
    q, box = G.lattice_periodic(k, rho, seed)
    qo, boxo = O.lattice_pbc(k, rho, seed)
    assert box == boxo
    np.testing.assert_array_equal(q, qo)


def test_makepair_reference_mode_matches_oracle():
    q = G.lattice_reference(0.5, 50.0)
    for kind, okind in ((G.LIST_HALF, "half"), (G.LIST_FULL, "full")):
        same_csr(G.makepair(q, 0.0, kind), O.makepair(q, 0.0, okind))


@pytest.mark.parametrize("k,rho", [(12, 1.0), (6, 1.0), (6, 0.5), (8, 0.73)])
def test_makepair_periodic_matches_oracle(k, rho):
    q, box = G.lattice_periodic(k, rho, 3)
    for kind, okind in ((G.LIST_HALF, "half"), (G.LIST_FULL, "full")):
        same_csr(G.makepair(q, box, kind), O.makepair(q, box, okind))


def test_makepair_config1_pair_count():
    q, box = G.lattice_periodic(12, 1.0, 2)
    half = G.makepair(q, box, G.LIST_HALF)
    assert half.n == 6912 and half.n_entries == 490038


@pytest.mark.slow
@pytest.mark.parametrize("rho,e_half", [(1.0, 8447585), (0.5, 4647396)])
def test_makepair_full_size_matches_oracle(rho, e_half):
    q, box = G.lattice_periodic(31, rho, 2)
    half = G.makepair(q, box, G.LIST_HALF)
    assert half.n == 119164 and half.n_entries == e_half
    same_csr(half, O.makepair(q, box, "half"))


def test_makepair_random_cloud_and_custom_search():
    rng = np.random.default_rng(0)
    q = np.zeros((3000, 4))
    q[:, :3] = rng.uniform(0, 14.0, size=(3000, 3))
    for box in (0.0, 14.0):
        for sl in (2.5, 3.3):
            for kind, okind in ((G.LIST_HALF, "half"), (G.LIST_FULL, "full")):
                same_csr(G.makepair(q, box, kind, sl), O.makepair(q, box, okind, sl * sl))


def test_makepair_edge_sizes():
    for n in (0, 1, 2):
        q = np.zeros((n, 4))
        q[:, 0] = np.arange(n) * 1.5
        c = G.makepair(q, 0.0, G.LIST_HALF)
        assert c.n == n and c.n_entries == (1 if n == 2 else 0)


def test_makepair_rejects_small_box_and_capacity():
    q, box = G.lattice_periodic(2, 1.0)
    with pytest.raises(G.LJGPUError) as e:
        G.makepair(q, box, G.LIST_HALF)
    assert e.value.code == G.ERR_INVALID_ARGUMENT
    q, box = G.lattice_periodic(6, 1.0)
    nop = np.zeros(q.shape[0], np.int32)
    ptr = np.zeros(q.shape[0], np.int32)
    lst = np.zeros(10, np.int32)
    import ctypes as C
    e = C.c_int64(0)
    rc = G.lib().ljgpu_makepair_host(q.shape[0], q.ctypes.data, box, 3.3, G.LIST_HALF,
                                     nop.ctypes.data, ptr.ctypes.data, lst.ctypes.data, 10,
                                     C.byref(e))
    assert rc == G.ERR_CAPACITY and e.value == 61292


def test_validate_csr_accepts_and_rejects():
    q, box = G.lattice_periodic(6, 1.0)
    half = G.makepair(q, box, G.LIST_HALF)
    full = G.makepair(q, box, G.LIST_FULL)
    G.validate_csr(half)
    G.validate_csr(full)

    def bad(csr, mutate, kind=None):
        c = G.CSR(csr.number_of_partners.copy(), csr.pointer.copy(), csr.sorted_list.copy(),
                  csr.kind if kind is None else kind)
        mutate(c)
        with pytest.raises(G.LJGPUError) as e:
            G.validate_csr(c)
        assert e.value.code == G.ERR_BAD_LIST

    bad(half, lambda c: c.sorted_list.__setitem__(5, c.n))          # j >= N
    bad(half, lambda c: c.sorted_list.__setitem__(5, -1))           # j < 0
    bad(half, lambda c: c.number_of_partners.__setitem__(3, -2))    # negative count
    bad(half, lambda c: c.pointer.__setitem__(0, c.n_entries))      # row leaves the list
    bad(full, lambda c: None, kind=G.LIST_HALF)                     # j < i in a half list
    bad(full, lambda c: c.sorted_list.__setitem__(c.pointer[7], 7))  # self pair


def test_random_shfl_matches_oracle():
    q, box = G.lattice_periodic(6, 0.5)
    full = G.makepair(q, box, G.LIST_FULL)
    ocsr = O.CSR(full.number_of_partners, full.pointer, full.sorted_list, "full")
    np.testing.assert_array_equal(G.random_shfl(full, 10).sorted_list,
                                  O.random_shfl(ocsr, 10).sorted_list)


def test_make_transposed_matches_oracle():
    q, box = G.lattice_periodic(6, 0.5)
    full = G.makepair(q, box, G.LIST_FULL)
    ocsr = O.CSR(full.number_of_partners, full.pointer, full.sorted_list, "full")
    w = int(full.number_of_partners.max()) + 2
    np.testing.assert_array_equal(G.make_transposed(full, w), O.make_transposed(ocsr, w))
    with pytest.raises(G.LJGPUError):
        G.make_transposed(full, w - 3)


def test_maxwell_boltzmann_statistics():
    """Seeded, zero mean momentum, per-component variance T."""
    n, t = 200_000, 1.3
    p = G.maxwell_boltzmann(n, t, 11)
    assert p.shape == (n, 4) and np.all(p[:, 3] == 0.0)
    np.testing.assert_array_equal(p, G.maxwell_boltzmann(n, t, 11))
    assert not np.array_equal(p, G.maxwell_boltzmann(n, t, 12))
    assert np.abs(p[:, :3].mean(axis=0)).max() < 1e-12
    var = p[:, :3].var(axis=0)
    # Sample variance of n normal draws: relative standard error sqrt(2/n).
    assert np.abs(var / t - 1.0).max() < 5 * np.sqrt(2.0 / n)
