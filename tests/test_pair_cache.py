"""Pair-list caches: the CPU programs' binary pair.dat
(cpu_ref/force_aos.cpp:603-620) and the GPU harness's text
.cache_pair_{half,all}.dat (cuda/force_cuda.cu:165-227).

pair.dat is pinned byte for byte: the file the library writes for the
reference's own init() + makepair() list must have the size and sha256 of the
file the reference's savepair() wrote (tests/golden/ref_pairdat.json, made by
tests/golden/make_golden.py from the compiled reference). The text cache has
no runnable reference writer (cuda/ needs nvcc); it is pinned by a hand-made
known answer of its fprintf layout and by round trips.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

import lj_gpu_amd as G
from tests.helpers import GOLDEN

PAIRDAT = json.load(open(os.path.join(GOLDEN, "ref_pairdat.json")))


def _same_csr(a: G.CSR, b: G.CSR) -> None:
    np.testing.assert_array_equal(a.number_of_partners, b.number_of_partners)
    np.testing.assert_array_equal(a.pointer, b.pointer)
    np.testing.assert_array_equal(a.sorted_list, b.sorted_list)


@pytest.mark.parametrize("name", sorted(PAIRDAT))
def test_pair_dat_is_byte_identical_to_savepair(tmp_path, name):
    g = PAIRDAT[name]
    q = G.lattice_reference(g["density"], g["box"])
    assert q.shape[0] == g["n"]
    half = G.makepair(q, 0.0, G.LIST_HALF)
    assert half.n_entries == g["number_of_pairs"]
    path = tmp_path / "pair.dat"
    G.write_pair_dat(str(path), half)
    blob = path.read_bytes()
    assert len(blob) == g["size"]
    assert hashlib.sha256(blob).hexdigest() == g["sha256"]
    # loadpair() + sortpair() of that file gives back the list.
    _same_csr(G.read_pair_dat(str(path), half.n), half)


def test_pair_dat_layout_with_small_capacities(tmp_path):
    half = G.CSR(np.array([2, 1, 0], np.int32), np.array([0, 2, 3], np.int32),
                 np.array([1, 2, 2], np.int32), G.LIST_HALF)
    path = tmp_path / "p.dat"
    G.write_pair_dat(str(path), half, n_capacity=4, pair_capacity=5)
    words = np.frombuffer(path.read_bytes(), dtype="<i4")
    np.testing.assert_array_equal(words, [3, 2, 1, 0, 0,  # number_of_pairs, nop[4]
                                          0, 0, 1, 0, 0,  # i_particles[5]
                                          1, 2, 2, 0, 0])  # j_particles[5]
    _same_csr(G.read_pair_dat(str(path), 3, 4, 5), half)


def test_pair_dat_keeps_stored_row_order(tmp_path):
    q, box = G.lattice_periodic(5, 1.0, 7)
    half = G.random_shfl(G.makepair(q, box, G.LIST_HALF), 10)
    path = tmp_path / "p.dat"
    G.write_pair_dat(str(path), half, n_capacity=600, pair_capacity=40000)
    _same_csr(G.read_pair_dat(str(path), half.n, 600, 40000), half)


def test_pair_dat_rejects_wrong_inputs(tmp_path):
    q, box = G.lattice_periodic(5, 1.0, 7)
    half = G.makepair(q, box, G.LIST_HALF)
    full = G.makepair(q, box, G.LIST_FULL)
    path = tmp_path / "p.dat"
    with pytest.raises(G.LJGPUError) as e:  # a full list has j < i
        G.write_pair_dat(str(path), full, 600, 40000)
    assert e.value.code == G.ERR_BAD_LIST
    with pytest.raises(G.LJGPUError) as e:
        G.write_pair_dat(str(path), half, 400, 40000)  # N above n_capacity
    assert e.value.code == G.ERR_CAPACITY
    G.write_pair_dat(str(path), half, 600, 40000)
    with pytest.raises(G.LJGPUError) as e:  # other capacities: size mismatch
        G.read_pair_dat(str(path), half.n, 600, 80000)
    assert e.value.code == G.ERR_BAD_LIST
    with pytest.raises(G.LJGPUError) as e:  # fewer particles than rows in the file
        G.read_pair_dat(str(path), half.n - 1, 600, 40000)
    assert e.value.code == G.ERR_BAD_LIST
    blob = bytearray(path.read_bytes())
    # Corrupt one j so that it is not above its i.
    off = 4 * (1 + 600 + 40000)
    blob[off:off + 4] = np.int32(0).tobytes()
    path.write_bytes(bytes(blob))
    with pytest.raises(G.LJGPUError) as e:
        G.read_pair_dat(str(path), half.n, 600, 40000)
    assert e.value.code == G.ERR_BAD_LIST
    with pytest.raises(G.LJGPUError) as e:
        G.read_pair_dat(str(tmp_path / "missing.dat"), half.n)
    assert e.value.code == G.ERR_IO


def test_text_cache_known_answer(tmp_path):
    half = G.CSR(np.array([2, 1, 0], np.int32), np.array([0, 2, 3], np.int32),
                 np.array([1, 2, 2], np.int32), G.LIST_HALF)
    path = tmp_path / ".cache_pair_half.dat"
    G.write_pair_cache(str(path), half)
    # makepaircache(): "%d %d\n" N E, N x "%d %d\n" nop ptr, E x "%d\n" j.
    assert path.read_text() == "3 3\n2 0\n1 2\n0 3\n1\n2\n2\n"
    _same_csr(G.read_pair_cache(str(path), 3, G.LIST_HALF), half)


@pytest.mark.parametrize("kind", [G.LIST_HALF, G.LIST_FULL])
def test_text_cache_round_trip(tmp_path, kind):
    q, box = G.lattice_periodic(4, 0.8, 3)
    csr = G.random_shfl(G.makepair(q, box, kind), 10)
    path = tmp_path / "cache.dat"
    G.write_pair_cache(str(path), csr)
    back = G.read_pair_cache(str(path), csr.n, kind)
    _same_csr(back, csr)
    assert back.kind == kind


def test_text_cache_rejects_wrong_inputs(tmp_path):
    q, box = G.lattice_periodic(5, 1.0, 7)
    full = G.makepair(q, box, G.LIST_FULL)
    path = tmp_path / "cache.dat"
    G.write_pair_cache(str(path), full)
    # cuda/force_cuda.cu:211-214: a cache of another particle number is refused.
    with pytest.raises(G.LJGPUError) as e:
        G.read_pair_cache(str(path), full.n + 1, G.LIST_FULL)
    assert e.value.code == G.ERR_BAD_LIST and "may be broken" in e.value.message
    # A full list read as a half list fails validation (j < i).
    with pytest.raises(G.LJGPUError) as e:
        G.read_pair_cache(str(path), full.n, G.LIST_HALF)
    assert e.value.code == G.ERR_BAD_LIST
    text = path.read_text()
    (tmp_path / "trunc.dat").write_text(text[: len(text) // 2])
    with pytest.raises(G.LJGPUError) as e:
        G.read_pair_cache(str(tmp_path / "trunc.dat"), full.n, G.LIST_FULL)
    assert e.value.code == G.ERR_BAD_LIST
    (tmp_path / "extra.dat").write_text(text + "5\n")
    with pytest.raises(G.LJGPUError):
        G.read_pair_cache(str(tmp_path / "extra.dat"), full.n, G.LIST_FULL)
    (tmp_path / "junk.dat").write_text(text.replace("\n", "\nx", 1))
    with pytest.raises(G.LJGPUError):
        G.read_pair_cache(str(tmp_path / "junk.dat"), full.n, G.LIST_FULL)
