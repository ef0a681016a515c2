"""The LDS tiling (ljgpu_make_tiles) re-expresses a full list exactly, stays
inside its staging areas, and the tile kernel's periodic unwrapping scheme,
emulated here in numpy, gives the oracle's forces."""
import numpy as np
import pytest

import lj_gpu_amd as G
import oracle as O
from helpers import normwise, periodic

MAX_STAGE = 65535 // 24


def check_tiles(csr, t, tile_particles):
    n = csr.n
    desc = t.desc
    assert desc[0, 0] == 0
    np.testing.assert_array_equal(desc[1:, 0], desc[:-1, 0] + desc[:-1, 1])
    assert desc[-1, 0] + desc[-1, 1] == n
    assert np.all((desc[:, 1] >= 1) & (desc[:, 1] <= tile_particles))
    assert np.all(desc[:, 3] <= MAX_STAGE) and t.max_stage == desc[:, 3].max()
    np.testing.assert_array_equal(desc[1:, 2], desc[:-1, 2] + desc[:-1, 3])
    # each block's entries: the contiguous range of its rows
    np.testing.assert_array_equal(desc[:, 4], csr.pointer[desc[:, 0]])
    last = desc[:, 0] + desc[:, 1] - 1
    np.testing.assert_array_equal(desc[:, 4] + desc[:, 5],
                                  csr.pointer[last] + csr.number_of_partners[last])
    lds = (desc[:, 3] * 24 + 15) // 16 * 16 + desc[:, 5] * 2
    assert lds.max() <= 65536
    assert np.all(t.local % 24 == 0)
    block_of = np.repeat(np.arange(desc.shape[0]), desc[:, 1])
    row_of_entry = np.repeat(np.arange(n), csr.number_of_partners)
    b = block_of[row_of_entry]
    slot = t.local.astype(np.int64) // 24
    assert np.all(slot < desc[b, 3])
    np.testing.assert_array_equal(t.stage[desc[b, 2] + slot], csr.sorted_list)
    for bi in range(0, desc.shape[0], max(1, desc.shape[0] // 50)):
        s = t.stage[desc[bi, 2]: desc[bi, 2] + desc[bi, 3]]
        assert np.all(np.diff(s) > 0)
        assert set(range(desc[bi, 0], desc[bi, 0] + desc[bi, 1])) <= set(s.tolist())


def emulate_tile_force(q, box, csr, t, cl2=O.CL2):
    """The tile kernel's arithmetic: positions shifted to the image nearest the
    block's middle particle, then raw differences (or the per-entry minimum
    image where some block particle is farther than box/2 - rc)."""
    n = csr.n
    F = np.zeros((n, 3))
    x = q[:, :3]
    rc = np.sqrt(cl2)
    for i0, ni, soff, ns, eoff, ne, _, _ in t.desc:
        c = x[i0 + ni // 2]
        idx = t.stage[soff: soff + ns]
        xs = x[idx].copy()
        xi = x[i0: i0 + ni].copy()
        if box > 0:
            xs -= box * np.rint((xs - c) / box)
            xi -= box * np.rint((xi - c) / box)
            far = np.any(np.abs(xi - c) >= 0.5 * box - rc)
        else:
            far = False
        for r in range(ni):
            i = i0 + r
            k0, k1 = csr.pointer[i], csr.pointer[i] + csr.number_of_partners[i]
            d = xs[t.local[k0:k1].astype(np.int64) // 24] - xi[r]
            if far:
                d -= box * np.rint(d / box)
            r2 = (d * d).sum(1)
            r6 = r2 ** 3
            df = (24.0 * r6 - 48.0) / (r6 * r6 * r2)
            df[r2 > cl2] = 0.0
            F[i] = (df[:, None] * d).sum(0)
    return F


@pytest.mark.parametrize("k,rho,tp", [(12, 1.0, 64), (12, 1.0, 32), (8, 0.5, 128), (12, 1.0, 1)])
def test_tiles_reexpress_the_list(k, rho, tp):
    q, box, _, full = periodic(k, rho)
    t = G.make_tiles(full, tp)
    check_tiles(full, t, tp)


def test_tiles_full_size_stage_counts():
    q, box, _, full = periodic(31, 1.0)
    t = G.make_tiles(full, 64)
    check_tiles(full, t, 64)
    assert t.desc.shape[0] == (full.n + 63) // 64
    assert t.max_stage < 1400  # about 1,300 staged positions per 64-particle tile


def test_tiles_split_when_partners_scatter():
    """A randomly permuted particle order scatters every row's partners over the
    whole index range; tiles then shrink to fit the LDS staging capacity."""
    q, box, _, _ = periodic(12, 1.0)
    perm = np.random.default_rng(5).permutation(q.shape[0])
    qp = np.ascontiguousarray(q[perm])
    full = G.makepair(qp, box, G.LIST_FULL)
    t = G.make_tiles(full, 64)
    check_tiles(full, t, 64)
    assert t.desc.shape[0] > (full.n + 63) // 64
    assert t.max_stage <= MAX_STAGE


@pytest.mark.parametrize("k,rho", [(12, 1.0), (6, 1.0), (8, 0.5)])
def test_unwrapped_frame_matches_oracle(k, rho):
    q, box, _, full = periodic(k, rho)
    t = G.make_tiles(full, 64)
    F = emulate_tile_force(q, box, full, t)
    p = np.zeros_like(q)
    O.force_aos(O.CSR(full.number_of_partners, full.pointer, full.sorted_list, "full"),
                q, p, 1.0, O.CL2, box, method="full")
    assert normwise(F, p[:, :3]) < 1e-13


def test_unwrapped_frame_open_box():
    q = G.lattice_reference(0.5, 20.0)
    full = G.makepair(q, 0.0, G.LIST_FULL)
    t = G.make_tiles(full, 64)
    F = emulate_tile_force(q, 0.0, full, t)
    p = np.zeros_like(q)
    O.force_aos(O.CSR(full.number_of_partners, full.pointer, full.sorted_list, "full"),
                q, p, 1.0, O.CL2, 0.0, method="full")
    assert normwise(F, p[:, :3]) < 1e-13


def test_make_tiles_rejects_bad_input():
    q, box, half, full = periodic(6, 1.0)
    with pytest.raises(G.LJGPUError):
        G.make_tiles(full, 0)
    bad = G.CSR(full.number_of_partners, full.pointer, full.sorted_list.copy(), G.LIST_FULL)
    bad.sorted_list[3] = full.n
    with pytest.raises(G.LJGPUError) as e:
        G.make_tiles(bad, 64)
    assert e.value.code == G.ERR_BAD_LIST
