"""The LDS tiling (ljgpu_make_tiles, and ljgpu_make_tiles_ordered over the
spatial slot order) re-expresses a full list exactly, stays inside its
staging areas, and the tile kernel's periodic unwrapping scheme, emulated
here in numpy, gives the oracle's forces."""
import numpy as np
import pytest

import lj_gpu_amd as G
import oracle as O
from helpers import column_order, normwise, periodic

MAX_STAGE = 65535 // 24


def slot_rows(csr, member):
    """Particle of each slot, and each slot's row start in the rows stored in
    slot order."""
    n = csr.n
    mem = np.arange(n) if member is None else np.asarray(member)
    sptr = np.zeros(n + 1, np.int64)
    sptr[1:] = np.cumsum(csr.number_of_partners[mem])
    return mem, sptr


def check_tiles(csr, t, tile_particles, member=None):
    n = csr.n
    mem, sptr = slot_rows(csr, member)
    desc = t.desc
    assert desc[0, 0] == 0
    np.testing.assert_array_equal(desc[1:, 0], desc[:-1, 0] + desc[:-1, 1])
    assert desc[-1, 0] + desc[-1, 1] == n
    assert np.all((desc[:, 1] >= 1) & (desc[:, 1] <= tile_particles))
    assert np.all(desc[:, 3] <= MAX_STAGE) and t.max_stage == desc[:, 3].max()
    np.testing.assert_array_equal(desc[1:, 2], desc[:-1, 2] + desc[:-1, 3])
    # each block's entries: the contiguous range of its slots' rows
    np.testing.assert_array_equal(desc[:, 4], sptr[desc[:, 0]])
    np.testing.assert_array_equal(desc[:, 4] + desc[:, 5], sptr[desc[:, 0] + desc[:, 1]])
    np.testing.assert_array_equal(desc[:, 7], mem[desc[:, 0] + desc[:, 1] // 2])
    lds = (desc[:, 3] * 24 + 15) // 16 * 16 + desc[:, 5] * 2
    assert lds.max() <= 65536
    assert np.all(t.local % 24 == 0)
    block_of = np.repeat(np.arange(desc.shape[0]), desc[:, 1])
    nop_slot = csr.number_of_partners[mem]
    b = block_of[np.repeat(np.arange(n), nop_slot)]
    # partners in slot-row order: row of slot s is the list row of mem[s]
    start = np.repeat(csr.pointer[mem].astype(np.int64) - sptr[:-1], nop_slot)
    partners = csr.sorted_list[start + np.arange(sptr[-1])]
    rank = t.local.astype(np.int64) // 24
    assert np.all(rank < desc[b, 3])
    np.testing.assert_array_equal(t.stage[desc[b, 2] + rank], partners)
    slot_of = np.empty(n, np.int64)
    slot_of[mem] = np.arange(n)
    for bi in range(0, desc.shape[0], max(1, desc.shape[0] // 50)):
        s = t.stage[desc[bi, 2]: desc[bi, 2] + desc[bi, 3]]
        assert np.all(np.diff(slot_of[s]) > 0)  # sorted by slot
        assert set(mem[desc[bi, 0]: desc[bi, 0] + desc[bi, 1]].tolist()) <= set(s.tolist())


def emulate_tile_force(q, box, csr, t, cl2=O.CL2, member=None):
    """The tile kernel's arithmetic: positions shifted to the image nearest the
    block's middle particle, then raw differences (or the per-entry minimum
    image where some block particle is farther than box/2 - rc)."""
    n = csr.n
    mem, sptr = slot_rows(csr, member)
    F = np.zeros((n, 3))
    x = q[:, :3]
    rc = np.sqrt(cl2)
    for i0, ni, soff, ns, eoff, ne, _, mid in t.desc:
        c = x[mid]
        idx = t.stage[soff: soff + ns]
        xs = x[idx].copy()
        xi = x[mem[i0: i0 + ni]].copy()
        if box > 0:
            xs -= box * np.rint((xs - c) / box)
            xi -= box * np.rint((xi - c) / box)
            far = np.any(np.abs(xi - c) >= 0.5 * box - rc)
        else:
            far = False
        for r in range(ni):
            s = i0 + r
            k0, k1 = sptr[s], sptr[s + 1]
            d = xs[t.local[k0:k1].astype(np.int64) // 24] - xi[r]
            if far:
                d -= box * np.rint(d / box)
            r2 = (d * d).sum(1)
            r6 = r2 ** 3
            df = (24.0 * r6 - 48.0) / (r6 * r6 * r2)
            df[r2 > cl2] = 0.0
            F[mem[s]] = (df[:, None] * d).sum(0)
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


@pytest.mark.parametrize("k,rho,tp", [(12, 1.0, 64), (8, 0.5, 32), (31, 1.0, 64)])
def test_spatial_order_tiles_reexpress_the_list(k, rho, tp):
    q, box, _, full = periodic(k, rho)
    member = column_order(q, box, tp)
    assert np.array_equal(np.sort(member), np.arange(full.n))
    t = G.make_tiles(full, tp, member)
    check_tiles(full, t, tp, member)
    if k == 31:
        # compact blocks: about 900 staged positions instead of about 1,300
        assert t.max_stage < 1000
        assert t.max_stage < 0.75 * G.make_tiles(full, tp).max_stage


def test_spatial_order_blocks_are_compact():
    """The 64-particle blocks of the order span a few units in every
    direction (columns about 3 wide at density 1, two at most per block),
    where a block of 64 consecutive lattice indices is a row of 16 unit cells
    that wraps around the box."""
    q, box, _, full = periodic(12, 1.0)

    def extents(order):
        x = q[order, :3]
        out = []
        for b0 in range(0, full.n - 63, 64):
            d = x[b0: b0 + 64] - x[b0 + 32]
            d -= box * np.rint(d / box)
            out.append(np.abs(d).max(0))
        return np.array(out)

    sp = extents(column_order(q, box, 64))
    ix = extents(np.arange(full.n))
    assert np.percentile(sp, 90, axis=0).max() < 5.0
    assert sp.max() < 0.5 * box - 3.0  # every block passes the kernel's frame check
    assert ix[:, 0].min() > 9.0


@pytest.mark.parametrize("k,rho", [(12, 1.0), (8, 0.5)])
def test_spatial_order_unwrapped_frame_matches_oracle(k, rho):
    q, box, _, full = periodic(k, rho)
    member = column_order(q, box, 64)
    t = G.make_tiles(full, 64, member)
    F = emulate_tile_force(q, box, full, t, member=member)
    p = np.zeros_like(q)
    O.force_aos(O.CSR(full.number_of_partners, full.pointer, full.sorted_list, "full"),
                q, p, 1.0, O.CL2, box, method="full")
    assert normwise(F, p[:, :3]) < 1e-13


def test_make_tiles_ordered_rejects_non_permutations():
    q, box, half, full = periodic(6, 1.0)
    member = np.arange(full.n, dtype=np.int32)
    member[5] = 4
    with pytest.raises(G.LJGPUError) as e:
        G.make_tiles(full, 64, member)
    assert e.value.code == G.ERR_INVALID_ARGUMENT
    member[5] = full.n
    with pytest.raises(G.LJGPUError):
        G.make_tiles(full, 64, member)
