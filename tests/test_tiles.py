"""The LDS tiling (ljgpu_make_tiles, and ljgpu_make_tiles_ordered over the
spatial slot order) re-expresses a full list exactly, stays inside its
staging areas, and the tile kernel's periodic unwrapping scheme, emulated
here in numpy, gives the oracle's forces."""
import numpy as np
import pytest

import lj_gpu_amd as G
import oracle as O
from helpers import bank_class_order, column_order, cut_inner, length_sorted, normwise, periodic

MAX_STAGE = 65535 // 24


def slot_rows(csr, member):
    """Particle of each slot, and each slot's row start in the rows stored in
    slot order."""
    n = csr.n
    mem = np.arange(n) if member is None else np.asarray(member)
    sptr = np.zeros(n + 1, np.int64)
    sptr[1:] = np.cumsum(csr.number_of_partners[mem])
    return mem, sptr


def check_tiles(csr, t, tile_particles, member=None, q=None, box=0.0, cl2=O.CL2):
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
    slot_of = np.empty(n, np.int64)
    slot_of[mem] = np.arange(n)
    # Each entry's partner rank in its block's staged set, in list order: the
    # staged sets, keyed (block, slot), are ascending over the whole stage.
    skey = np.repeat(np.arange(desc.shape[0]), desc[:, 3]) * (n + 1) + slot_of[t.stage]
    pkey = b * (n + 1) + slot_of[partners]
    at = np.searchsorted(skey, pkey)
    np.testing.assert_array_equal(skey[np.minimum(at, skey.size - 1)], pkey)
    list_offs = (at - desc[b, 2]) * 24
    # Rows in the bank-class order (checked on about 200 evenly spaced blocks),
    # each a permutation of its list row (checked on all). With positions the
    # rows are first split into cutoff classes (entries within the cutoff
    # first, both parts in list order), each part then in the bank-class
    # order with its classes continuing from the part's start.
    row_pos = np.arange(n) - desc[block_of, 0]
    pick = np.nonzero(block_of % max(1, desc.shape[0] // 200) == 0)[0]
    sel = np.repeat(sptr[pick] - np.cumsum(np.r_[0, nop_slot[pick]][:-1]), nop_slot[pick]) + \
        np.arange(int(nop_slot[pick].sum()))
    if q is None:
        expect = bank_class_order(list_offs, sptr[pick], nop_slot[pick], row_pos[pick])
    else:
        i_of = np.repeat(mem, nop_slot)
        inner = cut_inner(q, box, cl2, i_of, partners)
        parted = list_offs.copy()
        n_in = np.zeros(n, np.int64)
        for s_ in pick:
            a, m = sptr[s_], nop_slot[s_]
            if m > 256:
                continue
            f = inner[a: a + m]
            n_in[s_] = f.sum()
            parted[a: a + m] = np.concatenate([list_offs[a: a + m][f], list_offs[a: a + m][~f]])
        starts = np.concatenate([sptr[pick], sptr[pick] + n_in[pick]])
        lens = np.concatenate([n_in[pick], nop_slot[pick] - n_in[pick]])
        pos = np.concatenate([row_pos[pick], row_pos[pick]])
        k0 = np.concatenate([np.zeros(pick.size, np.int64), n_in[pick]])
        expect = bank_class_order(parted, starts, lens, pos, k0)
        # the picked rows hold entries of both classes
        assert n_in[pick].min() > 0 and (nop_slot[pick] - n_in[pick]).min() > 0
    np.testing.assert_array_equal(t.local[sel], expect[sel])
    row_id = np.repeat(np.arange(n, dtype=np.int64), nop_slot) << 16
    np.testing.assert_array_equal(np.sort(row_id + t.local), np.sort(row_id + list_offs))
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


def _bank_order_loop(ranks, r):
    """One row, literally: entry k takes the first remaining partner of class
    (4 * (r & 7) + k) mod 32, or of the next class up that the row holds."""
    pools = [[v for v in ranks if v % 32 == c] for c in range(32)]
    out = []
    for k in range(len(ranks)):
        c = (4 * (r % 8) + k) % 32
        while not pools[c]:
            c = (c + 1) % 32
        out.append(pools[c].pop(0))
    return out


def test_bank_class_order_hand_cases():
    # ranks 0, 32 (class 0) and 1, 33 (class 1); offsets are 24 * rank
    row = np.array([0, 32, 1, 33]) * 24
    one = lambda r: list(bank_class_order(row, [0], [4], [r]) // 24)
    assert one(0) == [0, 1, 32, 33]   # targets 0, 1, then 2 and 3 wrap to 0 and 1
    assert one(1) == [0, 32, 1, 33]   # targets 4..7 all wrap to class 0 first
    assert one(8) == one(0)           # only r & 7 matters
    # rows of one entry, and rows longer than 256, keep the list order
    long = np.arange(300)[::-1] * 24
    np.testing.assert_array_equal(bank_class_order(long, [0], [300], [3]), long)
    np.testing.assert_array_equal(bank_class_order(np.array([48]), [0], [1], [5]), [48])


def test_bank_class_order_matches_the_literal_rule():
    rng = np.random.default_rng(11)
    lengths = rng.integers(0, 258, 60)
    starts = np.concatenate([[0], np.cumsum(lengths)[:-1]])
    pos = rng.integers(0, 64, lengths.size)
    offs = rng.integers(0, 2700, int(lengths.sum())) * 24
    got = bank_class_order(offs, starts, lengths, pos)
    for a, m, r in zip(starts, lengths, pos):
        row = offs[a:a + m] // 24
        want = _bank_order_loop(list(row), r) if 2 <= m <= 256 else list(row)
        assert list(got[a:a + m] // 24) == want


def test_bank_class_order_gives_distinct_classes_per_half_wave():
    # Eight rows that each hold every class at least 4 times: for the first
    # 128 entries the 32 lanes of a half-wavefront (rows r = 0..7, entries
    # 4u..4u+3) read 32 distinct classes.
    rng = np.random.default_rng(5)
    rows = [rng.permutation(np.concatenate([np.arange(32) + 32 * rng.integers(0, 80, 32)
                                            for _ in range(5)])) for _ in range(8)]
    offs = np.concatenate(rows) * 24
    lengths = np.full(8, 160)
    got = bank_class_order(offs, np.arange(8) * 160, lengths, np.arange(8)).reshape(8, 160) // 24
    for base in range(0, 128, 4):
        cls = (got[:, base:base + 4] % 32).ravel()
        assert np.unique(cls).size == 32
    # and every row is a permutation of its list row
    for r in range(8):
        np.testing.assert_array_equal(np.sort(got[r]), np.sort(rows[r]))


def test_length_sorted_hand_cases():
    # blocks of 4 slots: lengths 9, 5, 2, 7 -> longest first; the last block
    # is partial
    nop = np.array([5, 7, 7, 2, 9, 1])
    assert list(length_sorted([4, 0, 3, 1, 2, 5], nop, 4)) == [4, 1, 0, 3, 2, 5]
    # ties keep their order, and blocks restart at each system boundary
    assert list(length_sorted([0, 1, 2, 3], np.array([3, 3, 5, 3]), 4, [0, 1, 4])) == [0, 2, 1, 3]


@pytest.mark.parametrize("k,rho,tp,spatial", [(12, 1.0, 64, True), (8, 0.5, 32, True),
                                              (12, 1.0, 64, False), (31, 1.0, 64, True)])
def test_cutoff_classes_of_tile_rows(k, rho, tp, spatial):
    """With positions, each tile row holds the entries within the cutoff
    first, then those beyond it (each part in list order, then bank-class
    ordered with its classes continuing from the part's start); the tiling
    still re-expresses the list and gives the oracle's forces."""
    q, box, _, full = periodic(k, rho)
    member = column_order(q, box, tp) if spatial else None
    t = G.make_tiles(full, tp, member, q, box, G.CL2)
    check_tiles(full, t, tp, member, q, box, G.CL2)
    if k <= 12:
        F = emulate_tile_force(q, box, full, t, member=member)
        p = np.zeros_like(q)
        O.force_aos(O.CSR(full.number_of_partners, full.pointer, full.sorted_list, "full"),
                    q, p, 1.0, O.CL2, box, method="full")
        assert normwise(F, p[:, :3]) < 1e-13
    # a wrong class boundary is caught: classes of another cutoff differ
    if k == 12 and spatial:
        t2 = G.make_tiles(full, tp, member, q, box, 2.8 ** 2)
        assert not np.array_equal(t2.local, t.local)
        with pytest.raises(AssertionError):
            check_tiles(full, t2, tp, member, q, box, G.CL2)


def test_cutoff_classes_hand_case():
    """Four particles on a line in an open box: partners of particle 0 at
    distances 1, 3.1 and 2 (list order 1, 2, 3); with cutoff 3 the row of
    particle 0 becomes partners 1, 3 (within) then 2 (beyond)."""
    q = np.zeros((4, 4))
    q[:, 0] = [0.0, 1.0, 3.1, 2.0]
    nop = np.array([3, 3, 3, 3], np.int32)
    ptr = np.array([0, 3, 6, 9], np.int32)
    lst = np.array([1, 2, 3, 0, 2, 3, 0, 1, 3, 0, 1, 2], np.int32)
    csr = G.CSR(nop, ptr, lst, G.LIST_FULL)
    t = G.make_tiles(csr, 64, None, q, 0.0, 9.0)
    stage = t.stage[t.desc[0, 2]: t.desc[0, 2] + t.desc[0, 3]]
    row0 = stage[t.local[0:3] // 24]
    assert sorted(row0[:2].tolist()) == [1, 3] and row0[2] == 2
    # without positions the row keeps all three partners in one class
    t0 = G.make_tiles(csr, 64)
    assert sorted(stage[t0.local[0:3] // 24].tolist()) == [1, 2, 3]


def test_bank_class_order_continues_from_the_part_start():
    row = np.array([0, 32, 1, 33]) * 24
    # a part starting at entry 1 of its row targets classes 1, 2, 3, 4 (r = 0):
    # class 1 first, then 2, 3 and 4 wrap round to class 0, then class 1 again
    got = list(bank_class_order(row, [0], [4], [0], [1]) // 24)
    assert got == [1, 0, 32, 33]
    # the literal loop with the targets shifted by the part start agrees
    rng = np.random.default_rng(3)
    ranks = list(rng.choice(300, 40, replace=False))
    for r, k0 in ((0, 0), (5, 17), (12, 33)):
        pools = [[v for v in ranks if v % 32 == c] for c in range(32)]
        want = []
        for k in range(len(ranks)):
            c = (4 * (r % 8) + k0 + k) % 32
            while not pools[c]:
                c = (c + 1) % 32
            want.append(pools[c].pop(0))
        got = bank_class_order(np.array(ranks) * 24, [0], [len(ranks)], [r], [k0]) // 24
        assert list(got) == want
