"""Batched replicas: independent periodic boxes of uneven size evaluated in
one launch of the tiled kernel (no reference counterpart; each replica is
pinned against the oracle and against a single-system run of the library).
"""
import numpy as np
import pytest

import lj_gpu_amd as G
import oracle as O
from helpers import FP64_TOL, PBC_DT, column_order, length_sorted, normwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    if G.device_count() < 1:
        pytest.skip("no HIP device")
    c = G.Context(0)
    yield c
    c.set_replicas([], [])
    c.close()


@pytest.fixture(scope="module")
def work():
    """Six replicas, k in 6..9, densities in [0.5, 1]."""
    return G.replica_workload(6, seed=11, cells=(6, 9))


def host_batch_list(q, off, boxes, kind):
    parts = [G.makepair(q[off[r]:off[r + 1]], boxes[r], kind) for r in range(len(boxes))]
    nop = np.concatenate([c.number_of_partners for c in parts])
    lst = np.concatenate([c.sorted_list + off[r] for r, c in enumerate(parts)])
    ptr = np.zeros_like(nop)
    ptr[1:] = np.cumsum(nop)[:-1]
    return G.CSR(nop, ptr, lst.astype(np.int32), kind)


def configure(ctx, lanes=8, unroll=4, tile=32, resident=0):
    ctx.set_kernel_config(lanes, 256, unroll)
    ctx.set_tiling(tile, 0, resident)
    ctx.set_half_mode(G.HALF_AUTO)


@pytest.mark.parametrize("kind", [G.LIST_HALF, G.LIST_FULL])
def test_batched_build_matches_host_lists(ctx, work, kind):
    q, off, boxes = work
    configure(ctx)
    ctx.set_replicas(off, boxes)
    ctx.set_half_mode(G.HALF_ATOMIC)
    ctx.build_list(ctx.array(q), q.shape[0], 0.0, kind)
    got = ctx.get_list()
    ref = host_batch_list(q, off, boxes, kind)
    np.testing.assert_array_equal(got.number_of_partners, ref.number_of_partners)
    np.testing.assert_array_equal(got.pointer, ref.pointer)
    np.testing.assert_array_equal(got.sorted_list, ref.sorted_list)
    ctx.set_half_mode(G.HALF_AUTO)


def test_batched_spatial_tile_order(ctx, work):
    """Each replica's particles are ordered within its own slot range (the
    numpy restatement over all replicas), so no tile mixes replicas."""
    q, off, boxes = work
    configure(ctx)
    ctx.set_replicas(off, boxes)
    ctx.build_list(ctx.array(q), q.shape[0], 0.0, G.LIST_FULL)
    order, member = ctx.get_tile_order()
    assert order == G.TILE_ORDER_SPATIAL
    nop = ctx.get_list().number_of_partners
    np.testing.assert_array_equal(
        member, length_sorted(column_order(q, 0.0, 32, off, boxes), nop, 32, off))
    t = ctx.get_tiles()
    for i0, ni, *_rest in t.desc:
        r = np.searchsorted(off, i0, side="right") - 1
        assert off[r] <= member[i0: i0 + ni].min() and member[i0: i0 + ni].max() < off[r + 1]
    ctx.set_replicas([], [])


def batched_momenta(ctx, q, off, boxes, steps=100, mixed=None, build=True, graphs=True):
    ctx.set_replicas(off, boxes)
    ctx.set_graphs(graphs)
    dq = ctx.array(q)
    if build:
        ctx.build_list(dq, q.shape[0], 0.0, G.LIST_FULL)
    else:
        ctx.set_list(host_batch_list(q, off, boxes, G.LIST_HALF))
    dp = ctx.zeros(q.shape)
    if mixed is None:
        ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, 0.0, steps)
    else:
        ctx.force_aos_mixed(dq, dp, PBC_DT, G.CL2, 0.0, steps, mixed)
    ctx.sync()
    ctx.set_graphs(True)
    return dp.download()


def test_batched_forces_match_oracle_per_replica(ctx, work):
    q, off, boxes = work
    configure(ctx)
    p = batched_momenta(ctx, q, off, boxes)
    for r in range(len(boxes)):
        qr = np.ascontiguousarray(q[off[r]:off[r + 1]])
        half = O.makepair(qr, boxes[r], "half")
        ref = np.zeros_like(qr)
        O.force_aos(half, qr, ref, PBC_DT, O.CL2, boxes[r], steps=100)
        assert normwise(p[off[r]:off[r + 1], :3], ref) < FP64_TOL, r


@pytest.mark.parametrize("lanes,unroll,tile,resident", [(8, 4, 32, 0), (16, 2, 64, 0),
                                                        (4, 4, 32, 2)])
def test_batched_forces_equal_single_system_runs(ctx, work, lanes, unroll, tile, resident):
    """Each replica's tiles are the tiles of that replica run alone, so the
    batched momenta equal the single-system ones bit for bit."""
    q, off, boxes = work
    configure(ctx, lanes, unroll, tile, resident)
    p = batched_momenta(ctx, q, off, boxes, steps=3)
    ctx.set_replicas([], [])
    for r in range(len(boxes)):
        qr = np.ascontiguousarray(q[off[r]:off[r + 1]])
        ctx.set_list(G.makepair(qr, boxes[r], G.LIST_FULL))
        dq, dp = ctx.array(qr), ctx.zeros(qr.shape)
        ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, boxes[r], 3)
        ctx.sync()
        np.testing.assert_array_equal(p[off[r]:off[r + 1]], dp.download())


def test_batched_host_list_and_graph_replay(ctx, work):
    q, off, boxes = work
    configure(ctx)
    a = batched_momenta(ctx, q, off, boxes, build=False, graphs=False)
    b = batched_momenta(ctx, q, off, boxes, build=True, graphs=True)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("accum", [G.ACCUM_F64, G.ACCUM_KAHAN_F32])
def test_batched_mixed_precision(ctx, work, accum):
    q, off, boxes = work
    configure(ctx)
    f = batched_momenta(ctx, q, off, boxes, steps=1)
    m = batched_momenta(ctx, q, off, boxes, steps=1, mixed=accum)
    for r in range(len(boxes)):
        s = slice(off[r], off[r + 1])
        assert normwise(m[s, :3], f[s, :3]) < 2e-6, r


def test_batched_rejections(ctx, work):
    q, off, boxes = work
    configure(ctx)
    ctx.set_replicas(off, boxes)
    bad = host_batch_list(q, off, boxes, G.LIST_FULL)
    lst = bad.sorted_list.copy()
    lst[bad.pointer[0]] = off[1]  # particle 0 paired with replica 1
    with pytest.raises(G.LJGPUError) as e:
        ctx.set_list(G.CSR(bad.number_of_partners, bad.pointer, lst, G.LIST_FULL))
    assert e.value.code == G.ERR_BAD_LIST
    # A list over fewer particles than the replicas hold.
    last = host_batch_list(q[:off[-2]], off[:-1], boxes[:-1], G.LIST_FULL)
    with pytest.raises(G.LJGPUError) as e:
        ctx.set_list(last)
    assert e.value.code == G.ERR_INVALID_ARGUMENT
    dq = ctx.array(q)
    with pytest.raises(G.LJGPUError):
        ctx.build_list(dq, q.shape[0], boxes[0], G.LIST_FULL)  # box must be 0
    ctx.build_list(dq, q.shape[0], 0.0, G.LIST_FULL)
    dp = ctx.zeros(q.shape)
    with pytest.raises(G.LJGPUError):
        ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, boxes[0], 1)
    ctx.set_kernel_config(1, 256, 2)
    with pytest.raises(G.LJGPUError) as e:
        ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, 0.0, 1)
    assert e.value.code == G.ERR_UNSUPPORTED
    configure(ctx)
    with pytest.raises(G.LJGPUError) as e:
        ctx.md_run(dq, dp, 0.0, 10)
    assert e.value.code == G.ERR_UNSUPPORTED
    ctx.set_replicas([], [])


@pytest.mark.parametrize("lanes,tile", [(8, 32), (4, 64)])
def test_batched_slot_copies_are_bitwise_per_step(ctx, work, lanes, tile):
    """A 20-step call over replicas runs on slot-ordered copies of q and p;
    20 one-step calls read the caller's arrays through the slot order. Same
    arithmetic with each replica's own box: equal bit for bit, from nonzero
    momenta whose fourth value per row comes back unchanged."""
    q, off, boxes = work
    configure(ctx, lanes=lanes, tile=tile)
    ctx.set_replicas(off, boxes)
    dq = ctx.array(q)
    ctx.build_list(dq, q.shape[0], 0.0, G.LIST_FULL)
    assert ctx.get_tile_order()[0] == G.TILE_ORDER_SPATIAL
    p0 = np.random.default_rng(7).standard_normal(q.shape)
    res = []
    for calls, steps in ((1, 20), (20, 1)):
        dp = ctx.array(p0)
        for _ in range(calls):
            ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, 0.0, steps)
        ctx.sync()
        res.append(dp.download())
    ctx.set_replicas([], [])
    np.testing.assert_array_equal(res[0], res[1])
    np.testing.assert_array_equal(res[0][:, 3], p0[:, 3])
    np.testing.assert_array_equal(dq.download(), q)
