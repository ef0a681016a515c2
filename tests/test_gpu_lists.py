"""Neighbour lists built and derived on the device.

ljgpu_build_list builds the Verlet list from device positions with a cell
list; the CSR it installs must be bit-identical to the host builder's
(ljgpu_makepair_host), which the CPU suite pins to the oracle's O(N^2)
makepair (tests/test_host_lists.py) and to the compiled reference's CSR
(tests/test_oracle_golden.py). Layouts derived on the device from a CSR (the
full list of a half list, row ordering, LDS tiles) must equal their host
restatements (expand_half_to_full via makepair FULL, ljgpu_make_tiles), and
the spatial slot order of the tiles the numpy restatement in helpers.py.
"""
import numpy as np
import pytest

import lj_gpu_amd as G
import oracle as O
from helpers import (FP64_TOL, PBC_DT, column_order, length_sorted, normwise, oracle_momenta,
                     periodic)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    if G.device_count() < 1:
        pytest.skip("no HIP device")
    c = G.Context(0)
    yield c
    c.close()


def _same(a: G.CSR, b: G.CSR) -> None:
    assert a.n == b.n
    np.testing.assert_array_equal(a.number_of_partners, b.number_of_partners)
    np.testing.assert_array_equal(a.pointer, b.pointer)
    np.testing.assert_array_equal(a.sorted_list, b.sorted_list)


def _device_build(ctx, q, box, kind, half_mode=G.HALF_AUTO, soa=False):
    ctx.set_half_mode(half_mode)
    if soa:
        dq = [ctx.array(np.ascontiguousarray(q[:, c])) for c in range(3)]
        ctx.build_list_soa(dq, q.shape[0], box, kind)
    else:
        dq = ctx.array(q)
        ctx.build_list(dq, q.shape[0], box, kind)
    return ctx.get_list(), ctx.list_info()


@pytest.mark.parametrize("k,rho", [(5, 1.0), (12, 1.0), (31, 1.0), (31, 0.5)])
@pytest.mark.parametrize("kind", [G.LIST_HALF, G.LIST_FULL])
def test_device_build_matches_host_periodic(ctx, k, rho, kind):
    q, box, half, full = periodic(k, rho)
    host = half if kind == G.LIST_HALF else full
    print(f"build k={k} rho={rho} kind={kind}", flush=True)
    # Half lists: the atomic mode evaluates the half list itself, the
    # default expands it; both are checked.
    got, info = _device_build(ctx, q, box, kind, G.HALF_ATOMIC)
    _same(got, host)
    assert info.n_entries == host.n_entries
    assert info.max_partners == host.number_of_partners.max()
    assert info.min_partners == host.number_of_partners.min()
    if kind == G.LIST_HALF:
        got, info = _device_build(ctx, q, box, kind, G.HALF_EXPAND)
        _same(got, full)
        assert info.n_entries == half.n_entries
        assert info.max_partners == half.number_of_partners.max()
        assert info.min_partners == half.number_of_partners.min()


@pytest.mark.parametrize("k,rho", [(12, 1.0), (31, 1.0), (31, 0.5)])
def test_device_build_wrapped_positions(ctx, k, rho):
    """Every position inside [0, box): cells whose stencil does not wrap
    skip the minimum image, which cannot shift any of their pairs."""
    q, box, _, _ = periodic(k, rho)
    q = q.copy()
    q[:, :3] -= box * np.floor(q[:, :3] / box)
    q[:, :3][q[:, :3] >= box] -= box
    assert q[:, :3].min() >= 0.0 and q[:, :3].max() < box
    for kind in (G.LIST_HALF, G.LIST_FULL):
        host = G.makepair(q, box, kind)
        got, _ = _device_build(ctx, q, box, kind, G.HALF_ATOMIC)
        _same(got, host)


def test_device_build_positions_outside_the_box(ctx):
    """Particles moved out by one box length (x in [box, 1.4 box)): the
    minimum image stays on in every cell, and the list equals the host's."""
    q, box, _, _ = periodic(12, 1.0)
    q = q.copy()
    sel = q[:, 0] < 0.4 * box
    q[sel, 0] += box
    for kind in (G.LIST_HALF, G.LIST_FULL):
        host = G.makepair(q, box, kind)
        got, _ = _device_build(ctx, q, box, kind, G.HALF_ATOMIC)
        _same(got, host)


@pytest.mark.parametrize("density", [0.5, 1.0])
def test_device_build_matches_host_reference_box(ctx, density):
    """The reference's own open-box problem (init() with L = 50)."""
    q = G.lattice_reference(density, 50.0)
    host = G.makepair(q, 0.0, G.LIST_HALF)
    got, _ = _device_build(ctx, q, 0.0, G.LIST_HALF, G.HALF_ATOMIC)
    _same(got, host)


def test_device_build_random_clouds(ctx):
    rng = np.random.default_rng(5)
    for n, side, box in [(3000, 20.0, 0.0), (3000, 20.0, 20.0), (500, 7.0, 7.0), (50, 100.0, 0.0),
                         (1, 3.0, 0.0), (2, 10.0, 10.0)]:
        q = np.zeros((n, 4))
        q[:, :3] = rng.uniform(0.0, side, size=(n, 3))
        for kind in (G.LIST_HALF, G.LIST_FULL):
            host = G.makepair(q, box, kind)
            got, _ = _device_build(ctx, q, box, kind, G.HALF_ATOMIC)
            _same(got, host)


@pytest.mark.parametrize("kind", [G.LIST_HALF, G.LIST_FULL])
def test_device_build_aos3_matches_host(ctx, kind):
    """Packed double[N][3] positions give the same list as double4 ones."""
    q, box, half, full = periodic(12, 1.0)
    ctx.set_half_mode(G.HALF_ATOMIC)
    dq = ctx.array(np.ascontiguousarray(q[:, :3]))
    ctx.build_list_aos3(dq, q.shape[0], box, kind)
    _same(ctx.get_list(), half if kind == G.LIST_HALF else full)
    ctx.set_half_mode(G.HALF_AUTO)


def test_device_build_soa_matches_aos(ctx):
    q, box, half, full = periodic(12, 1.0)
    got, _ = _device_build(ctx, q, box, G.LIST_FULL, soa=True)
    _same(got, full)


def test_device_build_rejects_bad_arguments(ctx):
    q, box, _, _ = periodic(5, 1.0)
    dq = ctx.array(q)
    with pytest.raises(G.LJGPUError):
        ctx.build_list(dq, q.shape[0], 5.0, G.LIST_FULL)  # box < 2 * search length
    with pytest.raises(G.LJGPUError):
        ctx.build_list(dq, q.shape[0], box, 7)


def test_forces_from_device_built_list(ctx):
    q, box, half, full = periodic(12, 1.0)
    ref = oracle_momenta(12, 1.0)
    ctx.set_kernel_config(8, 256, 4)
    ctx.set_tiling(32)
    ctx.set_half_mode(G.HALF_AUTO)
    dq = ctx.array(q)
    ctx.build_list(dq, q.shape[0], box, G.LIST_HALF)
    dp = ctx.zeros(q.shape)
    ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, 100)
    ctx.sync()
    assert normwise(dp.download()[:, :3], ref) < FP64_TOL


def test_empty_system(ctx):
    q = np.zeros((0, 4))
    dq = ctx.zeros((1, 4))
    ctx.build_list(dq, 0, 0.0, G.LIST_FULL)
    assert ctx.list_info().n_entries == 0


@pytest.mark.parametrize("k,rho", [(12, 1.0), (31, 1.0), (31, 0.5)])
@pytest.mark.parametrize("lanes,tile", [(8, 32), (16, 16), (4, 64)])
def test_device_tiles_match_host_tiles(ctx, k, rho, lanes, tile):
    q, box, half, full = periodic(k, rho)
    ctx.set_kernel_config(lanes, 256, 2)
    ctx.set_tiling(tile)
    ctx.set_half_mode(G.HALF_AUTO)
    ctx.set_tile_order(G.TILE_ORDER_INDEX)
    ctx.set_list(full)
    t = ctx.get_tiles()
    info = ctx.list_info()
    ctx.set_tile_order(G.TILE_ORDER_SPATIAL)
    assert t is not None
    h = G.make_tiles(full, info.tile_particles)
    # Identical whenever the host tiler did not have to shrink a block.
    if np.all(h.desc[:-1, 1] == info.tile_particles):
        np.testing.assert_array_equal(t.desc, h.desc)
        np.testing.assert_array_equal(t.stage, h.stage)
        np.testing.assert_array_equal(t.local, h.local)
    assert t.max_stage == info.tile_max_stage


def test_device_expansion_of_shuffled_half_list(ctx):
    """Half list with shuffled rows, installed from device memory: the
    expansion keeps lower partners ascending and the own row as stored."""
    q, box, half, full = periodic(6, 1.0)
    sh = G.random_shfl(half, 10)
    n = sh.n
    nop, ptr, lst = (ctx.array(a) for a in (sh.number_of_partners, sh.pointer, sh.sorted_list))
    ctx.set_half_mode(G.HALF_EXPAND)
    ctx.set_kernel_config(8, 256, 4)
    ctx.set_list_device(n, nop, ptr, lst, sh.n_entries, G.LIST_HALF)
    got = ctx.get_list()
    for i in range(n):
        row = got.sorted_list[got.pointer[i]: got.pointer[i] + got.number_of_partners[i]]
        lower = full.sorted_list[full.pointer[i]: full.pointer[i] + full.number_of_partners[i]]
        lower = lower[lower < i]
        own = sh.sorted_list[sh.pointer[i]: sh.pointer[i] + sh.number_of_partners[i]]
        np.testing.assert_array_equal(row, np.concatenate([lower, own]))


def test_device_validation_of_device_lists(ctx):
    q, box, half, full = periodic(5, 1.0)
    for mutate, kind in [(lambda l, p, c: l.__setitem__(3, -1), G.LIST_FULL),
                         (lambda l, p, c: l.__setitem__(p[5], 5), G.LIST_FULL),
                         (lambda l, p, c: p.__setitem__(0, 10 ** 6), G.LIST_FULL),
                         (lambda l, p, c: c.__setitem__(2, -3), G.LIST_FULL),
                         (lambda l, p, c: l.__setitem__(p[9], 1), G.LIST_HALF)]:
        src = full if kind == G.LIST_FULL else half
        lst, ptr, nop = src.sorted_list.copy(), src.pointer.copy(), src.number_of_partners.copy()
        mutate(lst, ptr, nop)
        d = [ctx.array(a) for a in (nop, ptr, lst)]
        with pytest.raises(G.LJGPUError) as e:
            ctx.set_list_device(src.n, d[0], d[1], d[2], src.n_entries, kind)
        assert e.value.code == G.ERR_BAD_LIST


@pytest.mark.slow
def test_device_build_large_system(ctx):
    """FCC 4*48^3 = 442,368 particles at density 0.8 (the dynamic-system
    density), full list."""
    q, box = G.lattice_periodic(48, 0.8, 2)
    host = G.makepair(q, box, G.LIST_FULL)
    got, _ = _device_build(ctx, q, box, G.LIST_FULL)
    _same(got, host)


def _check_spatial_tiles(ctx, q, box, full, tile):
    order, member = ctx.get_tile_order()
    assert order == G.TILE_ORDER_SPATIAL
    np.testing.assert_array_equal(
        member, length_sorted(column_order(q, box, tile), full.number_of_partners, tile))
    t = ctx.get_tiles()
    # tiles derived with positions carry the cutoff classes of those positions
    h = G.make_tiles(full, tile, member, q, box, G.CL2)
    assert t.max_stage == ctx.list_info().tile_max_stage
    # Identical whenever the host tiler did not have to shrink a block.
    if np.all(h.desc[:-1, 1] == tile):
        np.testing.assert_array_equal(t.desc, h.desc)
        np.testing.assert_array_equal(t.stage, h.stage)
        np.testing.assert_array_equal(t.local, h.local)
    return t


@pytest.mark.parametrize("k,rho", [(12, 1.0), (31, 0.5)])
def test_spatial_tiles_of_a_host_list(ctx, k, rho):
    """A host list gets index-order tiles at installation and the spatial
    order at its first force evaluation (which brings the positions)."""
    q, box, half, full = periodic(k, rho)
    ctx.set_kernel_config(4, 256, 4)
    ctx.set_tiling(64)
    ctx.set_half_mode(G.HALF_AUTO)
    ctx.set_list(full)
    order, member = ctx.get_tile_order()
    assert order == G.TILE_ORDER_INDEX
    np.testing.assert_array_equal(member, np.arange(full.n))
    dq, dp = ctx.array(q), ctx.zeros(q.shape)
    ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, 1)
    ctx.sync()
    t = _check_spatial_tiles(ctx, q, box, full, 64)
    if k == 31:
        assert t.max_stage < 0.8 * G.make_tiles(full, 64).max_stage


def test_spatial_tiles_of_a_device_list(ctx):
    """ljgpu_build_list orders the tiles from its own positions at once; the
    half list expanded on the device tiles the same way."""
    q, box, half, full = periodic(12, 1.0)
    ctx.set_kernel_config(8, 256, 2)
    ctx.set_tiling(32)
    for kind in (G.LIST_FULL, G.LIST_HALF):
        _device_build(ctx, q, box, kind)
        _check_spatial_tiles(ctx, q, box, full, 32)


def _force_once(ctx, q, box, full):
    ctx.set_kernel_config(4, 256, 4)
    ctx.set_tiling(64)
    ctx.set_list(full)
    dq, dp = ctx.array(q), ctx.zeros(q.shape)
    ctx.force_aos_f64(dq, dp, 1.0, G.CL2, box, 1)
    ctx.sync()
    return dp.download()


def test_spatial_order_of_a_sparse_system_in_a_large_box(ctx):
    """300 particles in a box of side 120: the column grid is capped at about
    two columns per particle, and forces still match the oracle."""
    rng = np.random.default_rng(8)
    box = 120.0
    q = np.zeros((300, 4))
    q[:, :3] = box * rng.random((300, 3))
    q[:150, :3] = 40.0 + 6.0 * rng.random((150, 3))  # a cluster with neighbours
    full = G.makepair(q, box, G.LIST_FULL)
    p = _force_once(ctx, q, box, full)
    order, member = ctx.get_tile_order()
    assert order == G.TILE_ORDER_SPATIAL
    np.testing.assert_array_equal(
        member, length_sorted(column_order(q, box, 64), full.number_of_partners, 64))
    ref = np.zeros_like(q)
    O.force_aos(O.CSR(full.number_of_partners, full.pointer, full.sorted_list, "full"),
                q, ref, 1.0, O.CL2, box, method="full")
    assert normwise(p, ref) < FP64_TOL


def test_crowded_bin_keeps_index_order(ctx):
    """More than 256 particles in one bin of the order: index order."""
    rng = np.random.default_rng(9)
    box = 20.0
    q = np.zeros((600, 4))
    q[:, :3] = box * rng.random((600, 3))
    q[:400, :3] = 7.0 + 0.01 * rng.random((400, 3))
    full = G.makepair(q, box, G.LIST_FULL)
    _force_once(ctx, q, box, full)
    order, member = ctx.get_tile_order()
    assert order == G.TILE_ORDER_INDEX


def test_open_box_keeps_index_order(ctx):
    q = G.lattice_reference(0.5, 20.0)
    full = G.makepair(q, 0.0, G.LIST_FULL)
    ctx.set_kernel_config(4, 256, 4)
    ctx.set_tiling(64)
    ctx.set_list(full)
    dq, dp = ctx.array(q), ctx.zeros(q.shape)
    ctx.force_aos_f64(dq, dp, 1.0, G.CL2, 0.0, 1)
    ctx.sync()
    order, member = ctx.get_tile_order()
    assert order == G.TILE_ORDER_INDEX
    np.testing.assert_array_equal(member, np.arange(full.n))


def test_tile_order_switch_and_untiled(ctx):
    q, box, half, full = periodic(8, 1.0)
    ctx.set_kernel_config(4, 256, 4)
    ctx.set_tiling(64)
    dq = ctx.array(q)
    ctx.build_list(dq, q.shape[0], box, G.LIST_FULL)
    assert ctx.get_tile_order()[0] == G.TILE_ORDER_SPATIAL
    ctx.set_tile_order(G.TILE_ORDER_INDEX)
    order, member = ctx.get_tile_order()
    assert order == G.TILE_ORDER_INDEX
    np.testing.assert_array_equal(member, np.arange(full.n))
    ctx.set_tile_order(G.TILE_ORDER_SPATIAL)
    dp = ctx.zeros(q.shape)
    ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, 1)
    assert ctx.get_tile_order()[0] == G.TILE_ORDER_SPATIAL
    with pytest.raises(G.LJGPUError) as e:
        ctx.set_tile_order(2)
    assert e.value.code == G.ERR_INVALID_ARGUMENT
    ctx.set_tiling(0)
    assert ctx.get_tile_order() == (-1, None)
    ctx.set_tiling(64)
