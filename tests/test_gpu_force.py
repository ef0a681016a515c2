"""GPU parity: the HIP kernels of libljgpu.so against the FP64 oracle.

Bar: FP64 kernels within 1e-12 normwise (max|dp| / max|p|) of the oracle's
force_sorted after 100 steps; mixed precision within its stated bound of the
FP64 oracle force; the reference's own problem (open box, L = 50) within
0.5e-10 of ref_data and 1e-12 of the compiled reference's 17-digit momenta.
"""
import numpy as np
import pytest

import lj_gpu_amd as G
from helpers import (FP64_TOL, PBC_DT, STEPS, REF_DATA_05, load_golden, normwise, oracle_force,
                     oracle_force_reference, oracle_momenta, periodic, reference_mode,
                     rms_normwise, to_soa)

pytestmark = pytest.mark.gpu

# Mixed precision bounds (normwise max|dF| / max|F| against the FP64 oracle):
# at density 1.0 the required 1e-6; at density 0.5, where the
# net forces are of the same size as single pair forces, 2e-6.
MIXED_TOL = {1.0: 1.0e-6, 0.5: 2.0e-6}


@pytest.fixture(scope="module")
def ctx():
    if G.device_count() < 1:
        pytest.skip("no HIP device")
    c = G.Context(0)
    yield c
    c.close()


def run_aos(ctx, q, csr, box, dt=PBC_DT, steps=STEPS, lanes=16, block=256,
            half_mode=G.HALF_AUTO, mixed=None, unroll=2, tile=64):
    ctx.set_kernel_config(lanes, block, unroll)
    ctx.set_tiling(tile)
    ctx.set_half_mode(half_mode)
    ctx.set_list(csr)
    info = ctx.list_info()
    tiled = tile > 0 and lanes > 1 and not (csr.kind == G.LIST_HALF and half_mode == G.HALF_ATOMIC)
    assert (info.tile_particles > 0) == tiled
    if tiled:
        # The requested size, capped at 1024 / lanes, halved while a block
        # overflows the LDS budget.
        want = min(tile, 1024 // lanes)
        assert info.tile_particles in [want >> s for s in range(8)]
    dq = ctx.array(q)
    dp = ctx.zeros(q.shape)
    if mixed is None:
        ctx.force_aos_f64(dq, dp, dt, G.CL2, box, steps)
    else:
        ctx.force_aos_mixed(dq, dp, dt, G.CL2, box, steps, mixed)
    ctx.sync()
    out = dp.download()
    dq.free()
    dp.free()
    return out


def run_soa(ctx, q, csr, box, dt=PBC_DT, steps=STEPS, lanes=16, mixed=None,
            half_mode=G.HALF_AUTO, tile=64):
    ctx.set_kernel_config(lanes, 256, 2)
    ctx.set_tiling(tile)
    ctx.set_half_mode(half_mode)
    ctx.set_list(csr)
    q3 = to_soa(q)
    dq = [ctx.array(np.ascontiguousarray(q3[c])) for c in range(3)]
    dp = [ctx.zeros(q.shape[0]) for _ in range(3)]
    if mixed is None:
        ctx.force_soa_f64(dq, dp, dt, G.CL2, box, steps)
    else:
        ctx.force_soa_mixed(dq, dp, dt, G.CL2, box, steps, mixed)
    ctx.sync()
    return np.stack([d.download() for d in dp], axis=1)


# ---------------------------------------------------------------------------
# FP64 parity on the periodic workloads
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("k,rho", [(12, 1.0), (31, 1.0), (31, 0.5)])
# Lanes 1 selects the ELL kernel, which is never tiled.
@pytest.mark.parametrize("lanes,tile", [(1, 0)] + [(l, t) for l in (4, 8, 16, 32, 64)
                                                   for t in (0, 64)])
@pytest.mark.parametrize("unroll", [1, 2, 4])
def test_f64_full_list_aos(ctx, k, rho, lanes, unroll, tile):
    q, box, _, full = periodic(k, rho)
    p = run_aos(ctx, q, full, box, lanes=lanes, unroll=unroll, tile=tile)
    assert normwise(p, oracle_momenta(k, rho)) < FP64_TOL


@pytest.mark.parametrize("k,rho", [(12, 1.0), (31, 1.0), (31, 0.5)])
@pytest.mark.parametrize("lanes,unroll", [(8, 4), (16, 2), (4, 1), (32, 2)])
def test_f64_tile_offsets_in_lds(ctx, k, rho, lanes, unroll):
    q, box, _, full = periodic(k, rho)
    ctx.set_tiling(64, True)
    ctx.set_kernel_config(lanes, 256, unroll)
    ctx.set_list(full)
    dq = ctx.array(q)
    dp = ctx.zeros(q.shape)
    ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, STEPS)
    p = dp.download()
    ctx.set_tiling(64, False)
    assert normwise(p, oracle_momenta(k, rho)) < FP64_TOL


@pytest.mark.parametrize("k,rho", [(12, 1.0), (31, 1.0)])
@pytest.mark.parametrize("resident", [1, 3, 8])
@pytest.mark.parametrize("stage_offsets", [False, True])
def test_f64_tile_persistent_grid(ctx, k, rho, resident, stage_offsets):
    q, box, _, full = periodic(k, rho)
    ctx.set_tiling(32, stage_offsets, resident)
    ctx.set_kernel_config(8, 256, 4)
    ctx.set_list(full)
    dq = ctx.array(q)
    dp = ctx.zeros(q.shape)
    ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, STEPS)
    p = dp.download()
    ctx.set_tiling(32, False, 0)
    assert normwise(p, oracle_momenta(k, rho)) < FP64_TOL


@pytest.mark.parametrize("k,rho", [(12, 1.0), (31, 0.5)])
@pytest.mark.parametrize("lanes,tile,unroll,mixed", [(4, 16, 4, None), (4, 16, 2, None),
                                                     (2, 32, 4, None),
                                                     (4, 16, 4, G.ACCUM_F64)])
def test_tile_staging_in_several_passes(ctx, k, rho, lanes, tile, unroll, mixed):
    """Tiles that stage more positions than one pass covers (4 per thread):
    the first pass runs ahead of the staging loop, the others in it."""
    q, box, _, full = periodic(k, rho)
    if mixed is None:
        p = run_aos(ctx, q, full, box, lanes=lanes, unroll=unroll, tile=tile)
        ref, tol = oracle_momenta(k, rho), FP64_TOL
    else:
        p = run_aos(ctx, q, full, box, dt=1.0, steps=1, lanes=lanes, unroll=unroll, tile=tile,
                    mixed=mixed)
        ref, tol = oracle_force(k, rho), MIXED_TOL[rho]
    info = ctx.list_info()
    threads = (info.tile_particles * lanes + 63) // 64 * 64
    assert info.tile_max_stage > 4 * threads
    assert normwise(p, ref) < tol


@pytest.mark.parametrize("tile", [16, 32, 128, 1024])
def test_f64_tile_sizes(ctx, tile):
    q, box, _, full = periodic(12, 1.0)
    p = run_aos(ctx, q, full, box, lanes=8, unroll=2, tile=tile)
    assert normwise(p, oracle_momenta(12, 1.0)) < FP64_TOL


def test_f64_tiles_in_small_periodic_box(ctx):
    """Box 9.5 < 2 * (tile extent + cutoff): every tile takes the per-entry
    minimum-image path."""
    q, box, _, full = periodic(6, 1.0)
    half = G.makepair(q, box, G.LIST_HALF)
    p_ref = np.zeros_like(q)
    import oracle as O
    O.force_aos(O.CSR(half.number_of_partners, half.pointer, half.sorted_list, "half"), q, p_ref,
                PBC_DT, O.CL2, box, steps=STEPS, method="sorted")
    for lanes in (8, 32):
        p = run_aos(ctx, q, full, box, lanes=lanes, unroll=2, tile=64)
        assert normwise(p, p_ref) < FP64_TOL


@pytest.mark.parametrize("order", [G.TILE_ORDER_INDEX, G.TILE_ORDER_SPATIAL])
def test_f64_tiles_with_permuted_particle_order(ctx, order):
    """Random particle order: index-order tiles shrink to fit LDS, the
    spatial order regroups the particles into full compact tiles; results
    unchanged either way."""
    q, box, _, _ = periodic(12, 1.0)
    perm = np.random.default_rng(5).permutation(q.shape[0])
    qp = np.ascontiguousarray(q[perm])
    full = G.makepair(qp, box, G.LIST_FULL)
    ctx.set_tile_order(order)
    p = run_aos(ctx, qp, full, box, lanes=8, unroll=2, tile=64)
    blocks = ctx.list_info().tile_blocks
    ctx.set_tile_order(G.TILE_ORDER_SPATIAL)
    if order == G.TILE_ORDER_INDEX:
        assert blocks > (full.n + 63) // 64
    else:
        assert blocks == (full.n + 63) // 64
    assert normwise(p, oracle_momenta(12, 1.0)[perm]) < FP64_TOL


@pytest.mark.parametrize("k,rho", [(12, 1.0), (31, 1.0), (31, 0.5)])
@pytest.mark.parametrize("mode", [G.HALF_ATOMIC, G.HALF_EXPAND])
@pytest.mark.parametrize("lanes,unroll", [(16, 2), (1, 4), (8, 1)])
def test_f64_half_list_aos(ctx, k, rho, mode, lanes, unroll):
    q, box, half, _ = periodic(k, rho)
    p = run_aos(ctx, q, half, box, half_mode=mode, lanes=lanes, unroll=unroll)
    assert ctx.list_info().half_mode == mode
    assert normwise(p, oracle_momenta(k, rho)) < FP64_TOL


@pytest.mark.parametrize("k,rho", [(12, 1.0), (31, 1.0), (31, 0.5)])
@pytest.mark.parametrize("kind", ["full", "half_atomic", "half_expand"])
def test_f64_soa(ctx, k, rho, kind):
    q, box, half, full = periodic(k, rho)
    csr = full if kind == "full" else half
    mode = G.HALF_ATOMIC if kind == "half_atomic" else G.HALF_EXPAND
    p = run_soa(ctx, q, csr, box, half_mode=mode)
    assert normwise(p, oracle_momenta(k, rho)) < FP64_TOL


def run_aos3(ctx, q, csr, box, dt=PBC_DT, steps=STEPS, lanes=4, unroll=4, tile=64, mixed=None,
             half_mode=G.HALF_AUTO):
    """As run_aos over packed double[N][3] q and p."""
    ctx.set_kernel_config(lanes, 256, unroll)
    ctx.set_tiling(tile)
    ctx.set_half_mode(half_mode)
    ctx.set_list(csr)
    dq = ctx.array(np.ascontiguousarray(q[:, :3]))
    dp = ctx.zeros((q.shape[0], 3))
    if mixed is None:
        ctx.force_aos3_f64(dq, dp, dt, G.CL2, box, steps)
    else:
        ctx.force_aos3_mixed(dq, dp, dt, G.CL2, box, steps, mixed)
    ctx.sync()
    out = dp.download()
    dq.free()
    dp.free()
    return out


@pytest.mark.parametrize("k,rho", [(12, 1.0), (31, 1.0), (31, 0.5)])
@pytest.mark.parametrize("kind", ["full", "half_atomic", "half_expand"])
def test_f64_aos3(ctx, k, rho, kind):
    """Packed double3 rows (the reference's double3 kernels): the oracle's
    momenta at 1e-12, and the double4 call's momenta bit for bit (same
    arithmetic), both over 100 steps (internal slot copies) and 3 steps
    (the double3 kernels themselves)."""
    q, box, half, full = periodic(k, rho)
    csr = full if kind == "full" else half
    mode = G.HALF_ATOMIC if kind == "half_atomic" else G.HALF_EXPAND
    p3 = run_aos3(ctx, q, csr, box, half_mode=mode)
    assert normwise(p3, oracle_momenta(k, rho)) < FP64_TOL
    p4 = run_aos(ctx, q, csr, box, lanes=4, unroll=4, half_mode=mode)
    if kind == "half_atomic":  # atomic scatter: the order of additions varies
        assert normwise(p3, p4) < 1e-14
    else:
        np.testing.assert_array_equal(p3, p4[:, :3])
    for lanes, unroll, tile in [(4, 4, 64), (16, 2, 0), (1, 2, 0)]:
        a = run_aos3(ctx, q, csr, box, steps=3, lanes=lanes, unroll=unroll, tile=tile,
                     half_mode=mode)
        b = run_aos(ctx, q, csr, box, steps=3, lanes=lanes, unroll=unroll, tile=tile,
                    half_mode=mode)
        if kind == "half_atomic":  # atomic scatter: the order of additions varies
            assert normwise(a, b) < 1e-14
        else:
            np.testing.assert_array_equal(a, b[:, :3])


@pytest.mark.parametrize("accum", [G.ACCUM_F64, G.ACCUM_KAHAN_F32])
@pytest.mark.parametrize("k,rho", [(12, 1.0), (31, 1.0)])
def test_mixed_aos3(ctx, k, rho, accum):
    q, box, _, full = periodic(k, rho)
    F = oracle_force(k, rho)
    m3 = run_aos3(ctx, q, full, box, dt=1.0, steps=1, mixed=accum)
    assert normwise(m3, F) < MIXED_TOL[rho]
    m4 = run_aos(ctx, q, full, box, dt=1.0, steps=1, lanes=4, unroll=4, mixed=accum)
    np.testing.assert_array_equal(m3, m4[:, :3])
    m3 = run_aos3(ctx, q, full, box, dt=0.01, steps=100, mixed=accum)
    m4 = run_aos(ctx, q, full, box, dt=0.01, steps=100, lanes=4, unroll=4, mixed=accum)
    np.testing.assert_array_equal(m3, m4[:, :3])


def test_aos3_rejects_misaligned_and_short_buffers(ctx):
    q, box, _, full = periodic(5, 1.0)
    ctx.set_list(full)
    n = q.shape[0]
    dq = ctx.array(np.ascontiguousarray(q[:, :3]))
    dp = ctx.zeros((n, 3))
    with pytest.raises(ValueError):
        ctx.force_aos3_f64(dq, ctx.zeros((n - 1, 3)), PBC_DT, G.CL2, box, 1)
    rc = G.lib().ljgpu_force_aos3_f64(ctx.handle, dq.ptr + 4, dp.ptr, PBC_DT, G.CL2, box, 1)
    assert rc == G.ERR_INVALID_ARGUMENT
    dq.free()
    dp.free()


@pytest.mark.parametrize("lanes,block", [(1, 64), (4, 1024), (64, 64), (16, 512), (32, 128)])
def test_f64_block_sizes(ctx, lanes, block):
    q, box, _, full = periodic(12, 1.0)
    p = run_aos(ctx, q, full, box, lanes=lanes, block=block)
    assert normwise(p, oracle_momenta(12, 1.0)) < FP64_TOL


def test_expanded_half_list_is_the_full_list(ctx):
    q, box, half, full = periodic(12, 1.0)
    ctx.set_half_mode(G.HALF_EXPAND)
    ctx.set_list(half)
    got = ctx.get_list()
    np.testing.assert_array_equal(got.number_of_partners, full.number_of_partners)
    np.testing.assert_array_equal(got.pointer, full.pointer)
    np.testing.assert_array_equal(got.sorted_list, full.sorted_list)


def test_shuffled_rows_and_transposed_list(ctx):
    q, box, _, full = periodic(12, 1.0)
    ref = oracle_momenta(12, 1.0)
    sh = G.random_shfl(full, 10)
    assert normwise(run_aos(ctx, q, sh, box, lanes=16), ref) < FP64_TOL
    w = int(full.number_of_partners.max())
    ell = G.make_transposed(full, w)
    ctx.set_kernel_config(1, 256)
    ctx.set_list_transposed(full.n, full.number_of_partners, ell, w, G.LIST_FULL)
    dq = ctx.array(q)
    dp = ctx.zeros(q.shape)
    ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, STEPS)
    assert normwise(dp.download(), ref) < FP64_TOL


def test_device_resident_list_input(ctx):
    q, box, _, full = periodic(12, 1.0)
    d_nop = ctx.array(full.number_of_partners)
    d_ptr = ctx.array(full.pointer)
    d_lst = ctx.array(full.sorted_list)
    ctx.set_kernel_config(16, 256)
    ctx.set_list_device(full.n, d_nop, d_ptr, d_lst, full.n_entries, G.LIST_FULL)
    dq = ctx.array(q)
    dp = ctx.zeros(q.shape)
    ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, STEPS)
    assert normwise(dp.download(), oracle_momenta(12, 1.0)) < FP64_TOL


def test_momenta_accumulate_across_calls(ctx):
    q, box, _, full = periodic(12, 1.0)
    ctx.set_kernel_config(16, 256)
    ctx.set_list(full)
    dq = ctx.array(q)
    dp = ctx.zeros(q.shape)
    for _ in range(4):
        ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, STEPS // 4)
    assert normwise(dp.download(), oracle_momenta(12, 1.0)) < FP64_TOL


# ---------------------------------------------------------------------------
# The reference's own problem: open box, L = 50, dt = 0.001, 100 steps
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode", [G.HALF_ATOMIC, G.HALF_EXPAND])
def test_reference_mode_matches_ref_data(ctx, mode):
    q, half = reference_mode(0.5)
    p = run_aos(ctx, q, half, 0.0, dt=G.REF_DT, steps=100, half_mode=mode)
    n = p.shape[0]
    rows = list(range(5)) + list(range(n - 5, n))
    ref = np.loadtxt(REF_DATA_05)
    assert np.abs(p[rows, :3] - ref).max() <= 0.51e-10
    g = load_golden("ref_aos_L50_rho0.5")
    r = g["rows"]
    for kind in ("pair", "sorted", "intrin", "mat_transpose"):
        assert np.abs(p[r, :3] - g[f"p_{kind}_rows"]).max() / float(g[f"p_{kind}_absmax"]) < 1e-12


def test_reference_kernel_abi_entry(ctx):
    """ljgpu_force_kernel_csr_f64 takes the reference kernel's argument list
    (device pointers, AoS double4, open box)."""
    q, half = reference_mode(0.5)
    full = G.makepair(q, 0.0, G.LIST_FULL)
    g = load_golden("ref_aos_L50_rho0.5")
    r = g["rows"]
    for csr in (full, half):
        dq = ctx.array(q)
        dp = ctx.zeros(q.shape)
        d_lst = ctx.array(csr.sorted_list)
        d_nop = ctx.array(csr.number_of_partners)
        d_ptr = ctx.array(csr.pointer)
        for _ in range(100):
            ctx.force_kernel_csr_f64(dq, dp, csr.n, G.REF_DT, G.CL2, d_lst, d_nop, d_ptr,
                                     csr.n_entries, csr.kind)
        ctx.sync()
        p = dp.download()
        assert np.abs(p[r, :3] - g["p_sorted_rows"]).max() / float(g["p_sorted_absmax"]) < 1e-12


# ---------------------------------------------------------------------------
# Mixed precision
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("k,rho", [(12, 1.0), (31, 1.0), (31, 0.5)])
@pytest.mark.parametrize("accum", [G.ACCUM_F64, G.ACCUM_KAHAN_F32])
@pytest.mark.parametrize("lanes,unroll", [(16, 2), (1, 4), (64, 1), (8, 4)])
def test_mixed_force_error_bound(ctx, k, rho, accum, lanes, unroll):
    q, box, _, full = periodic(k, rho)
    F = oracle_force(k, rho)
    p = run_aos(ctx, q, full, box, dt=1.0, steps=1, mixed=accum, lanes=lanes, unroll=unroll)
    err = normwise(p, F)
    rms = rms_normwise(p, F)
    print(f"mixed k={k} rho={rho} accum={accum}: max {err:.3e} rms {rms:.3e}")
    assert err < MIXED_TOL[rho]
    p_soa = run_soa(ctx, q, full, box, dt=1.0, steps=1, mixed=accum)
    assert normwise(p_soa, F) < MIXED_TOL[rho]


@pytest.mark.parametrize("accum", [G.ACCUM_F64, G.ACCUM_KAHAN_F32])
@pytest.mark.parametrize("lanes,unroll", [(4, 4), (16, 2), (1, 4)])
def test_mixed_open_box_reference_problem(ctx, accum, lanes, unroll):
    """Mixed precision on the reference's own problem (open box, L = 50,
    density 0.5): the open-box tiled kernels (4 and 16 lanes per particle)
    and the ELL kernel (1 lane) within the density-0.5 bound of the oracle
    force."""
    q, half = reference_mode(0.5)
    F = oracle_force_reference(0.5)
    p = run_aos(ctx, q, half, 0.0, dt=1.0, steps=1, mixed=accum, lanes=lanes, unroll=unroll)
    err = normwise(p, F)
    print(f"mixed open box accum={accum} lanes={lanes}: max {err:.3e}")
    assert err < MIXED_TOL[0.5]


def test_mixed_half_list_atomic(ctx):
    q, box, half, _ = periodic(12, 1.0)
    F = oracle_force(12, 1.0)
    p = run_aos(ctx, q, half, box, dt=1.0, steps=1, mixed=G.ACCUM_F64, half_mode=G.HALF_ATOMIC)
    assert normwise(p, F) < MIXED_TOL[1.0]


# ---------------------------------------------------------------------------
# Edge cases
# ---------------------------------------------------------------------------
def _two_body(ctx, r_vec, box=0.0, dt=1.0, origin=0.0):
    q = np.zeros((2, 4))
    q[0, :3] = origin
    q[1, :3] = q[0, :3] + np.asarray(r_vec, float)
    if box > 0:
        q[:, :3] %= box
    csr = G.CSR(np.array([1, 1], np.int32), np.array([0, 1], np.int32),
                np.array([1, 0], np.int32), G.LIST_FULL)
    return run_aos(ctx, q, csr, box, dt=dt, steps=1, lanes=16)


def test_two_body_known_answers(ctx):
    p = _two_body(ctx, [1.0, 0.0, 0.0], dt=0.5)
    np.testing.assert_allclose(p[:, 0], [-12.0, 12.0], rtol=1e-15)
    r0 = 2.0 ** (1.0 / 6.0)
    assert np.abs(_two_body(ctx, [r0, 0, 0])).max() < 1e-13
    p_in = _two_body(ctx, [3.0, 0.0, 0.0])  # r2 == CL2 exactly: kept
    assert p_in[0, 0] == pytest.approx(3.0 * (24.0 * 729.0 - 48.0) / 3.0 ** 14, rel=1e-14)
    assert np.abs(_two_body(ctx, [np.nextafter(3.0, 4.0), 0, 0])).max() == 0.0
    # periodic pair: stored at x = 0.5 and 9.4 in a box of 10, minimum image -1.1
    p_wrap = _two_body(ctx, [10.0 - 1.1, 0, 0], box=10.0, origin=0.5)
    d = (0.5 + (10.0 - 1.1)) - 0.5 - 10.0
    r2 = d * d
    expected = (24.0 * r2 ** 3 - 48.0) / (r2 ** 7) * d
    np.testing.assert_allclose(p_wrap[:, 0], [expected, -expected], rtol=1e-14)
    assert np.abs(p_wrap[:, 1:3]).max() == 0.0


def test_empty_rows_and_empty_list(ctx):
    q, box, _, full = periodic(12, 1.0)
    empty = G.CSR(np.zeros(full.n, np.int32), np.zeros(full.n, np.int32),
                  np.zeros(0, np.int32), G.LIST_FULL)
    p = run_aos(ctx, q, empty, box)
    assert np.abs(p).max() == 0.0
    # isolated particle: a row of length zero among normal rows
    q2 = np.vstack([q, [[box / 2, box / 2, box / 2, 0.0]]])
    q2[-1, :3] = [0.5 * box, 0.5 * box, 0.5 * box]
    c2 = G.makepair(q2, box, G.LIST_FULL)
    p2 = run_aos(ctx, q2, c2, box, steps=1)
    assert np.isfinite(p2).all()
    # N = 0
    ctx.set_list(G.CSR(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32),
                       G.LIST_FULL))
    dq = ctx.zeros((1, 4))
    ctx.force_aos_f64(dq, dq, 0.01, G.CL2, 0.0, 3)
    ctx.sync()


def test_bad_lists_are_rejected_before_any_launch(ctx):
    q, box, half, _ = periodic(12, 1.0)
    bad = G.CSR(half.number_of_partners.copy(), half.pointer.copy(), half.sorted_list.copy(),
                G.LIST_HALF)
    bad.sorted_list[10] = half.n + 3
    with pytest.raises(G.LJGPUError) as e:
        ctx.set_list(bad)
    assert e.value.code == G.ERR_BAD_LIST
    with pytest.raises(G.LJGPUError) as e:
        ctx.set_kernel_config(3, 256)
    assert e.value.code == G.ERR_INVALID_ARGUMENT
    with pytest.raises(G.LJGPUError) as e:
        ctx.set_kernel_config(16, 100)
    assert e.value.code == G.ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("layout", ["aos", "soa"])
@pytest.mark.parametrize("prec", ["f64", G.ACCUM_F64, G.ACCUM_KAHAN_F32])
@pytest.mark.parametrize("graphs", [True, False])
def test_slot_order_copies_are_bitwise_per_step(ctx, layout, prec, graphs):
    """A call of 20 steps over spatially ordered tiles runs over copies of q
    and p in slot order; 20 calls of one step run over the caller's arrays
    through the slot order. Same arithmetic: the momenta agree bit for bit
    (and the caller's positions are untouched)."""
    q, box, _, full = periodic(12, 1.0)
    ctx.set_graphs(graphs)
    ctx.set_kernel_config(4, 256, 4)
    ctx.set_tiling(64)
    ctx.set_half_mode(G.HALF_AUTO)
    ctx.set_list(full)
    res = []
    for calls, steps in ((1, 20), (20, 1)):
        if layout == "aos":
            dq, dp = ctx.array(q), ctx.zeros(q.shape)
        else:
            q3 = to_soa(q)
            dq = [ctx.array(np.ascontiguousarray(q3[c])) for c in range(3)]
            dp = [ctx.zeros(q.shape[0]) for _ in range(3)]
        for _ in range(calls):
            if layout == "aos" and prec == "f64":
                ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, steps)
            elif layout == "aos":
                ctx.force_aos_mixed(dq, dp, PBC_DT, G.CL2, box, steps, prec)
            elif prec == "f64":
                ctx.force_soa_f64(dq, dp, PBC_DT, G.CL2, box, steps)
            else:
                ctx.force_soa_mixed(dq, dp, PBC_DT, G.CL2, box, steps, prec)
        ctx.sync()
        if layout == "aos":
            np.testing.assert_array_equal(dq.download(), q)
            res.append(dp.download())
        else:
            res.append(np.stack([d.download() for d in dp], 1))
    ctx.set_graphs(True)
    assert ctx.get_tile_order()[0] == G.TILE_ORDER_SPATIAL
    np.testing.assert_array_equal(res[0], res[1])
    if prec == "f64":
        assert normwise(res[0][:, :3], oracle_momenta(12, 1.0, steps=20)) < FP64_TOL


@pytest.mark.parametrize("prec", ["f64", G.ACCUM_F64, G.ACCUM_KAHAN_F32])
@pytest.mark.parametrize("chains,graphs", [(2, True), (1, True), (1, False)])
def test_many_step_calls_carry_the_callers_momenta(ctx, prec, chains, graphs):
    """A call of 16 or more steps over spatially ordered tiles copies q and p
    into slot order and p back. From nonzero momenta whose fourth value per
    row no step changes: calls of 37, of 21 + 16 and of 1 step each end bit
    for bit alike, with the fourth values unchanged and (FP64) the oracle's
    momenta added to the starting ones."""
    q, box, _, full = periodic(12, 1.0)
    p0 = np.random.default_rng(5).standard_normal(q.shape)
    ctx.set_graphs(graphs)
    ctx.set_chains(chains)
    ctx.set_kernel_config(4, 256, 4)
    ctx.set_tiling(64)
    ctx.set_half_mode(G.HALF_AUTO)
    ctx.set_list(full)
    res = []
    for plan in ((37,), (21, 16), (1,) * 37):
        dq, dp = ctx.array(q), ctx.array(p0)
        for steps in plan:
            if prec == "f64":
                ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, steps)
            else:
                ctx.force_aos_mixed(dq, dp, PBC_DT, G.CL2, box, steps, prec)
        ctx.sync()
        np.testing.assert_array_equal(dq.download(), q)
        res.append(dp.download())
        dq.free()
        dp.free()
    ctx.set_chains(2)
    ctx.set_graphs(True)
    assert ctx.get_tile_order()[0] == G.TILE_ORDER_SPATIAL
    for r in res:
        np.testing.assert_array_equal(r[:, 3], p0[:, 3])
    np.testing.assert_array_equal(res[0], res[1])
    np.testing.assert_array_equal(res[0], res[2])
    if prec == "f64":
        ref = p0[:, :3] + oracle_momenta(12, 1.0, steps=37)[:, :3]
        assert normwise(res[0][:, :3], ref) < FP64_TOL


def test_graph_cache_reuses_graphs_across_step_counts_and_precisions(ctx):
    """Calls that alternate step counts and precisions replay cached graphs
    (64/16/4-launch chunks, captured once per argument set) and give the
    same momenta, bit for bit, as individual launches."""
    q, box, _, full = periodic(12, 1.0)
    ctx.set_kernel_config(4, 256, 4)
    ctx.set_tiling(64)
    ctx.set_half_mode(G.HALF_AUTO)
    ctx.set_list(full)
    plan = [("f64", 20), ("mixed", 20), ("f64", 50), ("f64", 20), ("mixed", 5), ("f64", 100),
            ("mixed", 20), ("f64", 50), ("f64", 3)]
    res = {}
    for graphs in (False, True):
        ctx.set_graphs(graphs)
        dq, dp = ctx.array(q), ctx.zeros(q.shape)
        counts = []
        for prec, steps in plan:
            if prec == "f64":
                ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, steps)
            else:
                ctx.force_aos_mixed(dq, dp, PBC_DT, G.CL2, box, steps, G.ACCUM_F64)
            counts.append(ctx.graph_stats()[0])
        ctx.sync()
        res[graphs] = dp.download()
        if graphs:
            # 20 = 16 + 4, 50 = 16 * 3 + 2, 100 = 64 + 16 * 2 + 4, 5 = 4 + 1:
            # the repeated (precision, chunk) pairs capture nothing new.
            new = np.diff([counts[0] - 2] + counts)
            assert list(new) == [2, 2, 0, 0, 1, 1, 0, 0, 0], counts
            assert ctx.graph_stats()[1] == 6
        dq.free()
        dp.free()
    ctx.set_graphs(True)
    np.testing.assert_array_equal(res[True], res[False])


def test_short_buffers_are_rejected_before_any_launch(ctx):
    q, box, _, full = periodic(6, 1.0)
    ctx.set_kernel_config(4, 256, 4)
    ctx.set_tiling(64)
    ctx.set_list(full)
    dq = ctx.array(q)
    short = ctx.zeros((q.shape[0] - 1, 4))
    with pytest.raises(ValueError):
        ctx.force_aos_f64(dq, short, PBC_DT, G.CL2, box, 1)
    with pytest.raises(ValueError):
        ctx.force_aos_mixed(short, dq, PBC_DT, G.CL2, box, 1)
    q3 = [ctx.zeros(q.shape[0]) for _ in range(3)]
    p3 = [ctx.zeros(q.shape[0]) for _ in range(2)] + [ctx.zeros(q.shape[0] - 1)]
    with pytest.raises(ValueError):
        ctx.force_soa_f64(q3, p3, PBC_DT, G.CL2, box, 1)
    d_nop = ctx.array(full.number_of_partners)
    d_ptr = ctx.array(full.pointer)
    d_lst = ctx.array(full.sorted_list[:-1].copy())
    with pytest.raises(ValueError):
        ctx.set_list_device(full.n, d_nop, d_ptr, d_lst, full.n_entries, G.LIST_FULL)
    with pytest.raises(ValueError):
        ctx.force_kernel_csr_f64(dq, dq, full.n, 0.01, G.CL2, d_lst, d_nop, d_ptr,
                                 full.n_entries, G.LIST_FULL)


@pytest.mark.parametrize("graphs", [True, False])
@pytest.mark.parametrize("prec", ["f64", G.ACCUM_F64])
def test_chains_are_bitwise_one_chain(ctx, graphs, prec):
    """A many-step call over tiles runs 2-4 parts of the tiles as concurrent
    launch chains; they update disjoint particles, so the momenta equal those
    of one chain bit for bit (spatial order, slot copies)."""
    q, box, _, full = periodic(12, 1.0)
    ctx.set_graphs(graphs)
    ctx.set_kernel_config(4, 256, 4)
    ctx.set_tiling(16)  # 432 tiles: both chains hold many
    ctx.set_half_mode(G.HALF_AUTO)
    ctx.set_list(full)
    res = []
    for chains in (1, 2, 3, 4):
        ctx.set_chains(chains)
        dq, dp = ctx.array(q), ctx.zeros(q.shape)
        for steps in (37, 5, 64):
            if prec == "f64":
                ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, steps)
            else:
                ctx.force_aos_mixed(dq, dp, PBC_DT, G.CL2, box, steps, prec)
        ctx.sync()
        res.append(dp.download())
    ctx.set_chains(2)
    ctx.set_graphs(True)
    ctx.set_tiling(64)
    for r in res[1:]:
        np.testing.assert_array_equal(res[0], r)
    if prec == "f64":
        assert normwise(res[1][:, :3], oracle_momenta(12, 1.0, steps=106)) < FP64_TOL


@pytest.mark.parametrize("prec", ["f64", G.ACCUM_F64])
def test_cutoff_classes_stay_exact_when_particles_move(ctx, prec):
    """The tile rows take their cutoff classes from the positions of the first
    evaluation. Positions moved afterwards (some pairs cross the cutoff both
    ways) still give the oracle's forces over the same list: every entry's r2
    is evaluated at every step, and a batch is skipped only when all its
    entries are beyond the cutoff at that step."""
    q0, box, _, full = periodic(12, 1.0)
    ctx.set_kernel_config(4, 256, 4)
    ctx.set_tiling(64)
    ctx.set_half_mode(G.HALF_AUTO)
    ctx.set_list(full)
    dq = ctx.array(q0)
    dp = ctx.zeros(q0.shape)
    ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, 1)  # tiles and classes from q0
    rng = np.random.default_rng(11)
    q1 = q0.copy()
    q1[:, :3] += rng.uniform(-0.12, 0.12, size=(q0.shape[0], 3))
    q1[:, :3] %= box
    # pairs of the list on both sides of the cutoff changed class
    i = np.repeat(np.arange(full.n), full.number_of_partners)
    j = full.sorted_list
    from helpers import cut_inner
    c0 = cut_inner(q0, box, G.CL2, i, j)
    c1 = cut_inner(q1, box, G.CL2, i, j)
    assert np.sum(c0 & ~c1) > 1000 and np.sum(~c0 & c1) > 1000
    dq1 = ctx.array(q1)
    for steps in (1, 40):
        dp1 = ctx.zeros(q1.shape)
        if prec == "f64":
            ctx.force_aos_f64(dq1, dp1, 1.0, G.CL2, box, steps)
        else:
            ctx.force_aos_mixed(dq1, dp1, 1.0, G.CL2, box, steps, prec)
        ctx.sync()
        import oracle as O
        F = np.zeros_like(q1)
        O.force_aos(O.CSR(full.number_of_partners, full.pointer, full.sorted_list, "full"), q1,
                    F, float(steps), O.CL2, box, method="full")
        tol = FP64_TOL if prec == "f64" else MIXED_TOL[1.0]
        assert normwise(dp1.download(), F) < tol
