"""Molecular dynamics around the force path (no reference counterpart: the
reference only repeats force evaluations on fixed positions).

Pins: the pair potential by two-body closed forms, by an all-pairs numpy sum
and by F = -dU/dq (central differences against the force kernel, itself
pinned to the oracle); the integrator by a numpy velocity-Verlet loop driven
by the oracle's forces and lists, by time reversibility and by energy
conservation; graph replay by bitwise equality with individual launches.
"""
import numpy as np
import pytest

import lj_gpu_amd as G
import oracle as O

pytestmark = pytest.mark.gpu

DT = 0.005


@pytest.fixture(scope="module")
def ctx():
    if G.device_count() < 1:
        pytest.skip("no HIP device")
    c = G.Context(0)
    yield c
    c.close()


def liquid(k, rho=0.8, seed=7, temperature=1.0):
    q, box = G.lattice_periodic(k, rho, 2)
    p = G.maxwell_boltzmann(q.shape[0], temperature, seed)
    return q, p, box


def u_pairs(r2, cl2=G.CL2):
    s3 = (1.0 / r2) ** 3
    return np.where(r2 <= cl2, 4.0 * s3 * (s3 - 1.0), 0.0)


def brute_potential(q, box):
    x = q[:, :3]
    u = 0.0
    for i in range(x.shape[0] - 1):
        d = x[i + 1:] - x[i]
        if box > 0:
            d -= box * np.rint(d / box)
        u += u_pairs((d * d).sum(axis=1)).sum()
    return u


def brute_pairs(q, box):
    x = q[:, :3]
    c = 0
    for i in range(x.shape[0] - 1):
        d = x[i + 1:] - x[i]
        if box > 0:
            d -= box * np.rint(d / box)
        c += int(((d * d).sum(axis=1) <= G.CL2).sum())
    return c


def test_energy_two_body_closed_forms(ctx):
    rc = 2.0 ** (1.0 / 6.0)
    for r, expected in [(1.0, 0.0), (rc, -1.0), (3.0, 4.0 * (3.0 ** -12 - 3.0 ** -6)),
                        (3.0 + 1e-9, 0.0), (1.5, 4.0 * (1.5 ** -12 - 1.5 ** -6))]:
        q = np.zeros((2, 4))
        q[1, 0] = r
        p = np.array([[1.0, 2.0, 2.0, 0.0], [0.0, 0.0, 0.0, 0.0]])
        csr = G.CSR(np.array([1, 0], np.int32), np.array([0, 1], np.int32),
                    np.array([1], np.int32), G.LIST_HALF)
        ctx.set_list(csr)
        u, k = ctx.energy(ctx.array(q), ctx.array(p), 0.0)
        assert u == pytest.approx(expected, abs=1e-15, rel=1e-13)
        assert k == 4.5


@pytest.mark.parametrize("kind", [G.LIST_HALF, G.LIST_FULL])
def test_energy_matches_all_pairs_sum(ctx, kind):
    q, p, box = liquid(6)
    ctx.set_half_mode(G.HALF_AUTO)
    ctx.set_list(G.makepair(q, box, kind))
    u, k, pairs = ctx.energy_pairs(ctx.array(q), ctx.array(p), box)
    assert u == pytest.approx(brute_potential(q, box), rel=1e-12)
    assert pairs == brute_pairs(q, box)
    assert k == pytest.approx(0.5 * (p[:, :3] ** 2).sum(), rel=1e-13)


@pytest.mark.parametrize("k,rho", [(6, 0.8), (12, 1.0), (31, 0.5)])
def test_tiled_energy_matches_row_energy(ctx, k, rho):
    """The potential over the LDS tiles (k_tile_energy) against the same
    sum over the CSR rows (k_potential, tiling off), and the pair counts
    exactly."""
    q, box = G.lattice_periodic(k, rho, 2)
    p = G.maxwell_boltzmann(q.shape[0], 1.0, 7)
    full = G.makepair(q, box, G.LIST_FULL)
    dq, dp = ctx.array(q), ctx.array(p)
    ctx.set_half_mode(G.HALF_AUTO)
    ctx.set_kernel_config(4, 256, 4)
    ctx.set_tiling(0)
    ctx.set_list(full)
    assert ctx.list_info().tile_particles == 0
    rows = ctx.energy_pairs(dq, dp, box)
    ctx.set_tiling(64)
    ctx.set_list(full)
    assert ctx.list_info().tile_particles > 0
    tiles = ctx.energy_pairs(dq, dp, box)
    ctx.force_aos_f64(dq, ctx.zeros(q.shape), 0.01, G.CL2, box, 1)  # spatial tiles
    spatial = ctx.energy_pairs(dq, dp, box)
    for got in (tiles, spatial):
        assert got[0] == pytest.approx(rows[0], rel=1e-13)
        assert got[1] == rows[1]
        assert got[2] == rows[2]


def test_force_is_minus_energy_gradient(ctx):
    q, p, box = liquid(6)
    full = G.makepair(q, box, G.LIST_FULL)
    ctx.set_list(full)
    dq, dp = ctx.array(q), ctx.zeros(q.shape)
    ctx.force_aos_f64(dq, dp, 1.0, G.CL2, box, 1)
    f = dp.download()
    zero = ctx.zeros(q.shape)
    # Central differences of the total energy (about -5,000): h = 1e-4 keeps
    # both the truncation and the rounding error near 1e-5.
    h = 1e-4
    for i, c in [(0, 0), (17, 1), (300, 2), (863, 0)]:
        up, dn = q.copy(), q.copy()
        up[i, c] += h
        dn[i, c] -= h
        gu, _ = ctx.energy(ctx.array(up), zero, box)
        gd, _ = ctx.energy(ctx.array(dn), zero, box)
        assert -(gu - gd) / (2 * h) == pytest.approx(f[i, c], rel=1e-4, abs=1e-3)


def numpy_verlet(q, p, box, steps, rebuild_every, dt=DT, skin=0.3):
    """Velocity Verlet with the oracle's full lists and forces."""
    q, p = q.copy(), p.copy()
    sl2 = (3.0 + skin) ** 2

    def force(q, csr):
        f = np.zeros_like(q)
        O.force_aos(csr, q, f, 1.0, O.CL2, box, steps=1, method="full")
        return f

    csr = O.makepair(q, box, "full", sl2=sl2)
    p += force(q, csr) * (0.5 * dt)
    for s in range(1, steps + 1):
        x = q[:, :3] + p[:, :3] * dt
        x -= box * np.floor(x / box)
        x[x >= box] -= box
        q[:, :3] = x
        if s % rebuild_every == 0:
            csr = O.makepair(q, box, "full", sl2=sl2)
        p += force(q, csr) * (dt if s < steps else 0.5 * dt)
    return q, p


def test_md_matches_numpy_verlet(ctx):
    q0, p0, box = liquid(6)
    steps, every = 40, 10
    qr, pr = numpy_verlet(q0, p0, box, steps, every)
    ctx.set_kernel_config(8, 256, 4)
    ctx.set_tiling(32)
    dq, dp = ctx.array(q0), ctx.array(p0)
    st, _ = ctx.md_run(dq, dp, box, steps, DT, rebuild_every=every)
    q, p = dq.download(), dp.download()
    d = q[:, :3] - qr[:, :3]
    d -= box * np.rint(d / box)
    assert np.abs(d).max() < 1e-10
    assert np.abs(p[:, :3] - pr[:, :3]).max() < 1e-9
    assert st.rebuilds == 1 + steps // every


def test_md_energy_conservation(ctx):
    """U + K with the shifted potential (the energy truncated forces conserve):
    the unshifted sum jumps by U(3) whenever a pair crosses the cutoff."""
    q, p, box = liquid(10)  # 4,000 particles
    dq, dp = ctx.array(q), ctx.array(p)
    st, samples = ctx.md_run(dq, dp, box, 1000, DT, rebuild_every=5, sample_every=100)
    e = samples[:, 3]
    print("E samples", e, "stats", st.as_dict())
    assert samples.shape[0] == 11
    assert st.max_abs_drift / abs(e[0]) < 1e-4


@pytest.mark.parametrize("every", [5, 0])
def test_md_time_reversibility(ctx, every):
    """Forward, momenta reversed, forward again: back to the start. Holds
    while every list is complete (no particle moved more than skin/2 between
    builds); a stale list breaks it (rebuild_every = 20 at T = 1 does)."""
    q0, p0, box = liquid(6)
    dq, dp = ctx.array(q0), ctx.array(p0)
    st, _ = ctx.md_run(dq, dp, box, 200, DT, rebuild_every=every)
    assert st.stale_lists == 0
    p = dp.download()
    p[:, :3] *= -1.0
    dp.upload(p)
    ctx.md_run(dq, dp, box, 200, DT, rebuild_every=every)
    q = dq.download()
    d = q[:, :3] - q0[:, :3]
    d -= box * np.rint(d / box)
    assert np.abs(d).max() < 1e-10
    assert np.abs(dp.download()[:, :3] + p0[:, :3]).max() < 1e-10


def test_md_adaptive_rebuild(ctx):
    q0, p0, box = liquid(6)
    dq, dp = ctx.array(q0), ctx.array(p0)
    st, _ = ctx.md_run(dq, dp, box, 200, DT, rebuild_every=0)
    assert st.max_displacement <= 0.15
    assert 5 < st.rebuilds < 100
    assert st.stale_lists == 0


def test_md_graph_replay_is_bitwise_eager(ctx):
    q0, p0, box = liquid(6)
    out = []
    for graphs in (True, False):
        ctx.set_graphs(graphs)
        dq, dp = ctx.array(q0), ctx.array(p0)
        ctx.md_run(dq, dp, box, 60, DT, rebuild_every=20, sample_every=30)
        out.append((dq.download(), dp.download()))
    ctx.set_graphs(True)
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_force_loop_graph_replay_is_bitwise_eager(ctx):
    q, _, box = liquid(8)
    ctx.set_list(G.makepair(q, box, G.LIST_FULL))
    dq = ctx.array(q)
    res = []
    for graphs in (True, False, True):
        ctx.set_graphs(graphs)
        dp = ctx.zeros(q.shape)
        ctx.force_aos_f64(dq, dp, DT, G.CL2, box, 250)  # two replays of 100 + 50 launches
        ctx.sync()
        res.append(dp.download())
    np.testing.assert_array_equal(res[0], res[1])
    np.testing.assert_array_equal(res[2], res[1])


def test_md_rebuild_interval_invariance(ctx):
    q0, p0, box = liquid(6)
    out = []
    for every in (1, 10):
        dq, dp = ctx.array(q0), ctx.array(p0)
        ctx.md_run(dq, dp, box, 50, DT, rebuild_every=every)
        out.append(dq.download())
    d = out[0][:, :3] - out[1][:, :3]
    d -= box * np.rint(d / box)
    assert np.abs(d).max() < 1e-11


def test_md_stale_lists_are_reported(ctx):
    q0, p0, box = liquid(6)
    dq, dp = ctx.array(q0), ctx.array(p0)
    st, _ = ctx.md_run(dq, dp, box, 40, DT, skin=0.0, rebuild_every=20)
    assert st.stale_lists == 2
    assert st.max_displacement > 0


def test_md_mixed_precision(ctx):
    q, p, box = liquid(10)
    drifts = {}
    for prec in (G.PREC_F64, G.PREC_MIXED):
        dq, dp = ctx.array(q), ctx.array(p)
        st, samples = ctx.md_run(dq, dp, box, 400, DT, rebuild_every=5, sample_every=50,
                                 precision=prec)
        drifts[prec] = st.max_abs_drift / abs(samples[0, 3])
    print("relative drifts", drifts)
    assert drifts[G.PREC_MIXED] < 1e-4


def test_md_zero_steps_changes_nothing(ctx):
    q0, p0, box = liquid(6)
    dq, dp = ctx.array(q0), ctx.array(p0)
    st, samples = ctx.md_run(dq, dp, box, 0, DT)
    np.testing.assert_array_equal(dq.download(), q0)
    np.testing.assert_array_equal(dp.download(), p0)
    assert st.samples == 1 and samples.shape[0] == 1


def test_md_crowded_cell_falls_back_to_sorted_rows(ctx):
    """A cell holding more particles than the per-cell ordering handles
    (here 294 on a 7x7x6 lattice of spacing 0.1) makes the MD rebuild sort
    its rows; the run is still deterministic and its list the exact full
    list. The time step is so small that no position changes (the largest
    force is about 1e15, so a drift moves a particle by about 1e-25), so both
    kicks of the one step use the list of the starting positions."""
    rng = np.random.default_rng(3)
    box = 20.0
    g = np.stack(np.meshgrid(np.arange(7), np.arange(7), np.arange(6), indexing="ij"), -1)
    crowd = 5.0 + 0.1 * g.reshape(-1, 3)
    spread = box * rng.random((706, 3))
    q = np.zeros((1000, 4))
    q[:, :3] = np.concatenate([crowd, spread])
    p = np.zeros_like(q)
    out = []
    for _ in range(2):
        dq, dp = ctx.array(q), ctx.array(p)
        st, _ = ctx.md_run(dq, dp, box, 1, 1e-20, skin=0.3, rebuild_every=1)
        out.append((dq.download(), dp.download(), st.pair_entries, st.rebuilds))
    np.testing.assert_array_equal(out[0][0], q)
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert np.all(np.isfinite(out[0][1])) and np.abs(out[0][1]).max() > 0
    full = G.makepair(q, box, G.LIST_FULL, search_length=3.3)
    assert out[0][3] == 2  # the first build and the one at step 1
    assert out[0][2] == 2 * full.n_entries  # two kicks over the same list


def test_md_slot_order_copies_return_to_caller_order(ctx):
    """In a periodic box the loop runs on spatially re-sorted copies of q and
    p and writes them back in the caller's order: a run with a rebuild at
    every step equals, particle by particle, one with a single build (the
    same dynamics, different slot orders), and no list stays installed."""
    q0, p0, box = liquid(6)
    out = []
    for every in (1, 1000):
        dq, dp = ctx.array(q0), ctx.array(p0)
        ctx.md_run(dq, dp, box, 5, DT, skin=0.3, rebuild_every=every)
        out.append((dq.download(), dp.download()))
    d = out[0][0][:, :3] - out[1][0][:, :3]
    d -= box * np.rint(d / box)
    assert np.abs(d).max() < 1e-12
    assert np.abs(out[0][1][:, :3] - out[1][1][:, :3]).max() < 1e-10
    # particles moved by about |p| dt per step: the caller's order is kept
    dq0 = out[0][0][:, :3] - q0[:, :3]
    dq0 -= box * np.rint(dq0 / box)
    assert np.abs(dq0).max() < 0.2
    with pytest.raises(G.LJGPUError) as e:
        ctx.list_info()
    assert e.value.code == G.ERR_NO_LIST
    # the caller's tile order is restored, and the bank-class row order with
    # it (the run tiles its own lists without it): the device tiles of a new
    # list equal the host model's
    full = G.makepair(q0, box, G.LIST_FULL)
    ctx.set_list(full)
    dq, dp = ctx.array(q0), ctx.zeros(q0.shape)
    ctx.force_aos_f64(dq, dp, DT, G.CL2, box, 1)
    order, member = ctx.get_tile_order()
    assert order == G.TILE_ORDER_SPATIAL
    t = ctx.get_tiles()
    info = ctx.list_info()
    h = G.make_tiles(full, info.tile_particles, member, q0, box, G.CL2)
    assert np.all(h.desc[:-1, 1] == info.tile_particles)
    np.testing.assert_array_equal(t.desc, h.desc)
    np.testing.assert_array_equal(t.local, h.local)
