"""Autotuner: every candidate timed, the fastest installed, results unchanged."""
import numpy as np
import pytest

import lj_gpu_amd as G
from helpers import FP64_TOL, PBC_DT, STEPS, normwise, oracle_momenta, periodic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    if G.device_count() < 1:
        pytest.skip("no HIP device")
    c = G.Context(0)
    yield c
    c.close()


def test_autotune_installs_the_fastest_configuration(ctx):
    q, box, half, full = periodic(12, 1.0)
    ctx.set_half_mode(G.HALF_AUTO)
    ctx.set_list(half)
    dq = ctx.array(q)
    best, table = ctx.autotune(dq, box, reps=16)
    assert len(table) == 40
    assert best["us_per_step"] == min(e["us_per_step"] for e in table)
    assert all(0 < e["us_min"] <= e["us_per_step"] for e in table)
    assert {e["chains"] for e in table} == {1, 2}
    info = ctx.list_info()
    assert (info.lanes_per_particle, info.unroll) == (best["lanes_per_particle"], best["unroll"])
    assert info.tile_particles == best["tile_used"]
    # The tuned configuration still meets the parity bar.
    dp = ctx.zeros(q.shape)
    ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, STEPS)
    ctx.sync()
    assert normwise(dp.download()[:, :3], oracle_momenta(12, 1.0)) < FP64_TOL


def test_autotune_mixed_and_replicas(ctx):
    q, off, boxes = G.replica_workload(8, seed=3, cells=(6, 10))
    ctx.set_replicas(off, boxes)
    dq = ctx.array(q)
    ctx.build_list(dq, q.shape[0], 0.0, G.LIST_FULL)
    best, table = ctx.autotune(dq, 0.0, precision=G.PREC_MIXED, reps=5)
    # Replicas run on the tiled kernel only: the untiled and ELL candidates
    # are skipped.
    assert len(table) == 37
    assert all(e["tile_used"] > 0 for e in table)
    assert best["tile_used"] > 0
    ctx.set_replicas([], [])


def test_autotune_requires_a_list(ctx):
    c = G.Context(0)
    with pytest.raises(G.LJGPUError) as e:
        c.autotune(c.zeros((4, 4)))
    assert e.value.code == G.ERR_NO_LIST
    c.close()
