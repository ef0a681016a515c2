"""ljgpu_energy_noalloc, the allocation-free variant of ljgpu_energy (its
partial sums live in a static per-GPU scratch array instead of context
memory), called through the C ABI and compared with ljgpu_energy: the same
doubles on consecutive calls, the later ones over smaller systems, so the
scratch still holds a larger call's partials when a smaller one runs."""
from __future__ import annotations

import numpy as np
import pytest

import lj_gpu_amd as G

pytestmark = pytest.mark.gpu


def _system(k, rho, seed):
    q, box = G.lattice_periodic(k, rho, 2)
    p = G.maxwell_boltzmann(q.shape[0], 1.0, seed)
    return q, p, box


def _brute_potential(q, box):
    x = q[:, :3]
    u, c = 0.0, 0
    for i in range(x.shape[0] - 1):
        d = x[i + 1:] - x[i]
        d -= box * np.rint(d / box)
        r2 = (d * d).sum(axis=1)
        s3 = (1.0 / r2) ** 3
        inside = r2 <= G.CL2
        u += float(np.where(inside, 4.0 * s3 * (s3 - 1.0), 0.0).sum())
        c += int(inside.sum())
    return u, c


def test_energy_noalloc_matches_energy_on_shrinking_calls():
    if G.device_count() < 1:
        pytest.skip("no HIP device")
    ctx = G.Context(0)
    try:
        # 32,000 particles (125 workgroups of partials), then 4,000 (16),
        # then 864 (4) over a half list.
        cases = [(20, 0.8, 7, G.LIST_FULL), (10, 1.0, 8, G.LIST_FULL), (6, 0.5, 9, G.LIST_HALF)]
        for idx, (k, rho, seed, kind) in enumerate(cases):
            q, p, box = _system(k, rho, seed)
            ctx.set_list(G.makepair(q, box, kind))
            dq, dp = ctx.array(q), ctx.array(p)
            # The allocation-free call first: on the first case it runs before
            # the context has allocated any energy buffer.
            got = ctx.energy_pairs_noalloc(dq, dp, box)
            ref = ctx.energy_pairs(dq, dp, box)
            print(f"case {idx}: N={q.shape[0]} noalloc={got} energy={ref}", flush=True)
            assert got == ref  # bit for bit: same partials, same order of addition
            again = ctx.energy_pairs_noalloc(dq, dp, box)
            assert again == ref
            if k <= 10:
                u, c = _brute_potential(q, box)
                assert got[0] == pytest.approx(u, rel=1e-12)
                assert got[2] == c
            assert got[1] == pytest.approx(0.5 * (p[:, :3] ** 2).sum(), rel=1e-13)
    finally:
        ctx.close()


def test_energy_noalloc_shares_scratch_across_contexts():
    """A second context on the same GPU uses the same static array; calls on
    it, one after the other, give ljgpu_energy's results."""
    if G.device_count() < 1:
        pytest.skip("no HIP device")
    a, b = G.Context(0), G.Context(0)
    try:
        q1, p1, box1 = _system(8, 0.9, 3)
        q2, p2, box2 = _system(5, 0.7, 4)
        a.set_list(G.makepair(q1, box1, G.LIST_FULL))
        b.set_list(G.makepair(q2, box2, G.LIST_FULL))
        a1, ap = a.array(q1), a.array(p1)
        b2, bp = b.array(q2), b.array(p2)
        ra, rb = a.energy_pairs(a1, ap, box1), b.energy_pairs(b2, bp, box2)
        assert a.energy_pairs_noalloc(a1, ap, box1) == ra
        assert b.energy_pairs_noalloc(b2, bp, box2) == rb
        assert a.energy_pairs_noalloc(a1, ap, box1) == ra
    finally:
        a.close()
        b.close()


def test_energy_noalloc_errors():
    if G.device_count() < 1:
        pytest.skip("no HIP device")
    ctx = G.Context(0)
    try:
        q, p, box = _system(5, 1.0, 1)
        dq, dp = ctx.array(q), ctx.array(p)
        with pytest.raises(G.LJGPUError):
            ctx.energy_pairs_noalloc(dq, dp, box)  # no list installed
        ctx.set_list(G.makepair(q, box, G.LIST_FULL))
        with pytest.raises(G.LJGPUError):
            ctx.energy_pairs_noalloc(dq, dp, box, cl2=-1.0)
    finally:
        ctx.close()
