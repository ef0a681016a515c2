"""Pin of the periodic lattice generator (lattice_periodic), which has no
counterpart in the reference (its init(), cpu_ref/force_aos.cpp:43-51,
142-169, builds the open box only).

The lattice is restated here from published algorithms, independently of the
C++ the product runs: an FCC cell of side s = (density / 4)^(-1/3) with the
basis (0,0,0), (0,s/2,s/2), (s/2,0,s/2), (s/2,s/2,0), cells in z, y, x order;
each site jittered by three draws (x, y, z) of std::uniform_real_distribution
over one std::mt19937(seed) stream, then wrapped into [0, L), L = cells * s.
The draws follow libstdc++ 11 (this image's, which the product links):
generate_canonical<double, 53> takes two 32-bit outputs, u = (x0 + x1 * 2^32)
/ 2^64 with the sum rounded to double (clamped below 1), and the distribution
returns u * (hi - lo) + lo, multiply and add rounded separately (the host
code is built with -ffp-contract=off). The MT19937 stream is numpy's
MT19937 with its legacy integer seeding (init_genrand, the seeding
std::mt19937 uses; pinned in test_random_shfl.py). A wrong seed, one 32-bit
output per draw, a swapped jitter axis or a missing wrap each fail. CPU only.
"""
import numpy as np
import pytest

import lj_gpu_amd as G


def restated_lattice(cells: int, density: float, seed: int, lo: float, hi: float,
                     variant: str = ""):
    """The restatement; `variant` plants a plausible mistake (the checks
    below show that each is caught): "one_output" (one 32-bit output per
    draw), "swap_xy" (the x and y draws swapped), "no_wrap"."""
    s = 1.0 / (density * 0.25) ** (1.0 / 3.0)
    hs = s * 0.5
    L = cells * s
    n = 4 * cells ** 3
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed)
    if variant == "one_output":
        u = bg.random_raw(3 * n).astype(np.float64) / 2.0 ** 32
    else:
        raw = bg.random_raw(6 * n).astype(np.uint64)
        x0 = raw[0::2].astype(np.float64)
        x1 = raw[1::2].astype(np.float64)
        u = (x0 + x1 * 2.0 ** 32) / 2.0 ** 64
        u[u >= 1.0] = np.nextafter(1.0, 0.0)
    jit = (u * (hi - lo) + lo).reshape(n, 3)
    if variant == "swap_xy":
        jit = jit[:, [1, 0, 2]]
    iz, iy, ix = np.meshgrid(np.arange(cells), np.arange(cells), np.arange(cells), indexing="ij")
    corner = np.stack([ix.ravel() * s, iy.ravel() * s, iz.ravel() * s], 1)
    basis = np.array([[0.0, 0.0, 0.0], [0.0, hs, hs], [hs, 0.0, hs], [hs, hs, 0.0]])
    sites = (corner[:, None, :] + basis[None, :, :]).reshape(n, 3)
    v = sites + jit
    if variant != "no_wrap":
        v = np.where(v < 0.0, v + L, v)
        v = np.where(v >= L, v - L, v)
    q = np.zeros((n, 4))
    q[:, :3] = v
    return q, L


@pytest.mark.parametrize("cells,density,seed,jitter", [
    (3, 1.0, 2, (-0.05, 0.05)),
    (5, 0.5, 2, (-0.05, 0.05)),
    (4, 0.8, 7, (0.0, 0.1)),
    (6, 1.0, 123456789, (-0.3, 0.3)),
    (31, 1.0, 2, (-0.05, 0.05)),  # the headline workload, 119,164 particles
])
def test_periodic_lattice_matches_restatement(cells, density, seed, jitter):
    q, box = G.lattice_periodic(cells, density, seed, jitter)
    ref, L = restated_lattice(cells, density, seed, *jitter)
    assert box == L
    np.testing.assert_array_equal(q, ref)
    assert q[:, :3].min() >= 0.0 and q[:, :3].max() < box


@pytest.mark.parametrize("variant", ["one_output", "swap_xy", "no_wrap"])
def test_restatement_detects_plausible_mistakes(variant):
    """A planted mistake in the restatement no longer matches the product
    (jitter U(-0.05, 0.05) wraps the sites at the origin below 0)."""
    q, _ = G.lattice_periodic(3, 1.0, 2, (-0.05, 0.05))
    bad, _ = restated_lattice(3, 1.0, 2, -0.05, 0.05, variant)
    assert not np.array_equal(q, bad)
    other_seed, _ = restated_lattice(3, 1.0, 3, -0.05, 0.05)
    assert not np.array_equal(q, other_seed)
