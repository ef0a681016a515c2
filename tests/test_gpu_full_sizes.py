"""GPU parity at the workloads' full sizes.

- FCC 4*31^3 = 119,164 particles (densities 1.0 and 0.5): the device-built
  list equals the oracle's brute-force makepair, pair for pair and in order.
- FCC 4*64^3 = 1,048,576 particles at density 1.0: FP64 and mixed-precision
  (FP64 and Kahan-FP32 accumulation) momenta after 100 steps against the
  oracle force over the same list.
- 256 batched replicas (FCC 4*k^3, k uniform in 6..20, density uniform in
  [0.5, 1]): a sample of replicas, among them the largest box near density 1
  and one near density 0.5, against the oracle; every replica bit for bit
  against its own single-system run with the same kernel configuration.
- FCC 4*128^3 = 8,388,608 particles at density 0.8: the step-0 force of a
  sample of particles against a brute-force sum over all particles, and
  1,000 velocity-Verlet steps checked through properties that hold at any
  size (total momentum, the shifted energy, stale-list reporting) and the
  force at the final positions against a brute-force sum.
"""
from __future__ import annotations

import numpy as np
import pytest

import lj_gpu_amd as G
import oracle as O
from helpers import FP64_TOL, PBC_DT, normwise, rms_normwise

pytestmark = pytest.mark.gpu

MIXED_TOL = 1.0e-6  # normwise max|dF| / max|F| at density 1.0


@pytest.fixture(scope="module")
def ctx():
    if G.device_count() < 1:
        pytest.skip("no HIP device")
    c = G.Context(0)
    yield c
    c.close()


def _oracle_csr(csr: G.CSR) -> O.CSR:
    kind = "half" if csr.kind == G.LIST_HALF else "full"
    return O.CSR(csr.number_of_partners, csr.pointer, csr.sorted_list, kind)


def _same_list(a, b) -> None:
    np.testing.assert_array_equal(a.number_of_partners, b.number_of_partners)
    np.testing.assert_array_equal(a.pointer, b.pointer)
    np.testing.assert_array_equal(a.sorted_list, b.sorted_list)


# ---------------------------------------------------------------------------
# 119,164 particles: device list against the oracle's makepair
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("rho,kind", [(1.0, G.LIST_FULL), (0.5, G.LIST_HALF)])
def test_device_list_equals_oracle_makepair_k31(ctx, rho, kind):
    q, box = G.lattice_periodic(31, rho, 2)
    print(f"device list k=31 rho={rho} kind={kind}", flush=True)
    ctx.set_half_mode(G.HALF_ATOMIC)  # the half list is installed as given
    dq = ctx.array(q)
    ctx.build_list(dq, q.shape[0], box, kind)
    got = ctx.get_list()
    ref = O.makepair(q, box, "half" if kind == G.LIST_HALF else "full")
    _same_list(got, ref)
    ctx.set_half_mode(G.HALF_AUTO)


# ---------------------------------------------------------------------------
# 1,048,576 particles: FP64 and mixed precision against the oracle
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def large(ctx):
    q, box = G.lattice_periodic(64, 1.0, 2)
    ctx.set_kernel_config(4, 256, 4)
    ctx.set_tiling(64)
    dq = ctx.array(q)
    ctx.build_list(dq, q.shape[0], box, G.LIST_FULL)
    csr = ctx.get_list()
    # Oracle F (dt = 1) over the same full list.
    F = np.zeros_like(q)
    O.force_aos(_oracle_csr(csr), q, F, 1.0, O.CL2, box, steps=1, method="full")
    yield q, box, dq, csr, F
    dq.free()


def test_large_device_list_equals_host_builder(large):
    q, box, _, csr, _ = large
    _same_list(csr, G.makepair(q, box, G.LIST_FULL))
    assert csr.n_entries == 148_669_692  # about 141.8 partners per particle


@pytest.mark.parametrize("prec", ["f64", G.ACCUM_F64, G.ACCUM_KAHAN_F32])
def test_large_forces_against_oracle(ctx, large, prec):
    """100 steps of dt = 0.01 from p = 0 against the oracle's F * (100 * 0.01):
    the step sum rounds differently from one product, by far less than
    either bound."""
    q, box, dq, _, F = large
    steps = 100
    dp = ctx.zeros(q.shape)
    print(f"N={q.shape[0]} prec={prec}", flush=True)
    if prec == "f64":
        ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, box, steps)
    else:
        ctx.force_aos_mixed(dq, dp, PBC_DT, G.CL2, box, steps, prec)
    ctx.sync()
    p = dp.download()
    dp.free()
    ref = F * (steps * PBC_DT)
    err, rms = normwise(p, ref), rms_normwise(p, ref)
    print(f"  max {err:.3e} rms {rms:.3e}", flush=True)
    assert err < (FP64_TOL if prec == "f64" else MIXED_TOL)


# ---------------------------------------------------------------------------
# 256 batched replicas
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def replicas(ctx):
    q, off, boxes = G.replica_workload(256, seed=4)
    ctx.set_replicas(off, boxes)
    dq = ctx.array(q)
    ctx.build_list(dq, q.shape[0], 0.0, G.LIST_FULL)
    best, _ = ctx.autotune(dq, 0.0, reps=5)
    print(f"replicas: N={q.shape[0]} tuned {best}", flush=True)
    csr = ctx.get_list()
    p = {}
    for prec in ("f64", G.ACCUM_F64):
        dp = ctx.zeros(q.shape)
        if prec == "f64":
            ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, 0.0, 100)
        else:
            ctx.force_aos_mixed(dq, dp, 1.0, G.CL2, 0.0, 1, prec)
        ctx.sync()
        p[prec] = dp.download()
        dp.free()
    dq.free()
    yield q, off, boxes, csr, p
    ctx.set_replicas([], [])


def _replica_sample(off, boxes):
    n = np.diff(off)
    rho = n / boxes ** 3
    big_dense = int(np.argmax(np.where(rho > 0.9, n, -1)))
    sparse = int(np.argmin(np.abs(rho - 0.5)))
    return sorted({big_dense, sparse, 0, len(boxes) - 1, int(np.argmax(n)), int(np.argmin(n))})


def test_replicas_sample_against_oracle(replicas):
    q, off, boxes, csr, p = replicas
    sample = _replica_sample(off, boxes)
    n = np.diff(off)
    for r in sample:
        s = slice(off[r], off[r + 1])
        qr = np.ascontiguousarray(q[s])
        half = O.makepair(qr, boxes[r], "half")
        ref = np.zeros_like(qr)
        O.force_aos(half, qr, ref, PBC_DT, O.CL2, boxes[r], steps=100)
        F = np.zeros_like(qr)
        O.force_aos(half, qr, F, 1.0, O.CL2, boxes[r], steps=1)
        e64 = normwise(p["f64"][s], ref)
        emx = normwise(p[G.ACCUM_F64][s], F)
        print(f"replica {r}: N={n[r]} rho={n[r] / boxes[r] ** 3:.3f} f64 {e64:.2e} "
              f"mixed {emx:.2e}", flush=True)
        assert e64 < FP64_TOL, r
        assert emx < 2e-6, r


def test_replicas_list_rows_are_each_replicas_own_list(replicas):
    """The batched device list restricted to a replica equals that replica's
    host-built list shifted by its offset."""
    q, off, boxes, csr, _ = replicas
    for r in _replica_sample(off, boxes):
        s = slice(off[r], off[r + 1])
        own = G.makepair(np.ascontiguousarray(q[s]), boxes[r], G.LIST_FULL)
        nop = csr.number_of_partners[s]
        np.testing.assert_array_equal(nop, own.number_of_partners)
        a = csr.pointer[off[r]]
        rows = csr.sorted_list[a:a + int(nop.sum())]
        np.testing.assert_array_equal(rows - off[r], own.sorted_list)


def test_replicas_equal_single_system_runs(ctx, replicas):
    """Each replica's tiles are the tiles of that replica run alone with the
    same (autotuned) kernel configuration, so all 256 replicas' momenta
    equal their single-system runs bit for bit."""
    q, off, boxes, _, p = replicas
    ctx.set_replicas([], [])
    for r in range(len(boxes)):
        s = slice(off[r], off[r + 1])
        qr = np.ascontiguousarray(q[s])
        dq, dp = ctx.array(qr), ctx.zeros(qr.shape)
        ctx.build_list(dq, qr.shape[0], boxes[r], G.LIST_FULL)
        ctx.force_aos_f64(dq, dp, PBC_DT, G.CL2, boxes[r], 100)
        ctx.sync()
        got = dp.download()
        dq.free()
        dp.free()
        np.testing.assert_array_equal(p["f64"][s], got, err_msg=f"replica {r}")


# ---------------------------------------------------------------------------
# 8,388,608 particles at density 0.8
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def md_system():
    q, box = G.lattice_periodic(128, 0.8, 2)
    p = G.maxwell_boltzmann(q.shape[0], 1.0, 7)
    return q, p, box


def _brute_force(q, box, i):
    d = q[:, :3] - q[i, :3]
    d -= box * np.rint(d / box)
    r2 = (d * d).sum(axis=1)
    m = (r2 <= G.CL2) & (np.arange(q.shape[0]) != i)
    r2m = r2[m]
    r6 = r2m ** 3
    df = (24.0 * r6 - 48.0) / (r6 * r6 * r2m)
    return (df[:, None] * d[m]).sum(axis=0)


def _default_config(ctx):
    ctx.set_kernel_config(4, 256, 4)
    ctx.set_tiling(64)


def test_md_size_step0_force_against_brute_force(ctx, md_system):
    q, _, box = md_system
    _default_config(ctx)
    dq = ctx.array(q)
    ctx.build_list(dq, q.shape[0], box, G.LIST_FULL)
    dp = ctx.zeros(q.shape)
    ctx.force_aos_f64(dq, dp, 1.0, G.CL2, box, 1)
    ctx.sync()
    F = dp.download()
    dq.free()
    dp.free()
    rng = np.random.default_rng(17)
    sample = rng.choice(q.shape[0], 24, replace=False)
    ref = np.array([_brute_force(q, box, int(i)) for i in sample])
    err = normwise(F[sample, :3], ref)
    print(f"N={q.shape[0]} step-0 force, {len(sample)} particles: {err:.3e}", flush=True)
    assert err < FP64_TOL


@pytest.mark.parametrize("policy", ["every20_skin0.3", "adaptive_skin0.8"])
def test_md_full_size_properties(ctx, md_system, policy):
    q, p, box = md_system
    every, skin = (20, 0.3) if policy == "every20_skin0.3" else (0, 0.8)
    _default_config(ctx)
    dq, dp = ctx.array(q), ctx.array(p)
    p0 = p[:, :3].sum(axis=0)
    st, samples = ctx.md_run(dq, dp, box, 1000, 0.005, skin=skin, rebuild_every=every,
                             precision=G.PREC_MIXED, sample_every=100)
    p1 = dp.download()[:, :3].sum(axis=0)
    q1 = dq.download()
    dq.free()
    dp.free()
    e0 = st.potential0 + st.kinetic0
    rel = st.max_abs_drift / abs(e0)
    print(f"{policy}: rebuilds {st.rebuilds} stale {st.stale_lists} max displacement "
          f"{st.max_displacement:.3f} E0 {e0:.6f} rel drift {rel:.3e} sum p {p0} -> {p1}",
          flush=True)
    assert st.steps == 1000 and st.samples == samples.shape[0] == 11
    # The force at the final positions (a fresh device list) against the
    # all-particle brute force, for a sample of particles.
    dq1 = ctx.array(q1)
    ctx.build_list(dq1, q1.shape[0], box, G.LIST_FULL)
    dF = ctx.zeros(q1.shape)
    ctx.force_aos_f64(dq1, dF, 1.0, G.CL2, box, 1)
    ctx.sync()
    F1 = dF.download()
    dq1.free()
    dF.free()
    sample = np.random.default_rng(23).choice(q1.shape[0], 12, replace=False)
    ref = np.array([_brute_force(q1, box, int(i)) for i in sample])
    assert normwise(F1[sample, :3], ref) < FP64_TOL
    # Each pair enters both rows of the full list, so the total momentum
    # moves only by rounding: of each particle's sum, and of the pair
    # differences, which the tiles take in a frame shifted by whole boxes
    # (positions up to 219 here, so a difference carries ~3e-14 absolute
    # error). A pair missing from one of its two rows would move it by about
    # |f| dt (~0.1 to 1) per step.
    scale = np.abs(p[:, :3]).sum()
    assert np.abs(p1 - p0).max() < 1e-13 * scale
    assert np.all(q1[:, :3] >= 0.0) and np.all(q1[:, :3] < box)
    # A stale list is one at whose rebuild some particle had moved more than
    # skin / 2 since the previous build.
    if every == 0:
        assert st.stale_lists == 0
        assert st.max_displacement <= skin / 2
        assert rel < 5e-5
    else:
        assert st.rebuilds == 1 + 1000 // 20  # the initial build and one every 20 steps
        assert (st.stale_lists > 0) == (st.max_displacement > skin / 2)
        assert rel < 1e-4
