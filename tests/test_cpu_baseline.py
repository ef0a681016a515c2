"""The timed CPU baselines (oracle/cpu_baseline.cpp) compute what the oracle
computes: force_sorted and the all-core full-list loop against the oracle on
periodic workloads, and the restated force_intrin_mat_transpose against the
compiled reference's own momenta on the reference problem (open box, L = 50).
CPU only."""
import numpy as np
import pytest

import oracle as O
from helpers import FP64_TOL, PBC_DT, load_golden, normwise, oracle_momenta, periodic, \
    reference_mode


def _run(kind, csr, q, dt, box, steps):
    qa = O.aligned_zeros(q.shape)
    qa[...] = q
    p = O.aligned_zeros(q.shape)
    O.run_baseline(kind, csr, qa, p, dt, O.CL2, box, steps)
    return p


@pytest.mark.parametrize("k,rho", [(6, 1.0), (12, 1.0), (12, 0.5)])
def test_baselines_match_the_oracle_on_periodic_workloads(k, rho):
    q, box, half, full = periodic(k, rho)
    h = O.CSR(half.number_of_partners, half.pointer, half.sorted_list, "half")
    f = O.CSR(full.number_of_partners, full.pointer, full.sorted_list, "full")
    ref = oracle_momenta(k, rho, steps=10)
    for kind, csr in ((O.BASELINE_SORTED, h), (O.BASELINE_MAT_TRANSPOSE, h),
                      (O.BASELINE_FULL_OMP, f)):
        p = _run(kind, csr, q, PBC_DT, box, 10)
        assert normwise(p, ref) < FP64_TOL, kind
        assert p[:, 3].max() == 0.0 and p[:, 3].min() == 0.0


def test_mat_transpose_matches_the_compiled_reference():
    """Reference problem (open box, L = 50, density 0.5, dt = 0.001, 100
    steps): the restated force_intrin_mat_transpose against the 17-digit
    momenta of the reference's own routine."""
    q, half = reference_mode(0.5)
    h = O.CSR(half.number_of_partners, half.pointer, half.sorted_list, "half")
    g = load_golden("ref_aos_L50_rho0.5")
    assert int(g["n"]) == q.shape[0]
    p = _run(O.BASELINE_MAT_TRANSPOSE, h, q, O.REF_DT, 0.0, 100)
    r = g["rows"]
    scale = float(g["p_mat_transpose_absmax"])
    assert np.abs(p[r, :3] - g["p_mat_transpose_rows"]).max() / scale < 1e-12
    np.testing.assert_allclose(p[:, :3].sum(0), g["p_mat_transpose_sum"], atol=1e-9 * scale)
    assert np.abs(p[:, :3]).max() == pytest.approx(scale, rel=1e-12)


def test_a_dropped_minimum_image_is_caught():
    """The periodic check above is sharp: the same baseline run without the
    minimum image (box = 0) differs from the oracle by far more than the bar."""
    q, box, half, _ = periodic(6, 1.0)
    h = O.CSR(half.number_of_partners, half.pointer, half.sorted_list, "half")
    p = _run(O.BASELINE_MAT_TRANSPOSE, h, q, PBC_DT, 0.0, 10)
    assert normwise(p, oracle_momenta(6, 1.0, steps=10)) > 1e-3


def test_unknown_kind_is_rejected():
    q, box, half, _ = periodic(6, 1.0)
    h = O.CSR(half.number_of_partners, half.pointer, half.sorted_list, "half")
    with pytest.raises(ValueError):
        O.run_baseline(7, h, O.aligned_zeros(q.shape), O.aligned_zeros(q.shape), 0.01)


def test_build_info_names_the_flags():
    info = O.baseline_lib().cb_build_info().decode()
    assert "-O3" in info and "x86-64-v3" in info
