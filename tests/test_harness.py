"""The benchmark program lj_gpu_amd/force_ljgpu: the reference's
cuda/force_cuda.cu main (init, makepair + random_shfl or the pair cache,
measure(), print_results) over the C ABI. Its stdout is the reference's
print_results format, so it is compared with the reference's own golden
output, ref_data/density{0.5,1}.dat, as a reference run would be.
"""
import os
import subprocess

import numpy as np
import pytest

import lj_gpu_amd as G
from helpers import REF_DATA_05, REF_DATA_1


def run(*args, cwd=None, timeout=300):
    assert os.path.exists(G.HARNESS_PATH), (
        f"{G.HARNESS_PATH} is missing: build it with `python -m lj_gpu_amd.build`")
    return subprocess.run([G.HARNESS_PATH, *map(str, args)], cwd=cwd, capture_output=True,
                          text=True, timeout=timeout)


def momenta(out):
    rows = np.loadtxt(out.stdout.splitlines())
    assert rows.shape == (10, 3)
    return rows


def ran_or_no_device(r):
    """rc 0, or rc 1 from ljgpu_create where no HIP device is visible."""
    assert r.returncode == 0 or (r.returncode == 1 and "no HIP device" in r.stderr), r.stderr


@pytest.mark.parametrize("args, message", [
    (["63"], "THREAD_BLOCK"), (["1088"], "THREAD_BLOCK"), (["100"], "THREAD_BLOCK"),
    (["--layout", "ell"], "usage"), (["--precision", "f16"], "usage"), (["--bogus"], "usage"),
    (["--loop"], "usage"),
])
def test_rejects_bad_command_lines(args, message):
    r = run(*args)
    assert r.returncode == 1 and message in r.stderr and r.stdout == ""


@pytest.mark.parametrize("half", [False, True])
def test_pair_cache_made_then_loaded(tmp_path, half):
    """First run: makepair + random_shfl(10) + makepaircache, as the reference
    does without a cache file; second run: loadpair, no rebuild."""
    cache = tmp_path / "cache.dat"
    extra = ["--half"] if half else []
    r = run("--box-length", 12, "--cache", cache, "--loop", 1, *extra)
    ran_or_no_device(r)
    assert "Now make pairlist" in r.stderr
    q = G.lattice_reference(0.5, 12.0)
    kind = G.LIST_HALF if half else G.LIST_FULL
    expect = tmp_path / "expect.dat"
    G.write_pair_cache(str(expect), G.random_shfl(G.makepair(q, 0.0, kind), 10))
    assert cache.read_bytes() == expect.read_bytes()
    r = run("--box-length", 12, "--cache", cache, "--loop", 1, *extra)
    ran_or_no_device(r)
    assert "Now make pairlist" not in r.stderr


def test_foreign_cache_is_rebuilt(tmp_path):
    """A cache for another particle number is refused and rebuilt
    (the reference's loadpair returns false for it)."""
    cache = tmp_path / "cache.dat"
    G.write_pair_cache(str(cache), G.makepair(G.lattice_reference(0.5, 10.0), 0.0, G.LIST_FULL))
    r = run("--box-length", 12, "--cache", cache, "--loop", 1)
    ran_or_no_device(r)
    assert "rebuilding" in r.stderr and "Now make pairlist" in r.stderr


# ---------------------------------------------------------------------------
# On the device: the reference's problem and its golden output
# ---------------------------------------------------------------------------
def ref_rows(path):
    return np.loadtxt(path)


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    [], ["256"], ["1024"], ["--half"], ["--no-shuffle"], ["--layout", "soa"],
    ["--layout", "soa", "--half"],
])
def test_reference_problem_density05(args):
    r = run(*args)
    assert r.returncode == 0, r.stderr
    assert "N=62500, ljgpu_force_" in r.stderr and "(without Host<->Device)" in r.stderr
    assert np.abs(momenta(r) - ref_rows(REF_DATA_05)).max() <= 0.51e-10


@pytest.mark.gpu
def test_reference_problem_density1():
    r = run("--density", 1.0)
    assert r.returncode == 0, r.stderr
    assert "N=119164" in r.stderr
    assert np.abs(momenta(r) - ref_rows(REF_DATA_1)).max() <= 0.51e-10


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed", "kahan"])
@pytest.mark.parametrize("layout", ["aos", "soa"])
def test_reference_problem_mixed(precision, layout):
    """Mixed precision keeps the reference's output to 2e-6 of its largest
    momentum (FP32 pair term, FP64 or compensated FP32 sums)."""
    r = run("--precision", precision, "--layout", layout)
    assert r.returncode == 0, r.stderr
    ref = ref_rows(REF_DATA_05)
    assert np.abs(momenta(r) - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.gpu
def test_cached_list_gives_the_same_output(tmp_path):
    cache = tmp_path / "cache.dat"
    a = run("--cache", cache)
    b = run("--cache", cache)
    assert a.returncode == 0 and b.returncode == 0, a.stderr + b.stderr
    assert "Now make pairlist" in a.stderr and "Now make pairlist" not in b.stderr
    assert a.stdout == b.stdout


@pytest.mark.gpu
def test_zero_loops_print_zero_momenta():
    r = run("--loop", 0)
    assert r.returncode == 0, r.stderr
    assert not momenta(r).any()
