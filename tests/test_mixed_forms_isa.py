"""VALU instructions per list entry of the mixed-precision forms on gfx950
(CPU only: the probe in tests/isa/mixed_forms.hip is compiled to gfx950
assembly with the library's own flags, never launched).

Background for the mixed-precision target: gfx950 issues a non-packed FP32
VALU instruction at the same rate as an FP64 one (packed FP32, the only
route to the doubled FP32 rate, is disabled in this build), so a form is
faster than the FP64 kernel only if it needs fewer instructions (or less LDS
traffic). The error study (test_mixed_error_study.py) admits, at density
1.0, the current form and the int32 fixed-point forms; this test counts what
each costs in the inner loop.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
from collections import Counter

import pytest

from lj_gpu_amd import build as B

PROBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "isa", "mixed_forms.hip")


def _largest_loop_ops(lines):
    labels = {}
    for idx, line in enumerate(lines):
        m = re.match(r"^(\.LBB\S+):", line)
        if m:
            labels[m.group(1)] = idx
    best = []
    for idx, line in enumerate(lines):
        m = re.search(r"s_c?branch\w*\s+(\.LBB\S+)", line)
        if m and m.group(1) in labels and labels[m.group(1)] < idx:
            body = lines[labels[m.group(1)]:idx + 1]
            ops = [b.split()[0] for b in body
                   if b.strip() and not b.strip().startswith((";", "."))]
            if len(ops) > len(best):
                best = ops
    return best


@pytest.fixture(scope="module")
def counts(tmp_path_factory):
    if not os.path.exists(B.HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("isa") / "mixed_forms.s"
    flags = [f for f in B.HIP_FLAGS if f not in ("-fPIC",)]
    subprocess.run([B.HIPCC, *flags, "--cuda-device-only", "-S", PROBE, "-o", str(out)],
                   check=True, capture_output=True)
    funcs, cur = {}, None
    for line in out.read_text().split("\n"):
        m = re.match(r"^_Z\d+(k_\w+?)P\w*:", line)
        if m:
            cur = m.group(1)
            funcs[cur] = []
            continue
        if cur:
            funcs[cur].append(line)
    res = {}
    for name, lines in funcs.items():
        ops = _largest_loop_ops(lines)
        valu = [o for o in ops if o.startswith("v_")]
        lds = Counter(o for o in ops if o.startswith("ds_"))
        res[name] = {"valu_per_entry": len(valu) / 4.0,
                     "trans_per_entry": sum("rcp" in o for o in valu) / 4.0,
                     "lds": dict(lds)}
    print()
    for name, r in res.items():
        print(f"  {name:20s} {r['valu_per_entry']:6.2f} VALU/entry  LDS {r['lds']}")
    return res


def test_probe_has_all_forms(counts):
    assert set(counts) == {"k_f64", "k_current", "k_int32_exact_cut", "k_int32_f32_cut"}
    for r in counts.values():
        assert r["trans_per_entry"] == 1.0  # one reciprocal per entry in every form


def test_current_mixed_form_costs_about_the_fp64_form(counts):
    f64, cur = counts["k_f64"]["valu_per_entry"], counts["k_current"]["valu_per_entry"]
    # Mixed replaces the FP64 reciprocal and powers by FP32 ones but adds
    # two conversions: no fewer instructions.
    assert cur >= f64


def test_int32_forms_cost_more_than_the_current_form(counts):
    cur = counts["k_current"]["valu_per_entry"]
    # Integer differences, int->float conversions, FP32->FP64 conversions
    # for the accumulation and (for the exact cutoff) int64 multiply-adds:
    # at least 20 % more instructions than the current form.
    assert counts["k_int32_f32_cut"]["valu_per_entry"] >= 1.2 * cur
    assert counts["k_int32_exact_cut"]["valu_per_entry"] >= 1.2 * cur
    # 12-byte entries halve the LDS bytes per entry (one b96 read instead of
    # b128 + b64).
    assert counts["k_int32_exact_cut"]["lds"] == {"ds_read_b96": 4}
