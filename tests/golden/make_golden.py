"""Write the golden fixtures under tests/golden/ from the reference's own CPU code.

Runs the reference's cpu_ref/force_aos.cpp and force_soa.cpp, compiled
unmodified into oracle/_ref/libljref_{aos,soa}.so by oracle/Makefile, and
stores what they compute as small .npz fixtures (plain arrays, no pickles):

  ref_aos_L50_rho0.5.npz   the reference's own init() (L=50, density 0.5,
                           N=62,500): position samples, CSR checksums of its
                           makepair()/sortpair(), momenta after 100 steps of
                           force_pair / force_sorted / force_intrin /
                           force_next / force_intrin_mat_transpose /
                           force_sorted_z_intrin (dt=0.001), sampled rows at
                           17 significant digits plus sums
  ref_aos_L12_rho0.5.npz   init() with L=12 (N=864): every momentum after
                           100 force_sorted steps, full CSR
  ref_soa_L50_rho0.5.npz   force_soa.cpp: force_pair / force_sorted /
                           force_intrin on its own init()
  ref_pairdat.json         size and sha256 of the binary pair.dat that the
                           reference's savepair() writes for init() with
                           L=12 and L=50 (density 0.5)
  ref_aos_open_rho1.npz    density 1.0 (N=119,164): positions from the
                           restated lattice generator (the reference's
                           density is a compile-time constant), pair list and
                           forces from the reference's makepair/force_sorted

Each configuration runs in its own process because the reference's
add_particle() keeps a static random stream. Usage:
    python tests/golden/make_golden.py            # all fixtures
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

KINDS_AOS = {"pair": 0, "sorted": 1, "intrin": 2, "next": 3, "mat_transpose": 4,
             "sorted_z_intrin": 6}
KINDS_SOA = {"pair": 0, "sorted": 1, "intrin": 2}


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def sample_rows(n: int) -> np.ndarray:
    """First 5, last 5 and every 997th particle."""
    rows = set(range(min(5, n))) | set(range(max(0, n - 5), n)) | set(range(0, n, 997))
    return np.array(sorted(rows), dtype=np.int64)


def run_config(name: str) -> None:
    import oracle as O

    layout = "soa" if name.startswith("ref_soa") else "aos"
    R = O.ref_lib(layout)
    if R is None:
        raise SystemExit("oracle/_ref not built: run `make -C oracle ref`")
    data = {}
    if name == "ref_aos_open_rho1":
        q = O.lattice_reference(1.0, 50.0)
        n = R.ljref_set_positions(q.shape[0], np.ascontiguousarray(q.reshape(-1)))
    else:
        box = 12.0 if "L12" in name else 50.0
        n = R.ljref_init(box)
    q4 = np.zeros(n * 4)
    R.ljref_get_positions(q4)
    q4 = q4.reshape(n, 4)
    npairs = R.ljref_makepair()
    nop = np.zeros(n, np.int32)
    ptr = np.zeros(n, np.int32)
    lst = np.zeros(npairs, np.int32)
    R.ljref_get_csr(nop, ptr, lst)
    rows = sample_rows(n)
    data["n"] = np.array(n)
    data["npairs"] = np.array(npairs)
    data["rows"] = rows
    data["q_rows"] = q4[rows, :3]
    data["q_sha256"] = np.array(sha(q4[:, :3]))
    data["nop_sha256"] = np.array(sha(nop))
    data["pointer_sha256"] = np.array(sha(ptr))
    data["sorted_list_sha256"] = np.array(sha(lst))
    data["nop_min"] = np.array(nop.min())
    data["nop_max"] = np.array(nop.max())
    if n <= 1000:
        data["number_of_partners"] = nop
        data["pointer"] = ptr
        data["sorted_list"] = lst
        data["q_all"] = q4[:, :3]
    kinds = KINDS_SOA if layout == "soa" else KINDS_AOS
    if name in ("ref_aos_open_rho1", "ref_aos_L12_rho0.5"):
        kinds = {"sorted": 1, "pair": 0}
    for kname, kid in kinds.items():
        R.ljref_zero_momenta()
        R.ljref_run(kid, 100)
        p4 = np.zeros(n * 4)
        R.ljref_get_momenta(p4)
        p4 = p4.reshape(n, 4)[:, :3]
        data[f"p_{kname}_rows"] = p4[rows]
        data[f"p_{kname}_sum"] = p4.sum(axis=0)
        data[f"p_{kname}_sumsq"] = np.array((p4 * p4).sum())
        data[f"p_{kname}_absmax"] = np.array(np.abs(p4).max())
        if n <= 1000:
            data[f"p_{kname}_all"] = p4
    meta = {"dt": 0.001, "steps": 100, "cutoff": 3.0, "search": 3.3, "layout": layout,
            "source": "reference cpu_ref compiled by oracle/Makefile (-O3 -mavx2 -mfma)"}
    data["meta_json"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **data)
    print(name, "n", n, "pairs", npairs)


def run_pairdat(box: float) -> dict:
    """The reference's savepair() on its own init() with side `box`, in a
    scratch directory: size and sha256 of the pair.dat it writes."""
    import tempfile

    import oracle as O

    R = O.ref_lib("aos")
    if R is None:
        raise SystemExit("oracle/_ref not built: run `make -C oracle ref`")
    n = R.ljref_init(box)
    with tempfile.TemporaryDirectory() as d:
        cwd = os.getcwd()
        os.chdir(d)
        try:
            npairs = R.ljref_savepair()
            blob = open("pair.dat", "rb").read()
        finally:
            os.chdir(cwd)
    return {"box": box, "density": 0.5, "n": n, "number_of_pairs": npairs,
            "size": len(blob), "sha256": hashlib.sha256(blob).hexdigest()}


def main() -> None:
    if len(sys.argv) > 1:
        if sys.argv[1].startswith("pairdat"):
            print(json.dumps(run_pairdat(float(sys.argv[1].split(":")[1]))))
        else:
            run_config(sys.argv[1])
        return
    for name in ("ref_aos_L50_rho0.5", "ref_aos_L12_rho0.5", "ref_soa_L50_rho0.5",
                 "ref_aos_open_rho1"):
        subprocess.run([sys.executable, os.path.abspath(__file__), name], check=True)
    # pair.dat written by the reference's savepair() (cpu_ref/force_aos.cpp:612-620).
    pairdat = {}
    for box in (12.0, 50.0):
        out = subprocess.run([sys.executable, os.path.abspath(__file__), f"pairdat:{box}"],
                             check=True, capture_output=True, text=True).stdout
        pairdat[f"L{int(box)}_rho0.5"] = json.loads(out.strip().splitlines()[-1])
    with open(os.path.join(OUT, "ref_pairdat.json"), "w") as f:
        json.dump(pairdat, f, indent=1)
    print("ref_pairdat", pairdat)


if __name__ == "__main__":
    main()
