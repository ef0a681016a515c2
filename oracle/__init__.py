"""Oracle: CPU restatement of kohnakagawa/lj_gpu's Lennard-Jones force path.

TEST INFRASTRUCTURE ONLY. Imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg, always as the checker (or the timed CPU
baseline), never as the thing measured or shipped. The product
(``lj_gpu_amd``) never imports this package.

Python bindings over ``oracle/liblj_oracle.so`` (built from oracle.cpp by
oracle/Makefile) and, where it was built, over the reference's own CPU code
linked into ``oracle/_ref/libljref_{aos,soa}.so``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblj_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

# Constants of the reference problem (cpu_ref/force_aos.cpp:10-14,31-33).
CUTOFF_LENGTH = 3.0
SEARCH_LENGTH = 3.3
CL2 = CUTOFF_LENGTH * CUTOFF_LENGTH
SL2 = SEARCH_LENGTH * SEARCH_LENGTH
REF_DT = 0.001
REF_L = 50.0

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle oracle`)")
        L = C.CDLL(LIB_PATH)
        L.orc_lattice_reference.argtypes = [C.c_double, C.c_double, _f64p, C.c_int32]
        L.orc_lattice_reference.restype = C.c_int32
        L.orc_lattice_pbc.argtypes = [C.c_int32, C.c_double, C.c_uint32, C.c_double, C.c_double,
                                      _f64p, C.c_int32, C.POINTER(C.c_double)]
        L.orc_lattice_pbc.restype = C.c_int32
        for name in ("orc_makepair_half", "orc_makepair_full"):
            f = getattr(L, name)
            f.argtypes = [C.c_int32, _f64p, C.c_double, C.c_double, _i32p, _i32p, _i32p, C.c_int64]
            f.restype = C.c_int64
        L.orc_sortpair.argtypes = [C.c_int32, _i32p, _i32p, _i32p, C.c_int64, _i32p, _i32p]
        L.orc_sortpair.restype = None
        L.orc_random_shfl.argtypes = [C.c_int32, _i32p, _i32p, _i32p, C.c_uint32]
        L.orc_random_shfl.restype = None
        L.orc_make_transposed.argtypes = [C.c_int32, _i32p, _i32p, _i32p, C.c_int32, _i32p]
        L.orc_make_transposed.restype = None
        L.orc_force_pair_aos.argtypes = [C.c_int64, _i32p, _i32p, _f64p, _f64p,
                                         C.c_double, C.c_double, C.c_double]
        L.orc_force_pair_soa.argtypes = [C.c_int64, _i32p, _i32p, _f64p, _f64p, _f64p,
                                         _f64p, _f64p, _f64p, C.c_double, C.c_double, C.c_double]
        for name in ("orc_force_sorted_aos", "orc_force_full_aos"):
            getattr(L, name).argtypes = [C.c_int32, _i32p, _i32p, _i32p, _f64p, _f64p,
                                         C.c_double, C.c_double, C.c_double]
        for name in ("orc_force_sorted_soa", "orc_force_full_soa"):
            getattr(L, name).argtypes = [C.c_int32, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p,
                                         _f64p, _f64p, _f64p, C.c_double, C.c_double, C.c_double]
        L.orc_time_force_sorted_aos.argtypes = [C.c_int32, _i32p, _i32p, _i32p, _f64p, _f64p,
                                                C.c_double, C.c_double, C.c_double, C.c_int32]
        L.orc_time_force_sorted_aos.restype = C.c_double
        _lib = L
    return _lib


# ---------------------------------------------------------------------------
# Lattices
# ---------------------------------------------------------------------------
def lattice_reference(density: float = 0.5, box_length: float = REF_L) -> np.ndarray:
    """The reference's open-box init(): returns q as an (N, 4) float64 array."""
    s = 1.0 / (density * 0.25) ** (1.0 / 3.0)
    sx = int(box_length / s) + 1
    cap = 4 * sx ** 3
    q = np.zeros(cap * 4, dtype=np.float64)
    n = lib().orc_lattice_reference(density, box_length, q, cap)
    if n < 0:
        raise RuntimeError("lattice capacity")
    return q[: 4 * n].reshape(n, 4).copy()


def lattice_pbc(k: int, density: float, seed: int = 2, jitter: tuple = (-0.05, 0.05)):
    """Periodic FCC lattice of 4*k^3 particles; returns (q (N,4), box length)."""
    n = 4 * k ** 3
    q = np.zeros(n * 4, dtype=np.float64)
    box = C.c_double(0.0)
    got = lib().orc_lattice_pbc(k, density, seed, jitter[0], jitter[1], q, n, C.byref(box))
    if got != n:
        raise RuntimeError("lattice capacity")
    return q.reshape(n, 4), box.value


# ---------------------------------------------------------------------------
# Neighbour lists (CSR)
# ---------------------------------------------------------------------------
@dataclass
class CSR:
    number_of_partners: np.ndarray
    pointer: np.ndarray
    sorted_list: np.ndarray
    kind: str  # "half" or "full"

    @property
    def n(self) -> int:
        return int(self.number_of_partners.shape[0])

    @property
    def n_entries(self) -> int:
        return int(self.sorted_list.shape[0])

    def pairs(self):
        """Flat (i, j) arrays in CSR order (the reference's i_particles/j_particles)."""
        i = np.repeat(np.arange(self.n, dtype=np.int32), self.number_of_partners)
        return i, self.sorted_list.copy()


def makepair(q: np.ndarray, box: float = 0.0, kind: str = "half", sl2: float = SL2,
             cap: int | None = None) -> CSR:
    """Brute-force O(N^2) neighbour list in the reference's order."""
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 4)
    n = q.shape[0]
    if cap is None:
        cap = max(1, n * (200 if kind == "full" else 100))
    nop = np.zeros(max(n, 1), dtype=np.int32)
    ptr = np.zeros(max(n, 1), dtype=np.int32)
    lst = np.zeros(cap, dtype=np.int32)
    fn = lib().orc_makepair_half if kind == "half" else lib().orc_makepair_full
    e = fn(n, q.reshape(-1), box, sl2, nop, ptr, lst, cap)
    if e < 0:
        raise RuntimeError("pair capacity exceeded")
    return CSR(nop[:n].copy(), ptr[:n].copy(), lst[:e].copy(), kind)


def random_shfl(csr: CSR, seed: int = 10) -> CSR:
    lst = csr.sorted_list.copy()
    lib().orc_random_shfl(csr.n, csr.number_of_partners, csr.pointer, lst, seed)
    return CSR(csr.number_of_partners.copy(), csr.pointer.copy(), lst, csr.kind)


def make_transposed(csr: CSR, width: int | None = None) -> np.ndarray:
    if width is None:
        width = int(csr.number_of_partners.max()) if csr.n else 0
    out = np.zeros(max(1, csr.n * width), dtype=np.int32)
    lib().orc_make_transposed(csr.n, csr.number_of_partners, csr.pointer, csr.sorted_list,
                              width, out)
    return out[: csr.n * width]


# ---------------------------------------------------------------------------
# Forces (p is updated in place; AoS arrays are (N, 4), SoA (3, N))
# ---------------------------------------------------------------------------
def _lst(csr: CSR) -> np.ndarray:
    return csr.sorted_list if csr.sorted_list.size else np.zeros(1, dtype=np.int32)


def force_aos(csr: CSR, q: np.ndarray, p: np.ndarray, dt: float, cl2: float = CL2,
              box: float = 0.0, steps: int = 1, method: str = "sorted") -> None:
    """p += F(q) dt, `steps` times. method: "sorted" (half CSR, force_sorted),
    "pair" (half, force_pair over flat pairs) or "full" (full CSR)."""
    L = lib()
    qf, pf = q.reshape(-1), p.reshape(-1)
    assert qf.flags.c_contiguous and pf.flags.c_contiguous
    if method == "pair":
        ip, jp = csr.pairs()
        ip = ip if ip.size else np.zeros(1, np.int32)
        jp = jp if jp.size else np.zeros(1, np.int32)
        for _ in range(steps):
            L.orc_force_pair_aos(csr.n_entries, ip, jp, qf, pf, dt, cl2, box)
        return
    fn = L.orc_force_sorted_aos if method == "sorted" else L.orc_force_full_aos
    nop = csr.number_of_partners if csr.n else np.zeros(1, np.int32)
    ptr = csr.pointer if csr.n else np.zeros(1, np.int32)
    for _ in range(steps):
        fn(csr.n, nop, ptr, _lst(csr), qf, pf, dt, cl2, box)


def force_soa(csr: CSR, q3: np.ndarray, p3: np.ndarray, dt: float, cl2: float = CL2,
              box: float = 0.0, steps: int = 1, method: str = "sorted") -> None:
    """SoA form of force_aos: q3, p3 are (3, N) C-contiguous arrays."""
    L = lib()
    qx, qy, qz = (np.ascontiguousarray(q3[c]) for c in range(3))
    px, py, pz = (np.ascontiguousarray(p3[c]) for c in range(3))
    if method == "pair":
        ip, jp = csr.pairs()
        for _ in range(steps):
            L.orc_force_pair_soa(csr.n_entries, ip, jp, qx, qy, qz, px, py, pz, dt, cl2, box)
    else:
        fn = L.orc_force_sorted_soa if method == "sorted" else L.orc_force_full_soa
        for _ in range(steps):
            fn(csr.n, csr.number_of_partners, csr.pointer, _lst(csr), qx, qy, qz,
               px, py, pz, dt, cl2, box)
    p3[0], p3[1], p3[2] = px, py, pz


def time_force_sorted_aos(csr: CSR, q: np.ndarray, dt: float, cl2: float = CL2,
                          box: float = 0.0, steps: int = 1) -> float:
    """Wall seconds for `steps` force_sorted calls on one core (CPU baseline)."""
    p = np.zeros_like(q)
    return lib().orc_time_force_sorted_aos(csr.n, csr.number_of_partners, csr.pointer,
                                           _lst(csr), q.reshape(-1), p.reshape(-1), dt, cl2,
                                           box, steps)


# ---------------------------------------------------------------------------
# Timed CPU baselines (oracle/liblj_cpubase.so, built from cpu_baseline.cpp)
# ---------------------------------------------------------------------------
BASELINE_PATH = os.path.join(HERE, "liblj_cpubase.so")
BASELINE_SORTED, BASELINE_MAT_TRANSPOSE, BASELINE_FULL_OMP = 0, 1, 2
_cb = None


def baseline_lib() -> C.CDLL:
    global _cb
    if _cb is None:
        if not os.path.exists(BASELINE_PATH):
            raise RuntimeError(f"CPU baseline library missing: {BASELINE_PATH} "
                               "(run `make -C oracle oracle`)")
        L = C.CDLL(BASELINE_PATH)
        L.cb_run.argtypes = [C.c_int, C.c_int32, _i32p, _i32p, _i32p, _f64p, _f64p, C.c_double,
                             C.c_double, C.c_double, C.c_int32]
        L.cb_run.restype = C.c_double
        L.cb_threads.restype = C.c_int
        L.cb_build_info.restype = C.c_char_p
        _cb = L
    return _cb


def aligned_zeros(shape, align: int = 64) -> np.ndarray:
    """float64 zeros whose data starts on an `align`-byte boundary."""
    n = int(np.prod(shape))
    raw = np.zeros(n + align // 8, dtype=np.float64)
    off = (-raw.ctypes.data % align) // 8
    return raw[off: off + n].reshape(shape)


def run_baseline(kind: int, csr: CSR, q: np.ndarray, p: np.ndarray, dt: float,
                 cl2: float = CL2, box: float = 0.0, steps: int = 1) -> float:
    """`steps` evaluations of CPU baseline `kind` on p (in place, AoS (N, 4),
    32-byte aligned); returns wall seconds."""
    assert q.ctypes.data % 32 == 0 and p.ctypes.data % 32 == 0
    nop = csr.number_of_partners if csr.n else np.zeros(1, np.int32)
    ptr = csr.pointer if csr.n else np.zeros(1, np.int32)
    t = baseline_lib().cb_run(kind, csr.n, nop, ptr, _lst(csr), q.reshape(-1), p.reshape(-1),
                              dt, cl2, box, steps)
    if t < 0:
        raise ValueError(f"unknown baseline kind {kind}")
    return t


def format_reference_output(p: np.ndarray) -> str:
    """The reference's print_results(): first and last 5 momenta at %.10f
    (cpu_ref/force_aos.cpp:622-630)."""
    n = p.shape[0]
    rows = list(range(min(5, n))) + list(range(max(n - 5, 0), n))
    return "".join("%.10f %.10f %.10f\n" % (p[i, 0], p[i, 1], p[i, 2]) for i in rows)


# ---------------------------------------------------------------------------
# The reference's own CPU code (present only where oracle/_ref was built)
# ---------------------------------------------------------------------------
def ref_lib(layout: str = "aos") -> C.CDLL | None:
    path = os.path.join(REF_DIR, f"libljref_{layout}.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.ljref_init.argtypes = [C.c_double]
    L.ljref_set_positions.argtypes = [C.c_int, _f64p]
    L.ljref_get_positions.argtypes = [_f64p]
    L.ljref_get_momenta.argtypes = [_f64p]
    L.ljref_get_csr.argtypes = [_i32p, _i32p, _i32p]
    L.ljref_get_pairs.argtypes = [_i32p, _i32p]
    L.ljref_run.argtypes = [C.c_int, C.c_int]
    L.ljref_run.restype = C.c_double
    return L
