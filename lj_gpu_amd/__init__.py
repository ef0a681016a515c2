"""lj_gpu_amd — MI355X-native Lennard-Jones force evaluation over a Verlet list.

Python host mirror of the reference harness (kohnakagawa/lj_gpu,
cuda/force_cuda.cu and cpu_ref/force_aos.cpp) over the C ABI of
``libljgpu.so`` (include/ljgpu.h). The reference's host steps map as:

    init()                       -> lattice_reference() / lattice_periodic()
    makepair() + sortpair()      -> makepair()            (cell list, exact pair set)
    random_shfl()                -> random_shfl()
    make_transposed_pairlist()   -> make_transposed()
    cuda_ptr<T>                  -> DeviceArray
    measure(kernel, ...)         -> Context.force_* (+ Context.timer_*)
    check_loadedpair()           -> validate_csr()
    savepair() / loadpair()      -> write_pair_dat() / read_pair_dat()
    makepaircache() / loadpair() -> write_pair_cache() / read_pair_cache()

All compute runs in the HIP kernels of libljgpu.so; if the library is
missing, importing the device API raises (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os
import re
from dataclasses import dataclass

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libljgpu.so")
# The benchmark program (harness/force_ljgpu.cpp), built beside the library.
HARNESS_PATH = os.path.join(PKG, "force_ljgpu")
HEADER = os.path.join(os.path.dirname(PKG), "include", "ljgpu.h")

# Status codes and enums (include/ljgpu.h).
OK = 0
ERR_INVALID_ARGUMENT = -1
ERR_OUT_OF_MEMORY = -2
ERR_HIP = -3
ERR_BAD_LIST = -4
ERR_NO_LIST = -5
ERR_CAPACITY = -6
ERR_NO_DEVICE = -7
ERR_IO = -8
ERR_UNSUPPORTED = -9

LIST_HALF = 0
LIST_FULL = 1
MEM_HOST = 0
MEM_DEVICE = 1
ACCUM_F64 = 0
ACCUM_KAHAN_F32 = 1
HALF_AUTO = 0
HALF_ATOMIC = 1
HALF_EXPAND = 2
TILE_ORDER_INDEX = 0
TILE_ORDER_SPATIAL = 1

# The reference problem's constants (cpu_ref/force_aos.cpp:10-14,31-33).
CUTOFF_LENGTH = 3.0
SEARCH_LENGTH = 3.3
CL2 = CUTOFF_LENGTH * CUTOFF_LENGTH
REF_DT = 0.001
REF_BOX_LENGTH = 50.0


class LJGPUError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ljgpu error {code}: {msg}")
        self.code = code
        self.message = msg


class ListInfo(C.Structure):
    _fields_ = [("particle_number", C.c_int32), ("kind", C.c_int32), ("n_entries", C.c_int64),
                ("n_pairs", C.c_int64), ("max_partners", C.c_int32), ("min_partners", C.c_int32),
                ("half_mode", C.c_int32), ("lanes_per_particle", C.c_int32),
                ("unroll", C.c_int32), ("block_size", C.c_int32),
                ("tile_particles", C.c_int32), ("tile_blocks", C.c_int32),
                ("tile_max_stage", C.c_int32)]


_lib = None
_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_d = C.c_double


PREC_F64 = 0
PREC_MIXED = 1


class TuneEntry(C.Structure):
    """ljgpu_tune_entry (include/ljgpu.h)."""
    _fields_ = [("lanes_per_particle", C.c_int32), ("unroll", C.c_int32),
                ("tile_particles", C.c_int32), ("resident_per_cu", C.c_int32),
                ("tile_used", C.c_int32), ("us_per_step", C.c_float), ("chains", C.c_int32),
                ("us_min", C.c_float)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class MDConfig(C.Structure):
    """ljgpu_md_config (include/ljgpu.h)."""
    _fields_ = [("dt", C.c_double), ("cutoff", C.c_double), ("skin", C.c_double),
                ("rebuild_every", C.c_int32), ("precision", C.c_int32), ("accum", C.c_int32),
                ("sample_every", C.c_int32)]


class MDStats(C.Structure):
    """ljgpu_md_stats (include/ljgpu.h)."""
    _fields_ = [("steps", C.c_int32), ("rebuilds", C.c_int32), ("samples", C.c_int32),
                ("stale_lists", C.c_int32), ("potential0", C.c_double), ("kinetic0", C.c_double),
                ("potential1", C.c_double), ("kinetic1", C.c_double),
                ("max_abs_drift", C.c_double), ("max_displacement", C.c_double),
                ("pair_entries", C.c_int64), ("ms_total", C.c_float), ("ms_rebuild", C.c_float)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib() -> C.CDLL:
    """Load libljgpu.so (built by lj_gpu_amd/build.py); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m lj_gpu_amd.build`")
    L = C.CDLL(LIB_PATH)
    sig = {
        "ljgpu_device_count": [C.POINTER(C.c_int)],
        "ljgpu_create": [C.c_int, C.POINTER(_vp)],
        "ljgpu_destroy": [_vp],
        "ljgpu_sync": [_vp],
        "ljgpu_malloc": [_vp, C.c_size_t, C.POINTER(_vp)],
        "ljgpu_free": [_vp, _vp],
        "ljgpu_memcpy_htod": [_vp, _vp, _vp, C.c_size_t],
        "ljgpu_memcpy_dtoh": [_vp, _vp, _vp, C.c_size_t],
        "ljgpu_memcpy_dtod": [_vp, _vp, _vp, C.c_size_t],
        "ljgpu_memset": [_vp, _vp, C.c_int, C.c_size_t],
        "ljgpu_timer_start": [_vp],
        "ljgpu_timer_stop": [_vp, C.POINTER(C.c_float)],
        "ljgpu_set_list_csr": [_vp, _i32, _vp, _vp, _vp, _i64, C.c_int, C.c_int],
        "ljgpu_set_list_transposed": [_vp, _i32, _vp, _vp, _i32, C.c_int, C.c_int],
        "ljgpu_get_list_info": [_vp, C.POINTER(ListInfo)],
        "ljgpu_get_list_csr": [_vp, _vp, _vp, _vp, _i64],
        "ljgpu_set_half_mode": [_vp, C.c_int],
        "ljgpu_set_tiling": [_vp, _i32, C.c_int, _i32],
        "ljgpu_set_kernel_config": [_vp, C.c_int, C.c_int, C.c_int],
        "ljgpu_force_aos_f64": [_vp, _vp, _vp, _d, _d, _d, _i32],
        "ljgpu_force_aos3_f64": [_vp, _vp, _vp, _d, _d, _d, _i32],
        "ljgpu_force_aos3_mixed": [_vp, _vp, _vp, _d, _d, _d, _i32, C.c_int],
        "ljgpu_build_list_aos3": [_vp, _vp, _i32, _d, _d, C.c_int],
        "ljgpu_force_soa_f64": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _d, _i32],
        "ljgpu_force_aos_mixed": [_vp, _vp, _vp, _d, _d, _d, _i32, C.c_int],
        "ljgpu_force_soa_mixed": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _d, _i32, C.c_int],
        "ljgpu_force_kernel_csr_f64": [_vp, _vp, _vp, _i32, _d, _d, _vp, _vp, _vp, _i64, C.c_int],
        "ljgpu_lattice_reference": [_d, _d, _vp, _i32, C.POINTER(_i32)],
        "ljgpu_lattice_periodic": [_i32, _d, C.c_uint32, _d, _d, _vp, _i32, C.POINTER(_d)],
        "ljgpu_makepair_host": [_i32, _vp, _d, _d, C.c_int, _vp, _vp, _vp, _i64, C.POINTER(_i64)],
        "ljgpu_validate_csr": [_i32, _vp, _vp, _vp, _i64, C.c_int],
        "ljgpu_random_shfl": [_i32, _vp, _vp, _vp, C.c_uint32],
        "ljgpu_make_transposed": [_i32, _vp, _vp, _vp, _i32, _vp],
        "ljgpu_make_tiles": [_i32, _vp, _vp, _vp, _i64, _i32, _vp, _i64, _vp, _i64, _vp,
                             C.POINTER(_i32), C.POINTER(_i64), C.POINTER(_i32)],
        "ljgpu_make_tiles_ordered": [_i32, _vp, _vp, _vp, _i64, _i32, _vp, _vp, _i64, _vp, _i64,
                                     _vp, C.POINTER(_i32), C.POINTER(_i64), C.POINTER(_i32)],
        "ljgpu_make_tiles_cut": [_i32, _vp, _vp, _vp, _i64, _i32, _vp, _vp, _d, _d, _vp, _i64,
                                 _vp, _i64, _vp, C.POINTER(_i32), C.POINTER(_i64),
                                 C.POINTER(_i32)],
        "ljgpu_set_tile_order": [_vp, C.c_int],
        "ljgpu_get_tile_order": [_vp, C.POINTER(C.c_int), _vp, _i64],
        "ljgpu_build_list": [_vp, _vp, _i32, _d, _d, C.c_int],
        "ljgpu_build_list_soa": [_vp, _vp, _vp, _vp, _i32, _d, _d, C.c_int],
        "ljgpu_get_tiles": [_vp, _vp, _i64, _vp, _i64, _vp, _i64, C.POINTER(_i32),
                            C.POINTER(_i64)],
        "ljgpu_set_replicas": [_vp, _i32, _vp, _vp],
        "ljgpu_autotune": [_vp, _vp, _d, _d, C.c_int, C.c_int, _i32, C.POINTER(TuneEntry)],
        "ljgpu_query_autotune": [_vp, C.POINTER(TuneEntry), _i32, C.POINTER(_i32)],
        "ljgpu_md_run": [_vp, _vp, _vp, _i32, _d, C.POINTER(MDConfig), _i32, C.POINTER(MDStats),
                         _vp, _i32],
        "ljgpu_energy": [_vp, _vp, _vp, _d, _d, C.POINTER(_d), C.POINTER(_d), C.POINTER(_i64)],
        "ljgpu_energy_noalloc": [_vp, _vp, _vp, _d, _d, C.POINTER(_d), C.POINTER(_d),
                                 C.POINTER(_i64)],
        "ljgpu_maxwell_boltzmann": [_i32, _d, C.c_uint32, _vp],
        "ljgpu_set_graphs": [_vp, C.c_int],
        "ljgpu_graph_stats": [_vp, C.POINTER(_i64), C.POINTER(_i32)],
        "ljgpu_set_chains": [_vp, C.c_int],
        "ljgpu_pair_dat_write": [C.c_char_p, _i32, _vp, _vp, _vp, _i64, _i32, _i64],
        "ljgpu_pair_dat_read": [C.c_char_p, _i32, _i32, _i64, _vp, _vp, _vp, _i64,
                                C.POINTER(_i64)],
        "ljgpu_pair_cache_write": [C.c_char_p, _i32, _vp, _vp, _vp, _i64],
        "ljgpu_pair_cache_read": [C.c_char_p, _i32, C.c_int, _vp, _vp, _vp, _i64,
                                  C.POINTER(_i64)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.ljgpu_last_error.restype = C.c_char_p
    L.ljgpu_last_error.argtypes = []
    L.ljgpu_version.restype = C.c_char_p
    L.ljgpu_version.argtypes = []
    _lib = L
    return L


def header_functions() -> list[str]:
    """Names of every function declared in include/ljgpu.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(ljgpu_\w+)\s*\(", text, re.M)))


def _check(rc: int) -> None:
    if rc != OK:
        msg = lib().ljgpu_last_error()
        raise LJGPUError(rc, msg.decode() if msg else "")


def _need(arr, nbytes: int, what: str) -> None:
    """Raise before a launch when a device buffer is shorter than the kernel
    reads or writes (the C ABI takes bare pointers and cannot check)."""
    if arr is None or arr.ptr is None:
        raise ValueError(f"{what}: buffer missing or freed")
    if arr.nbytes < nbytes:
        raise ValueError(f"{what}: {arr.nbytes} bytes, need at least {nbytes}")


def _ptr(a: np.ndarray | None) -> int | None:
    if a is None:
        return None
    assert a.flags.c_contiguous
    return a.ctypes.data if a.size else None


# ---------------------------------------------------------------------------
# Host-side setup (the reference's init / makepair / sortpair / random_shfl)
# ---------------------------------------------------------------------------
@dataclass
class CSR:
    """A Verlet list in the reference's CSR form (number_of_partners,
    pointer, sorted_list; cuda/force_cuda.cu:28,146-162)."""
    number_of_partners: np.ndarray
    pointer: np.ndarray
    sorted_list: np.ndarray
    kind: int

    @property
    def n(self) -> int:
        return int(self.number_of_partners.shape[0])

    @property
    def n_entries(self) -> int:
        return int(self.sorted_list.shape[0])

    @property
    def n_pairs(self) -> int:
        return self.n_entries if self.kind == LIST_HALF else self.n_entries // 2


def lattice_reference(density: float = 0.5, box_length: float = REF_BOX_LENGTH) -> np.ndarray:
    """The reference's init(): open-box FCC lattice, (N, 4) float64."""
    s = 1.0 / (density * 0.25) ** (1.0 / 3.0)
    cap = 4 * (int(box_length / s) + 1) ** 3
    q = np.zeros((cap, 4), dtype=np.float64)
    n = _i32(0)
    _check(lib().ljgpu_lattice_reference(density, box_length, _ptr(q), cap, C.byref(n)))
    return q[: n.value].copy()


def lattice_periodic(cells: int, density: float, seed: int = 2,
                     jitter: tuple[float, float] = (-0.05, 0.05)) -> tuple[np.ndarray, float]:
    """Periodic FCC lattice of 4*cells^3 particles; returns (q (N, 4), box)."""
    n = 4 * cells ** 3
    q = np.zeros((n, 4), dtype=np.float64)
    box = _d(0.0)
    _check(lib().ljgpu_lattice_periodic(cells, density, seed, jitter[0], jitter[1], _ptr(q), n,
                                        C.byref(box)))
    return q, box.value


def makepair(q: np.ndarray, box: float = 0.0, kind: int = LIST_HALF,
             search_length: float = SEARCH_LENGTH) -> CSR:
    """Verlet list with the reference's exact pair set and CSR order."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    n = q.shape[0]
    nop = np.zeros(n, np.int32)
    ptr = np.zeros(n, np.int32)
    cap = _i64(0)
    rc = lib().ljgpu_makepair_host(n, _ptr(q), box, search_length, kind, _ptr(nop), _ptr(ptr),
                                   None, 0, C.byref(cap))
    if rc not in (OK, ERR_CAPACITY):
        _check(rc)
    lst = np.zeros(max(cap.value, 1), np.int32)
    e = _i64(0)
    _check(lib().ljgpu_makepair_host(n, _ptr(q), box, search_length, kind, _ptr(nop), _ptr(ptr),
                                     _ptr(lst), lst.size, C.byref(e)))
    return CSR(nop, ptr, lst[: e.value].copy(), kind)


def validate_csr(csr: CSR) -> None:
    _check(lib().ljgpu_validate_csr(csr.n, _ptr(csr.number_of_partners), _ptr(csr.pointer),
                                    _ptr(csr.sorted_list), csr.n_entries, csr.kind))


def random_shfl(csr: CSR, seed: int = 10) -> CSR:
    lst = csr.sorted_list.copy()
    _check(lib().ljgpu_random_shfl(csr.n, _ptr(csr.number_of_partners), _ptr(csr.pointer),
                                   _ptr(lst), seed))
    return CSR(csr.number_of_partners.copy(), csr.pointer.copy(), lst, csr.kind)


def make_transposed(csr: CSR, width: int | None = None) -> np.ndarray:
    if width is None:
        width = int(csr.number_of_partners.max()) if csr.n else 0
    out = np.zeros(max(csr.n * width, 1), np.int32)
    _check(lib().ljgpu_make_transposed(csr.n, _ptr(csr.number_of_partners), _ptr(csr.pointer),
                                       _ptr(csr.sorted_list), width, _ptr(out)))
    return out[: csr.n * width]


def replica_workload(n_replicas: int = 256, seed: int = 4, cells: tuple[int, int] = (6, 20),
                     density: tuple[float, float] = (0.5, 1.0)):
    """Batched-replica workload: replica b is a periodic FCC lattice of
    4*k_b^3 particles at density rho_b (k_b uniform in cells, inclusive;
    rho_b uniform in density, both drawn from numpy's default_rng(seed)),
    jittered with seed 1000 + b. Returns (q (N, 4), offsets, boxes)."""
    rng = np.random.default_rng(seed)
    ks = rng.integers(cells[0], cells[1] + 1, size=n_replicas)
    rhos = rng.uniform(density[0], density[1], size=n_replicas)
    qs, boxes = [], []
    for b in range(n_replicas):
        q, box = lattice_periodic(int(ks[b]), float(rhos[b]), 1000 + b)
        qs.append(q)
        boxes.append(box)
    offsets = np.zeros(n_replicas + 1, np.int32)
    offsets[1:] = np.cumsum([q.shape[0] for q in qs])
    q = np.concatenate(qs) if qs else np.zeros((0, 4))
    return q, offsets, np.array(boxes, np.float64)


def maxwell_boltzmann(n: int, temperature: float, seed: int) -> np.ndarray:
    """Seeded Maxwell-Boltzmann momenta (unit masses), (N, 4), zero mean."""
    p = np.zeros((n, 4), np.float64)
    _check(lib().ljgpu_maxwell_boltzmann(n, temperature, seed, _ptr(p)))
    return p


def write_pair_dat(path: str, csr: CSR, n_capacity: int = 0, pair_capacity: int = 0) -> None:
    """Write a half list as the CPU programs' binary pair.dat
    (cpu_ref/force_aos.cpp:612-620); 0 capacities are the reference's."""
    _check(lib().ljgpu_pair_dat_write(os.fsencode(path), csr.n, _ptr(csr.number_of_partners),
                                      _ptr(csr.pointer), _ptr(csr.sorted_list), csr.n_entries,
                                      n_capacity, pair_capacity))


def read_pair_dat(path: str, particle_number: int, n_capacity: int = 0,
                  pair_capacity: int = 0) -> CSR:
    """Read pair.dat and rebuild the CSR as loadpair() + sortpair() do
    (cpu_ref/force_aos.cpp:100-120,603-610)."""
    n = particle_number
    nop = np.zeros(n, np.int32)
    ptr = np.zeros(n, np.int32)
    e = _i64(0)
    args = (os.fsencode(path), n, n_capacity, pair_capacity, _ptr(nop), _ptr(ptr))
    _check(lib().ljgpu_pair_dat_read(*args, None, 0, C.byref(e)))
    lst = np.zeros(max(e.value, 1), np.int32)
    _check(lib().ljgpu_pair_dat_read(*args, _ptr(lst), lst.size, C.byref(e)))
    return CSR(nop, ptr, lst[: e.value].copy(), LIST_HALF)


def write_pair_cache(path: str, csr: CSR) -> None:
    """Write the GPU harness's text cache (.cache_pair_{half,all}.dat,
    cuda/force_cuda.cu:165-177)."""
    _check(lib().ljgpu_pair_cache_write(os.fsencode(path), csr.n, _ptr(csr.number_of_partners),
                                        _ptr(csr.pointer), _ptr(csr.sorted_list), csr.n_entries))


def read_pair_cache(path: str, particle_number: int, kind: int) -> CSR:
    """Read a text pair cache (cuda/force_cuda.cu:203-227); refuses a cache
    for another particle number and validates the list as `kind`."""
    n = particle_number
    nop = np.zeros(n, np.int32)
    ptr = np.zeros(n, np.int32)
    e = _i64(0)
    _check(lib().ljgpu_pair_cache_read(os.fsencode(path), n, kind, _ptr(nop), _ptr(ptr), None, 0,
                                       C.byref(e)))
    lst = np.zeros(max(e.value, 1), np.int32)
    _check(lib().ljgpu_pair_cache_read(os.fsencode(path), n, kind, _ptr(nop), _ptr(ptr),
                                       _ptr(lst), lst.size, C.byref(e)))
    return CSR(nop, ptr, lst[: e.value].copy(), kind)


@dataclass
class Tiles:
    """LDS tiling of a full list (what ljgpu_set_tiling installs)."""
    desc: np.ndarray      # (n_blocks, 8): first slot, count, staging offset,
                          # staging count, first entry, entry count, replica,
                          # middle particle (of slot first + count // 2)
    stage: np.ndarray     # staged particle indices
    local: np.ndarray     # uint16 LDS byte offsets of the rows in slot order
    max_stage: int


def make_tiles(csr: CSR, tile_particles: int = 64, member=None, q=None, box: float = 0.0,
               cl2: float = CL2) -> Tiles:
    """Host model of the LDS tiling; member (one particle per slot) orders
    the particles the blocks cover, None for index order; positions q (N, 4)
    put the rows in cutoff classes (entries within the cutoff cl2 first)."""
    nb, ns, ms = _i32(0), _i64(0), _i32(0)
    mem = None if member is None else np.ascontiguousarray(member, dtype=np.int32)
    qq = None if q is None else np.ascontiguousarray(q, dtype=np.float64)
    args = (csr.n, _ptr(csr.number_of_partners), _ptr(csr.pointer), _ptr(csr.sorted_list),
            csr.n_entries, tile_particles, None if mem is None else _ptr(mem),
            None if qq is None else _ptr(qq), box, cl2)
    _check(lib().ljgpu_make_tiles_cut(*args, None, 0, None, 0, None, C.byref(nb),
                                      C.byref(ns), C.byref(ms)))
    desc = np.zeros(max(8 * nb.value, 1), np.int32)
    stage = np.zeros(max(ns.value, 1), np.int32)
    local = np.zeros(max(csr.n_entries, 1), np.uint16)
    _check(lib().ljgpu_make_tiles_cut(*args, _ptr(desc), desc.size, _ptr(stage), stage.size,
                                      _ptr(local), C.byref(nb), C.byref(ns), C.byref(ms)))
    return Tiles(desc[: 8 * nb.value].reshape(-1, 8), stage[: ns.value],
                 local[: csr.n_entries], ms.value)


# ---------------------------------------------------------------------------
# Device side
# ---------------------------------------------------------------------------
class DeviceArray:
    """Device buffer owned by a Context (the reference's cuda_ptr<T>)."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        _check(lib().ljgpu_malloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        _check(lib().ljgpu_memcpy_htod(self.ctx.handle, self.ptr, _ptr(a), self.nbytes))
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _check(lib().ljgpu_memcpy_dtoh(self.ctx.handle, _ptr(out), self.ptr, self.nbytes))
        return out

    def zero(self) -> None:
        _check(lib().ljgpu_memset(self.ctx.handle, self.ptr, 0, self.nbytes))

    def free(self) -> None:
        if self.ptr is not None and self.ctx.handle is not None:
            lib().ljgpu_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One device context (one HIP stream) of libljgpu."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib().ljgpu_create(device, C.byref(h)))
        self.handle = h.value

    def close(self) -> None:
        if self.handle is not None:
            lib().ljgpu_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # memory
    def array(self, a: np.ndarray) -> DeviceArray:
        return DeviceArray(self, a.shape, a.dtype).upload(a)

    def empty(self, shape, dtype=np.float64) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def zeros(self, shape, dtype=np.float64) -> DeviceArray:
        d = DeviceArray(self, shape, dtype)
        d.zero()
        return d

    def sync(self) -> None:
        _check(lib().ljgpu_sync(self.handle))

    # configuration
    def set_kernel_config(self, lanes: int = 0, block: int = 0, unroll: int = 0) -> None:
        """Lanes per i-particle (1: ELL kernel), workgroup size, entries per
        lane per pipelined batch; 0 keeps the current value."""
        _check(lib().ljgpu_set_kernel_config(self.handle, lanes, block, unroll))

    def set_half_mode(self, mode: int) -> None:
        _check(lib().ljgpu_set_half_mode(self.handle, mode))

    def set_tiling(self, tile_particles: int, stage_offsets: bool = False,
                   resident_per_cu: int = 0) -> None:
        """Particles per LDS tile (0: untiled global-gather kernels); with
        stage_offsets the tile's entry offsets are staged in LDS as well;
        resident_per_cu > 0 runs a persistent grid of that many workgroups per
        compute unit (0: one workgroup per tile)."""
        _check(lib().ljgpu_set_tiling(self.handle, tile_particles, int(stage_offsets),
                                      resident_per_cu))

    # lists
    def set_list(self, csr: CSR) -> None:
        _check(lib().ljgpu_set_list_csr(self.handle, csr.n, _ptr(csr.number_of_partners),
                                        _ptr(csr.pointer), _ptr(csr.sorted_list),
                                        csr.n_entries, csr.kind, MEM_HOST))

    def set_list_device(self, n: int, nop: DeviceArray, ptr: DeviceArray, lst: DeviceArray,
                        n_entries: int, kind: int) -> None:
        _need(nop, 4 * n, "number_of_partners")
        _need(ptr, 4 * n, "pointer")
        _need(lst, 4 * n_entries, "sorted_list")
        _check(lib().ljgpu_set_list_csr(self.handle, n, nop.ptr, ptr.ptr, lst.ptr, n_entries,
                                        kind, MEM_DEVICE))

    def set_list_transposed(self, n: int, nop: np.ndarray, transposed: np.ndarray, width: int,
                            kind: int) -> None:
        _check(lib().ljgpu_set_list_transposed(self.handle, n, _ptr(nop), _ptr(transposed),
                                               width, kind, MEM_HOST))

    def build_list(self, q: DeviceArray, n: int, box: float = 0.0, kind: int = LIST_HALF,
                   search_length: float = SEARCH_LENGTH) -> None:
        """Build the Verlet list on the device from device AoS positions
        (the reference's makepair + sortpair, exact pair set and order)."""
        _need(q, 32 * n, "q (AoS double4)")
        _check(lib().ljgpu_build_list(self.handle, q.ptr, n, box, search_length, kind))

    def build_list_aos3(self, q: DeviceArray, n: int, box: float = 0.0, kind: int = LIST_HALF,
                        search_length: float = SEARCH_LENGTH) -> None:
        """build_list from packed double[N][3] device positions."""
        _need(q, 24 * n, "q (AoS double3)")
        _check(lib().ljgpu_build_list_aos3(self.handle, q.ptr, n, box, search_length, kind))

    def build_list_soa(self, q3: list, n: int, box: float = 0.0, kind: int = LIST_HALF,
                       search_length: float = SEARCH_LENGTH) -> None:
        for c in range(3):
            _need(q3[c], 8 * n, f"q[{c}] (SoA)")
        _check(lib().ljgpu_build_list_soa(self.handle, q3[0].ptr, q3[1].ptr, q3[2].ptr, n, box,
                                          search_length, kind))

    def get_tiles(self) -> Tiles | None:
        """The LDS tiling in use, copied back (None when untiled)."""
        nb, ns = _i32(0), _i64(0)
        _check(lib().ljgpu_get_tiles(self.handle, None, 0, None, 0, None, 0, C.byref(nb),
                                     C.byref(ns)))
        if nb.value == 0:
            return None
        e = int(self.get_list().n_entries)
        desc = np.zeros(8 * nb.value, np.int32)
        stage = np.zeros(max(ns.value, 1), np.int32)
        local = np.zeros(max(e, 1), np.uint16)
        _check(lib().ljgpu_get_tiles(self.handle, _ptr(desc), desc.size, _ptr(stage), stage.size,
                                     _ptr(local), local.size, C.byref(nb), C.byref(ns)))
        d = desc.reshape(-1, 8)
        return Tiles(d, stage[: ns.value], local[:e], int(d[:, 3].max()))

    def set_tile_order(self, order: int) -> None:
        """TILE_ORDER_SPATIAL (default): tiles of nearby particles;
        TILE_ORDER_INDEX: tiles of consecutive indices."""
        _check(lib().ljgpu_set_tile_order(self.handle, order))

    def get_tile_order(self):
        """(order in use, particle of each slot), or (-1, None) when untiled."""
        order = C.c_int(0)
        _check(lib().ljgpu_get_tile_order(self.handle, C.byref(order), None, 0))
        if order.value < 0:
            return -1, None
        n = self.list_info().particle_number
        member = np.zeros(max(n, 1), np.int32)
        _check(lib().ljgpu_get_tile_order(self.handle, C.byref(order), _ptr(member), member.size))
        return order.value, member[:n]

    def set_replicas(self, offsets, boxes) -> None:
        """Batched replicas (offsets: n_replicas + 1 particle offsets; boxes:
        one side per replica); empty arguments return to one system."""
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        bx = np.ascontiguousarray(boxes, dtype=np.float64)
        nrep = max(off.size - 1, 0)
        _check(lib().ljgpu_set_replicas(self.handle, nrep, _ptr(off) if nrep else None,
                                        _ptr(bx) if nrep else None))

    def autotune(self, q: DeviceArray, box: float = 0.0, cl2: float = CL2,
                 precision: int = PREC_F64, accum: int = ACCUM_F64,
                 reps: int = 20) -> tuple[dict, list[dict]]:
        """Time every candidate kernel configuration on the installed list,
        install the fastest; returns (best, table)."""
        _need(q, 32 * self._particles(), "q (AoS double4)")
        best = TuneEntry()
        _check(lib().ljgpu_autotune(self.handle, q.ptr, box, cl2, precision, accum, reps,
                                    C.byref(best)))
        cnt = _i32(0)
        _check(lib().ljgpu_query_autotune(self.handle, None, 0, C.byref(cnt)))
        table = (TuneEntry * max(cnt.value, 1))()
        _check(lib().ljgpu_query_autotune(self.handle, table, cnt.value, C.byref(cnt)))
        return best.as_dict(), [table[i].as_dict() for i in range(cnt.value)]

    def set_graphs(self, enable: bool) -> None:
        _check(lib().ljgpu_set_graphs(self.handle, 1 if enable else 0))

    def set_chains(self, chains: int) -> None:
        """Concurrent launch chains of a many-step call over tiles (1 or 2)."""
        _check(lib().ljgpu_set_chains(self.handle, chains))

    def graph_stats(self) -> tuple[int, int]:
        """(graphs captured over the context's life, graphs cached now)."""
        cap, cached = _i64(0), _i32(0)
        _check(lib().ljgpu_graph_stats(self.handle, C.byref(cap), C.byref(cached)))
        return cap.value, cached.value

    def energy(self, q: DeviceArray, p: DeviceArray, box: float = 0.0,
               cl2: float = CL2) -> tuple[float, float]:
        """(potential, kinetic) over the installed list (unshifted potential)."""
        return self.energy_pairs(q, p, box, cl2)[:2]

    def energy_pairs(self, q: DeviceArray, p: DeviceArray, box: float = 0.0,
                     cl2: float = CL2) -> tuple[float, float, int]:
        """(potential, kinetic, pairs inside the cutoff)."""
        self._check_aos(q, p)
        u, k, c = _d(0.0), _d(0.0), _i64(0)
        _check(lib().ljgpu_energy(self.handle, q.ptr, p.ptr, box, cl2, C.byref(u), C.byref(k),
                                  C.byref(c)))
        return u.value, k.value, c.value

    def energy_pairs_noalloc(self, q: DeviceArray, p: DeviceArray, box: float = 0.0,
                             cl2: float = CL2) -> tuple[float, float, int]:
        """energy_pairs through ljgpu_energy_noalloc: no device allocation,
        partial sums in the GPU's static scratch (calls overlapping on one GPU
        race)."""
        self._check_aos(q, p)
        u, k, c = _d(0.0), _d(0.0), _i64(0)
        _check(lib().ljgpu_energy_noalloc(self.handle, q.ptr, p.ptr, box, cl2, C.byref(u),
                                          C.byref(k), C.byref(c)))
        return u.value, k.value, c.value

    def md_run(self, q: DeviceArray, p: DeviceArray, box: float, steps: int, dt: float = 0.005,
               cutoff: float = CUTOFF_LENGTH, skin: float = 0.3, rebuild_every: int = 20,
               precision: int = PREC_F64, accum: int = ACCUM_F64, sample_every: int = 0,
               max_samples: int = 0) -> tuple[MDStats, np.ndarray]:
        """Velocity-Verlet run in place on device AoS q, p; returns the stats
        and the energy samples (rows: step, U, K, U + K). In a periodic box
        no list stays installed afterwards (the run's lists index its
        internal, spatially sorted copies)."""
        _need(q, 32 * q.shape[0], "q (AoS double4)")
        _need(p, 32 * q.shape[0], "p (AoS double4)")
        cfg = MDConfig(dt, cutoff, skin, rebuild_every, precision, accum, sample_every)
        st = MDStats()
        cap = max_samples if max_samples > 0 else (steps // sample_every + 2 if sample_every else 2)
        buf = np.zeros((cap, 4), np.float64)
        _check(lib().ljgpu_md_run(self.handle, q.ptr, p.ptr, q.shape[0], box, C.byref(cfg), steps,
                                  C.byref(st), _ptr(buf), cap))
        return st, buf[: min(st.samples, cap)].copy()

    def list_info(self) -> ListInfo:
        info = ListInfo()
        _check(lib().ljgpu_get_list_info(self.handle, C.byref(info)))
        return info

    def get_list(self) -> CSR:
        info = self.list_info()
        n = info.particle_number
        # The evaluated list can be the expanded full list of a half input.
        nop = np.zeros(n, np.int32)
        ptr = np.zeros(n, np.int32)
        cap = max(2 * info.n_entries, 1)
        lst = np.zeros(cap, np.int32)
        _check(lib().ljgpu_get_list_csr(self.handle, _ptr(nop), _ptr(ptr), _ptr(lst), cap))
        total = int(nop.sum())
        kind = LIST_FULL if (info.kind == LIST_FULL or info.half_mode == HALF_EXPAND) else LIST_HALF
        return CSR(nop, ptr, lst[:total].copy(), kind)

    # forces
    def _particles(self) -> int:
        return int(self.list_info().particle_number)

    def _check_aos(self, q: DeviceArray, p: DeviceArray) -> None:
        n = self._particles()
        _need(q, 32 * n, "q (AoS double4)")
        _need(p, 32 * n, "p (AoS double4)")

    def _check_soa(self, q3: list, p3: list) -> None:
        n = self._particles()
        for c in range(3):
            _need(q3[c], 8 * n, f"q[{c}] (SoA)")
            _need(p3[c], 8 * n, f"p[{c}] (SoA)")

    def force_aos_f64(self, q: DeviceArray, p: DeviceArray, dt: float, cl2: float = CL2,
                      box: float = 0.0, steps: int = 1) -> None:
        self._check_aos(q, p)
        _check(lib().ljgpu_force_aos_f64(self.handle, q.ptr, p.ptr, dt, cl2, box, steps))

    def _check_aos3(self, q: DeviceArray, p: DeviceArray) -> None:
        n = self._particles()
        _need(q, 24 * n, "q (AoS double3)")
        _need(p, 24 * n, "p (AoS double3)")

    def force_aos3_f64(self, q: DeviceArray, p: DeviceArray, dt: float, cl2: float = CL2,
                       box: float = 0.0, steps: int = 1) -> None:
        """force_aos_f64 over packed double[N][3] arrays."""
        self._check_aos3(q, p)
        _check(lib().ljgpu_force_aos3_f64(self.handle, q.ptr, p.ptr, dt, cl2, box, steps))

    def force_aos3_mixed(self, q: DeviceArray, p: DeviceArray, dt: float, cl2: float = CL2,
                         box: float = 0.0, steps: int = 1, accum: int = ACCUM_F64) -> None:
        self._check_aos3(q, p)
        _check(lib().ljgpu_force_aos3_mixed(self.handle, q.ptr, p.ptr, dt, cl2, box, steps,
                                            accum))

    def force_soa_f64(self, q3: list, p3: list, dt: float, cl2: float = CL2, box: float = 0.0,
                      steps: int = 1) -> None:
        self._check_soa(q3, p3)
        _check(lib().ljgpu_force_soa_f64(self.handle, q3[0].ptr, q3[1].ptr, q3[2].ptr, p3[0].ptr,
                                         p3[1].ptr, p3[2].ptr, dt, cl2, box, steps))

    def force_aos_mixed(self, q: DeviceArray, p: DeviceArray, dt: float, cl2: float = CL2,
                        box: float = 0.0, steps: int = 1, accum: int = ACCUM_F64) -> None:
        self._check_aos(q, p)
        _check(lib().ljgpu_force_aos_mixed(self.handle, q.ptr, p.ptr, dt, cl2, box, steps, accum))

    def force_soa_mixed(self, q3: list, p3: list, dt: float, cl2: float = CL2, box: float = 0.0,
                        steps: int = 1, accum: int = ACCUM_F64) -> None:
        self._check_soa(q3, p3)
        _check(lib().ljgpu_force_soa_mixed(self.handle, q3[0].ptr, q3[1].ptr, q3[2].ptr, p3[0].ptr,
                                           p3[1].ptr, p3[2].ptr, dt, cl2, box, steps, accum))

    def force_kernel_csr_f64(self, q: DeviceArray, p: DeviceArray, n: int, dt: float, cl2: float,
                             sorted_list: DeviceArray, nop: DeviceArray, ptr: DeviceArray,
                             n_entries: int, kind: int) -> None:
        _need(q, 32 * n, "q (AoS double4)")
        _need(p, 32 * n, "p (AoS double4)")
        _need(nop, 4 * n, "number_of_partners")
        _need(ptr, 4 * n, "pointer")
        _need(sorted_list, 4 * n_entries, "sorted_list")
        _check(lib().ljgpu_force_kernel_csr_f64(self.handle, q.ptr, p.ptr, n, dt, cl2,
                                                sorted_list.ptr, nop.ptr, ptr.ptr, n_entries,
                                                kind))

    # timing
    def timer_start(self) -> None:
        _check(lib().ljgpu_timer_start(self.handle))

    def timer_stop(self) -> float:
        ms = C.c_float(0.0)
        _check(lib().ljgpu_timer_stop(self.handle, C.byref(ms)))
        return float(ms.value)


def device_count() -> int:
    c = C.c_int(0)
    rc = lib().ljgpu_device_count(C.byref(c))
    return c.value if rc == OK else 0


def measure(ctx: Context, q: DeviceArray, p: DeviceArray, dt: float, steps: int = 100,
            cl2: float = CL2, box: float = 0.0, mixed: bool = False,
            accum: int = ACCUM_F64) -> float:
    """The reference's measure(): `steps` force calls on the installed list,
    timed on the context stream; returns milliseconds (cuda/force_cuda.cu:319-342)."""
    ctx.timer_start()
    if mixed:
        ctx.force_aos_mixed(q, p, dt, cl2, box, steps, accum)
    else:
        ctx.force_aos_f64(q, p, dt, cl2, box, steps)
    return ctx.timer_stop()
