"""Shared workloads and oracle results for the test suite (cached per process)."""
from __future__ import annotations

import functools
import os

import numpy as np

import lj_gpu_amd as G
import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_DATA_05 = os.path.join(GOLDEN, "ref_data_density0.5.dat")
REF_DATA_1 = os.path.join(GOLDEN, "ref_data_density1.dat")

# Periodic workloads: FCC 4*k^3 particles, seeded jitter U(-0.05, 0.05),
# cutoff 3.0, search length 3.3, dt = 0.01, 100 steps.
PBC_DT = 0.01
STEPS = 100
FP64_TOL = 1e-12  # normwise max|dp| / max|p| against the FP64 oracle


def normwise(a: np.ndarray, ref: np.ndarray) -> float:
    a = np.asarray(a)[..., :3]
    ref = np.asarray(ref)[..., :3]
    scale = np.abs(ref).max()
    return float(np.abs(a - ref).max() / scale) if scale > 0 else float(np.abs(a).max())


def rms_normwise(a: np.ndarray, ref: np.ndarray) -> float:
    a = np.asarray(a)[..., :3]
    ref = np.asarray(ref)[..., :3]
    return float(np.sqrt(((a - ref) ** 2).mean()) / np.abs(ref).max())


def load_golden(name: str) -> dict:
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@functools.lru_cache(maxsize=None)
def periodic(k: int, density: float, seed: int = 2):
    """(q (N,4), box, half CSR, full CSR) from the product's host builder."""
    q, box = G.lattice_periodic(k, density, seed)
    half = G.makepair(q, box, G.LIST_HALF)
    full = G.makepair(q, box, G.LIST_FULL)
    return q, box, half, full


@functools.lru_cache(maxsize=None)
def oracle_momenta(k: int, density: float, steps: int = STEPS, dt: float = PBC_DT):
    """Oracle force_sorted (half list, j scatter) after `steps` steps from p = 0."""
    q, box, half, _ = periodic(k, density)
    csr = O.CSR(half.number_of_partners, half.pointer, half.sorted_list, "half")
    p = np.zeros_like(q)
    O.force_aos(csr, q, p, dt, O.CL2, box, steps=steps, method="sorted")
    return p


@functools.lru_cache(maxsize=None)
def oracle_force_reference(density: float = 0.5):
    """Oracle F (one step with dt = 1) of the reference's own problem (open
    box), from its half list."""
    q, half = reference_mode(density)
    csr = O.CSR(half.number_of_partners, half.pointer, half.sorted_list, "half")
    p = np.zeros_like(q)
    O.force_aos(csr, q, p, 1.0, O.CL2, 0.0, steps=1, method="sorted")
    return p


@functools.lru_cache(maxsize=None)
def oracle_force(k: int, density: float):
    """Oracle F (one step with dt = 1) from the full list."""
    q, box, _, full = periodic(k, density)
    csr = O.CSR(full.number_of_partners, full.pointer, full.sorted_list, "full")
    p = np.zeros_like(q)
    O.force_aos(csr, q, p, 1.0, O.CL2, box, steps=1, method="full")
    return p


@functools.lru_cache(maxsize=None)
def reference_mode(density: float = 0.5):
    """The reference's own problem: open box L=50, U(0,0.1) jitter, half list."""
    q = G.lattice_reference(density, 50.0)
    half = G.makepair(q, 0.0, G.LIST_HALF)
    return q, half


def to_soa(q4: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(q4[:, :3].T)


def column_order(q: np.ndarray, box: float, tile_particles: int = 64, offsets=None,
                 boxes=None) -> np.ndarray:
    """Numpy restatement of the first step of the library's spatial tile
    order: each periodic system cut into columns along x about
    0.75 * cbrt(T / density) wide in y and z (the count per side rounded to
    the nearest integer), each column into x bins of about one particle (the
    bin of a coordinate is floor(wrapped / width), clamped), particles ordered
    by (system, z column, y column, x bin, index). length_sorted applies the
    second step."""
    n = q.shape[0]
    off = [0, n] if offsets is None else [int(v) for v in offsets]
    bx = [box] if offsets is None else [float(v) for v in boxes]
    keys = np.zeros(n, np.int64)
    key0 = 0
    for r in range(len(off) - 1):
        a, b, L = off[r], off[r + 1], bx[r]
        m = max(b - a, 1)
        rho = m / (L * L * L)
        side = 0.75 * np.cbrt(tile_particles / rho)
        nc = [1, 1, 1]
        for c in (1, 2):
            v = np.floor(L / side + 0.5)
            nc[c] = 1024 if v >= 1024.0 else max(1, int(v))
        while nc[1] * nc[2] > 2 * m + 16:  # at most about 2 columns per particle
            big = 1 if nc[1] >= nc[2] else 2
            nc[big] = (nc[big] + 1) // 2
        nc[0] = max(1, m // (nc[1] * nc[2]))
        cell = []
        for c in range(3):
            x = q[a:b, c] - L * np.floor(q[a:b, c] / L)
            v = np.floor(x / (L / nc[c])).astype(np.int64)
            cell.append(np.clip(v, 0, nc[c] - 1))
        keys[a:b] = key0 + (cell[2] * nc[1] + cell[1]) * nc[0] + cell[0]
        key0 += nc[0] * nc[1] * nc[2]
    return np.lexsort((np.arange(n), keys)).astype(np.int32)


def bank_class_order(offs: np.ndarray, starts: np.ndarray, lengths: np.ndarray,
                     pos: np.ndarray, k0=None) -> np.ndarray:
    """Numpy restatement of the library's entry order within tile rows.

    offs: LDS byte offsets (24 * staged rank) in list order; row i spans
    offs[starts[i] : starts[i] + lengths[i]] and sits at position pos[i] of
    its tile. An entry's class is rank mod 32. Entry k of a row at position r
    takes the first entry (in list order) of class (4 * (r & 7) + k0 + k)
    mod 32 that the row still holds, or of the next class up, cyclically,
    that it does (k0: where the row starts within its tile row, 0 by
    default). Rows shorter than 2 or longer than 256 entries keep the list
    order. All rows advance together, one entry index per step."""
    offs = np.asarray(offs)
    out = offs.copy()
    starts, lengths, pos = (np.asarray(a, np.int64) for a in (starts, lengths, pos))
    k0 = np.zeros_like(pos) if k0 is None else np.asarray(k0, np.int64)
    sel = np.nonzero((lengths >= 2) & (lengths <= 256))[0]
    if sel.size == 0:
        return out
    st, ln, ps, kk = starts[sel], lengths[sel], pos[sel], k0[sel]
    R = sel.size
    row_base = np.concatenate([[0], np.cumsum(ln)[:-1]])
    row_of = np.repeat(np.arange(R), ln)
    e = np.repeat(st - row_base, ln) + np.arange(int(ln.sum()))
    vals = offs[e].astype(np.int64)
    cls = (vals // 24) % 32
    key = row_of * 32 + cls
    grouped = vals[np.argsort(key, kind="stable")]  # by row, class, list position
    cnt = np.bincount(key, minlength=R * 32).reshape(R, 32)
    cstart = row_base[:, None] + np.concatenate(
        [np.zeros((R, 1), np.int64), np.cumsum(cnt, 1)[:, :-1]], 1)
    taken = np.zeros((R, 32), np.int64)
    rows = np.arange(R)
    for k in range(int(ln.max())):
        live = rows[ln > k]
        c = (4 * (ps[live] % 8) + kk[live] + k) % 32
        miss = np.nonzero(cnt[live, c] == taken[live, c])[0]
        if miss.size:  # target class used up: the next class up that is left
            lm = live[miss]
            cand = (c[miss][:, None] + np.arange(32)) % 32
            left = cnt[lm[:, None], cand] - taken[lm[:, None], cand]
            c[miss] = cand[np.arange(miss.size), (left > 0).argmax(1)]
        out[st[live] + k] = grouped[cstart[live, c] + taken[live, c]]
        taken[live, c] += 1
    return out


def length_sorted(member: np.ndarray, nop: np.ndarray, tile_particles: int,
                  offsets=None) -> np.ndarray:
    """The library's spatial order after its second step: within each block of
    tile_particles consecutive slots (restarted at every system boundary),
    slots reordered by the row length nop of their particles, longest first,
    ties kept in order."""
    member = np.array(member, dtype=np.int32)
    nop = np.asarray(nop)
    off = [0, member.size] if offsets is None else [int(v) for v in offsets]
    for a, b in zip(off[:-1], off[1:]):
        for i0 in range(a, b, tile_particles):
            seg = member[i0:min(i0 + tile_particles, b)].copy()
            member[i0:i0 + seg.size] = seg[np.argsort(-nop[seg], kind="stable")]
    return member


def cut_inner(q: np.ndarray, box: float, cl2: float, i: np.ndarray, j: np.ndarray) -> np.ndarray:
    """The tile rows' cutoff classes: r2 <= cl2 with r2 = (dx*dx + dy*dy) +
    dz*dz (separately rounded) over the threshold minimum image of
    d = q_j - q_i (box > 0), as the library evaluates it."""
    d = q[j, :3] - q[i, :3]
    if box > 0:
        h = 0.5 * box
        d = np.where(d > h, d - box, np.where(d < -h, d + box, d))
    r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    return r2 <= cl2
