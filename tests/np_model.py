"""Independent numpy formulation of the CAVE cube step, used to pin the oracle.

It is written from the mathematics, not from the reference's code: the
neighbourhood is {d in {-1,0,1}^3 : |d|_1 in selected classes} (class 1 =
faces, 2 = edges, 3 = corners), counted by slicing a zero-padded live array,
and the transition is the three-case table of the search kernel
(searchSimulationCube.slang:178-211).  Cells are a dense (Z, Y, X) uint8
array; pack/unpack convert to the reference row-padded nibble layout.
"""
from __future__ import annotations

import itertools

import numpy as np


def unpack(words: np.ndarray, X: int, Y: int, Z: int) -> np.ndarray:
    upr = (X + 7) // 8
    w = np.asarray(words, dtype=np.uint32).reshape(Z, Y, upr)
    nib = np.stack([(w >> (4 * i)) & 0xF for i in range(8)], axis=-1).reshape(Z, Y, upr * 8)
    return nib[:, :, :X].astype(np.uint8)


def pack(cells: np.ndarray) -> np.ndarray:
    Z, Y, X = cells.shape
    upr = (X + 7) // 8
    pad = np.zeros((Z, Y, upr * 8), dtype=np.uint32)
    pad[:, :, :X] = cells
    pad = pad.reshape(Z, Y, upr, 8)
    w = np.zeros((Z, Y, upr), dtype=np.uint32)
    for i in range(8):
        w |= pad[..., i] << np.uint32(4 * i)
    return w.reshape(-1)


def offsets(flags: int):
    for d in itertools.product((-1, 0, 1), repeat=3):
        c = sum(1 for v in d if v)
        if c and (flags >> (c - 1)) & 1:
            yield d


def neighbor_counts(live: np.ndarray, flags: int, wrap: bool = False) -> np.ndarray:
    Z, Y, X = live.shape
    if wrap:
        # toroidal: one rolled copy per offset, so a short axis counts the same
        # cell (or the cell itself) once per offset that reaches it
        n = np.zeros((Z, Y, X), dtype=np.int32)
        for dx, dy, dz in offsets(flags):
            n += np.roll(live, shift=(-dz, -dy, -dx), axis=(0, 1, 2))
        return n
    p = np.zeros((Z + 2, Y + 2, X + 2), dtype=np.int32)
    p[1:-1, 1:-1, 1:-1] = live
    n = np.zeros((Z, Y, X), dtype=np.int32)
    for dx, dy, dz in offsets(flags):
        n += p[1 + dz:1 + dz + Z, 1 + dy:1 + dy + Y, 1 + dx:1 + dx + X]
    return n


def step(cells: np.ndarray, M: int, flags: int, birth: int, survival: int, wrap: bool = False):
    """Returns (new cells, (dead, dying, born, surviving))."""
    live = (cells == M).astype(np.int32)
    n = neighbor_counts(live, flags, wrap)
    bit = lambda mask: np.where(n > 0, (np.uint64(mask) >> (np.maximum(n, 1) - 1).astype(np.uint64)) & np.uint64(1), 0).astype(bool)
    surv = (cells == M) & bit(survival)
    born = (cells == 0) & bit(birth)
    out = np.where(cells == 0, 0, cells.astype(np.int32) - 1)
    out = np.where(surv | born, M, out).astype(np.int64) & 0xF
    dead = int(((cells == 0) & ~born).sum())
    dying = int(((cells != 0) & ~surv).sum())
    return out.astype(np.uint8), (dead, dying, int(born.sum()), int(surv.sum()))


# ---- elongated rhombic dodecahedron cells (ERD) ----
# Cell centres of the reference's ERD honeycomb (src/shapes/elongatedRhombicDodecahedron.h:
# ComputePosition): x spacing sqrt(3), y spacing 2, z spacing sqrt(3), odd y rows
# shifted by sqrt(3)/2 in x and z.
def erd_position(x, y, z):
    h = np.sqrt(3.0) / 2.0
    odd = (np.asarray(y) % 2 == 1).astype(np.float64)
    return np.stack([np.asarray(x) * np.sqrt(3.0) + odd * h, np.asarray(y) * 2.0,
                     np.asarray(z) * np.sqrt(3.0) + odd * h], axis=-1)


def erd_geometric_neighbors(parity):
    """Offsets d in {-1,0,1}^3 from a cell on a row of the given y parity,
    sorted by the distance between cell centres: the 12 nearest share a face,
    the next 4 share an edge."""
    base = np.array([4, 4 + parity, 4])
    ds = [np.array(d) for d in itertools.product((-1, 0, 1), repeat=3) if any(d)]
    dist = [float(np.linalg.norm(erd_position(*(base + d)) - erd_position(*base))) for d in ds]
    order = np.argsort(dist, kind="stable")
    return [tuple(int(v) for v in ds[i]) for i in order], [dist[i] for i in order]


def erd_step(cells: np.ndarray, M: int, flags: int, birth: int, survival: int, wrap: bool = False):
    """ERD transition with neighbour sets from the geometry: faces = the 12
    nearest centres, edges = the next 4 (selected per y parity of the cell)."""
    Z, Y, X = cells.shape
    live = (cells == M).astype(np.int32)
    n = np.zeros((Z, Y, X), dtype=np.int32)
    for parity in (0, 1):
        nb, _ = erd_geometric_neighbors(parity)
        sel = (nb[:12] if flags & 1 else []) + (nb[12:16] if flags & 2 else [])
        cnt = np.zeros((Z, Y, X), dtype=np.int32)
        if wrap:
            for dx, dy, dz in sel:
                cnt += np.roll(live, shift=(-dz, -dy, -dx), axis=(0, 1, 2))
        else:
            p = np.zeros((Z + 2, Y + 2, X + 2), dtype=np.int32)
            p[1:-1, 1:-1, 1:-1] = live
            for dx, dy, dz in sel:
                cnt += p[1 + dz:1 + dz + Z, 1 + dy:1 + dy + Y, 1 + dx:1 + dx + X]
        rows = (np.arange(Y) % 2 == parity)
        n[:, rows, :] = cnt[:, rows, :]
    bit = lambda mask: np.where(n > 0, (np.uint64(mask) >> (np.maximum(n, 1) - 1).astype(np.uint64)) & np.uint64(1), 0).astype(bool)
    surv = (cells == M) & bit(survival)
    born = (cells == 0) & bit(birth)
    out = np.where(cells == 0, 0, cells.astype(np.int32) - 1)
    out = np.where(surv | born, M, out).astype(np.int64) & 0xF
    dead = int(((cells == 0) & ~born).sum())
    dying = int(((cells != 0) & ~surv).sum())
    return out.astype(np.uint8), (dead, dying, int(born.sum()), int(surv.sum()))


def pack_linear(cells: np.ndarray) -> np.ndarray:
    flat = cells.reshape(-1).astype(np.uint32)
    n = (flat.size + 7) // 8
    pad = np.zeros(n * 8, np.uint32)
    pad[:flat.size] = flat
    pad = pad.reshape(n, 8)
    w = np.zeros(n, np.uint32)
    for i in range(8):
        w |= pad[:, i] << np.uint32(4 * i)
    return w


def unpack_linear(words: np.ndarray, X: int, Y: int, Z: int) -> np.ndarray:
    w = np.asarray(words, dtype=np.uint32)
    nib = np.stack([(w >> (4 * i)) & 0xF for i in range(8)], axis=-1).reshape(-1)
    return nib[: X * Y * Z].astype(np.uint8).reshape(Z, Y, X)
