"""Seeded synthetic tetrahedral meshes for the VBD workloads.

All meshes use one fixed 6-tet Kuhn (Freudenthal) split of each hexahedral cell: the six
tets share the cell diagonal v000-v111, tet for axis order (a, b, c) is
(v000, v000+e_a, v000+e_a+e_b, v111), and odd permutations swap their middle two corners so
every tet has positive orientation.  Vertex ids of a grid are i + (nx+1) * (j + (ny+1) * k)
(i fastest).  Cells are visited k-major, then j, then i, and each cell emits its six tets
in lexicographic permutation order.

The reference repository ships no meshes (its main.cpp:180-183 loads TetGen files from
hard-coded Windows paths), so every workload here is generated.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

# (axis order, parity) of the six Kuhn tets
_PERMS = [((0, 1, 2), 1), ((0, 2, 1), -1), ((1, 0, 2), -1), ((1, 2, 0), 1), ((2, 0, 1), 1), ((2, 1, 0), -1)]


def _kuhn_local() -> np.ndarray:
    """Six tets as corner offsets (dx, dy, dz) in {0,1}^3, encoded as dx + 2 dy + 4 dz."""
    out = []
    for perm, parity in _PERMS:
        p0 = np.zeros(3, dtype=np.int64)
        p1 = p0.copy(); p1[perm[0]] = 1
        p2 = p1.copy(); p2[perm[1]] = 1
        p3 = np.ones(3, dtype=np.int64)
        corners = [p0, p1, p2, p3]
        if parity < 0:
            corners[1], corners[2] = corners[2], corners[1]
        out.append([int(c[0] + 2 * c[1] + 4 * c[2]) for c in corners])
    return np.asarray(out, dtype=np.int64)


KUHN = _kuhn_local()


@dataclass
class Mesh:
    """A tetrahedral mesh plus the boundary conditions and material of one workload."""

    name: str
    rest: np.ndarray                 # (N, 3) float32 rest positions
    x0: np.ndarray                   # (N, 3) float32 initial positions
    tets: np.ndarray                 # (T, 4) int32
    fixed: np.ndarray                # (N,) uint8
    mu: float = 1e5
    lam: float = 4e5
    density: float = 1000.0
    tet_mu: Optional[np.ndarray] = None      # (T,) float32 or None
    tet_lambda: Optional[np.ndarray] = None  # (T,) float32 or None
    gravity: tuple = (0.0, -9.8, 0.0)
    v0: Optional[np.ndarray] = None          # (N, 3) float32 initial velocity or None
    meta: dict = field(default_factory=dict)

    @property
    def num_verts(self) -> int:
        return int(self.rest.shape[0])

    @property
    def num_tets(self) -> int:
        return int(self.tets.shape[0])


def grid_tets(nx: int, ny: int, nz: int, cell_mask: Optional[np.ndarray] = None) -> np.ndarray:
    """Kuhn-split tets of an nx*ny*nz cell grid; cell_mask (nz, ny, nx) keeps a subset."""
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    if cell_mask is not None:
        sel = cell_mask.astype(bool)
        i, j, k = i[sel], j[sel], k[sel]
    else:
        i, j, k = i.ravel(), j.ravel(), k.ravel()
    sx, sy = nx + 1, (nx + 1) * (ny + 1)
    base = i + sx * j + sy * k
    corner_off = np.array([dx + sx * dy + sy * dz for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)], dtype=np.int64)
    tets = base[:, None, None] + corner_off[KUHN][None, :, :]
    return tets.reshape(-1, 4).astype(np.int32)


def grid_positions(nx: int, ny: int, nz: int, cell: float, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    p = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64) * cell
    return (p + np.asarray(origin, dtype=np.float64)).astype(np.float32)


def beam(nx: int = 20, ny: int = 4, nz: int = 4, cell: float = 0.05, mu: float = 1e5, lam: float = 4e5,
         density: float = 1000.0, name: str = "beam") -> Mesh:
    """Cantilever beam, x = 0 face pinned, gravity -y."""
    rest = grid_positions(nx, ny, nz, cell)
    tets = grid_tets(nx, ny, nz)
    i = np.arange(rest.shape[0]) % (nx + 1)
    fixed = (i == 0).astype(np.uint8)
    return Mesh(name=name, rest=rest, x0=rest.copy(), tets=tets, fixed=fixed, mu=mu, lam=lam, density=density,
                gravity=(0.0, -9.8, 0.0), meta=dict(nx=nx, ny=ny, nz=nz, cell=cell))


def noisy_cube(n: int = 128, side: float = 1.0, noise: float = 0.3, seed: int = 0, mu: float = 1e5,
               lam: float = 4e5, density: float = 1000.0) -> Mesh:
    """n^3-cell cube of side `side`; bottom (min-y) face pinned; free vertices start at rest
    plus seeded uniform +-noise*cell per component."""
    cell = side / n
    rest = grid_positions(n, n, n, cell)
    tets = grid_tets(n, n, n)
    j = (np.arange(rest.shape[0]) // (n + 1)) % (n + 1)
    fixed = (j == 0).astype(np.uint8)
    rng = np.random.default_rng(seed)
    jit = rng.uniform(-noise, noise, size=rest.shape) * cell
    x0 = rest.astype(np.float64) + jit * (1 - fixed[:, None])
    return Mesh(name=f"cube{n}", rest=rest, x0=x0.astype(np.float32), tets=tets, fixed=fixed, mu=mu, lam=lam,
                density=density, gravity=(0.0, -9.8, 0.0), meta=dict(n=n, cell=cell, seed=seed))


def voxel_sphere(n: int = 96, radius: float = 1.0, jitter: float = 0.2, seed: int = 0, mu: float = 5e4,
                 lam: float = 2e5, density: float = 1000.0, shuffle: bool = True) -> Mesh:
    """Sphere voxelised on an n^3 grid over [-r, r]^3 (cells whose centre lies inside kept),
    Kuhn split, interior vertices jittered by seeded uniform +-jitter*cell, vertex ids shuffled
    by seed.  No pins; gravity -z (the contact plane is a solver parameter)."""
    cell = 2.0 * radius / n
    c = (np.arange(n) + 0.5) * cell - radius
    kz, jy, ix = np.meshgrid(c, c, c, indexing="ij")
    mask = (ix ** 2 + jy ** 2 + kz ** 2) < radius ** 2
    tets_full = grid_tets(n, n, n, cell_mask=mask)
    grid = grid_positions(n, n, n, cell, origin=(-radius, -radius, -radius))
    used = np.unique(tets_full)
    remap = -np.ones(grid.shape[0], dtype=np.int64)
    remap[used] = np.arange(used.size)
    tets = remap[tets_full].astype(np.int32)
    rest = grid[used].astype(np.float64)
    # interior vertex = all 8 incident cells kept
    cnt = np.zeros(grid.shape[0], dtype=np.int64)
    k, j, i = np.nonzero(mask)
    sx, sy = n + 1, (n + 1) * (n + 1)
    base = i + sx * j + sy * k
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                np.add.at(cnt, base + dx + sx * dy + sy * dz, 1)
    interior = cnt[used] == 8
    rng = np.random.default_rng(seed)
    jit = rng.uniform(-jitter, jitter, size=rest.shape) * cell
    rest = rest + jit * interior[:, None]
    nv = rest.shape[0]
    if shuffle:
        perm = rng.permutation(nv)          # new id of old vertex v is inv[v]
        inv = np.empty(nv, dtype=np.int64)
        inv[perm] = np.arange(nv)
        rest = rest[perm]
        tets = inv[tets].astype(np.int32)
    rest = rest.astype(np.float32)
    return Mesh(name=f"sphere{n}", rest=rest, x0=rest.copy(), tets=tets, fixed=np.zeros(nv, dtype=np.uint8), mu=mu,
                lam=lam, density=density, gravity=(0.0, 0.0, -9.8),
                meta=dict(n=n, cell=cell, seed=seed, cells=int(mask.sum())))


def batch_beams(count: int = 256, seed: int = 0, cell: float = 0.05, nx_range=(8, 64), nyz_range=(2, 8),
                density: float = 1000.0) -> List[Mesh]:
    """Independent beams of uneven size: nx ~ U{8..64}, ny = nz ~ U{2..8}, per-beam mu
    log-uniform in [1e4, 1e6] Pa and Poisson ratio nu ~ U[0.3, 0.45], lambda = 2 mu nu / (1 - 2 nu)."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        nyz = int(rng.integers(nyz_range[0], nyz_range[1] + 1))
        mu = float(10.0 ** rng.uniform(4.0, 6.0))
        nu = float(rng.uniform(0.3, 0.45))
        lam = 2.0 * mu * nu / (1.0 - 2.0 * nu)
        m = beam(nx, nyz, nyz, cell=cell, mu=mu, lam=lam, density=density, name=f"beam{b}")
        m.tet_mu = np.full(m.num_tets, mu, dtype=np.float32)
        m.tet_lambda = np.full(m.num_tets, lam, dtype=np.float32)
        m.meta.update(nu=nu)
        out.append(m)
    return out


def smooth_field(points: np.ndarray, seed: int, lo: float, hi: float, terms: int = 4) -> np.ndarray:
    """Sum of low-frequency sines over `points` (M, 3), mapped affinely onto [lo, hi]."""
    rng = np.random.default_rng(seed)
    acc = np.zeros(points.shape[0], dtype=np.float64)
    for _ in range(terms):
        kvec = rng.uniform(-2.0 * np.pi, 2.0 * np.pi, size=3)
        phase = rng.uniform(0, 2 * np.pi)
        acc += np.sin(points @ kvec + phase) * rng.uniform(0.5, 1.0)
    a, b = acc.min(), acc.max()
    if b - a < 1e-12:
        return np.full(points.shape[0], 0.5 * (lo + hi))
    return lo + (acc - a) / (b - a) * (hi - lo)


def diff_cube(n: int = 16, side: float = 1.0, seed: int = 0, lo: float = 5e4, hi: float = 2e5,
              density: float = 1000.0) -> Mesh:
    """Cube for the differentiable workload: per-tet mu from a smooth seeded field in [lo, hi],
    lambda = 4 mu, bottom (min-y) face pinned, gravity -y."""
    cell = side / n
    rest = grid_positions(n, n, n, cell)
    tets = grid_tets(n, n, n)
    j = (np.arange(rest.shape[0]) // (n + 1)) % (n + 1)
    fixed = (j == 0).astype(np.uint8)
    cent = rest[tets].astype(np.float64).mean(axis=1)
    mu = smooth_field(cent, seed, lo, hi).astype(np.float32)
    return Mesh(name=f"diffcube{n}", rest=rest, x0=rest.copy(), tets=tets, fixed=fixed, mu=float(mu.mean()),
                lam=float(4 * mu.mean()), density=density, tet_mu=mu, tet_lambda=(4.0 * mu).astype(np.float32),
                gravity=(0.0, -9.8, 0.0), meta=dict(n=n, cell=cell, seed=seed))


def diff_workload(n: int = 16, seed: int = 0, v_scale: float = 0.05):
    """The differentiable workload: (mesh with per-tet mu field (seed), lambda = 4 mu,
    ground-truth mu field (seed + 1), its lambda = 4 mu, initial velocity v0 (N,3)).
    v0 is a seeded smooth field of magnitude ~v_scale m/s on free vertices."""
    m = diff_cube(n, seed=seed)
    cent = m.rest[m.tets].astype(np.float64).mean(axis=1)
    mu_true = smooth_field(cent, seed + 1, 5e4, 2e5).astype(np.float32)
    lam_true = (4.0 * mu_true).astype(np.float32)
    p = m.rest.astype(np.float64)
    v0 = np.stack([smooth_field(p, seed + 10 + c, -v_scale, v_scale) for c in range(3)], axis=1)
    v0[m.fixed == 1] = 0.0
    return m, mu_true, lam_true, v0.astype(np.float32)


def concat(meshes: List[Mesh], name: str = "batch") -> tuple:
    """Disjoint union of meshes: returns (Mesh with per-tet materials, vert_offsets, tet_offsets)."""
    vo = np.zeros(len(meshes) + 1, dtype=np.int64)
    to = np.zeros(len(meshes) + 1, dtype=np.int64)
    for i, m in enumerate(meshes):
        vo[i + 1] = vo[i] + m.num_verts
        to[i + 1] = to[i] + m.num_tets
    rest = np.concatenate([m.rest for m in meshes]).astype(np.float32)
    x0 = np.concatenate([m.x0 for m in meshes]).astype(np.float32)
    tets = np.concatenate([m.tets.astype(np.int64) + vo[i] for i, m in enumerate(meshes)]).astype(np.int32)
    fixed = np.concatenate([m.fixed for m in meshes]).astype(np.uint8)
    tmu = np.concatenate([m.tet_mu if m.tet_mu is not None else np.full(m.num_tets, m.mu, np.float32)
                          for m in meshes]).astype(np.float32)
    tlam = np.concatenate([m.tet_lambda if m.tet_lambda is not None else np.full(m.num_tets, m.lam, np.float32)
                           for m in meshes]).astype(np.float32)
    u = Mesh(name=name, rest=rest, x0=x0, tets=tets, fixed=fixed, mu=meshes[0].mu, lam=meshes[0].lam,
             density=meshes[0].density, tet_mu=tmu, tet_lambda=tlam, gravity=meshes[0].gravity)
    return u, vo, to


def write_tetgen(prefix: str, positions: np.ndarray, tets: np.ndarray, one_based: bool = True,
                 chunk: int = 65536) -> None:
    """Write TetGen .node/.ele files (the input format of the reference loader, main.cpp:14-168):
    coordinates as the shortest decimal that reads back as the same double (Python repr), one
    row per node / tet, formatted in chunks of `chunk` rows (a 128^3 cube's 2.1M nodes and 12.6M
    tets take seconds)."""
    base = 1 if one_based else 0
    pos = np.asarray(positions, dtype=np.float64)
    tt = np.asarray(tets, dtype=np.int64)
    with open(prefix + ".node", "w") as f:
        f.write(f"{pos.shape[0]} 3 0 0\n")
        for a in range(0, pos.shape[0], chunk):
            b = min(a + chunk, pos.shape[0])
            rows = np.empty((b - a, 4), dtype=object)
            rows[:, 0] = np.arange(a + base, b + base).tolist()
            rows[:, 1:] = np.array(pos[a:b].tolist(), dtype=object)
            f.write(("%d %r %r %r\n" * (b - a)) % tuple(rows.ravel().tolist()))
    with open(prefix + ".ele", "w") as f:
        f.write(f"{tt.shape[0]} 4 0\n")
        for a in range(0, tt.shape[0], chunk):
            b = min(a + chunk, tt.shape[0])
            rows = np.empty((b - a, 5), dtype=np.int64)
            rows[:, 0] = np.arange(a + base, b + base)
            rows[:, 1:] = tt[a:b] + base
            f.write(("%d %d %d %d %d\n" * (b - a)) % tuple(rows.ravel().tolist()))


def _tetgen_rows(path: str, dtype):
    """Header fields and the first n body rows of a TetGen file as an (n, columns) array; empty
    lines and lines starting with '#' are skipped (as include/vbd/TetGenIO.h does)."""
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines() if ln.strip() and not ln.startswith("#")]
    head = lines[0].split()
    n = int(head[0])
    body = lines[1:1 + n]
    if len(body) != n:
        raise ValueError(f"{path}: {len(body)} of {n} rows")
    tokens = " ".join(body).split()
    if n and len(tokens) % n == 0:  # the usual case: every row has the same column count
        return head, np.array(tokens, dtype=dtype).reshape(n, -1)
    width = min(len(r.split()) for r in body)
    return head, np.array([r.split()[:width] for r in body], dtype=dtype)


def read_tetgen(prefix: str):
    """Read TetGen .node/.ele files: (positions float32 (N,3), tets int32 (T,4)); 0- or 1-based
    ids detected from the first node index, as main.cpp:38-39 and :112-114 do.  Coordinates go
    decimal -> float64 -> float32 (the C++ loader uses strtof; they can differ in the last bit
    only for decimals within a float64 ulp of a float32 rounding midpoint)."""
    _, nodes = _tetgen_rows(prefix + ".node", np.float64)
    start0 = nodes.shape[0] > 0 and int(nodes[0, 0]) == 0
    pos = nodes[:, 1:4].astype(np.float32)
    _, ele = _tetgen_rows(prefix + ".ele", np.int64)
    tets = ele[:, 1:5]
    if not start0:
        tets = tets - 1
    return pos, tets.astype(np.int32)