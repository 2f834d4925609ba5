"""Synthetic AMG Galerkin problems (R, A, P) on sorted CSR.

Every generator is deterministic.  Random draws use numpy's PCG64 bit
generator (``numpy.random.Generator(numpy.random.PCG64(seed))``) in the order
documented on each function, so the CPU oracle and the GPU path see identical
inputs.  Grid points are numbered lexicographically, ``i = x + nx*(y + ny*z)``,
and aggregates of a structured aggregation are numbered the same way on the
coarse grid, so a Galerkin product ``R*A*P`` is again a stencil operator on the
coarse grid (which is what makes the hierarchy of ``hierarchy_3d`` possible).

Smoothed aggregation follows the problem statements: ``P = (I - w D^-1 A) P0``
with ``w = 2/3`` and ``R = P^T``.  Row ``i`` of ``S = I - w D^-1 A`` is
``S_ik = delta_ik - (w / d_i) * a_ik``; ``P_ic`` is the sum of ``S_ik`` over the
stored entries ``k`` of row ``i`` with ``agg(k) = c``, added in increasing ``k``.
Entries are kept even when they sum to zero (the reference keeps every
touched index, code/src/include/sparseMatrix.h:85-104).

Names of the workloads:
  poisson2d_sa(64)            2D 5-point Poisson, 64x64 grid, 2x2 aggregates
  poisson3d_sa(128)           3D 7-point Poisson, 128^3 grid, 2x2x2 aggregates
  stencil27_sa(100, seed)     3D 27-point random SPD-like stencil, 3x3x3 aggregates
  irregular_sa(n, seed)       lognormal row lengths, random aggregation
  hierarchy_3d(96, levels)    5-level 2x2x2 SA hierarchy from the 96^3 7-point operator
  poisson2d_batch(256, seed)  256 independent 2D Poisson problems of uneven size
"""
from __future__ import annotations

import numpy as np

from .csr import CSR, from_coo

OMEGA = 2.0 / 3.0


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


# --------------------------------------------------------------------------
# structured stencils
# --------------------------------------------------------------------------
def _stencil_offsets(dims, full: bool):
    """Sorted (dz, dy, dx) offsets; full=True gives the 3^d box, else the star."""
    nd = len(dims)
    offs = []
    rng = (-1, 0, 1)
    if nd == 2:
        for dy in rng:
            for dx in rng:
                if full or abs(dx) + abs(dy) <= 1:
                    offs.append((0, dy, dx))
    else:
        for dz in rng:
            for dy in rng:
                for dx in rng:
                    if full or abs(dx) + abs(dy) + abs(dz) <= 1:
                        offs.append((dz, dy, dx))
    return offs


def _grid_coords(dims):
    nx = dims[0]
    ny = dims[1]
    nz = dims[2] if len(dims) == 3 else 1
    i = np.arange(nx * ny * nz, dtype=np.int64)
    x = i % nx
    y = (i // nx) % ny
    z = i // (nx * ny)
    return i, x, y, z, nx, ny, nz


def _stencil_layout(dims, full: bool):
    """Slot layout of a stencil: (cols[M,S], valid[M,S], slot offsets)."""
    i, x, y, z, nx, ny, nz = _grid_coords(dims)
    offs = _stencil_offsets(dims, full)
    S = len(offs)
    M = i.size
    cols = np.empty((M, S), dtype=np.int64)
    valid = np.empty((M, S), dtype=bool)
    for s, (dz, dy, dx) in enumerate(offs):
        cols[:, s] = i + dx + nx * (dy + ny * dz)
        valid[:, s] = ((x + dx >= 0) & (x + dx < nx) & (y + dy >= 0) & (y + dy < ny)
                       & (z + dz >= 0) & (z + dz < nz))
    return cols, valid, offs


def _csr_from_slots(M, cols, valid, vals) -> CSR:
    irp = np.zeros(M + 1, dtype=np.int64)
    np.cumsum(valid.sum(axis=1), out=irp[1:])
    return CSR(M, M, irp, cols[valid], vals[valid])


def poisson_operator(dims) -> CSR:
    """Star-stencil Poisson operator with Dirichlet boundary (rows drop missing
    neighbours): 2D 5-point (diagonal 4) or 3D 7-point (diagonal 6), neighbours -1."""
    cols, valid, offs = _stencil_layout(dims, full=False)
    M = cols.shape[0]
    diag = 2.0 * len(dims)
    vals = np.full(cols.shape, -1.0)
    center = offs.index((0, 0, 0))
    vals[:, center] = diag
    return _csr_from_slots(M, cols, valid, vals)


def stencil27_random(n: int, seed: int) -> CSR:
    """3D 27-point operator on an n^3 grid.

    Draw order: one Uniform[0.5, 1.5) value per (row, slot) of the full n^3 x 27
    slot array, row-major, from PCG64(seed).  Off-diagonal a_ij = -u_ij, then
    symmetrised a_ij <- (a_ij + a_ji) / 2; the diagonal is 1.001 times the
    absolute off-diagonal row sum.
    """
    dims = (n, n, n)
    cols, valid, offs = _stencil_layout(dims, full=True)
    M, S = cols.shape
    u = _rng(seed).uniform(0.5, 1.5, size=(M, S))
    a = -u
    # slot s pairs with slot S-1-s (offsets are sorted and symmetric)
    center = offs.index((0, 0, 0))
    sym = np.zeros_like(a)
    for s in range(S):
        if s == center:
            continue
        rows = np.flatnonzero(valid[:, s])
        j = cols[rows, s]
        sym[rows, s] = 0.5 * (a[rows, s] + a[j, S - 1 - s])
    offabs = np.where(valid, np.abs(sym), 0.0)
    offabs[:, center] = 0.0
    sym[:, center] = 1.001 * offabs.sum(axis=1)
    return _csr_from_slots(M, cols, valid, sym)


def structured_aggregation(dims, block):
    """Aggregate id of every grid point for block aggregation (partial blocks at
    the far faces when a side is not a multiple of the block).  Returns
    (agg[M], coarse_dims)."""
    i, x, y, z, nx, ny, nz = _grid_coords(dims)
    cdims = tuple(-(-d // b) for d, b in zip(dims, block))
    cx = x // block[0]
    cy = y // block[1]
    if len(dims) == 3:
        cz = z // block[2]
        agg = cx + cdims[0] * (cy + cdims[1] * cz)
    else:
        agg = cx + cdims[0] * cy
    return agg.astype(np.int64), cdims


def random_aggregation(n: int, seed: int, lo: int = 4, hi: int = 12):
    """Random aggregation: a PCG64(seed) permutation of the n rows is cut into
    consecutive aggregates whose sizes are drawn Uniform{lo..hi} from the same
    generator (the last aggregate takes what is left).  Returns (agg[n], n_agg)."""
    g = _rng(seed)
    perm = g.permutation(n)
    sizes = g.integers(lo, hi + 1, size=n // lo + 2)
    ends = np.cumsum(sizes)
    k = int(np.searchsorted(ends, n))
    sizes = sizes[: k + 1].copy()
    sizes[k] -= ends[k] - n
    agg = np.empty(n, dtype=np.int64)
    agg[perm] = np.repeat(np.arange(k + 1, dtype=np.int64), sizes)
    return agg, k + 1


def smoothed_prolongator(A: CSR, agg: np.ndarray, nagg: int, omega: float = OMEGA,
                         diag_mode: str = "diag", absolute: bool = False) -> CSR:
    """P = (I - omega D^-1 A) P0 with P0 the 0/1 aggregation matrix.

    diag_mode "diag": d_i = a_ii;  "absplus1": d_i = |a_ii| + 1 (a missing
    diagonal counts as 0).  absolute=True sums |terms| instead (the magnitude
    scale of every entry, for rounding-error bounds)."""
    rows = A.row_ids()
    isdiag = A.ja == rows
    d = np.zeros(A.M)
    d[rows[isdiag]] = A.vals[isdiag]
    if diag_mode == "absplus1":
        d = np.abs(d) + 1.0
    elif diag_mode != "diag":
        raise ValueError(diag_mode)
    if np.any(d == 0):
        raise ValueError("zero diagonal scaling")
    w = omega / d
    s = -(w[rows] * A.vals)
    s[isdiag] += 1.0
    if absolute:
        s = np.abs(s)
    r = rows
    c = agg[A.ja]
    v = s
    has_diag = np.zeros(A.M, dtype=bool)
    has_diag[rows[isdiag]] = True
    if not has_diag.all():
        miss = np.flatnonzero(~has_diag)
        r = np.concatenate([r, miss])
        c = np.concatenate([c, agg[miss]])
        v = np.concatenate([v, np.ones(miss.size)])
    # sum duplicates per (row, aggregate); entries of a row are in increasing k order
    return from_coo(A.M, nagg, r, c, v, sum_duplicates=True)


def galerkin_inputs(A: CSR, agg, nagg, diag_mode="diag"):
    P = smoothed_prolongator(A, agg, nagg, diag_mode=diag_mode)
    R = P.transpose()
    return R, A, P


# --------------------------------------------------------------------------
# named workloads
# --------------------------------------------------------------------------
def poisson2d_sa(n: int = 64):
    """2D 5-point Poisson on n x n with 2x2 smoothed aggregation."""
    A = poisson_operator((n, n))
    agg, cdims = structured_aggregation((n, n), (2, 2))
    R, A, P = galerkin_inputs(A, agg, int(np.prod(cdims)))
    return dict(R=R, A=A, P=P, name=f"poisson2d_5pt_{n}x{n}_sa2x2", coarse_dims=cdims)


def poisson3d_sa(n: int = 128):
    """3D 7-point Poisson on n^3 with 2x2x2 smoothed aggregation."""
    dims = (n, n, n)
    A = poisson_operator(dims)
    agg, cdims = structured_aggregation(dims, (2, 2, 2))
    R, A, P = galerkin_inputs(A, agg, int(np.prod(cdims)))
    return dict(R=R, A=A, P=P, name=f"poisson3d_7pt_{n}^3_sa2x2x2", coarse_dims=cdims)


def stencil27_sa(n: int = 100, seed: int = 3):
    """3D 27-point random stencil on n^3 with 3x3x3 smoothed aggregation."""
    dims = (n, n, n)
    A = stencil27_random(n, seed)
    agg, cdims = structured_aggregation(dims, (3, 3, 3))
    R, A, P = galerkin_inputs(A, agg, int(np.prod(cdims)))
    return dict(R=R, A=A, P=P, name=f"stencil27_{n}^3_sa3x3x3_seed{seed}", coarse_dims=cdims)


def lognormal_matrix(n: int, seed: int, mu: float = 2.5, sigma: float = 1.0,
                     clip=(1, 4096)) -> CSR:
    """n x n matrix with per-row off-diagonal counts round(lognormal(mu, sigma))
    clipped to `clip`, columns uniform without replacement (excluding the
    diagonal), plus the diagonal; values Uniform[-1, 1).

    Draw order from PCG64(seed): the n counts; then, for all rows at once, the
    candidate columns (with replacement, row-major); rows that lost entries to
    duplicates or to the diagonal are refilled in increasing row order by
    further draws; finally one value per stored entry in CSR order."""
    g = _rng(seed)
    cnt = np.rint(g.lognormal(mu, sigma, size=n)).astype(np.int64)
    cnt = np.clip(cnt, clip[0], min(clip[1], n - 1))
    rows = np.repeat(np.arange(n, dtype=np.int64), cnt)
    cols = g.integers(0, n, size=rows.size, dtype=np.int64)
    key = np.unique(rows * n + cols)
    r = key // n
    c = key - r * n
    keep = c != r
    r, c = r[keep], c[keep]
    have = np.bincount(r, minlength=n)
    short = np.flatnonzero(have < cnt)
    if short.size:
        extra_r, extra_c = [], []
        present = {}
        for row in short:
            lo, hi = np.searchsorted(r, [row, row + 1])
            s = set(c[lo:hi].tolist())
            need = int(cnt[row] - (hi - lo))
            while need:
                cand = int(g.integers(0, n))
                if cand != row and cand not in s:
                    s.add(cand)
                    extra_r.append(row)
                    extra_c.append(cand)
                    need -= 1
        r = np.concatenate([r, np.array(extra_r, np.int64)])
        c = np.concatenate([c, np.array(extra_c, np.int64)])
    r = np.concatenate([r, np.arange(n, dtype=np.int64)])
    c = np.concatenate([c, np.arange(n, dtype=np.int64)])
    M = from_coo(n, n, r, c, np.zeros(r.size))
    M.vals = g.uniform(-1.0, 1.0, size=M.nnz)
    return M


def irregular_sa(n: int = 1_000_000, seed: int = 4):
    """Irregular rows: lognormal row lengths, random aggregation of sizes
    Uniform{4..12} (PCG64(seed+1)), P = (I - 2/3 D^-1 A) P0 with D = |diag| + 1."""
    A = lognormal_matrix(n, seed)
    agg, nagg = random_aggregation(n, seed + 1)
    R, A, P = galerkin_inputs(A, agg, nagg, diag_mode="absplus1")
    return dict(R=R, A=A, P=P, name=f"irregular_lognormal_{n}_seed{seed}", coarse_dims=(nagg,))


def hierarchy_3d(n: int = 96, levels: int = 5, rap=None):
    """Smoothed-aggregation hierarchy from the 3D 7-point Poisson operator on n^3.

    Level l+1's operator is R_l A_l P_l computed by `rap(R, A, P) -> CSR`; the
    aggregation of every level is 2x2x2 on that level's grid.  Returns the list
    of `levels` (R, A, P) triples (the last coarse operator is not included)."""
    if rap is None:
        raise ValueError("hierarchy_3d needs a rap(R, A, P) callable")
    dims = (n, n, n)
    A = poisson_operator(dims)
    out = []
    for _ in range(levels):
        agg, cdims = structured_aggregation(dims, (2, 2, 2))
        R, A, P = galerkin_inputs(A, agg, int(np.prod(cdims)))
        out.append(dict(R=R, A=A, P=P, dims=dims))
        A = rap(R, A, P)
        dims = cdims
    return out


def poisson2d_batch(count: int = 256, seed: int = 5, lo: int = 16, hi: int = 256):
    """`count` independent 2D 5-point Poisson problems with side n ~ Uniform{lo..hi}
    (PCG64(seed)), each with 2x2 smoothed aggregation."""
    sides = _rng(seed).integers(lo, hi + 1, size=count)
    probs = []
    for n in sides:
        p = poisson2d_sa(int(n))
        probs.append(dict(R=p["R"], A=p["A"], P=p["P"], side=int(n)))
    return probs


def mac_counts(R: CSR, A: CSR, P: CSR, RA_irp=None, RA_ja=None):
    """Multiply-add counts of the pair form: MAC1 = sum over R entries of nnz(A row),
    MAC2 = sum over RA entries of nnz(P row) (needs the RA pattern)."""
    alen = np.diff(A.irp)
    mac1 = int(alen[R.ja].sum())
    mac2 = None
    if RA_ja is not None:
        mac2 = int(np.diff(P.irp)[RA_ja].sum())
    return mac1, mac2
