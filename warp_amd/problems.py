"""Seeded synthetic workloads that feed the sparse path.

These stand in for the ``warp.fem`` callers (warp/_src/fem/integrate.py:1640-1870
writes COO triplets, then calls ``bsr_set_from_triplets``):

* :func:`poisson2d_q1` -- -Lap u = f, bilinear Q1 on an n x n grid of the unit
  square, zero Dirichlet boundary, scalar (1x1) triplets.
* :func:`elasticity3d_q1` -- linear elasticity, trilinear hex elements on an
  n^3 grid, one clamped face, 3x3 triplets from the element loop.
* :func:`elasticity_triplets_shuffled` / :func:`synthetic_triplets` -- the
  assembly workloads (element-loop pattern with random values, shuffled; and
  random keys with a chosen mean multiplicity).
* :func:`gaussian_random_field`, :func:`fv7_poisson` -- -div(k grad u) = f on
  an n^3 cell grid, 7-point finite volumes, harmonic face coefficients,
  k = exp(g) with g a Gaussian random field.

Large generators run on the given torch device with ``torch.Generator``
streams; small ones use ``numpy.random.default_rng``.
"""

from __future__ import annotations

import math

import numpy as np
import torch

# ---------------------------------------------------------------------------
# 2D Poisson, bilinear Q1
# ---------------------------------------------------------------------------

_Q1_STIFF = np.array(
    [[4.0, -1.0, -2.0, -1.0], [-1.0, 4.0, -1.0, -2.0], [-2.0, -1.0, 4.0, -1.0], [-1.0, -2.0, -1.0, 4.0]]
) / 6.0
_Q1_MASS = np.array([[4.0, 2.0, 1.0, 2.0], [2.0, 4.0, 2.0, 1.0], [1.0, 2.0, 4.0, 2.0], [2.0, 1.0, 2.0, 4.0]]) / 36.0


def poisson2d_q1(n: int = 128, seed: int = 0):
    """Triplets and right-hand side of -Lap u = f on an n x n Q1 grid.

    Interior nodes (n-1)^2 are the dofs (node (i, j) -> (j-1)(n-1) + i-1);
    boundary nodes map to index -1 so assembly drops their triplets (zero
    Dirichlet).  f ~ U[-1, 1] per grid node (default_rng(seed)) and
    b = M f with the consistent Q1 mass matrix, restricted to the dofs.

    Returns numpy (rows, cols, values[float32], b[float32], ndof).
    """
    h = 1.0 / n
    m = n - 1
    ix, iy = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    ex, ey = ix.ravel(), iy.ravel()
    # local nodes counter-clockwise: (0,0) (1,0) (1,1) (0,1)
    lx = np.array([0, 1, 1, 0])
    ly = np.array([0, 0, 1, 1])
    nx = ex[:, None] + lx[None, :]
    ny = ey[:, None] + ly[None, :]
    interior = (nx > 0) & (nx < n) & (ny > 0) & (ny < n)
    dof = np.where(interior, (ny - 1) * m + (nx - 1), -1).astype(np.int32)
    rows = np.repeat(dof, 4, axis=1).ravel()
    cols = np.tile(dof, (1, 4)).ravel()
    vals = np.broadcast_to(_Q1_STIFF.ravel(), (ex.size, 16)).ravel().astype(np.float32)

    rng = np.random.default_rng(seed)
    f = rng.uniform(-1.0, 1.0, size=(n + 1) * (n + 1))
    node = ny * (n + 1) + nx
    fe = f[node]  # (ne, 4)
    be = (fe @ _Q1_MASS.T) * (h * h)
    b = np.zeros(m * m)
    mask = dof >= 0
    np.add.at(b, dof[mask], be[mask])
    return rows, cols, vals, b.astype(np.float32), m * m


# ---------------------------------------------------------------------------
# 3D linear elasticity, trilinear hexahedra
# ---------------------------------------------------------------------------


def hex8_elasticity_stiffness(h: float, E: float = 1.0, nu: float = 0.3) -> np.ndarray:
    """24 x 24 stiffness of a cube of side h (2x2x2 Gauss), dof order (node, xyz).

    Local node l = a + 2 b + 4 c at corner (a, b, c) in {0, 1}^3.
    """
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[3:, 3:] = np.eye(3) * mu
    corners = np.array([[a, b, c] for c in (0, 1) for b in (0, 1) for a in (0, 1)], dtype=float)
    g = 0.5 / math.sqrt(3.0)
    pts = [0.5 - g, 0.5 + g]
    K = np.zeros((24, 24))
    for qx in pts:
        for qy in pts:
            for qz in pts:
                q = np.array([qx, qy, qz])
                # gradients of trilinear shape functions on [0,1]^3, scaled to physical size
                grad = np.zeros((8, 3))
                for l, (a, b, c) in enumerate(corners):
                    fx = q[0] if a else 1 - q[0]
                    fy = q[1] if b else 1 - q[1]
                    fz = q[2] if c else 1 - q[2]
                    sx = 1.0 if a else -1.0
                    sy = 1.0 if b else -1.0
                    sz = 1.0 if c else -1.0
                    grad[l] = np.array([sx * fy * fz, fx * sy * fz, fx * fy * sz]) / h
                B = np.zeros((6, 24))
                for l in range(8):
                    gx, gy, gz = grad[l]
                    B[0, 3 * l] = gx
                    B[1, 3 * l + 1] = gy
                    B[2, 3 * l + 2] = gz
                    B[3, 3 * l], B[3, 3 * l + 1] = gy, gx
                    B[4, 3 * l + 1], B[4, 3 * l + 2] = gz, gy
                    B[5, 3 * l], B[5, 3 * l + 2] = gz, gx
                K += B.T @ D @ B * (h**3 / 8.0)
    return K


def _hex_element_nodes(n: int, device) -> torch.Tensor:
    """(n^3, 8) global node ids of every element, node = i + (n+1) j + (n+1)^2 k."""
    p = n + 1
    e = torch.arange(n**3, device=device, dtype=torch.int64)
    ex, ey, ez = e % n, (e // n) % n, e // (n * n)
    loc = torch.tensor([[a, b, c] for c in (0, 1) for b in (0, 1) for a in (0, 1)], device=device)
    nodes = (ex[:, None] + loc[None, :, 0]) + p * (ey[:, None] + loc[None, :, 1]) + p * p * (ez[:, None] + loc[None, :, 2])
    return nodes


def elasticity3d_q1(n: int = 64, seed_modulus: int = 1, seed_force: int = 2, nu: float = 0.3, device="cpu"):
    """Element-loop triplets of 3D Q1 elasticity with the x = 0 face clamped.

    E_e = exp(N(0,1)) per element (generator seed ``seed_modulus``), body
    force N(0,1) per node and component (``seed_force``), b = h^3 f.  Blocks
    touching a clamped node are zeroed (kept as explicit zero blocks, so the
    pattern is the full 27-neighbour one) and one identity block per clamped
    node is appended; b is zero on clamped dofs.  Assemble with
    ``prune_numerical_zeros=False`` to get (3n+1)^3 blocks.

    Returns (rows[int32], cols[int32], values[T,3,3] float32, b[(n+1)^3, 3]
    float32, nnodes) as torch tensors on ``device``.
    """
    device = torch.device(device)
    h = 1.0 / n
    p = n + 1
    nnodes = p**3
    K0 = torch.tensor(hex8_elasticity_stiffness(h, 1.0, nu), dtype=torch.float32, device=device)
    K0 = K0.view(8, 3, 8, 3).permute(0, 2, 1, 3).contiguous()  # (A, B, 3, 3)
    gen = torch.Generator(device=device).manual_seed(seed_modulus)
    E = torch.exp(torch.randn(n**3, generator=gen, device=device, dtype=torch.float32))
    nodes = _hex_element_nodes(n, device)  # (ne, 8)
    clamped = (nodes % p) == 0  # x = 0 face
    keep = (~clamped).to(torch.float32)
    vals = E[:, None, None, None, None] * K0[None]  # (ne, 8, 8, 3, 3)
    vals = vals * (keep[:, :, None, None, None] * keep[:, None, :, None, None])
    rows = nodes[:, :, None].expand(-1, 8, 8).reshape(-1)
    cols = nodes[:, None, :].expand(-1, 8, 8).reshape(-1)
    vals = vals.reshape(-1, 3, 3)
    cn = torch.arange(0, nnodes, p, device=device, dtype=torch.int64)  # clamped nodes
    eye = torch.eye(3, dtype=torch.float32, device=device).expand(cn.numel(), 3, 3)
    rows = torch.cat([rows, cn]).to(torch.int32)
    cols = torch.cat([cols, cn]).to(torch.int32)
    vals = torch.cat([vals, eye]).contiguous()
    genf = torch.Generator(device=device).manual_seed(seed_force)
    f = torch.randn((nnodes, 3), generator=genf, device=device, dtype=torch.float32)
    b = f * (h**3)
    b[cn] = 0.0
    return rows, cols, vals, b.contiguous(), nnodes


def elasticity_triplets_shuffled(n: int = 64, seed_values: int = 3, seed_perm: int = 4, device="cpu"):
    """Assembly workload: the (row, col) stream of the n^3 hex element loop
    (64 blocks per element), values N(0,1) 3x3, randomly permuted."""
    device = torch.device(device)
    nodes = _hex_element_nodes(n, device)
    rows = nodes[:, :, None].expand(-1, 8, 8).reshape(-1)
    cols = nodes[:, None, :].expand(-1, 8, 8).reshape(-1)
    T = rows.numel()
    gv = torch.Generator(device=device).manual_seed(seed_values)
    vals = torch.randn((T, 3, 3), generator=gv, device=device, dtype=torch.float32)
    gp = torch.Generator(device=device).manual_seed(seed_perm)
    perm = torch.randperm(T, generator=gp, device=device)
    return rows[perm].to(torch.int32), cols[perm].to(torch.int32), vals[perm].contiguous(), (n + 1) ** 3


def synthetic_triplets(n: int = 64, T: int = 20_000_000, U: int = 2_500_000, seed: int = 7, device="cpu"):
    """Assembly workload: U distinct (row, col) keys drawn from the 27-point
    pattern of an (n+1)^3 node grid, then T triplets drawn uniformly among
    them (mean multiplicity T/U), values N(0,1) 3x3."""
    device = torch.device(device)
    p = n + 1
    g = torch.Generator(device=device).manual_seed(seed)
    # all 27-neighbour keys, sampled without replacement via a random permutation prefix
    node = torch.arange(p**3, device=device, dtype=torch.int64)
    i, j, k = node % p, (node // p) % p, node // (p * p)
    d = torch.tensor([[a, b, c] for c in (-1, 0, 1) for b in (-1, 0, 1) for a in (-1, 0, 1)], device=device)
    ni, nj, nk = i[:, None] + d[None, :, 0], j[:, None] + d[None, :, 1], k[:, None] + d[None, :, 2]
    ok = (ni >= 0) & (ni < p) & (nj >= 0) & (nj < p) & (nk >= 0) & (nk < p)
    rows_all = node[:, None].expand(-1, 27)[ok]
    cols_all = (ni + p * nj + p * p * nk)[ok]
    pick = torch.randperm(rows_all.numel(), generator=g, device=device)[:U]
    key_r, key_c = rows_all[pick], cols_all[pick]
    which = torch.randint(0, U, (T,), generator=g, device=device)
    vals = torch.randn((T, 3, 3), generator=g, device=device, dtype=torch.float32)
    return key_r[which].to(torch.int32), key_c[which].to(torch.int32), vals, p**3


# ---------------------------------------------------------------------------
# 7-point finite volumes with a log-normal coefficient
# ---------------------------------------------------------------------------


def gaussian_random_field(n: int, sigma: float, corr_len: float, seed: int, device="cpu") -> torch.Tensor:
    """Periodic Gaussian random field on n^3 cells by FFT spectral synthesis.

    Covariance exp(-r^2 / (2 l^2)) with l = ``corr_len`` cells: white noise
    (N(0,1), torch generator ``seed``) is filtered by sqrt of the spectral
    density exp(-l^2 |kappa|^2 / 2), then shifted / scaled to mean 0 and
    standard deviation ``sigma``.  Returns float32 (n, n, n).
    """
    device = torch.device(device)
    g = torch.Generator(device=device).manual_seed(seed)
    w = torch.randn((n, n, n), generator=g, device=device, dtype=torch.float32)
    kappa = 2.0 * math.pi * torch.fft.fftfreq(n, device=device)
    k2 = kappa[:, None, None] ** 2 + kappa[None, :, None] ** 2 + kappa[None, None, :] ** 2
    filt = torch.exp(-0.25 * (corr_len**2) * k2)  # sqrt of the spectral density
    field = torch.fft.ifftn(torch.fft.fftn(w) * filt).real
    field = field - field.mean()
    field = field * (sigma / field.std().clamp_min(1e-30))
    return field.to(torch.float32)


def fv7_poisson(k: torch.Tensor):
    """CSR of -div(k grad u) on the unit cube, cell-centred, zero Dirichlet.

    Face transmissibility between cells i, j: 2 k_i k_j / (k_i + k_j); a
    boundary face contributes 2 k_i to the diagonal.  The operator is scaled
    by h^2 (rows sum the face terms directly), so the matching right-hand
    side is h^2 f.  Cell (x, y, z) -> x + n y + n^2 z; columns sorted.

    Returns (offsets int32[N+1], columns int32[nnz], values float32[nnz]).
    """
    n = k.shape[0]
    N = n**3
    dev = k.device
    kk = k.to(torch.float64).reshape(-1)
    idx = torch.arange(N, device=dev, dtype=torch.int64)
    x, y, z = idx % n, (idx // n) % n, idx // (n * n)
    # neighbour order by column: -z, -y, -x, self, +x, +y, +z
    steps = [(-n * n, z > 0), (-n, y > 0), (-1, x > 0), (0, None), (1, x < n - 1), (n, y < n - 1), (n * n, z < n - 1)]
    cols = []
    vals = []
    valid = []
    diag = torch.zeros(N, dtype=torch.float64, device=dev)
    for step, ok in steps:
        if step == 0:
            cols.append(idx)
            vals.append(None)
            valid.append(torch.ones(N, dtype=torch.bool, device=dev))
            continue
        j = torch.where(ok, idx + step, idx)
        t = 2.0 * kk * kk[j] / (kk + kk[j])
        t_face = torch.where(ok, t, 2.0 * kk)  # boundary face: half-cell distance
        diag += t_face
        cols.append(j)
        vals.append(-t)
        valid.append(ok)
    vals[3] = diag
    C = torch.stack(cols, 1)
    V = torch.stack(vals, 1)
    M = torch.stack(valid, 1)
    counts = M.sum(1).to(torch.int32)
    offsets = torch.zeros(N + 1, dtype=torch.int32, device=dev)
    offsets[1:] = torch.cumsum(counts, 0).to(torch.int32)
    return offsets, C[M].to(torch.int32).contiguous(), V[M].to(torch.float32).contiguous()


def fv7_problem(n: int, sigma: float, corr_len: float, seed_field: int, seed_rhs: int, device="cpu"):
    """GRF coefficient, 7-point operator and b = h^2 f with f ~ N(0,1)."""
    g = gaussian_random_field(n, sigma, corr_len, seed_field, device)
    k = torch.exp(g)
    offsets, columns, values = fv7_poisson(k)
    gen = torch.Generator(device=torch.device(device)).manual_seed(seed_rhs)
    f = torch.randn(n**3, generator=gen, device=torch.device(device), dtype=torch.float32)
    b = f / (n * n)
    return offsets, columns, values, b, k
