"""Synthetic test matrices and right-hand sides for the benchmark workloads.

All generators return scipy CSC matrices (original, lexicographic numbering,
x fastest) and draw random numbers from a portable splitmix64 stream so the
same data can be produced in any language: for stream `seed`, draw k
(k = 0, 1, ...) is z = mix(seed + (k + 1) * 0x9E3779B97F4A7C15) and
u = (z >> 11) * 2^-53 in [0, 1).
"""
import numpy as np
import scipy.sparse as sp

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, offset=0):
    """uint64 draws offset .. offset+count-1 of stream `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, count, a=0.0, b=1.0, offset=0):
    u = (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return a + (b - a) * u


def normal(seed, count):
    """N(0,1) by Box-Muller on consecutive pairs of uniforms."""
    m = (count + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = np.maximum(u[0::2], 2.0 ** -53)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    out = np.empty(2 * m)
    out[0::2] = z[:m]
    out[1::2] = z[m:]
    return out[:count]


def _stencil(shape, coeffs, diag):
    """Assemble a matrix from per-direction neighbour coefficient arrays.

    shape: grid sizes (x fastest); coeffs: list of (offset vector, array of
    coefficients per cell); neighbours outside the grid are dropped
    (Dirichlet)."""
    dims = len(shape)
    n = int(np.prod(shape))
    idx = np.arange(n).reshape(shape[::-1])  # [z][y][x] or [y][x]
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [diag.ravel()]
    for off, c in coeffs:
        sl_src, sl_dst = [], []
        for ax in range(dims):
            o = off[dims - 1 - ax]
            if o > 0:
                sl_src.append(slice(0, -o))
                sl_dst.append(slice(o, None))
            elif o < 0:
                sl_src.append(slice(-o, None))
                sl_dst.append(slice(0, o))
            else:
                sl_src.append(slice(None))
                sl_dst.append(slice(None))
        r = idx[tuple(sl_src)].ravel()
        cidx = idx[tuple(sl_dst)].ravel()
        cv = np.broadcast_to(c, idx.shape)[tuple(sl_src)].ravel()
        rows.append(r)
        cols.append(cidx)
        vals.append(cv)
    A = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return A


def laplace2d(nx, ny=None, shift=0.0):
    """5-point Laplacian on an nx x ny interior grid (diag 4 + shift, -1)."""
    ny = nx if ny is None else ny
    shape = (nx, ny)
    one = -np.ones((ny, nx))
    coeffs = [((1, 0), one), ((-1, 0), one), ((0, 1), one), ((0, -1), one)]
    return _stencil(shape, coeffs, np.full((ny, nx), 4.0 + shift))


def convdiff3d(nx, ny=None, nz=None, diffusivity=1.0, velocity=(1.0, 0.5, 0.25), h=1.0):
    """7-point first-order upwind convection-diffusion on an nx x ny x nz grid
    (Dirichlet). Row stencil for v >= 0: west -(D/h^2 + vx/h), east -D/h^2,
    likewise for y and z; diagonal 6 D/h^2 + (vx + vy + vz)/h."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    shape = (nx, ny, nz)
    g = (nz, ny, nx)
    D = diffusivity / (h * h)
    coeffs = []
    diag = np.full(g, 6.0 * D)
    for ax, v in enumerate(velocity):
        vv = v / h
        up = np.full(g, -(D + max(vv, 0.0)))
        dn = np.full(g, -(D + max(-vv, 0.0)))
        diag += abs(vv)
        o = [0, 0, 0]
        o[ax] = -1
        coeffs.append((tuple(o), up))
        o = [0, 0, 0]
        o[ax] = 1
        coeffs.append((tuple(o), dn))
    return _stencil(shape, coeffs, diag)


def convdiff3d_varcoef(n, kappa, velocity=(1.0, 0.0, 0.0)):
    """7-point upwind convection-diffusion on an n^3 grid with per-cell
    diffusivity kappa (array n^3, x fastest). Face diffusivity = harmonic mean
    of the two cells; a boundary face uses the cell's own value."""
    g = (n, n, n)
    k = np.asarray(kappa, dtype=np.float64).reshape(g)
    diag = np.zeros(g)
    coeffs = []
    for ax, v in enumerate(velocity):
        axis = 2 - ax
        for sgn in (-1, 1):
            kn = np.roll(k, -sgn, axis=axis)
            face = 2.0 * k * kn / (k + kn)
            # boundary faces: no neighbour, own diffusivity on the diagonal
            edge = [slice(None)] * 3
            edge[axis] = slice(0, 1) if sgn < 0 else slice(n - 1, n)
            face_b = face.copy()
            face_b[tuple(edge)] = k[tuple(edge)]
            diag += face_b
            conv = max(v, 0.0) if sgn < 0 else max(-v, 0.0)
            c = -(face + conv)
            o = [0, 0, 0]
            o[ax] = sgn
            coeffs.append((tuple(o), c))
        diag += abs(v)
    return _stencil((n, n, n), coeffs, diag)


def elasticity3d(ne, E=1.0, nu=0.3, clamp=True):
    """Trilinear hexahedral (Q1) linear elasticity on an ne^3 unit-cube mesh,
    2x2x2 Gauss, 3 dof per node (node-major, x fastest). Clamped z = 0 dofs
    are kept as identity rows/columns (decoupled) so n = 3 (ne+1)^3."""
    h = 1.0 / ne
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[np.arange(3), np.arange(3)] += 2 * mu
    C[np.arange(3, 6), np.arange(3, 6)] = mu
    # element stiffness (same for every element)
    corners = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0],
                        [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]], dtype=float)
    gp = np.array([-1, 1]) / np.sqrt(3.0)
    Ke = np.zeros((24, 24))
    for a in gp:
        for b in gp:
            for c in gp:
                xi = np.array([a, b, c])
                s = 2 * corners - 1
                dN = np.zeros((8, 3))
                for i in range(8):
                    f = (1 + s[i] * xi)
                    dN[i, 0] = s[i, 0] * f[1] * f[2] / 8
                    dN[i, 1] = s[i, 1] * f[0] * f[2] / 8
                    dN[i, 2] = s[i, 2] * f[0] * f[1] / 8
                dN *= 2.0 / h
                Bm = np.zeros((6, 24))
                for i in range(8):
                    Bm[0, 3 * i] = dN[i, 0]
                    Bm[1, 3 * i + 1] = dN[i, 1]
                    Bm[2, 3 * i + 2] = dN[i, 2]
                    Bm[3, 3 * i] = dN[i, 1]
                    Bm[3, 3 * i + 1] = dN[i, 0]
                    Bm[4, 3 * i + 1] = dN[i, 2]
                    Bm[4, 3 * i + 2] = dN[i, 1]
                    Bm[5, 3 * i] = dN[i, 2]
                    Bm[5, 3 * i + 2] = dN[i, 0]
                Ke += Bm.T @ C @ Bm * (h / 2) ** 3
    nn = ne + 1
    ex, ey, ez = np.meshgrid(np.arange(ne), np.arange(ne), np.arange(ne), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    nodes = np.stack([(ex + cx) + nn * (ey + cy) + nn * nn * (ez + cz)
                      for cx, cy, cz in corners.astype(int)], axis=1)  # (nel, 8)
    dofs = (3 * nodes[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24)
    rows = np.repeat(dofs, 24, axis=1).ravel()
    cols = np.tile(dofs, (1, 24)).ravel()
    vals = np.tile(Ke.ravel(), dofs.shape[0])
    n = 3 * nn ** 3
    A = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    if clamp:
        z0 = np.arange(nn * nn)  # nodes with z = 0
        fixed = np.zeros(n, dtype=bool)
        fixed[(3 * z0[:, None] + np.arange(3)).ravel()] = True
        A = A.tocoo()
        keep = ~(fixed[A.row] | fixed[A.col])
        r = np.concatenate([A.row[keep], np.nonzero(fixed)[0]])
        c = np.concatenate([A.col[keep], np.nonzero(fixed)[0]])
        v = np.concatenate([A.data[keep], np.ones(fixed.sum())])
        A = sp.csc_matrix((v, (r, c)), shape=(n, n))
    A.sort_indices()
    return A
