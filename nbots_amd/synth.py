"""Synthetic structured-grid problems for the FEM solve path.

The reference has no generator for these workloads (its mesher is a Delaunay
tessellator), and its random numbers are unsuitable: ``nb_statistics_runif``
leaves [min, max] for many seeds (``sources/nb/statistics_bot.c:47-52``).
This module defines the inputs that the tests, the bench and the oracle
share:

* grid: node id ``j*(nx+1)+i`` at ``(i*h, j*h)``; square ``(i, j)`` is cut
  along its ``(i,j)->(i+1,j+1)`` diagonal into ``(n00, n10, n11)`` and
  ``(n00, n11, n01)``, both counter-clockwise, element id
  ``2*(j*nx+i) + {0, 1}``;
* "model" vertices and segments, the way the reference addresses boundary
  conditions (``nb_mesh2D_get_invtx`` / ``nb_mesh2D_insgm_get_node``):
  vertices 0..3 are the corners (0,0), (nx,0), (nx,ny), (0,ny), vertex 4 is
  the right mid-edge node (nx, ny//2); segments 0..3 are bottom, right, top,
  left, each listed from its start corner to its end corner;
* random streams: counter-based splitmix64, one stream per quantity, so a
  value depends only on (seed, stream, index).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
STREAM_JITTER = 1
STREAM_MODULUS = 2
STREAM_DENSITY = 3
STREAM_PROBE = 4
STREAM_PROJECTION = 16      # streams 16.. : test projections w_k, one per k
DEFAULT_SEED = 20261202


def splitmix64(seed: int, stream: int, n: int) -> np.ndarray:
    """n outputs of splitmix64 started at state seed*2^20 + stream."""
    with np.errstate(over="ignore"):
        s0 = np.uint64((int(seed) * (1 << 20) + int(stream)) & ((1 << 64) - 1))
        z = s0 + (np.arange(1, n + 1, dtype=np.uint64) * GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """U[0,1) doubles with 53 random bits."""
    return (splitmix64(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """N(0,1) by Box-Muller, one value per pair of uniforms."""
    u = uniform(seed, stream, 2 * n)
    u1 = 1.0 - u[0::2]          # (0, 1]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


@dataclass
class Problem:
    """A 2-dof T3 problem in the reference's vocabulary (flat arrays)."""
    nx: int
    ny: int
    nod: np.ndarray                     # float64 [2*N_nod]
    conn: np.ndarray                    # uint32  [3*N_elem]
    vtx: np.ndarray                     # uint32  model vertex -> node
    sgm_ptr: np.ndarray                 # uint32  [N_sgm+1]
    sgm_nodes: np.ndarray               # uint32
    E: float = 1.0
    nu: float = 0.3
    thickness: float = 1.0
    elem_E: np.ndarray | None = None    # float64 [N_elem] or None
    rho: np.ndarray | None = None       # SIMP densities, if any
    # boundary-condition lists: (ids uint32, masks uint8 [2n], values f64 [2n])
    dv: tuple = field(default_factory=lambda: _empty_bc())
    nv: tuple = field(default_factory=lambda: _empty_bc())
    ds: tuple = field(default_factory=lambda: _empty_bc())
    ns: tuple = field(default_factory=lambda: _empty_bc())
    # function-valued Neumann segments, applied after ns: (segment id,
    # (mask_x, mask_y), f) with f(x, y) -> (value_x, value_y), a traction
    # (nb_bcond_push_function with NB_NEUMANN, NB_BC_ON_SEGMENT)
    ns_fn: tuple = ()
    name: str = ""

    @property
    def N_nod(self) -> int:
        return self.nod.size // 2

    @property
    def N_elem(self) -> int:
        return self.conn.size // 3

    @property
    def N(self) -> int:
        return 2 * self.N_nod


def _empty_bc():
    return (np.zeros(0, np.uint32), np.zeros(0, np.uint8), np.zeros(0, np.float64))


def _bc(ids, masks, vals):
    return (np.asarray(ids, np.uint32), np.asarray(masks, np.uint8).reshape(-1),
            np.asarray(vals, np.float64).reshape(-1))


def grid(nx: int, ny: int, h: float = 1.0, jitter: float = 0.0,
         seed: int = DEFAULT_SEED):
    """Nodes, triangles, model vertices and segments of an nx x ny grid."""
    i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
    x = (i * h).astype(np.float64).ravel()
    y = (j * h).astype(np.float64).ravel()
    if jitter > 0.0:
        n = x.size
        u = uniform(seed, STREAM_JITTER, 2 * n).reshape(n, 2)
        inner = ((i > 0) & (i < nx) & (j > 0) & (j < ny)).ravel()
        x = x + inner * (2.0 * u[:, 0] - 1.0) * jitter * h
        y = y + inner * (2.0 * u[:, 1] - 1.0) * jitter * h
    nod = np.empty(2 * x.size, np.float64)
    nod[0::2] = x
    nod[1::2] = y
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny))
    n00 = (jj * (nx + 1) + ii).ravel()
    n10 = n00 + 1
    n01 = n00 + nx + 1
    n11 = n01 + 1
    conn = np.empty((nx * ny, 2, 3), np.uint32)
    conn[:, 0] = np.stack([n00, n10, n11], 1)
    conn[:, 1] = np.stack([n00, n11, n01], 1)
    conn = conn.reshape(-1)
    W = nx + 1
    vtx = np.array([0, nx, ny * W + nx, ny * W, (ny // 2) * W + nx], np.uint32)
    bottom = np.arange(0, nx + 1)
    right = np.arange(0, ny + 1) * W + nx
    top = ny * W + np.arange(nx, -1, -1)
    left = np.arange(ny, -1, -1) * W
    segs = [bottom, right, top, left]
    sgm_ptr = np.zeros(5, np.uint32)
    sgm_ptr[1:] = np.cumsum([s.size for s in segs])
    sgm_nodes = np.concatenate(segs).astype(np.uint32)
    return nod, conn, vtx, sgm_ptr, sgm_nodes


def cantilever_point(nx: int = 64, ny: int = 32, jitter: float = 0.1,
                     seed: int = DEFAULT_SEED) -> Problem:
    """Jittered cantilever, left edge clamped, unit downward load at the
    right mid-edge node, E=1, nu=0.3, thickness 1 (64x32 by default)."""
    nod, conn, vtx, sp, sn = grid(nx, ny, jitter=jitter, seed=seed)
    return Problem(nx, ny, nod, conn, vtx, sp, sn,
                   ds=_bc([3], [1, 1], [0.0, 0.0]),
                   nv=_bc([4], [0, 1], [0.0, -1.0]),
                   name=f"cantilever_point_{nx}x{ny}")


def cantilever_traction(nx: int, ny: int, seed: int = DEFAULT_SEED,
                        sigma: float = 1.0, traction: float = 1.0) -> Problem:
    """Cantilever with lognormal(0, sigma) per-element modulus, left edge
    clamped, uniform shear traction on the right edge (direction -y).  The
    traction is a constant-valued Neumann segment whose value is the total
    force traction*length, as the reference splits it by sub-segment length
    (set_bconditions.c:133-154)."""
    nod, conn, vtx, sp, sn = grid(nx, ny)
    ne = conn.size // 3
    elem_E = np.exp(sigma * normal(seed, STREAM_MODULUS, ne))
    return Problem(nx, ny, nod, conn, vtx, sp, sn, elem_E=elem_E,
                   ds=_bc([3], [1, 1], [0.0, 0.0]),
                   ns=_bc([1], [0, 1], [0.0, -traction * ny]),
                   name=f"cantilever_traction_{nx}x{ny}")


def damage_beam(nx: int = 24, ny: int = 12, jitter: float = 0.1, deflection: float = -0.6,
                seed: int = DEFAULT_SEED) -> Problem:
    """Jittered beam for the quasi-static damage solver: left edge clamped,
    right edge given the vertical displacement `deflection` (a Dirichlet
    segment the solver scales by the load factor of each step), E=1,
    nu=0.3, thickness 1."""
    nod, conn, vtx, sp, sn = grid(nx, ny, jitter=jitter, seed=seed)
    return Problem(nx, ny, nod, conn, vtx, sp, sn,
                   ds=_bc([3, 1], [1, 1, 0, 1], [0.0, 0.0, 0.0, deflection]),
                   name=f"damage_{nx}x{ny}")


def simp_modulus(rho: np.ndarray, penal: float = 3.0, E0: float = 1.0,
                 Emin: float = 1e-9) -> np.ndarray:
    return Emin + rho ** penal * (E0 - Emin)


def mbb(nx: int, ny: int, seed: int = DEFAULT_SEED, rho=None,
        penal: float = 3.0, E0: float = 1.0, Emin: float = 1e-9) -> Problem:
    """MBB half-beam: u_x = 0 on the left edge (symmetry), u_y = 0 at the
    bottom-right corner, unit downward load at the top-left corner.  rho:
    None -> U(0.3, 1.0) from the seed; a float -> uniform density."""
    nod, conn, vtx, sp, sn = grid(nx, ny)
    ne = conn.size // 3
    if rho is None:
        rho = 0.3 + 0.7 * uniform(seed, STREAM_DENSITY, ne)
    elif np.isscalar(rho):
        rho = np.full(ne, float(rho))
    rho = np.asarray(rho, np.float64)
    return Problem(nx, ny, nod, conn, vtx, sp, sn, E=E0,
                   elem_E=simp_modulus(rho, penal, E0, Emin), rho=rho,
                   ds=_bc([3], [1, 0], [0.0, 0.0]),
                   dv=_bc([1], [0, 1], [0.0, 0.0]),
                   nv=_bc([3], [0, 1], [0.0, -1.0]),
                   name=f"mbb_{nx}x{ny}")


def projections(n: int, k: int = 8, seed: int = DEFAULT_SEED) -> np.ndarray:
    """k vectors of length n with entries U[-1, 1), row i from stream
    STREAM_PROJECTION + i.  With a solution's norm, F.u and its samples, the
    values w_i.u pin the whole vector: a difference e changes each w_i.u by
    a random amount of standard deviation ||e|| / sqrt(3), the same scale as
    w_i.u itself has for u, so |w_i.(u - v)| / ||u|| measures ||u - v|| / ||u||
    along random directions."""
    return np.stack([2.0 * uniform(seed, STREAM_PROJECTION + i, n) - 1.0 for i in range(k)])


def elem_centroids(p: Problem):
    """(x0 + x1 + x2) / 3 per element, summed left to right."""
    x = p.nod[0::2][p.conn.reshape(-1, 3)]
    y = p.nod[1::2][p.conn.reshape(-1, 3)]
    return (x[:, 0] + x[:, 1] + x[:, 2]) / 3.0, (y[:, 0] + y[:, 1] + y[:, 2]) / 3.0


def elem_areas(p: Problem) -> np.ndarray:
    c = p.conn.reshape(-1, 3)
    x = p.nod[0::2][c]
    y = p.nod[1::2][c]
    return 0.5 * ((x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) -
                  (x[:, 2] - x[:, 0]) * (y[:, 1] - y[:, 0]))
