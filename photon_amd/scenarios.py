"""Seeded synthetic scenarios for the pressure projection path.

Every array the tests and the benchmark feed to the GPU library and to the CPU
oracle comes from here, so both see identical bits.  Random numbers come from
SplitMix64 (a counter-based generator: output i is mix(seed + (i+1)*gamma), so
a whole field is generated vectorised) and Box-Muller for normal deviates.

Layouts follow the reference Solver (src/sim/solver.h:216-220): cell
c = i + nx*(j + ny*k); x faces (nx+1)*ny*nz, y faces nx*(ny+1)*nz, z faces
nx*ny*(nz+1).  Face open fractions follow Solver::updateFace
(src/sim/solver.cpp:66-101): 0 next to a solid cell and at the ground, 1
otherwise; with no cracks the matrix coefficient equals the open fraction.

Workloads (sizes in cells at the stated cell size):
  room      20x15x12 at 0.2 m: solid walls, one 0.8x2.0 m door in the x=0
            wall (cells j=6..9, k=0..9) open to ambient, burner cells at the
            floor centre (i=9..10, j=6..7).
  house     16x10x8 m at 0.2 m (80x50x40) or 0.1 m (160x100x80): closed domain,
            40 seeded axis-aligned wall boxes 0.2 m thick and full height, 10
            openings 0.8-1.6 m wide and 2.0 m high cut into them.
  batch     grids uniform between 20x15x12 and 80x50x40 cells, walls scaled to
            size, for the batched solves.
  cracked   the room's 20x15x12 box split at x = 2.0 m by a thin partition:
            every wall, the ceiling and the partition are intact barriers
            with crack leakage, the partition's door is closed (undercut
            leakage), the right-hand room has an open window and a
            free-standing leaky panel; the left room is a closed zone whose
            only exits are cracks.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n outputs of SplitMix64 started at `seed`, skipping `offset` outputs."""
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + idx * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class Rng:
    """Scalar draws from one SplitMix64 stream (geometry decisions)."""

    def __init__(self, seed: int):
        self.seed = seed & 0xFFFFFFFFFFFFFFFF
        self.count = 0

    def u64(self) -> int:
        v = int(splitmix64(self.seed, 1, self.count)[0])
        self.count += 1
        return v

    def uniform(self) -> float:
        return (self.u64() >> 11) * (1.0 / 9007199254740992.0)

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + min(hi - lo, int(self.uniform() * (hi - lo + 1)))


def normal_field(seed: int, n: int) -> np.ndarray:
    """n N(0,1) deviates (float64) by Box-Muller on SplitMix64 pairs."""
    m = (n + 1) // 2
    raw = splitmix64(seed, 2 * m)
    u1 = ((raw[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u2 = (raw[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(2.0 * math.pi * u2)
    out[1::2] = r * np.sin(2.0 * math.pi * u2)
    return out[:n]


@dataclasses.dataclass
class Grid:
    """Geometry of one scenario: the inputs of the reference's matrix."""

    nx: int
    ny: int
    nz: int
    dx: float
    blocked: np.ndarray          # uint8 [N]
    fx: np.ndarray               # float32 face open fractions
    fy: np.ndarray
    fz: np.ndarray
    ax: np.ndarray               # float32 matrix coefficients (ax_, ay_, az_)
    ay: np.ndarray
    az: np.ndarray
    burner_cells: np.ndarray | None = None
    # crack (leak) faces (uint8 per face: lx_, ly_, lz_), or None for none
    lx: np.ndarray | None = None
    ly: np.ndarray | None = None
    lz: np.ndarray | None = None
    # (zone int32 [N], zone_open uint8 [nzones]), formed on first use
    _zones: tuple | None = dataclasses.field(default=None, init=False, repr=False, compare=False)

    @property
    def zone(self) -> np.ndarray:
        if self._zones is None:
            compute_zones(self)
        return self._zones[0]

    @property
    def zone_open(self) -> np.ndarray:
        if self._zones is None:
            compute_zones(self)
        return self._zones[1]

    @property
    def n(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def leaks(self):
        """(lx, ly, lz) crack flags, or None when the grid has no cracks."""
        if self.lx is None:
            return None
        return self.lx, self.ly, self.lz

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)


def faces_from_mask(blocked3: np.ndarray, open_sides=(), door=None):
    """Face open fractions from a solid mask [nz, ny, nx] (Solver::updateFace).

    Domain boundary faces are closed except the sides listed in `open_sides`
    ("x0", "x1", "y0", "y1", "top") and the x=0 door cells given as a boolean
    [nz, ny] mask in `door`.  The ground (k = 0 z faces) is always closed.
    """
    nz, ny, nx = blocked3.shape
    g = ~blocked3.astype(bool)
    fx = np.zeros((nz, ny, nx + 1), np.float32)
    fx[:, :, 1:nx] = g[:, :, 1:] & g[:, :, :-1]
    if "x0" in open_sides:
        fx[:, :, 0] = g[:, :, 0]
    if "x1" in open_sides:
        fx[:, :, nx] = g[:, :, -1]
    if door is not None:
        fx[:, :, 0] = np.where(door & g[:, :, 0], 1.0, fx[:, :, 0])
    fy = np.zeros((nz, ny + 1, nx), np.float32)
    fy[:, 1:ny, :] = g[:, 1:, :] & g[:, :-1, :]
    if "y0" in open_sides:
        fy[:, 0, :] = g[:, 0, :]
    if "y1" in open_sides:
        fy[:, ny, :] = g[:, -1, :]
    fz = np.zeros((nz + 1, ny, nx), np.float32)
    fz[1:nz] = g[1:] & g[:-1]
    if "top" in open_sides:
        fz[nz] = g[-1]
    return fx.ravel(), fy.ravel(), fz.ravel()


def compute_zones(grid: Grid):
    """Pressure zones: gas cells joined through faces with a > 0; a zone is
    open if it reaches a side or top boundary face with a > 0
    (src/sim/gasphase.cpp:111-158).  Labels are numbered by lowest cell."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    nx, ny, nz = grid.shape
    n = grid.n
    ax = grid.ax.reshape(nz, ny, nx + 1)
    ay = grid.ay.reshape(nz, ny + 1, nx)
    az = grid.az.reshape(nz + 1, ny, nx)
    gas = grid.blocked.reshape(nz, ny, nx) == 0
    idx = np.arange(n).reshape(nz, ny, nx)
    rows, cols = [], []
    ex = (ax[:, :, 1:nx] > 0) & gas[:, :, 1:] & gas[:, :, :-1]
    rows.append(idx[:, :, :-1][ex]); cols.append(idx[:, :, 1:][ex])
    ey = (ay[:, 1:ny, :] > 0) & gas[:, 1:, :] & gas[:, :-1, :]
    rows.append(idx[:, :-1, :][ey]); cols.append(idx[:, 1:, :][ey])
    ez = (az[1:nz] > 0) & gas[1:] & gas[:-1]
    rows.append(idx[:-1][ez]); cols.append(idx[1:][ez])
    r = np.concatenate(rows); c = np.concatenate(cols)
    m = coo_matrix((np.ones(len(r)), (r, c)), shape=(n, n))
    _, lab = connected_components(m, directed=False)
    lab = lab.reshape(nz, ny, nx)
    lab = np.where(gas, lab, -1)
    # renumber by first (lowest) cell of each zone, as the reference's seed order
    flat = lab.ravel()
    first = {}
    order = []
    uniq, first_idx = np.unique(flat[flat >= 0], return_index=True)
    gas_idx = np.nonzero(flat >= 0)[0]
    first_cell = gas_idx[first_idx]
    order = np.argsort(first_cell)
    remap = np.full(uniq.max() + 1 if len(uniq) else 1, -1, np.int64)
    remap[uniq[order]] = np.arange(len(uniq))
    zone = np.where(flat >= 0, remap[np.maximum(flat, 0)], -1).astype(np.int32)
    nzones = len(uniq)
    zopen = np.zeros(max(nzones, 1), np.uint8)
    z3 = zone.reshape(nz, ny, nx)
    for sel, a in ((z3[:, :, 0], ax[:, :, 0]), (z3[:, :, -1], ax[:, :, nx]),
                   (z3[:, 0, :], ay[:, 0, :]), (z3[:, -1, :], ay[:, ny, :]),
                   (z3[-1], az[nz])):
        hit = (sel >= 0) & (a > 0)
        zopen[np.unique(sel[hit])] = 1
    grid._zones = (zone, zopen[:nzones] if nzones else zopen[:0])
    return grid


def _finish(blocked3, dx, open_sides=(), door=None) -> Grid:
    nz, ny, nx = blocked3.shape
    fx, fy, fz = faces_from_mask(blocked3, open_sides, door)
    g = Grid(nx, ny, nz, float(dx), blocked3.astype(np.uint8).ravel(), fx, fy, fz,
             fx.copy(), fy.copy(), fz.copy())
    return g


def room(dx: float = 0.2, window: bool = False, door: bool = True) -> Grid:
    """Single 4.0 x 3.0 x 2.4 m room: walls closed, a 0.8 x 2.0 m door in the
    x = 0 wall (j = 6..9 and k = 0..9 at 0.2 m) open to ambient, burner cells
    (0.4 x 0.4 m) at the floor centre; `door` = False seals it.  With
    `window`, also a 1.2 x 1.0 m
    window open to ambient in the y = 3.0 m wall (x 1.4..2.6 m, z 1.0..2.0 m;
    i = 7..12, k = 5..9 at 0.2 m)."""
    nx, ny, nz = round(4.0 / dx), round(3.0 / dx), round(2.4 / dx)
    blocked = np.zeros((nz, ny, nx), np.uint8)
    yc = (np.arange(ny) + 0.5) * dx
    zc = (np.arange(nz) + 0.5) * dx
    opening = (np.abs(zc[:, None] - 1.0) < 1.0) & (yc[None, :] > 1.2) & (yc[None, :] < 2.0)
    g = _finish(blocked, dx, door=opening if door else None)
    xc = (np.arange(nx) + 0.5) * dx
    if window:
        win = (np.abs(zc[:, None] - 1.5) < 0.5) & (np.abs(xc[None, :] - 2.0) < 0.6)
        fy = g.fy.reshape(nz, ny + 1, nx)
        fy[:, ny, :] = np.where(win, 1.0, fy[:, ny, :])
        g.ay = g.fy.copy()
    bi = np.nonzero(np.abs(xc - 2.0) < 0.2)[0]
    bj = np.nonzero(np.abs(yc - 1.4) < 0.2)[0]
    g.burner_cells = np.array([i + nx * j for j in bj for i in bi], np.int32)
    return g


def house(seed: int, nx: int = 80, ny: int = 50, nz: int = 40, dx: float = 0.2,
          n_walls: int = 40, n_open: int = 10) -> Grid:
    """Closed box with seeded interior walls (0.2 m thick, full height) and
    door-height openings 0.8-1.6 m wide.  Geometry is drawn in metres on a
    0.2 m lattice, then rasterised at dx, so the 0.1 m grid refines the 0.2 m
    one."""
    rng = Rng(seed)
    Lx, Ly, Lz = nx * dx, ny * dx, nz * dx
    gx, gy = int(round(Lx / 0.2)), int(round(Ly / 0.2))
    walls = []
    for _ in range(n_walls):
        if rng.randint(0, 1) == 0:              # x-normal wall: plane at x, spans y
            p = rng.randint(1, gx - 2)
            length = rng.randint(min(10, gy - 2), max(10, int(0.6 * gy)))
            a = rng.randint(0, max(0, gy - length))
            walls.append((0, p, a, min(gy, a + length)))
        else:                                    # y-normal wall
            p = rng.randint(1, gy - 2)
            length = rng.randint(min(10, gx - 2), max(10, int(0.6 * gx)))
            a = rng.randint(0, max(0, gx - length))
            walls.append((1, p, a, min(gx, a + length)))
    openings = []
    for _ in range(n_open):
        wi = rng.randint(0, n_walls - 1)
        o, p, a, b = walls[wi]
        width = rng.randint(4, 8)               # 0.8 .. 1.6 m in 0.2 m units
        s = rng.randint(a, max(a, b - width))
        openings.append((o, p, s, min(b, s + width)))
    xc = (np.arange(nx) + 0.5) * dx
    yc = (np.arange(ny) + 0.5) * dx
    zc = (np.arange(nz) + 0.5) * dx
    blocked = np.zeros((nz, ny, nx), bool)
    for o, p, a, b in walls:
        if o == 0:
            sx = (xc > p * 0.2) & (xc < p * 0.2 + 0.2)
            sy = (yc > a * 0.2) & (yc < b * 0.2)
        else:
            sy = (yc > p * 0.2) & (yc < p * 0.2 + 0.2)
            sx = (xc > a * 0.2) & (xc < b * 0.2)
        blocked |= (sy[:, None] & sx[None, :])[None, :, :]
    for o, p, a, b in openings:
        if o == 0:
            sx = (xc > p * 0.2) & (xc < p * 0.2 + 0.2)
            sy = (yc > a * 0.2) & (yc < b * 0.2)
        else:
            sy = (yc > p * 0.2) & (yc < p * 0.2 + 0.2)
            sx = (xc > a * 0.2) & (xc < b * 0.2)
        sz = zc < 2.0
        blocked &= ~(sz[:, None, None] & (sy[:, None] & sx[None, :])[None, :, :])
    return _finish(blocked.astype(np.uint8), dx)


LEAK_OPEN = 0.05   # open fraction at or above which an intact barrier is an opening (solver.cpp:20)


def barrier_face(leak: float, intact: bool = True):
    """(open fraction, matrix coefficient, crack flag) of a gas-gas face that
    carries a barrier whose open fraction while intact is `leak`
    (Solver::updateFace, src/sim/solver.cpp:82-100): a failed or opened
    barrier is fully open; an intact one is open by `leak`, and below
    LEAK_OPEN that opening is a crack whose flow is the orifice equation's,
    so it leaves the pressure matrix (coefficient 0)."""
    f = 1.0 if not intact else float(np.float32(leak))
    crack = intact and 0.0 < f < LEAK_OPEN
    a = f if f > 0 and not crack else 0.0
    return f, a, crack


def cracked(dx: float = 0.2, wall_leak: float = 0.0003, ceiling_leak: float = 0.0002, door_leak: float = 0.01,
            panel_leak: float = 0.0005) -> Grid:
    """Two rooms joined only by cracks, one of them open to ambient.

    Box 4.0 x 3.0 x 2.4 m; the ground is closed.  Walls (x = 0, x = 4, y = 0,
    y = 3) and ceiling are intact barriers of leakage `wall_leak` and
    `ceiling_leak`; a partition at x = 2.0 m (leakage `wall_leak`) has a
    closed 0.8 x 2.0 m door (y 1.2..2.0 m, z 0..2.0 m, leakage `door_leak`).
    The right room has a 1.0 x 1.0 m window open to ambient in the y = 3 wall
    (x 2.6..3.6 m, z 1.0..2.0 m) and a free-standing 0.8 x 1.0 m panel on the
    y = 0.8 m plane (x 2.6..3.4 m, leakage `panel_leak`) whose two sides are
    in the same zone.  Burner cells: 0.4 x 0.4 m at the left room's floor
    centre."""
    nx, ny, nz = round(4.0 / dx), round(3.0 / dx), round(2.4 / dx)
    n = nx * ny * nz
    fx = np.ones((nz, ny, nx + 1), np.float32)
    fy = np.ones((nz, ny + 1, nx), np.float32)
    fz = np.ones((nz + 1, ny, nx), np.float32)
    ax, ay, az = fx.copy(), fy.copy(), fz.copy()
    lx = np.zeros(fx.shape, np.uint8)
    ly = np.zeros(fy.shape, np.uint8)
    lz = np.zeros(fz.shape, np.uint8)

    def put(f, a, l, sel, leak):
        fv, av, cv = barrier_face(leak)
        f[sel], a[sel], l[sel] = fv, av, cv

    xc = (np.arange(nx) + 0.5) * dx
    yc = (np.arange(ny) + 0.5) * dx
    zc = (np.arange(nz) + 0.5) * dx
    put(fx, ax, lx, (slice(None), slice(None), 0), wall_leak)
    put(fx, ax, lx, (slice(None), slice(None), nx), wall_leak)
    put(fy, ay, ly, (slice(None), 0, slice(None)), wall_leak)
    put(fy, ay, ly, (slice(None), ny, slice(None)), wall_leak)
    put(fz, az, lz, (nz,), ceiling_leak)
    fz[0], az[0], lz[0] = 0.0, 0.0, 0                       # ground
    ip = round(2.0 / dx)
    put(fx, ax, lx, (slice(None), slice(None), ip), wall_leak)
    door = (zc[:, None] < 2.0) & (yc[None, :] > 1.2) & (yc[None, :] < 2.0)
    fv, av, cv = barrier_face(door_leak)
    fx[:, :, ip] = np.where(door, fv, fx[:, :, ip])
    ax[:, :, ip] = np.where(door, av, ax[:, :, ip])
    lx[:, :, ip] = np.where(door, np.uint8(cv), lx[:, :, ip])
    win = (np.abs(zc[:, None] - 1.5) < 0.5) & (np.abs(xc[None, :] - 3.1) < 0.5)
    fy[:, ny, :] = np.where(win, 1.0, fy[:, ny, :])
    ay[:, ny, :] = np.where(win, 1.0, ay[:, ny, :])
    ly[:, ny, :] = np.where(win, np.uint8(0), ly[:, ny, :])
    jp = round(0.8 / dx)
    panel = (zc[:, None] < 1.0) & (np.abs(xc[None, :] - 3.0) < 0.4)
    fv, av, cv = barrier_face(panel_leak)
    fy[:, jp, :] = np.where(panel, fv, fy[:, jp, :])
    ay[:, jp, :] = np.where(panel, av, ay[:, jp, :])
    ly[:, jp, :] = np.where(panel, np.uint8(cv), ly[:, jp, :])
    g = Grid(nx, ny, nz, float(dx), np.zeros(n, np.uint8), fx.ravel(), fy.ravel(), fz.ravel(), ax.ravel(),
             ay.ravel(), az.ravel(), lx=lx.ravel(), ly=ly.ravel(), lz=lz.ravel())
    bi = np.nonzero(np.abs(xc - 1.0) < 0.2)[0]
    bj = np.nonzero(np.abs(yc - 1.4) < 0.2)[0]
    g.burner_cells = np.array([i + nx * j for j in bj for i in bi], np.int32)
    return g


def face_updates(grid: Grid, axis: int, mask, leak: float, intact: bool = True):
    """Face changes (axis, face, open fraction, coefficient, crack flag) for
    every face of an axis selected by a boolean mask of that axis's face
    array shape ([nz, ny, nx+1], [nz, ny+1, nx] or [nz+1, ny, nx]), each now
    carrying a barrier of open fraction `leak` (barrier_face); leak = 0 with
    intact closes the faces."""
    face = np.flatnonzero(np.asarray(mask, bool).ravel()).astype(np.int32)
    f, a, c = barrier_face(leak, intact)
    n = len(face)
    return (np.full(n, axis, np.int32), face, np.full(n, f, np.float32), np.full(n, a, np.float32),
            np.full(n, 1 if c else 0, np.uint8))


def cracked_events(grid: Grid):
    """The face changes of the cracked two-room grid's fire: (1) the
    partition door fails (fully open), joining the closed left room to the
    open right one; (2) the window shuts to a crack (leakage 0.002) and the
    panel becomes solid (open fraction 0), which leaves one closed zone."""
    nx, ny, nz = grid.shape
    dx = grid.dx
    yc = (np.arange(ny) + 0.5) * dx
    zc = (np.arange(nz) + 0.5) * dx
    xc = (np.arange(nx) + 0.5) * dx
    ip = round(2.0 / dx)
    door = np.zeros((nz, ny, nx + 1), bool)
    door[:, :, ip] = (zc[:, None] < 2.0) & (yc[None, :] > 1.2) & (yc[None, :] < 2.0)
    first = face_updates(grid, 0, door, 1.0, intact=False)
    win = np.zeros((nz, ny + 1, nx), bool)
    win[:, ny, :] = (np.abs(zc[:, None] - 1.5) < 0.5) & (np.abs(xc[None, :] - 3.1) < 0.5)
    panel = np.zeros((nz, ny + 1, nx), bool)
    panel[:, round(0.8 / dx), :] = (zc[:, None] < 1.0) & (np.abs(xc[None, :] - 3.0) < 0.4)
    w_ = face_updates(grid, 1, win, 0.002)
    p_ = face_updates(grid, 1, panel, 0.0)
    second = tuple(np.concatenate([x, y]) for x, y in zip(w_, p_))
    return [first, second]


def apply_face_updates(grid: Grid, upd) -> Grid:
    """A copy of grid with face changes applied (as Solver::updateFace)."""
    g = dataclasses.replace(grid, fx=grid.fx.copy(), fy=grid.fy.copy(), fz=grid.fz.copy(), ax=grid.ax.copy(),
                            ay=grid.ay.copy(), az=grid.az.copy(),
                            lx=None if grid.lx is None else grid.lx.copy(),
                            ly=None if grid.ly is None else grid.ly.copy(),
                            lz=None if grid.lz is None else grid.lz.copy())
    axis, face, f, a, c = upd
    nf = (len(g.fx), len(g.fy), len(g.fz))
    if g.lx is None and np.any(c):
        g.lx, g.ly, g.lz = (np.zeros(k, np.uint8) for k in nf)
    for d, F, A, L in ((0, g.fx, g.ax, g.lx), (1, g.fy, g.ay, g.ly), (2, g.fz, g.az, g.lz)):
        sel = axis == d
        F[face[sel]] = f[sel]
        A[face[sel]] = a[sel]
        if L is not None:
            L[face[sel]] = c[sel]
    g._zones = None
    return g


def batch_grid(seed: int) -> Grid:
    """One scenario of the batched workload: grid uniform between 20x15x12
    and 80x50x40 cells, walls and openings scaled to its floor area."""
    rng = Rng(seed ^ 0x5DEECE66D)
    nx, ny, nz = rng.randint(20, 80), rng.randint(15, 50), rng.randint(12, 40)
    scale = (nx * ny) / (80 * 50)
    nw = max(2, int(round(40 * scale)))
    no = max(1, int(round(10 * scale)))
    return house(seed, nx, ny, nz, 0.2, nw, no)


def burner_field(grid: Grid, seed: int, rho: float = 1.2, cp: float = 1005.0, T: float = 293.15) -> np.ndarray:
    """Divergence source of a square gas burner: a 2x2-cell burner on a
    seeded gas floor cell, heat release uniform in [10, 500] kW, expanding the
    gas of its cells at S = HRR / (cells rho cp T V) (the heat term of
    energyAndExpansion, src/sim/gasphase.cpp:251-271, for one step), then zero
    mean per closed zone (a closed room's pressure rises uniformly), FP32."""
    rng = Rng(seed ^ 0xB0B0)
    nx, ny = grid.nx, grid.ny
    blk = grid.blocked.reshape(grid.nz, ny, nx)
    floor = [(i, j) for j in range(ny - 1) for i in range(nx - 1)
             if not blk[0, j:j + 2, i:i + 2].any()]
    i0, j0 = floor[rng.randint(0, len(floor) - 1)]
    hrr = 10e3 + (500e3 - 10e3) * rng.uniform()
    V = grid.dx ** 3
    s = np.zeros(grid.n, np.float64)
    for j in (j0, j0 + 1):
        for i in (i0, i0 + 1):
            s[i + nx * j] = hrr / (4 * rho * cp * T * V)
    if grid.zone is not None and len(grid.zone_open):
        z = grid.zone
        sel = (grid.blocked == 0) & (z >= 0)
        nzn = len(grid.zone_open)
        sums = np.bincount(z[sel], weights=s[sel], minlength=nzn)
        cnts = np.bincount(z[sel], minlength=nzn)
        mean = np.where(cnts > 0, sums / np.maximum(cnts, 1), 0.0)
        s[sel] -= np.where(grid.zone_open == 0, mean, 0.0)[z[sel]]
    return s.astype(np.float32)


def divergence_field(grid: Grid, seed: int) -> np.ndarray:
    """S ~ N(0,1) on gas cells, zero mean per closed zone (FP64), 0 in solids;
    returned as float32.  This is the right-hand side of the solves up to the
    factor dx^2/dt applied by the RHS kernel."""
    s = normal_field(seed, grid.n)
    gas = grid.blocked == 0
    s[~gas] = 0.0
    if grid.zone is not None:
        z = grid.zone
        nzn = len(grid.zone_open)
        if nzn:
            sel = gas & (z >= 0)
            sums = np.bincount(z[sel], weights=s[sel], minlength=nzn)
            cnts = np.bincount(z[sel], minlength=nzn)
            mean = np.where(cnts > 0, sums / np.maximum(cnts, 1), 0.0)
            closed = grid.zone_open == 0
            adj = np.where(closed, mean, 0.0)
            s[sel] -= adj[z[sel]]
    return s.astype(np.float32)


def rhs_from_divergence(grid: Grid, S: np.ndarray, dt: float) -> np.ndarray:
    """b = S * dx^2/dt in float32, the RHS the projection builds from S when
    the predicted velocity is zero (src/sim/gasphase.cpp:538)."""
    inv = np.float32(np.float32(grid.dx) * np.float32(grid.dx) / np.float32(dt))
    b = (S.astype(np.float32) * inv).astype(np.float32)
    b[grid.blocked != 0] = 0
    return b


def vented(grid: Grid, seed: int, frac: float = 0.15) -> Grid:
    """The same geometry with a seeded share of its open interior faces made
    partial openings (open fraction U[0.05, 0.95], as vents and partly open
    doors are, src/sim/solver.cpp:82-100): a matrix with fractional
    coefficients."""
    def perturb(f, salt):
        u = (splitmix64(seed ^ salt, len(f)) >> np.uint64(11)).astype(np.float64) / 9007199254740992.0
        v = (splitmix64(seed ^ (salt * 7 + 1), len(f)) >> np.uint64(11)).astype(np.float64) / 9007199254740992.0
        out = f.copy()
        sel = (f > 0) & (u < frac)
        out[sel] = (0.05 + 0.9 * v[sel]).astype(np.float32)
        return out
    fx, fy, fz = perturb(grid.fx, 11), perturb(grid.fy, 13), perturb(grid.fz, 17)
    return Grid(grid.nx, grid.ny, grid.nz, grid.dx, grid.blocked.copy(), fx, fy, fz, fx.copy(), fy.copy(), fz.copy())
