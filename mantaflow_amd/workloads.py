"""Synthetic inputs for the batched-solve and serving measurements, built
with numpy and the GPU module only.

batched_problems(): independent pressure problems of uneven size - 2D grids
with sides drawn from {64..512} and 3D grids with sides from {32..160},
walls on all sides plus seeded box/sphere obstacles covering 0-20 % of the
interior, and a random divergent velocity (components N(0,1) smoothed by
three binomial passes per axis) with setWallBcs applied, so every problem's
right-hand side is compatible.
"""
import numpy as np

from . import manta as m

FLUID, OBSTACLE = 1, 2


def obstacle_flags_np(shape, frac, rng):
    """Flag grid (int32, x fastest): obstacle boundary layer (initDomain with
    walls on every side), fluid interior (fillGrid), then boxes and spheres
    stamped into the interior until `frac` of it is obstacle."""
    nx, ny, nz = shape
    is3d = nz > 1
    g = np.full((nz, ny, nx), FLUID, np.int32)
    g[:, :, 0] = g[:, :, -1] = OBSTACLE
    g[:, 0, :] = g[:, -1, :] = OBSTACLE
    if is3d:
        g[0] = g[-1] = OBSTACLE
    inner = np.zeros(g.shape, bool)
    if is3d:
        inner[1:-1, 1:-1, 1:-1] = True
    else:
        inner[:, 1:-1, 1:-1] = True
    target = frac * inner.sum()
    rmax = max(2.0, min(a for a in shape if a > 1) / 8.0)
    count = 0
    while count < target:
        c = rng.uniform(1, np.array([nx - 1, ny - 1, max(nz - 1, 1.0001)]))
        box = rng.random() < 0.5
        h = rng.uniform(1.0, rmax, 3) if box else np.full(3, rng.uniform(1.5, rmax))
        # evaluate the shape on its bounding box only (cell centres i + 0.5)
        lo = [max(0, int(np.floor(c[a] - h[a] - 1))) for a in range(3)]
        hi = [min(n, int(np.ceil(c[a] + h[a] + 1))) for a, n in enumerate((nx, ny, nz))]
        if not is3d:
            lo[2], hi[2] = 0, 1
        z, y, x = np.meshgrid(np.arange(lo[2], hi[2]) + 0.5, np.arange(lo[1], hi[1]) + 0.5,
                              np.arange(lo[0], hi[0]) + 0.5, indexing="ij")
        dz = (z - c[2]) if is3d else 0.0
        if box:
            msk = (np.abs(x - c[0]) <= h[0]) & (np.abs(y - c[1]) <= h[1]) & (np.abs(dz) <= h[2])
        else:
            msk = (x - c[0]) ** 2 + (y - c[1]) ** 2 + dz ** 2 < h[0] * h[0]
        sub = (slice(lo[2], hi[2]), slice(lo[1], hi[1]), slice(lo[0], hi[0]))
        new = msk & inner[sub] & (g[sub] != OBSTACLE)
        g[sub][new] = OBSTACLE
        count += int(new.sum())
    return g.ravel()


def smooth_field(rng, shape, passes=3):
    """Face velocities N(0,1), `passes` binomial [1/4,1/2,1/4] passes per
    axis (AoS, float32; z component 0 in 2D)."""
    nx, ny, nz = shape
    v = rng.standard_normal((nz, ny, nx, 3))
    axes = (0, 1, 2) if nz > 1 else (1, 2)
    for _ in range(passes):
        for ax in axes:
            v = 0.25 * np.roll(v, 1, ax) + 0.5 * v + 0.25 * np.roll(v, -1, ax)
    if nz == 1:
        v[..., 2] = 0
    return np.ascontiguousarray(v.reshape(-1, 3).astype(np.float32))


class Problem:
    """One pressure problem on its own solver."""

    def __init__(self, shape, frac, seed, rel_tol=1e-4):
        rng = np.random.default_rng(seed)
        nx, ny, nz = shape
        self.shape = shape
        self.s = m.Solver(name="problem", gridSize=m.vec3(nx, ny, nz), dim=3 if nz > 1 else 2)
        self.s.timestep = 1.0
        self.s.mfxRelTol = rel_tol
        self.flags = self.s.create(m.FlagGrid)
        self.flags.fromNumpy(obstacle_flags_np(shape, frac, rng))
        self.vel0 = smooth_field(rng, shape)
        self.vel = self.s.create(m.MACGrid)
        self.pressure = self.s.create(m.RealGrid)
        self.reset()

    def reset(self):
        self.vel.fromNumpy(self.vel0)
        m.setWallBcs(flags=self.flags, vel=self.vel)

    def problem(self):
        return (self.vel, self.pressure, self.flags)

    @property
    def cells(self):
        return self.shape[0] * self.shape[1] * self.shape[2]


def batched_shapes(n2d=32, n3d=32, seed2d=5, seed3d=6, seedfrac=7):
    r2, r3, rf = np.random.default_rng(seed2d), np.random.default_rng(seed3d), np.random.default_rng(seedfrac)
    shapes = [(int(a), int(a), 1) for a in r2.integers(64, 513, n2d)]
    shapes += [(int(a), int(a), int(a)) for a in r3.integers(32, 161, n3d)]
    fracs = [float(x) for x in rf.uniform(0.0, 0.2, len(shapes))]
    return shapes, fracs


def batched_problems(n2d=32, n3d=32, rel_tol=1e-4):
    shapes, fracs = batched_shapes(n2d, n3d)
    return [Problem(sh, fr, 1000 + i, rel_tol) for i, (sh, fr) in enumerate(zip(shapes, fracs))]
