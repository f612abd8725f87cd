"""Shared helpers for the parity tests: seeded inputs in the reference's host
layout and GPU <-> oracle round trips."""
import numpy as np

from oracle import EMPTY, FLUID, OBSTACLE, OUTFLOW


def make_flags(o, shape, bw=0, open_bound="", obstacles=0, rng=None, stick=False):
    """initDomain + fillGrid (+ setOpenBound) on the oracle, plus optional
    obstacle boxes stamped into the interior (reference flag semantics)."""
    f = o.init_domain(shape, boundaryWidth=bw)
    o.fill_grid(shape, f)
    if open_bound:
        o.set_open_bound(shape, f, bw, open_bound)
    nx, ny, nz = shape
    g = f.reshape(nz, ny, nx)
    for _ in range(obstacles):
        x0, y0 = rng.integers(2, max(3, nx - 4)), rng.integers(2, max(3, ny - 4))
        sx, sy = rng.integers(1, 4), rng.integers(1, 4)
        if nz > 1:
            z0, sz = rng.integers(2, max(3, nz - 4)), rng.integers(1, 4)
        else:
            z0, sz = 0, 1
        g[z0:z0 + sz, y0:y0 + sy, x0:x0 + sx] = OBSTACLE | (64 if stick else 0)
    return f


def rand_vel(rng, n, scale=1.0):
    return (rng.uniform(-1, 1, (n, 3)) * scale).astype(np.float32)


def smooth_vel(rng, shape, passes=2, scale=1.0):
    nx, ny, nz = shape
    v = rng.standard_normal((nz, ny, nx, 3)).astype(np.float64)
    for _ in range(passes):
        for ax in ((0, 1, 2) if nz > 1 else (1, 2)):
            v = 0.25 * np.roll(v, 1, ax) + 0.5 * v + 0.25 * np.roll(v, -1, ax)
    v *= scale / max(np.abs(v).max(), 1e-12)
    if nz == 1:
        v[..., 2] = 0
    return np.ascontiguousarray(v.reshape(-1, 3).astype(np.float32))


class GpuScene:
    """A manta Solver with helpers to move numpy arrays in and out."""

    def __init__(self, manta, shape, dt=1.0):
        self.m = manta
        self.shape = shape
        nx, ny, nz = shape
        self.s = manta.Solver(name="t", gridSize=manta.vec3(nx, ny, nz), dim=3 if nz > 1 else 2)
        self.s.timestep = dt

    def flags(self, arr):
        g = self.s.create(self.m.FlagGrid)
        g.fromNumpy(arr)
        return g

    def real(self, arr=None):
        g = self.s.create(self.m.RealGrid)
        if arr is not None:
            g.fromNumpy(arr)
        return g

    def mac(self, arr=None):
        g = self.s.create(self.m.MACGrid)
        if arr is not None:
            g.fromNumpy(arr)
        return g

    def vec(self, arr=None):
        g = self.s.create(self.m.VecGrid)
        if arr is not None:
            g.fromNumpy(arr)
        return g


def assert_bitexact(a, b, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    same = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else (a == b)
    if not np.all(same):
        idx = np.argwhere(~same)[:5]
        raise AssertionError("%s: %d of %d differ, first %s: gpu=%s oracle=%s" % (
            what, int((~same).sum()), same.size, idx.tolist(),
            [a[tuple(i)] for i in idx], [b[tuple(i)] for i in idx]))


def obstacle_flags(o, shape, frac, rng):
    """initDomain() walls plus seeded boxes and spheres stamped into the
    interior until `frac` of the interior cells are obstacles, then fillGrid
    (the 256^3 mixed-precision projection workload, at any size)."""
    f = o.init_domain(shape)
    nx, ny, nz = shape
    g = f.reshape(nz, ny, nx)
    zz, yy, xx = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    is3d = nz > 1
    inner = np.zeros(g.shape, bool)
    if is3d:
        inner[1:-1, 1:-1, 1:-1] = True
    else:
        inner[:, 1:-1, 1:-1] = True
    target = frac * inner.sum()
    rmax = max(2.0, min(s for s in shape if s > 1) / 8.0)
    while ((g & OBSTACLE) != 0)[inner].sum() < target:
        c = rng.uniform(1, np.array([nx - 1, ny - 1, max(nz - 1, 1.0001)]))
        dz = (zz - c[2]) if is3d else 0.0
        if rng.random() < 0.5:
            h = rng.uniform(1.0, rmax, 3)
            m = (np.abs(xx - c[0]) <= h[0]) & (np.abs(yy - c[1]) <= h[1]) & (np.abs(dz) <= h[2])
        else:
            r = rng.uniform(1.5, rmax)
            m = (xx - c[0]) ** 2 + (yy - c[1]) ** 2 + dz ** 2 < r * r
        g[m & inner] = OBSTACLE
    o.fill_grid(shape, f)
    return f


def component_demean(p, flags, shape):
    """Pressure modulo one constant per connected fluid region (6-connected)."""
    from scipy import ndimage
    nx, ny, nz = shape
    fl = (flags.reshape(nz, ny, nx) & FLUID) != 0
    lab, n = ndimage.label(fl)
    q = p.reshape(nz, ny, nx).astype(np.float64).copy()
    for c in range(1, n + 1):
        m = lab == c
        q[m] -= q[m].mean()
    q[~fl] = 0.0
    return q
