"""Grid::save / Grid::load of device grids in the reference's .uni and .raw
formats: every grid type round-trips exactly, the .uni payload is the
grid's host-layout data, and a file of another type or size is refused
with the reference's error text."""
import numpy as np
import pytest

from mantaflow_amd import gridio, manta

pytestmark = pytest.mark.gpu

SHAPE = (12, 10, 7)


def _grids(s, rng):
    n = int(np.prod(SHAPE))
    out = []
    for cls, comp, dt in ((manta.RealGrid, 1, np.float32), (manta.IntGrid, 1, np.int32),
                          (manta.VecGrid, 3, np.float32), (manta.MACGrid, 3, np.float32),
                          (manta.LevelsetGrid, 1, np.float32)):
        g = s.create(cls)
        a = (rng.standard_normal((n, comp)) * 50).astype(dt)
        g.fromNumpy(a if comp == 3 else a[:, 0])
        out.append((cls, g))
    f = s.create(manta.FlagGrid)
    f.initDomain(boundaryWidth=1)
    out.append((manta.FlagGrid, f))
    return out


@pytest.mark.parametrize("ext", [".uni", ".raw"])
def test_device_grid_round_trip(tmp_path, ext):
    s = manta.Solver(name="io", gridSize=manta.vec3(*SHAPE), dim=3)
    for cls, g in _grids(s, np.random.default_rng(5)):
        p = str(tmp_path / ("%s%s" % (cls.__name__, ext)))
        g.save(p)
        h = s.create(cls)
        h.load(p)
        np.testing.assert_array_equal(h.toNumpy(), g.toNumpy(), err_msg=cls.__name__)
        if ext == ".uni":
            head, a = gridio.read_uni(p)
            assert head["dims"] == SHAPE and head["gridType"] == g.getGridType()
            np.testing.assert_array_equal(a, g.toNumpy())


def test_uni_load_refuses_other_type_or_size(tmp_path):
    s = manta.Solver(name="io", gridSize=manta.vec3(*SHAPE), dim=3)
    r = s.create(manta.RealGrid)
    p = str(tmp_path / "r.uni")
    r.save(p)
    with pytest.raises(RuntimeError, match="grid type doesn't match"):
        s.create(manta.IntGrid).load(p)
    # real and levelset grids read each other's files (unifyGridType)
    s.create(manta.LevelsetGrid).load(p)
    t = manta.Solver(name="io2", gridSize=manta.vec3(8, 8, 8), dim=3)
    with pytest.raises(RuntimeError, match="grid dim doesn't match"):
        t.create(manta.RealGrid).load(p)


@pytest.mark.parametrize("name,dims,cls,dim", [("real_5x4x3", (5, 4, 3), "RealGrid", 3),
                                                ("vec3_4x3x2", (4, 3, 2), "MACGrid", 3),
                                                ("level_6x5x1", (6, 5, 1), "LevelsetGrid", 2)])
def test_device_grid_load_reference_file(name, dims, cls, dim):
    """Grid.load of files written by the reference's own .uni writer
    (tests/golden): the device grid holds the file's cells exactly."""
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    ref = np.load(os.path.join(here, "uniio_read.npz"))
    s = manta.Solver(name="gold", gridSize=manta.vec3(*dims), dim=dim)
    g = s.create(getattr(manta, cls))
    g.load(os.path.join(here, name + ".uni"))
    want = ref[name + "_data"].reshape(-1, 3) if cls == "MACGrid" else ref[name + "_data"].reshape(-1)
    np.testing.assert_array_equal(g.toNumpy(), want)
