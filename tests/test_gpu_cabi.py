"""C++ host code on the C ABI alone (tests/cabi/plume_host.cpp, built by
__graft_entry__.build()): the 3D plume stepped by direct plugin calls, and
the same plume stepped by the serving engine beside a second session, both
equal the Python face's plume (serve.PlumeScene.run_single, the scene
script's calls) bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from mantaflow_amd import serve

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cabi", "plume_host")


@pytest.mark.parametrize("shape,steps", [((40, 48, 36), 4), ((33, 20, 27), 3)])
def test_cpp_host_plume_matches_python_face(tmp_path, shape, steps):
    assert os.path.exists(EXE), "tests/cabi/plume_host is not built (python __graft_entry__.py builds it)"
    seed = 5
    n = int(np.prod(shape))
    # the source values serve.PlumeScene draws for this seed
    val = np.random.default_rng(seed).uniform(0.8, 1.0, n).astype(np.float32)
    val.tofile(str(tmp_path / "val.f32"))
    r = subprocess.run([EXE, *[str(a) for a in shape], str(steps), str(tmp_path / "val.f32"), str(tmp_path)],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    assert "largest batch 2" in r.stdout  # the two sessions' steps shared ticks
    ps = serve.PlumeScene(shape, seed=seed)
    for _ in range(steps):
        ps.run_single(serve.REQ_STEP)
    for tag in ("direct", "served"):
        for name, g in (("density", ps.density), ("vel", ps.vel), ("pressure", ps.pressure)):
            a = np.fromfile(str(tmp_path / ("%s_%s.f32" % (tag, name))), np.float32)
            np.testing.assert_array_equal(a, g.toNumpy().ravel(), err_msg="%s %s" % (tag, name))
