"""Writes the .uni golden files of tests/test_gridio.py with the reference's
own Python .uni writer (tensorflow/tools/uniio.py:writeUni, read from the
reference checkout given as the first argument; run once in a container that
has it).  The files hold small seeded grids:

  real_5x4x3.uni  Real grid  (gridType 1,  elementType 1), values seeded
  vec3_4x3x2.uni  MAC grid   (gridType 12, MAC|Vec3, elementType 2), xyz per cell
  level_6x5x1.uni Levelset   (gridType 17, elementType 1), a 2D grid

and uniio_read.npz: what uniio.readUni returns for each (header fields and
content), so the test can check this package's reader and writer against
the reference's reader and writer byte for byte without importing it.
uniio.py writes every grid as float32 (:103-104), so no int grid is made."""
import os
import sys

import numpy as np


def main(ref_root, out_dir):
    sys.path.insert(0, os.path.join(ref_root, "tensorflow", "tools"))
    import uniio  # noqa: E402  (the reference's module)

    rng = np.random.default_rng(2024)
    cases = {
        "real_5x4x3": ((5, 4, 3), 1, 1, 1),
        "vec3_4x3x2": ((4, 3, 2), 12, 2, 3),
        "level_6x5x1": ((6, 5, 1), 17, 1, 1),
    }
    out = {}
    for name, ((nx, ny, nz), gtype, etype, nc) in cases.items():
        data = rng.uniform(-2, 2, (nz, ny, nx, nc)).astype(np.float32)
        header = {"dimX": nx, "dimY": ny, "dimZ": nz, "gridType": gtype, "elementType": etype,
                  "bytesPerElement": 4 * nc, "info": b"golden vector written by uniio.writeUni",
                  "dimT": 0, "timestamp": 1234567890123456}
        path = os.path.join(out_dir, name + ".uni")
        uniio.writeUni(path, header, data)
        h, content = uniio.readUni(path)
        out[name + "_content"] = np.asarray(content, np.float32)
        out[name + "_header"] = np.array([h["dimX"], h["dimY"], h["dimZ"], h["gridType"], h["elementType"],
                                          h["bytesPerElement"], h["dimT"], h["timestamp"]], np.int64)
        out[name + "_data"] = data
    np.savez(os.path.join(out_dir, "uniio_read.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", os.path.dirname(os.path.abspath(__file__)))
