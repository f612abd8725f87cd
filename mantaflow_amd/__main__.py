"""Run a mantaflow scene script on the GPU path: python -m mantaflow_amd scene.py [args]

The script's `from manta import *` resolves to mantaflow_amd.manta.
"""
import runpy
import sys

from . import manta


def main():
    if len(sys.argv) < 2:
        print("usage: python -m mantaflow_amd scene.py [args...]")
        return 1
    sys.modules["manta"] = manta
    script = sys.argv[1]
    sys.argv = sys.argv[1:]
    runpy.run_path(script, run_name="__main__")
    return 0


if __name__ == "__main__":
    sys.exit(main())
