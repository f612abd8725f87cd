"""The oracle's pinning checks must reject plausible mistakes: each build of
the oracle with one planted mistake (-DOG_MUTANT=k, the list at the top of
oracle/mfx_oracle.c: swapped MAC-accessor neighbours, a dropped MacCormack
skip rule, swapped clamp bounds, an unshifted MAC lookup, outflow and wall
rule slips, cubic-lookup and midpoint-trace slips, diffusion-matrix and
wave-RHS slips) must fail at least one
check of tests/oracle_pins.py.  CPU only;
the builds go to a temporary directory."""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from oracle import Oracle
from tests.oracle_pins import all_pins

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
MUTANTS = list(range(1, 21))


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    out = tmp_path_factory.mktemp("oracle_mutants")

    def build(k):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR, "mutant", "MUT=%d" % k, "OUT=%s" % out],
                       check=True, capture_output=True)
        return k, str(out / ("liboracle_m%d.so" % k))

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        return dict(ex.map(build, MUTANTS))


@pytest.mark.parametrize("k", MUTANTS)
def test_planted_mistake_is_rejected(mutant_libs, k):
    o = Oracle(path=mutant_libs[k])
    failed = []
    for name, pin in all_pins():
        try:
            pin(o)
        except AssertionError:
            failed.append(name)
    assert failed, "planted mistake %d passes every pin" % k
