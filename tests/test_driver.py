"""The C++ drop-in (include/h2w_initialisation.hpp) through its command-line
driver: same arguments and summary lines as the reference's main()."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "h2weak_amd", "h2w_main")


def _run(args, timeout=300):
    return subprocess.run([DRV] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)


def test_driver_usage_and_no_gpu():
    from conftest import gpu_available
    r = _run([])
    assert r.returncode == 2 and "usage" in r.stderr
    if not gpu_available():
        r = _run([3, 16, 3, 1, 8, 1])
        assert r.returncode == 1 and "no HIP device" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("args,iters_max", [((3, 16, 3, 1, 8, 1), 40), ((2, 48, 6, 1, 10, 0, 1), 80),
                                            ((1, 4000, 64, 1, 12, 1, 0, 3), 60)])
def test_driver_solves(args, iters_max):
    r = _run(args)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    it = int(re.search(r"total number of iteration is: (\d+)", out).group(1))
    err = float(re.search(r"relative error in solution: ([0-9.eE+-]+)", out).group(1))
    assert 0 < it <= iters_max
    assert err < 1e-6
