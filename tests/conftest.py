import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs (not part of the default CPU suite)")


def gpu_available():
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so.7")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


def pytest_collection_modifyitems(config, items):
    if "not slow" in (config.getoption("-m") or "") or config.getoption("-m") in ("gpu", "not gpu"):
        pass
    run_slow = os.environ.get("H2W_RUN_SLOW") == "1" or "slow" in (config.getoption("-m") or "").replace("not slow", "")
    for it in items:
        if "slow" in it.keywords and not run_slow:
            it.add_marker(pytest.mark.skip(reason="slow oracle run: set H2W_RUN_SLOW=1 or -m slow"))
