"""The library's diagnostic level (CPU: no device call).

bda_set_debug_level / bda_get_debug_level replace the reference's
compile-time BDA_DEBUG_LEVEL switch (src/sda_app/common/bda_utils.hpp:46-57);
the initial level comes from the environment variable of the same name.
The lines themselves are checked on the GPU (test_gpu_debug_level.py).
"""
import os
import subprocess
import sys

import pytest

import fpga_amd as fa


def test_set_and_get_round_trip():
    prev = fa.set_debug_level(2)
    try:
        assert fa.debug_level() == 2
        assert fa.set_debug_level(0) == 2
        assert fa.debug_level() == 0
        with pytest.raises(fa.BdaError):
            fa.set_debug_level(-1)
        assert fa.debug_level() == 0
    finally:
        fa.set_debug_level(prev)


@pytest.mark.parametrize("env,expect", [("2", 2), ("0", 0), (None, 0), ("-3", 0)])
def test_initial_level_from_environment(env, expect):
    e = dict(os.environ)
    e.pop("BDA_DEBUG_LEVEL", None)
    if env is not None:
        e["BDA_DEBUG_LEVEL"] = env
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", "import fpga_amd as fa; print(fa.debug_level())"], cwd=root, env=e,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert int(r.stdout.strip().splitlines()[-1]) == expect
