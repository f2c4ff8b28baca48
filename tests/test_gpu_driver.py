"""The test_bfs-style C++ driver (back40computing_amd/test_bfs, built from
csrc/test_bfs.cpp over include/b40c_bfs.hpp) run end to end on the device, as
the reference's driver is run by its scripts (test/bfs/test_bfs.cu:584-783).

The driver validates every search against its own serial traversal and exits
nonzero on any INCORRECT result (the reference only printed it)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "back40computing_amd", "test_bfs")

pytestmark = pytest.mark.gpu


def run_driver(*args, timeout=240):
    assert os.path.exists(DRIVER), "test_bfs driver is not built"
    print("test_bfs", " ".join(args), flush=True)
    p = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, (p.returncode, p.stdout[-3000:], p.stderr[-3000:])
    assert "INCORRECT" not in p.stdout
    return p.stdout


@pytest.mark.parametrize("graph", [["grid2d", "300"], ["rmat", "16", "16"], ["random", "20000", "80000"]])
@pytest.mark.parametrize("extra", [[], ["--mark-pred"], ["--iterative"], ["--instrumented"]])
def test_driver_every_strategy(gpu, graph, extra):
    iters = 2
    out = run_driver(*graph, "--strategy=0,1,2,3,4", "--i=%d" % iters, *extra)
    # one validated result per strategy per iteration (plus the warm-up)
    assert out.count("CORRECT") == 5 * (iters + 1), out[-2000:]
    if "--instrumented" in extra:
        duties = [float(x) for x in re.findall(r"avg cta duty: ([0-9.]+)%", out)]
        # the persistent strategies (expand-contract, contract-expand,
        # two-phase fused, hybrid, direction-optimizing) report their duty
        assert len(duties) >= 4 * iters, out[-2000:]
        assert all(0.0 < d <= 100.0 for d in duties), duties


def test_driver_rejects_unknown_strategy(gpu):
    p = subprocess.run([DRIVER, "grid2d", "10", "--strategy=9"], capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
