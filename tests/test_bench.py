"""bench.py's host-side logic (no GPU): the refusal of more than one GPU and
the two-step launch a batch gets for each SP_OPT_FUSE value, as the library
chooses it (strips of one or two cells per lane, column groups of one or two
cells per lane, and the fallbacks for spin and odd nz)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from seispie_amd import _lib  # noqa: E402


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in p.stderr + p.stdout


@pytest.mark.parametrize("fuse,physics,nzs,want", [
    (0, _lib.PSV, [500], 0),
    (1, _lib.PSV, [500], 1),
    (2, _lib.PSV, [500, 400], 2),
    (2, _lib.PSV, [500, 401], 1),   # odd nz: one cell per lane
    (2, _lib.SPIN, [600], 1),       # spin: one cell per lane
    (3, _lib.SPIN, [600], 3),       # column groups, one cell per lane: any column
    (3, _lib.PSV, [1040], 3),
    (4, _lib.PSV, [1000], 4),
    (4, _lib.PSV, [1031], 1),       # two cells per lane need even nz
    (4, _lib.SPIN, [600], 1),
    (4, _lib.SH, [600, 500, 400], 4),
])
def test_step_mode_follows_the_library(fuse, physics, nzs, want):
    assert bench.step_mode(fuse, physics, nzs) == want
