"""The autotuner's shape cache persists across processes: a first process
times every candidate tile shape of every translation at plan creation and
appends the winners to the cache file (FMBEM_TUNE_CACHE); a second process
creating the same plan reads them and runs no timing launch, so it gets the
same shapes with their recorded times, faster.  Lines of another device
tag, of an unknown format or naming shapes this build cannot launch are
ignored."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

CHILD = r"""
import json, sys, time
import torch
import fmbem_gpu_c_amd as F
m = F.Mesh.icosphere(4, 0.1)
t = F.Tree(m, depth=5)
torch.cuda.synchronize()
t0 = time.perf_counter()
pl = F.Plan(t, 2 * 3.141592653589793 * 4000.0 / 343.0, precision=F.FMB_MIXED, nrhs_max=3,
            autotune=1, p_rule=F.PRULE_EXPLICIT, p_levels=[0, 0, 62, 44, 30, 20])
torch.cuda.synchronize()
dt = time.perf_counter() - t0
shapes = {"%d/%d/%d" % (kind, l, prec): list(pl.tile_shape_bf(kind, l, prec))
          for kind in range(3) for l in range(2 if kind == 0 else 3, 6)
          for prec in (F.FMB_FP64, F.FMB_MIXED)}
print(json.dumps({"setup_s": dt, "shapes": shapes}))
"""


def _child(cache):
    env = dict(os.environ, FMBEM_TUNE_CACHE=cache)
    p = subprocess.run([sys.executable, "-c", CHILD], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    return json.loads(p.stdout.strip().splitlines()[-1])


def test_shapes_persist_across_processes(tmp_path):
    cache = str(tmp_path / "sub" / "tiles.txt")
    # a foreign device's line and junk lines are skipped
    os.makedirs(os.path.dirname(cache))
    with open(cache, "w") as f:
        f.write("fmbem-rcr-tile-v1 gfx000/1cu 0 0 62 62 1 3 0 1 2 6 4 0 0.1\n")
        f.write("not a cache line\n")
        f.write("fmbem-rcr-tile-v1 gfx950/256cu 0 0 62 62 1 3 0 9 9 9 9 0 0.1\n")
    first = _child(cache)
    lines = open(cache).read().splitlines()
    n_tuned = len(lines) - 3
    assert n_tuned == len(first["shapes"])
    second = _child(cache)
    assert len(open(cache).read().splitlines()) == len(lines)  # nothing re-tuned
    for key, s1 in first["shapes"].items():
        s2 = second["shapes"][key]
        assert s1[:5] == s2[:5], key
        assert s2[5] == pytest.approx(s1[5], rel=1e-3), key
        assert s1[5] > 0
    print("setup with tuning %.2fs, from the cache %.2fs"
          % (first["setup_s"], second["setup_s"]), flush=True)
    assert second["setup_s"] < first["setup_s"]
