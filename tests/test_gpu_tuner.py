"""The autotuner's choices persist per GPU architecture: a first process
times the candidates of a float matmult list and of a float product's
block shape and appends them to the cache file; a second process with the
same file reuses every choice without timing; both get the same results."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = r"""
import json, sys
import numpy as np
sys.path.insert(0, %r)
from pdl_amd import core as C
C.init(0)
rng = np.random.default_rng(5)
ns = rng.integers(8, 160, 96)
As = [C.ndarray.from_numpy(rng.uniform(-1, 1, (n, n)).astype(np.float32)) for n in ns]
Bs = [C.ndarray.from_numpy(rng.uniform(-1, 1, (n, n)).astype(np.float32)) for n in ns]
ml = C.MatmultList(As, Bs)
ml.launch("queue")
a = C.ndarray.from_numpy(rng.standard_normal((700, 600)).astype(np.float32))
b = C.ndarray.from_numpy(rng.standard_normal((600, 900)).astype(np.float32))
c = C.matmult(a, b)
digest = float(sum(float(o.numpy().astype(np.float64).sum()) for o in ml.outputs)) + float(c.numpy().sum())
out = C.tune_stats()
out["digest"] = digest
print(json.dumps(out))
""" % ROOT


def run(env):
    r = subprocess.run([sys.executable, "-c", SCRIPT], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_second_process_reuses_the_tuning_cache(tmp_path):
    path = tmp_path / "tune.txt"
    env = dict(os.environ, PDLHIP_TUNE_FILE=str(path))
    env.pop("PDLHIP_TUNE", None)
    first = run(env)
    assert first["measured"] >= 2 and first["reused"] == 0, first
    lines = path.read_text().splitlines()
    assert len(lines) == first["measured"]
    assert any(l.startswith("list:float:") for l in lines) and any(l.startswith("matmult:float:") for l in lines)
    second = run(env)
    assert second["measured"] == 0 and second["reused"] >= 2, second
    assert path.read_text().splitlines() == lines
    assert second["digest"] == first["digest"]
