"""Every kernel the autotuner may pick for a matmult list, forced in turn
through the tuning cache file (the key a first process wrote, set to one
candidate), checked per product against the exact product: float lists
through the SIMT 32/64 tiles, the FP32 matrix-core kernel with its default
and its half k stage, and the BF16 three-way split; double lists through the
SIMT tiles and both FP64 matrix-core variants.  Each forced run must reuse
the cached choice (no timing) and give the same result in its queue and
per-product launches, and products over an empty inner dimension must come
out as zeros in both."""
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
t = int(sys.argv[1])
dt = np.float32 if t == C.L.F else np.float64
rng = np.random.default_rng(77 + t)
ns = rng.integers(8, 256, size=64, endpoint=True)
ah = [rng.uniform(-1, 1, (n, n)).astype(dt) for n in ns]
bh = [rng.uniform(-1, 1, (n, n)).astype(dt) for n in ns]
ml = C.MatmultList([C.ndarray.from_numpy(a, t) for a in ah], [C.ndarray.from_numpy(b, t) for b in bh])
worst, same = 0.0, True
for mode in ("queue", "each"):
    ml.launch(mode)
    outs = [o.numpy() for o in ml.outputs]
    if mode == "queue":
        first = outs
    else:
        same = all(np.array_equal(x, y) for x, y in zip(first, outs))
    for k in range(len(ns)):
        truth = ah[k].astype(np.longdouble) @ bh[k].astype(np.longdouble)
        err = (np.linalg.norm((outs[k].astype(np.longdouble) - truth).astype(np.float64), axis=1) /
               np.linalg.norm(truth.astype(np.float64), axis=1))
        worst = max(worst, float(err.max()))
ml.free()
# products over an empty inner dimension (h x 0 times 0 x w) are zeros in
# both launch modes, per-product first
eh = [np.zeros((h, 0), dt) for h in (5, 70)]
eb = [np.zeros((0, w), dt) for w in (7, 65)]
ml = C.MatmultList([C.ndarray.from_numpy(a, t) for a in eh], [C.ndarray.from_numpy(b, t) for b in eb])
empty_zero = True
for mode in ("each", "queue"):
    ml.outputs[1].assign(C.ndarray.from_numpy(np.ones((70, 65), dt), t))
    ml.launch(mode)
    outs = [o.numpy() for o in ml.outputs]
    empty_zero = empty_zero and [o.shape for o in outs] == [(5, 7), (70, 65)] and all(not o.any() for o in outs)
ml.free()
out = C.tune_stats()
out.update({"worst": worst, "queue_equals_each": same, "empty_zero": empty_zero})
print(json.dumps(out))
""" % ROOT


def run(env, t):
    r = subprocess.run([sys.executable, "-c", SCRIPT, str(t)], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("t,cands,tol", [(9, [32, 64, 0, 2, 1], 4e-6), (10, [32, 64, 0, 2], 2e-14)],
                         ids=["float", "double"])
def test_every_list_kernel_candidate(tmp_path, t, cands, tol):
    path = tmp_path / "tune.txt"
    env = dict(os.environ, PDLHIP_TUNE_FILE=str(path))
    env.pop("PDLHIP_TUNE", None)
    first = run(env, t)
    assert first["measured"] >= 1 and first["worst"] < tol, first
    keys = [line.split()[0] for line in path.read_text().splitlines() if line.startswith("list:")]
    assert len(keys) == 2, keys                     # the random list, the empty-k list
    for c in cands:
        path.write_text("".join("%s %d\n" % (k, c) for k in keys))
        got = run(env, t)
        assert got["measured"] == 0 and got["reused"] >= 1, (c, got)
        assert got["worst"] < tol, (c, got)
        assert got["queue_equals_each"] and got["empty_zero"], (c, got)
