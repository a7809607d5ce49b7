"""bench.py's contract pieces that need no GPU: the multi-GPU request exits
with a message before any device work, and the roofline numerator (the kick's
algorithmic bytes) follows from its stated per-particle and per-cell terms."""
import json
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=REPO)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_kick_algorithmic_bytes():
    import bench
    n, nm = 10_000_000, (127, 127, 127)
    nc = 127 ** 3
    total, kick = bench.algorithmic_bytes(n, nm, "mixed")
    # six FP64 rows read and written, three FP32 field grids read once
    assert kick == 96 * n + 12 * nc
    # particle passes A-D: 24 + 56 + 56 + 96 bytes per particle, then the mesh
    assert total - (24 + 56 + 56 + 96) * n > 0
    total64, kick64 = bench.algorithmic_bytes(n, nm, "fp64")
    assert kick64 == 96 * n + 24 * nc and total64 > total
