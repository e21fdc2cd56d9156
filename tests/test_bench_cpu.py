"""The benchmark's contract, on the CPU: one GPU only (exit 2 with a message
before any device work), the metric naming its workload and accuracy bar, and the
algorithmic flop count of the G contraction (brute-force count of the distinct
G values on a small case; the deck figure quoted with the roofline)."""
import os
import subprocess
import sys
from pathlib import Path
from types import SimpleNamespace

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


@pytest.mark.parametrize("args,env", [(["--gpus", "2"], {}), ([], {"WORLD_SIZE": "4"})])
def test_more_than_one_gpu_exits(args, env):
    e = dict(os.environ, **env)
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], cwd=ROOT, env=e,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_metric_names_workload_and_accuracy_bar():
    """The metric string carries the deck sizes, the skewness bar and the
    comparison it is quoted on."""
    for part in ("evals/s", "NN=1000", "NM=30", "1e-4", "mixed vs FP64"):
        assert part in bench.METRIC, part


def _brute_force_flops(ntc, nnl, nm, nh):
    # one dot product of length NTC*NNL per distinct value G[x](y,u; z,v) with
    # (y,u) <= (z,v) in the row order of the table (y major), u, v in 0..nh-1
    rows = [(y, u) for y in range(nm) for u in range(nh)]
    n = 0
    for x in range(nm):
        for i, _ in enumerate(rows):
            for _ in rows[i:]:
                n += 1
    return 2.0 * ntc * nnl * n


@pytest.mark.parametrize("ntc,nnl,nm,nh", [(1, 3, 2, 3), (2, 5, 3, 4), (3, 7, 1, 6)])
def test_flop_count_small(ntc, nnl, nm, nh):
    prob = SimpleNamespace(ntc=ntc, nnl=nnl, nfreqs=nh)
    assert bench.gsym_flops(prob, nm) == _brute_force_flops(ntc, nnl, nm, nh)


def test_flop_count_deck():
    prob = SimpleNamespace(ntc=2, nnl=1000, nfreqs=513)
    assert bench.gsym_flops(prob, 30) == pytest.approx(1.4212e13, rel=1e-4)
