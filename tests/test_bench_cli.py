"""bench.py's contract checks that need no GPU."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_bench_imports_and_refuses_multi_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bench_algorithmic_bytes():
    sys.path.insert(0, str(ROOT))
    import bench
    H, _, cfg = bench.build_workload("C1", 0)
    N = H.n_rows * H.b_len
    b = bench.algorithmic_bytes(H, N, 16, 16)
    vec = N * 16 * 16
    h_bytes = H.n_blocks() * 16 + 4 * H.n_blocks() + 4 * (H.n_rows + 1)
    assert b["moments"] == 3 * vec + h_bytes
    assert b["autodiff"] == 6 * vec + h_bytes + 2 * H.n_blocks() * 16
