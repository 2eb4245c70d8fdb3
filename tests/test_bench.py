"""Host-side checks of bench.py: its refusal of multi-GPU runs and the algorithmic-bytes model."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_multi_gpu_request_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_algorithmic_bytes_of_the_chain():
    sys.path.insert(0, ROOT)
    import bench
    from spice21_amd import workloads as W

    nl, _ = W.chain_mc(4)
    cn = nl.to_c()
    s = dict(nnz_a=328, nnz_lu=364)
    # 128 Mos1 x (10 params x 8 B + 4 terminals x 4 B) + 2 Vsrc x (8 B + 3 x 4 B) + 8 x 328 + 2w x 364 + 32 x 68
    assert bench.algorithmic_bytes(cn, s, 8) == 128 * 96 + 2 * 20 + 8 * 328 + 16 * 364 + 32 * 68 == 22952
    assert bench.algorithmic_bytes(cn, s, 4) == 20040
