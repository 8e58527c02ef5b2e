"""bench.py's host-side helpers (no GPU): the first-in-first-out replay of
the plain serving loop and the refusal of more than one GPU."""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_fifo_replay_queues_behind_earlier_requests():
    # request 0 arrives at 0 and takes 3 s (frames end 1 s and 3 s in);
    # request 1 arrives at 1 s, waits until 3 s, takes 2 s (one frame at 2 s);
    # request 2 arrives at 10 s to an idle server, takes 1 s.
    first, inter, span = bench.fifo_replay([0.0, 1.0, 10.0], [3.0, 2.0, 1.0], [[1.0, 3.0], [2.0], [0.5, 1.0]])
    assert np.allclose(first, [1000.0, 4000.0, 500.0])
    assert np.allclose(inter, [2000.0, 500.0])
    assert span == 11.0


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
