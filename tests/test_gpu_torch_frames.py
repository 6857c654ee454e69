"""torch.autograd over the frame evaluator: the gradients of a force-matching
loss in the tunables and in the positions, against the evaluator's own
force-matching call and FP64 central differences of the oracle
(tests/torch_frames_check.py), with CPU tensors in this process and with
CUDA tensors in a process of their own, where torch's HIP runtime is
initialised before the library loads."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_autograd_cpu_tensors():
    print("case: autograd, 2 frames x 960 atoms, CPU tensors", flush=True)
    sys.path.insert(0, HERE)
    from torch_frames_check import check
    check("cpu")


def test_autograd_cuda_tensors():
    print("case: autograd, 2 frames x 960 atoms, CUDA tensors (own process)", flush=True)
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "torch_frames_check.py"),
                        "cuda:0"], capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-3000:], flush=True)
    assert r.returncode == 0
