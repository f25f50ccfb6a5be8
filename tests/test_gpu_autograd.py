"""Gradient-check harness for the checkpointed rollout through a
torch.autograd.Function (tests/autograd_rollout.py), run in a process of its
own so that torch is imported before the library and the process holds one
HIP runtime."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_autograd_rollout_gradient():
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "autograd_rollout.py")], capture_output=True,
                       text=True, timeout=240)
    print(r.stdout[-4000:], flush=True)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "[autograd] ok" in r.stdout
