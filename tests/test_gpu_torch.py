"""The library inside a process that uses PyTorch on the same GPU.

PyTorch ships its own HIP runtime; when torch is imported first, libffrs_gpu.so binds to
that runtime (same SONAME) and both share the device. The check runs in a child process
so that the test process keeps a single runtime.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
sys.path.insert(0, ROOT)
import numpy as np
import torch
x = torch.arange(1024, device="cuda", dtype=torch.int32)
assert int(x.sum()) == 1024 * 1023 // 2
import libffrs_amd as L
from libffrs_amd import workloads as W
from oracle import oracle as O
msgs, counts, s = W.rs256_plain(512, 16, 8, seed=9)
rs = L.RS256(ecc_len=16)
par = np.frombuffer(bytes(rs.encode(msgs.tobytes())), np.uint8).reshape(-1, 16)
assert (par == O.RS256(16).encode_batch(msgs)).all()
y = torch.from_numpy(par.astype(np.int32)).cuda()
assert int(y.sum()) == int(par.astype(np.int64).sum())
print("shared runtime ok", flush=True)
"""


def test_torch_imported_first_shares_the_runtime():
    r = subprocess.run([sys.executable, "-c", CHILD.replace("ROOT", repr(ROOT))], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "shared runtime ok" in r.stdout
