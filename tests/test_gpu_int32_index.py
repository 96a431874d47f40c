"""The int32 index dtype (YGGDRAX_INDEX_PRECISION=int32, yggdrax/dtypes.py)
on the GPU: the index dtype is fixed at import, so the case runs in one
child process (tests/gpu_int32_index_case.py)."""

from __future__ import annotations

import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_int32_index_outputs_match_oracle():
    env = dict(os.environ, YGGDRAX_INDEX_PRECISION="int32")
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "gpu_int32_index_case.py")],
                       env=env, capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-3000:], flush=True)
    assert r.returncode == 0
    assert "INT32 OK" in r.stdout
