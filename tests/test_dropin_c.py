"""Drop-in proof: the reference-style C test program (tests/c/test_dropin.c),
built against the unchanged include/mofft.h and linked to libmofft.so, passes
on the GPU."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).parent / "c" / "build" / "test_dropin"


@pytest.mark.gpu
def test_c_dropin_suite():
    assert BIN.exists(), "run python -m mofft_amd.build first"
    res = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=110)
    print(res.stdout)
    print(res.stderr)
    assert res.returncode == 0, res.stderr[-4000:]
