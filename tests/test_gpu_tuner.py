"""The autotuner (mofft_amd/tune/mofft_tune, the role of the reference's
tools/plan_search.cpp) on a few lengths: it must exit 0 (3 means a timed
plan disagreed with the PRETUNED plan's output), and its JSON must pass the
wisdom compiler's validation (tools/compile_wisdom.py:12-153 rules: radices
from the supported set multiplying to the length, a mask naming only level
boundaries, a plain kernel bank, one entry per (length, precision)); each
recorded plan then runs against the oracle.  The pass-cost mode's table must
cover every compiled kernel it can place and give positive times."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import seeded_complex

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TUNE = os.path.join(ROOT, "mofft_amd", "tune", "build", "mofft_tune")


def test_tuner_writes_valid_wisdom(gpu, tmp_path):
    from mofft_compiler import wisdom
    out = tmp_path / "w.json"
    r = subprocess.run([TUNE, "--precision", "both", "--lengths", "1024,4096,65536",
                        "--points", str(1 << 20), "--candidates", "4", "--samples", "2",
                        "-o", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    entries, formal, source = wisdom.load_entries(out)
    assert source == "mofft_tune" and not formal
    assert sorted((e["length"], e["precision"]) for e in entries) == \
        sorted((n, b) for n in (1024, 4096, 65536) for b in (32, 64))
    raw = json.loads(out.read_text())["entries"]
    for j, e in zip(sorted(raw, key=lambda j: (j["length"], j["precision"])), entries):
        assert j["median_seconds"] > 0 and j["batch"] >= 1
        n, bits = e["length"], e["precision"]
        x = seeded_complex((2, n), seed=n + bits).astype(np.complex64 if bits == 32 else np.complex128)
        with gpu.Plan(n, 2, bits, radices=e["radices"], pass_mask=e["mask"]) as p:
            if e["bank"]:
                getattr(gpu.lib(), f"mofft_plan_set_kernel_bank_f{bits}")(p.handle, e["bank"])
            y = np.empty_like(x)
            p.execute(x, y)
        tol = (1e-6 if bits == 32 else 1e-15) * math.log2(n)
        assert oracle.rel_l2(y, oracle.fft_dd(x, -1)) <= tol, (n, bits, e)


def test_tuner_pass_cost_table(gpu, tmp_path):
    out = tmp_path / "p.json"
    r = subprocess.run([TUNE, "--pass-costs", str(out), "--precision", "32",
                        "--points", str(1 << 20), "--samples", "1"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    t = json.loads(out.read_text())
    assert t["schema_version"] == 2 and t["target"] == "gfx950"
    ks = [k for k in gpu.kernels() if k.precision_bits == 32]
    rows = [e for e in t["entries"] if e["role"] == "rows"]
    later = [e for e in t["entries"] if e["role"] == "later"]
    assert len(rows) == sum(1 for k in ks if k.rows)
    # every strided kernel as a later pass at every span its pairing fits in
    # the 2^20 points of this run (L x span <= 2^20)
    spans = (16, 64, 256, 1024)
    assert len(later) == sum(1 for k in ks if not k.rows
                             for s in spans if k.length * s <= 1 << 20)
    for e in t["entries"]:
        assert e["seconds"] > 0 and e["bytes"] > 0
        assert math.prod(e["stages"]) == e["length"]
        assert e["role"] == "rows" or e["span"] in spans
