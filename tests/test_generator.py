"""The gfx950 kernel generator (CPU only).

Mirrors the reference's compiler tests (tests/test_compiler.py:37-49,
611-643): every emitted codelet form equals the direct DFT, emission is
deterministic, and the registry covers the reference radix set.
"""
import json

import numpy as np
import pytest

from mofft_compiler import codelets, emit, graph, registry, wisdom


def eval_codelet(r, form, x, dtype):
    g, y = codelets.build(r, form)
    inputs = {f"r{j}": x.real[j].astype(dtype) for j in range(r)}
    inputs.update({f"i{j}": x.imag[j].astype(dtype) for j in range(r)})
    outs = graph.evaluate(g, [n for c in y for n in (c.re, c.im)], inputs, dtype)
    return np.array(outs[0::2]) + 1j * np.array(outs[1::2])


@pytest.mark.parametrize("r", codelets.RADICES)
@pytest.mark.parametrize("form", ["auto", "direct"])
def test_codelet_equals_dft(r, form):
    rng = np.random.default_rng(8137 + r)
    x = rng.standard_normal((r, 9)) + 1j * rng.standard_normal((r, 9))
    got = eval_codelet(r, form, x, np.float64)
    j = np.arange(r)
    w = np.exp(-2j * np.pi * np.outer(j, j) / r)
    want = w.T @ x
    assert np.max(np.abs(got - want)) < 1e-13 * r


@pytest.mark.parametrize("r", codelets.RADICES)
def test_codelet_fp32_accuracy(r):
    rng = np.random.default_rng(r)
    x = (rng.uniform(-1, 1, (r, 64)) + 1j * rng.uniform(-1, 1, (r, 64))).astype(np.complex64)
    got = eval_codelet(r, "auto", x, np.float32)
    want = np.fft.fft(x.astype(np.complex128), axis=0)
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert rel < 2e-7 * max(1.0, np.log2(r))


def test_factored_form_cheaper_than_direct():
    for r in codelets.RADICES:
        g, y = codelets.build(r)
        fast = graph.op_count([n for c in y for n in (c.re, c.im)])
        g, y = codelets.build(r, "direct")
        slow = graph.op_count([n for c in y for n in (c.re, c.im)])
        assert fast["add"] + fast["mul"] <= slow["add"] + slow["mul"]
    g, y = codelets.build(16)
    ops = graph.op_count([n for c in y for n in (c.re, c.im)])
    assert ops == {"add": 144, "mul": 24, "neg": ops["neg"]}


def test_emission_is_deterministic(tmp_path):
    a = emit.emit(tmp_path / "a")
    b = emit.emit(tmp_path / "b")
    assert a["digest"] == b["digest"]
    for f in (tmp_path / "a").iterdir():
        assert f.read_bytes() == (tmp_path / "b" / f.name).read_bytes()


def test_registry_covers_reference_radices():
    es = registry.entries()
    registry.check(es)
    for bits in (32, 64):
        for r in codelets.RADICES:
            assert any(e.precision == bits and not e.rows and e.length == r for e in es)
        pow2_rows = {e.length for e in es if e.precision == bits and e.rows
                     and e.length & (e.length - 1) == 0}
        assert {1 << k for k in range(1, 13)} <= pow2_rows
    for e in es:
        assert e.threads <= 1024 and e.lds_bytes <= registry.LDS_CAP


def test_chain_for():
    assert registry.chain_for(1728) == (12, 12, 12)
    assert registry.chain_for(45) in ((9, 5), (5, 9))
    assert registry.chain_for(2197) == (13, 13, 13)
    with pytest.raises(ValueError):
        registry.chain_for(17)


def _wisdom(tmp_path, entries):
    p = tmp_path / "w.json"
    p.write_text(json.dumps({"schema_version": 1, "target": "gfx950",
                             "entries": entries}))
    return p


def test_wisdom_compiler_validates(tmp_path):
    good = {"length": 1024, "precision": "fp32", "radices": [4, 16, 16],
            "blocked_transpose_mask": 0}
    text = wisdom.emit_c(_wisdom(tmp_path, [good]))
    assert "1024u, 32" in text
    assert "4u, 3" in wisdom.emit_c(_wisdom(tmp_path, [{**good, "kernel_bank": 4}]))
    bad_cases = [
        {**good, "radices": [17, 16]},
        {**good, "radices": [4, 16, 8]},
        {**good, "blocked_transpose_mask": 4},
        {**good, "precision": "fp16"},
        {**good, "kernel_bank": 9},
        {**good, "kernel_bank": 3},      # matrix-core split precision: never PRETUNED
    ]
    for bad in bad_cases:
        with pytest.raises((ValueError, KeyError)):
            wisdom.emit_c(_wisdom(tmp_path, [bad]))
    with pytest.raises(ValueError):
        wisdom.emit_c(_wisdom(tmp_path, [good, good]))
