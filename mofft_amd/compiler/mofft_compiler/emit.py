"""Emit the generated HIP/C sources for gfx950.

Outputs (all deterministic for a given registry):

* ``mofft_codelets.h``    - ``mofft::Dft<R>::run<T>(re, im)`` straight-line
  forward codelets for R in the reference radix set (emitter.py:283-466 is
  the SME counterpart);
* ``mofft_registry.c``    - the host-side registry table (plain C);
* ``mofft_kernels_pN.hip`` - explicit instantiations of the pass kernel and
  one launcher per registry entry, split into parts for parallel builds;
* ``mofft_launch_table.cpp`` - the host-only launcher table indexed by registry index
  (the counterpart of the reference's generated dispatchers,
  emitter.py:557-620);
* ``manifest.json``       - entry list and op counts.
"""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

from . import codelets, registry
from .graph import Const, live_nodes, op_count

HEADER = "/* Generated by mofft_amd/compiler/mofft_compiler/emit.py - do not edit. */\n"


def _lit(c: Const) -> str:
    return f"kc<T>({float(c.f64).hex()}, {float(c.f32).hex()}f)"


def emit_codelet(r: int) -> tuple[str, dict]:
    g, y = codelets.build(r)
    outs = [n for c in y for n in (c.re, c.im)]
    lines = [f"template <> struct Dft<{r}> {{",
             "  template <typename T>",
             f"  static __device__ __forceinline__ void run(T (&re)[{r}], T (&im)[{r}]) {{"]
    names: dict[int, str] = {}
    for node in live_nodes(outs):
        if node.kind == "in":
            kind, idx = node.name[0], int(node.name[1:])
            var = f"x{kind}{idx}"
            arr = "re" if kind == "r" else "im"
            lines.append(f"    const T {var} = {arr}[{idx}];")
            names[node.index] = var
        elif node.kind == "const":
            names[node.index] = _lit(node.const)
        else:
            var = f"v{node.index}"
            a = [names[x.index] for x in node.args]
            if node.kind == "add":
                expr = f"{a[0]} + {a[1]}"
            elif node.kind == "sub":
                expr = f"{a[0]} - {a[1]}"
            elif node.kind == "mul":
                expr = f"{a[0]} * {a[1]}"
            else:
                expr = f"-{a[0]}"
            lines.append(f"    const T {var} = {expr};")
            names[node.index] = var
    for k, c in enumerate(y):
        lines.append(f"    re[{k}] = {names[c.re.index]};")
        lines.append(f"    im[{k}] = {names[c.im.index]};")
    lines += ["  }", "};", ""]
    return "\n".join(lines), op_count(outs)


def mfma_table(radix: int) -> list[int]:
    """FP16 bit patterns of the matrix-core radix-16 / radix-32 operand
    (csrc/hip/fft_pass.h, MStage): the real 2R x 2R form M of the forward
    R-point DFT matrix (rows 0..R-1 real and R..2R-1 imaginary outputs,
    columns 0..R-1 real and R..2R-1 imaginary inputs), split M = Mh + Ml
    into two FP16 values (Mh = fp16(M), Ml = fp16(M - Mh), both rounded to
    nearest from the exact cos/sin values), in A-operand lane order of
    R / 16 row tiles of 32 and R / 8 K-blocks of 16: entry
    [part][rt][kb][lane][e] = M{h,l}[32 rt + lane % 32][16 kb + 8 (lane // 32) + e]
    for part 0 (hi) / 1 (lo)."""
    import mpmath
    import numpy as np
    mpmath.mp.prec = 113
    r2 = 2 * radix

    def w(q, m):   # exp(-2 pi i q m / R)
        a = mpmath.mpf(2) * mpmath.pi * ((q * m) % radix) / radix
        return mpmath.cos(a), -mpmath.sin(a)

    M = [[mpmath.mpf(0)] * r2 for _ in range(r2)]
    for q in range(radix):
        for m in range(radix):
            wr, wi = w(q, m)
            M[q][m] = wr
            M[q][radix + m] = -wi
            M[radix + q][m] = wi
            M[radix + q][radix + m] = wr
    out = []
    for part in (0, 1):
        for rt in range(radix // 16):
            for kb in range(radix // 8):
                for lane in range(64):
                    for e in range(8):
                        v = M[32 * rt + lane % 32][16 * kb + 8 * (lane // 32) + e]
                        hi = np.float16(float(v))
                        val = hi if part == 0 else np.float16(float(v - mpmath.mpf(float(hi))))
                        out.append(int(np.array(val, dtype=np.float16).view(np.uint16)))
    return out


def emit_codelets_header() -> tuple[str, dict]:
    parts = [HEADER, "#pragma once", "#include <hip/hip_runtime.h>", "",
             "namespace mofft {", "",
             "template <typename T>",
             "__device__ __forceinline__ constexpr T kc(double d, float f) {",
             "  if constexpr (sizeof(T) == 4) return f; else return d;",
             "}", "",
             "template <int R> struct Dft;", ""]
    for radix in (16, 32):
        tab = mfma_table(radix)
        parts.append(f"/* matrix-core radix-{radix} operand: split FP16 DFT matrix, A-operand")
        parts.append("   lane order [part][row tile][K-block][lane][e] (emit.py mfma_table) */")
        parts.append("static __constant__ __attribute__((aligned(16))) unsigned short "
                     f"mofft_mfma{radix}_a[{len(tab)}] = {{")
        for i in range(0, len(tab), 16):
            parts.append("  " + ", ".join(f"0x{v:04x}" for v in tab[i:i + 16]) + ",")
        parts += ["};", ""]
    counts = {}
    for r in codelets.RADICES:
        if r > 16:
            continue  # radix 32/64 run as two-stage L = 32/64 pass kernels
        text, oc = emit_codelet(r)
        parts.append(text)
        counts[r] = oc
    parts += ["}  // namespace mofft", ""]
    return "\n".join(parts), counts


def _targs(e: registry.Entry) -> str:
    t = "float" if e.precision == 32 else "double"
    # matrix-core kernels (bank 3) mark their radix-16 / radix-32 sub-stages
    # kMfma16 / kMfma32
    stages = ", ".join(f"kMfma{r}" if (e.variant == registry.MFMA_VARIANT and r in (16, 32))
                       else str(r) for r in e.stages)
    return f"{t}, {e.length}, {e.tpt}, {e.cols}, {'true' if e.rows else 'false'}, {stages}"


def _registry_rows(es: list[registry.Entry]) -> list[str]:
    lines = []
    for e in es:
        st = ", ".join(str(r) for r in e.stages)
        lines.append(f"  {{{e.precision}, {e.length}, {e.rows}, {e.tpt}, {e.cols}, "
                     f"{len(e.stages)}, {{{st}}}, {e.lds_bytes}, {e.variant}}},")
    return lines


def emit_registry_c(es: list[registry.Entry], bs: list[registry.Entry]) -> str:
    lines = [HEADER, '#include "../mofft_abi.h"', "",
             "const mofft_registry_entry mofft_registry[] = {"]
    lines += _registry_rows(es)
    lines += ["};", f"const int mofft_registry_count = {len(es)};", "",
              "/* tile kinds of the persistent batch kernels (index = kind) */",
              "const mofft_registry_entry mofft_batch_registry[] = {"]
    lines += _registry_rows(bs)
    lines += ["};", f"const int mofft_batch_registry_count = {len(bs)};", ""]
    return "\n".join(lines)


PASS_ROLES = {"rows": 0, "first": 1, "later": 2}
PASS_SPANS = (16, 64, 256, 1024)   # strides / spans the strided roles are timed at


def codelet_flops(counts: dict) -> dict[int, int]:
    """Real flops of one radix-r DFT codelet (adds + multiplies of the
    generated straight-line code; negations are free), and of the radix
    32 / 64 levels, which run as two-stage L = 32 (8 x 4) / 64 (8 x 8)
    kernels: their codelets plus one complex twiddle product (6 flops) per
    non-trivial point of the second stage."""
    out = {r: c["add"] + c["mul"] for r, c in counts.items()}
    for L, (a, b) in ((32, (8, 4)), (64, (8, 8))):
        out[L] = (L // a) * out[a] + (L // b) * out[b] + 6 * (L - L // b)
    return out


def emit_costs_c(es: list[registry.Entry], counts: dict,
                 passes_json: Path | None) -> str:
    """The analytical planner's measured data: generated codelet flop
    counts by radix, and the pass-cost table (ns per byte moved) of every
    registry entry in each role it was timed in, from `mofft_tune
    --pass-costs` output: column 0 as the single pass of N = L (rows),
    columns 1-4 as a strided first pass at stride 16 / 64 / 256 / 1024,
    columns 5-8 as a strided pass with pass twiddles at span 16 / 64 / 256
    / 1024; 0 = not measured."""
    flops = codelet_flops(counts)
    index = {(e.precision, e.rows, e.length, tuple(e.stages), e.variant): i
             for i, e in enumerate(es)}
    ns = [[0.0] * 9 for _ in es]
    source = "none"
    if passes_json is not None and Path(passes_json).exists():
        payload = json.loads(Path(passes_json).read_text())
        if payload.get("schema_version") != 2 or payload.get("target") != "gfx950":
            raise ValueError("pass-cost table of another schema or target")
        source = str(payload.get("source", "none"))
        for e in payload["entries"]:
            key = (int(e["precision"]), int(e["rows"]), int(e["length"]),
                   tuple(int(r) for r in e["stages"]), int(e["variant"]))
            i = index.get(key)
            if i is None:
                continue   # a kernel no longer compiled
            if e["bytes"] <= 0 or e["seconds"] <= 0:
                raise ValueError(f"pass cost must be positive: {e}")
            role = PASS_ROLES[e["role"]]
            col = 0 if role == 0 else (1 if role == 1 else 5) + PASS_SPANS.index(int(e["span"]))
            ns[i][col] = 1e9 * float(e["seconds"]) / float(e["bytes"])
    lines = [HEADER, '#include "../mofft_internal.h"', "",
             "/* flops of one radix-r codelet, index r (0: no codelet) */",
             "const double mofft_codelet_flops[65] = {"]
    lines.append("  " + ", ".join(f"{float(flops.get(r, 0))}" for r in range(65)))
    lines += ["};", "",
              "/* ns per byte moved (one read + one write of the data) of every",
              "   registry entry: [rows, first at stride 16, 64, 256, 1024,",
              "   later at span 16, 64, 256, 1024]; 0 = not measured */",
              f"const double mofft_pass_cost_ns[{max(1, len(es))}][9] = {{"]
    for i, row in enumerate(ns):
        lines.append("  {" + ", ".join(f"{v:.5e}" for v in row) + f"}},  /* {es[i].symbol} */")
    if not es:
        lines.append("  {0.0},")
    lines += ["};", f"const char *const mofft_pass_cost_source = {json.dumps(source)};", ""]
    return "\n".join(lines)


def emit_batch_part(bs: list[registry.Entry], precision: int) -> str:
    """One precision's persistent batch kernel: a switch over its kinds."""
    t = "float" if precision == 32 else "double"
    asserts: list[str] = []
    lines = [HEADER, '#include "../hip/batch.h"', "", "namespace mofft {",
             f"struct BatchDispatchF{precision} {{",
             f"  static __device__ __forceinline__ void run(int kind, const mofft_batch_segment &sg,",
             f"                                             int64_t block, {t} *lds) {{",
             "    switch (kind) {"]
    for i, e in enumerate(bs):
        if e.precision != precision:
            continue
        lines.append(f"      case {i}: batch_tile<{_targs(e)}>(sg, block, lds); return;")
        asserts.append(f"static_assert(StageTw<{_targs(e)}>::BYTES == {e.lds_bytes},"
                       f' "registry LDS bytes of batch kind {i}");')
    lines += ["      default: return;", "    }", "  }", "};", ""] + asserts + ["",
              f"__global__ void __launch_bounds__(kBatchThreads) MOFFT_BATCH_WAVES_F{precision}",
              f"batch_kernel_f{precision}(mofft_batch_header h) {{",
              f"  batch_loop<{t}, BatchDispatchF{precision}>(h);", "}", "",
              f"int batch_occupancy_f{precision}(int lds, int *bpc, int *cus) {{",
              f"  return BatchLauncher<batch_kernel_f{precision}>::occupancy(lds, bpc, cus);", "}",
              f"int batch_launch_f{precision}(const mofft_batch_header *h, int grid, int lds, void *s) {{",
              f"  return BatchLauncher<batch_kernel_f{precision}>::launch(h, grid, lds, (hipStream_t)s);",
              "}", "}  // namespace mofft", ""]
    return "\n".join(lines)


def emit_kernel_part(es: list[registry.Entry], indices: list[int]) -> str:
    lines = [HEADER, '#include "../hip/fft_pass.h"', "", "namespace mofft {"]
    for i in indices:
        e = es[i]
        lines.append(f"int launch_{e.symbol}(const mofft_pass_args *a, void *s) {{"
                     f" return PassKernel<{_targs(e)}>::launch(a, (hipStream_t)s); }}")
        lines.append(f"static_assert(PassKernel<{_targs(e)}>::LDS == {e.lds_bytes},"
                     f' "registry LDS bytes of {e.symbol}");')
    lines += ["}  // namespace mofft", ""]
    return "\n".join(lines)


def kernel_part_of(e: registry.Entry, parts: int) -> int:
    """Source file of an entry's launcher: a hash of its symbol, so that
    adding or removing registry entries recompiles only the files those
    entries land in (a file holds ~70 instantiations, ~4 s each)."""
    return int(hashlib.sha256(e.symbol.encode()).hexdigest()[:8], 16) % parts


def emit_launch_table(es: list[registry.Entry]) -> str:
    lines = [HEADER, '#include "../mofft_abi.h"', "",
             "namespace mofft {"]
    for e in es:
        lines.append(f"int launch_{e.symbol}(const mofft_pass_args *, void *);")
    lines += ["", "typedef int (*launch_fn)(const mofft_pass_args *, void *);",
              "extern const launch_fn launch_table[] = {"]
    for e in es:
        lines.append(f"  launch_{e.symbol},")
    lines += ["};", f"extern const int launch_table_size = {len(es)};",
              "}  // namespace mofft", ""]
    return "\n".join(lines)


def emit(out_dir: str | Path, parts: int = 24,
         passes_json: str | Path | None = None) -> dict:
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    es = registry.entries()
    registry.check(es)
    files: dict[str, str] = {}
    files["mofft_codelets.h"], counts = emit_codelets_header()
    bs = registry.batch_entries()
    files["mofft_registry.c"] = emit_registry_c(es, bs)
    for precision in (32, 64):
        files[f"mofft_batch_f{precision}.hip"] = emit_batch_part(bs, precision)
    buckets: list[list[int]] = [[] for _ in range(parts)]
    for i, e in enumerate(es):
        buckets[kernel_part_of(e, parts)].append(i)
    for stale in out.glob("mofft_kernels_p*.hip"):
        if int(stale.stem.rsplit("_p", 1)[1]) >= parts:
            stale.unlink()
    for k, idx in enumerate(buckets):
        files[f"mofft_kernels_p{k}.hip"] = emit_kernel_part(es, sorted(idx))
    files["mofft_launch_table.cpp"] = emit_launch_table(es)
    files["mofft_costs.c"] = emit_costs_c(es, counts, Path(passes_json) if passes_json else None)
    digest = hashlib.sha256()
    for name in sorted(files):
        digest.update(name.encode())
        digest.update(files[name].encode())
    manifest = {
        "schema": 1,
        "target": "gfx950",
        "entries": [{"precision": e.precision, "length": e.length, "rows": e.rows,
                     "threads_per_column": e.tpt, "columns_per_block": e.cols,
                     "stages": list(e.stages), "lds_bytes": e.lds_bytes,
                     "variant": e.variant}
                    for e in es],
        "batch_kinds": [{"precision": e.precision, "length": e.length, "rows": e.rows,
                         "threads_per_column": e.tpt, "columns_per_block": e.cols,
                         "stages": list(e.stages), "lds_bytes": e.lds_bytes}
                        for e in bs],
        "codelet_ops": {str(k): v for k, v in counts.items()},
        "digest": digest.hexdigest(),
    }
    files["manifest.json"] = json.dumps(manifest, indent=1, sort_keys=True) + "\n"
    for name, text in files.items():
        path = out / name
        if not path.exists() or path.read_text() != text:
            path.write_text(text)
    return manifest
