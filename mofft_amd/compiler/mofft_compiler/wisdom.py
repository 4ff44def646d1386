"""Compile GPU plan wisdom (JSON written by the autotuner) into a C table.

Role of the reference's tools/compile_wisdom.py:12-153: validate the entries
(radices from the supported set multiplying to the length, a mask that only
names level boundaries, one entry per (length, precision)) and emit
``mofft_generated_wisdom_lookup``.  The table is data, searched linearly.

JSON schema (version 1)::

  {"schema_version": 1, "target": "gfx950", "formal_measurement": bool,
   "source": "<label>", "kernel_digest": "<manifest digest>",
   "entries": [{"length": N, "precision": "fp32"|"fp64",
                "radices": [outer, ..., inner],
                "blocked_transpose_mask": int,   # pass boundaries
                "direct_input": bool, "stage_layout": "transpose"|"section",
                "kernel_bank": int, "median_seconds": float}]}
"""
from __future__ import annotations

import json
from pathlib import Path

RADICES = {2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 32, 64}


PLAIN_BANKS = (0, 1, 2, 4, 5)


def load_entries(path: str | Path) -> tuple[list[dict], bool, str]:
    payload = json.loads(Path(path).read_text())
    if payload.get("schema_version") != 1:
        raise ValueError("unsupported wisdom schema")
    if payload.get("target") != "gfx950":
        raise ValueError("wisdom was tuned for another target")
    entries = payload.get("entries")
    if not isinstance(entries, list):
        raise ValueError("wisdom entries must be a list")
    seen = set()
    out = []
    for entry in entries:
        length = int(entry["length"])
        precision = {"fp32": 32, "fp64": 64}.get(entry["precision"])
        if precision is None:
            raise ValueError(f"invalid precision {entry['precision']!r}")
        radices = [int(r) for r in entry["radices"]]
        prod = 1
        for r in radices:
            if r not in RADICES:
                raise ValueError(f"unsupported wisdom radix {r}")
            prod *= r
        if prod != length or not radices or len(radices) > 64:
            raise ValueError(f"wisdom radices do not multiply to {length}")
        mask = int(entry.get("blocked_transpose_mask", 0))
        if mask < 0 or mask >> max(0, len(radices) - 1):
            raise ValueError(f"invalid pass mask {mask}")
        key = (length, precision)
        if key in seen:
            raise ValueError(f"duplicate wisdom entry {key}")
        seen.add(key)
        direct = entry.get("direct_input", False)
        if not isinstance(direct, bool):
            raise ValueError("direct_input must be a JSON boolean")
        layout = entry.get("stage_layout", "transpose")
        if layout not in ("transpose", "section"):
            raise ValueError(f"invalid stage layout {layout!r}")
        bank = entry.get("kernel_bank", 0)
        # tile-shape banks of the registry; bank 3 (matrix-core split
        # precision) rounds differently and is never a PRETUNED choice
        if isinstance(bank, bool) or not isinstance(bank, int) or bank not in PLAIN_BANKS:
            raise ValueError(f"kernel_bank must be one of {PLAIN_BANKS}")
        out.append({"length": length, "precision": precision, "radices": radices,
                    "mask": mask, "direct": int(direct),
                    "layout": int(layout == "section"), "bank": bank})
    out.sort(key=lambda e: (e["length"], e["precision"]))
    return out, bool(payload.get("formal_measurement")), str(payload.get("source", "none"))


def emit_c(path_json: str | Path) -> str:
    entries, formal, source = load_entries(path_json)
    lines = ["/* Generated by mofft_compiler.wisdom - do not edit. */",
             '#include "../mofft_wisdom.h"', "",
             "typedef struct { size_t length; int precision; uint64_t mask;",
             "  int direct; int layout; unsigned bank; int depth; int radices[64]; } wisdom_row;",
             "",
             f"static const wisdom_row rows[] = {{"]
    for e in entries:
        rs = ", ".join(str(r) for r in e["radices"])
        lines.append(f"  {{{e['length']}u, {e['precision']}, UINT64_C({e['mask']}), "
                     f"{e['direct']}, {e['layout']}, {e['bank']}u, {len(e['radices'])}, {{{rs}}}}},")
    if not entries:
        lines.append("  {0u, 0, UINT64_C(0), 0, 0, 0u, 0, {0}},")
    lines += ["};", "",
              "size_t mofft_generated_wisdom_lookup(size_t length, int precision_bits,",
              "                                     int *radices, size_t capacity,",
              "                                     uint64_t *blocked_transpose_mask,",
              "                                     int *direct_input, int *stage_layout,",
              "                                     unsigned int *kernel_bank) {",
              "  for (size_t i = 0; i < sizeof(rows) / sizeof(rows[0]); ++i) {",
              "    const wisdom_row *r = &rows[i];",
              "    if (r->depth == 0 || r->length != length || r->precision != precision_bits) continue;",
              "    if ((size_t)r->depth > capacity) return 0;",
              "    for (int k = 0; k < r->depth; ++k) radices[k] = r->radices[k];",
              "    *blocked_transpose_mask = r->mask;",
              "    *direct_input = r->direct;",
              "    *stage_layout = r->layout;",
              "    *kernel_bank = r->bank;",
              "    return (size_t)r->depth;",
              "  }",
              "  return 0;",
              "}", "",
              f"int mofft_generated_wisdom_is_formal(void) {{ return {int(formal)}; }}", "",
              f"const char *mofft_generated_wisdom_source(void) {{ return {json.dumps(source)}; }}",
              ""]
    return "\n".join(lines)
