"""Extracts the reference's known-answer vectors into tests/golden/*.json.

Reads (as text) the reference's own device tests:
  core/src/common/mont/tests/main.cu   -- bn254 Fr instances 1, 2, the
                                          to/from-Montgomery pair (lines 238-525)
  core/src/common/curve/tests/bn254.cu -- G1 points p1, p2 and p1+p2,
                                          p1+100*p2, 128*p1 (lines 174-203)
Values are written as big-endian hex strings of the integers the
BIG_INTEGER_CHUNKS8/16 macros spell (field.cuh:10-13: arguments run from the
most significant 32-bit chunk to the least).

Usage: python tests/golden/make_kats.py [/path/to/reference]
"""
from __future__ import annotations

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def chunks_to_int(body: str) -> int:
    words = [int(w.strip(), 16) for w in body.split(",") if w.strip()]
    v = 0
    for w in words:  # most significant first
        v = (v << 32) | w
    return v


DECL = re.compile(
    r"const\s+u32\s+(\w+)\s*\[[^\]]*\]\s*=\s*BIG_INTEGER_CHUNKS(?:8|16)\(([^)]*)\);", re.S
)


def parse_decls(text: str) -> dict[str, int]:
    return {m.group(1): chunks_to_int(m.group(2)) for m in DECL.finditer(text)}


def split_namespaces(text: str) -> dict[str, str]:
    out = {}
    for name in ("instance1", "instance2", "instance3"):
        i = text.find(f"namespace {name}")
        if i < 0:
            continue
        j = text.find("namespace instance", i + 10)
        out[name] = text[i : j if j > 0 else len(text)]
    return out


def main(ref: str) -> None:
    mont = open(os.path.join(ref, "core/src/common/mont/tests/main.cu")).read()
    field = {"source": "core/src/common/mont/tests/main.cu", "field": "bn254_fr", "instances": {}}
    for name, body in split_namespaces(mont).items():
        field["instances"][name] = {k: hex(v) for k, v in parse_decls(body).items()}
    conv = mont[mont.find('TEST_CASE("Convert to and from Montgomery")') :]
    conv = conv[: conv.find("TEST_CASE", 10)]
    field["convert"] = {k: hex(v) for k, v in parse_decls(conv).items()}
    with open(os.path.join(HERE, "field_kats.json"), "w") as f:
        json.dump(field, f, indent=1, sort_keys=True)

    curve = open(os.path.join(ref, "core/src/common/curve/tests/bn254.cu")).read()
    pts = {k: hex(v) for k, v in parse_decls(curve).items()}
    with open(os.path.join(HERE, "curve_kats.json"), "w") as f:
        json.dump(
            {
                "source": "core/src/common/curve/tests/bn254.cu",
                "note": "coordinates are Fq Montgomery-form words; pr = p1 + p2, "
                "pr100 = p1 + 100*p2, p1r128 = 128*p1",
                "points": pts,
            },
            f,
            indent=1,
            sort_keys=True,
        )


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
