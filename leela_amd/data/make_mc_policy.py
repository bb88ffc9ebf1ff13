"""Writes leela_amd/data/mc_policy.bin: the tables of the reference's
Monte-Carlo playout policy, read from its headers, for the library's pattern
playouts (leela_amd/csrc/mc_policy.cpp).

    python leela_amd/data/make_mc_policy.py [REFERENCE_DIR]

Tables (the reference's data; the playouts are defined by them):
  Patterns.h       Matcher::Patterns, the 181,203 valid 3x3 patterns with
                   liberty bits (int)
  PatternHash.h    Matcher::G, the 8,192 displacements of the minimal
                   perfect hash (int)
  PatternsLive.h   PolicyWeights::live_patterns, the 23,518 canonical
                   patterns that carry a weight (int)
  PolicyWeights.h  PolicyWeights::feature_weights (22) and pattern_weights
                   (23,518) (float)
  PolicyWeightsSL.h  the same for the SL-trained gammas (float)

File layout, little-endian: the magic "LZMCPOL1", then five int32 counts
(patterns, G, live, features, pattern weights), then the arrays in the order
above (Patterns, G, live_patterns as int32; feature_weights,
feature_weights_sl, pattern_weights, pattern_weights_sl as float32).
"""
from __future__ import annotations

import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
OUT = HERE / "mc_policy.bin"


def _array_body(text: str, name: str) -> str:
    m = re.search(re.escape(name) + r"\s*\{(.*?)\}\s*;", text, re.S)
    if not m:
        raise ValueError(f"{name} not found")
    # Drop comments.
    return re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)


def ints(text: str, name: str) -> np.ndarray:
    body = _array_body(text, name)
    return np.array([int(t) for t in re.findall(r"-?\d+", body)], np.int32)


def floats(text: str, name: str) -> np.ndarray:
    body = _array_body(text, name)
    vals = re.findall(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?", body)
    return np.array([float(v) for v in vals], np.float64).astype(np.float32)


def main(ref: Path) -> int:
    rd = lambda f: (ref / f).read_text()
    patterns = ints(rd("Patterns.h"), "Matcher::Patterns")
    g = ints(rd("PatternHash.h"), "Matcher::G")
    live = ints(rd("PatternsLive.h"), "PolicyWeights::live_patterns")
    pw = rd("PolicyWeights.h")
    sl = rd("PolicyWeightsSL.h")
    fw = floats(pw, "PolicyWeights::feature_weights")
    pat_w = floats(pw, "PolicyWeights::pattern_weights")
    fw_sl = floats(sl, "PolicyWeights::feature_weights_sl")
    pat_w_sl = floats(sl, "PolicyWeights::pattern_weights_sl")
    assert len(patterns) == 181203 and len(g) == 8192, (len(patterns), len(g))
    assert len(live) == 23518 and len(pat_w) == 23518 == len(pat_w_sl)
    assert len(fw) == 22 == len(fw_sl)
    with open(OUT, "wb") as f:
        f.write(b"LZMCPOL1")
        f.write(np.array([len(patterns), len(g), len(live), len(fw),
                          len(pat_w)], "<i4").tobytes())
        for a in (patterns, g, live):
            f.write(a.astype("<i4").tobytes())
        for a in (fw, fw_sl, pat_w, pat_w_sl):
            f.write(a.astype("<f4").tobytes())
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")
    return 0


if __name__ == "__main__":
    sys.exit(main(Path(sys.argv[1] if len(sys.argv) > 1
                       else "/root/reference")))
