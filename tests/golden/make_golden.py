"""Writes tests/golden/vectors.json.

Expected values come from CPython's ``bytes.replace`` -- the reference's stated
contract (cuda_replace.cu:2-3) -- on the reference's own generated cases
(test/test.py:44-99), its README and wrapper self-test outputs, regression
inputs, and the seeded benchmark inputs.  The reference library itself cannot
run here (it needs nvcc and an NVIDIA GPU).

Run: ``python tests/golden/make_golden.py`` (about a minute; writes ~10 KB).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from cases import KNOWN_ANSWERS, REFERENCE_CASES, REGRESSIONS  # noqa: E402
from cuda_replace_amd import workloads  # noqa: E402


def sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


def main() -> None:
    out = {"python": sys.version.split()[0], "known_answers": [], "reference_cases": [],
           "regressions": [], "seeded": [], "configs": []}
    for name, src, pat, rep, expected in KNOWN_ANSWERS:
        assert src.replace(pat, rep) == expected, name
        out["known_answers"].append({"name": name, "src": src.hex(), "pat": pat.hex(), "rep": rep.hex(),
                                     "out": expected.hex()})
    for name, make, pat, rep in REFERENCE_CASES:
        data = make()
        res = data.replace(pat, rep)
        out["reference_cases"].append({"name": name, "n_in": len(data), "sha_in": sha(data),
                                       "pat": pat.hex(), "rep": rep.hex(), "n_out": len(res),
                                       "sha_out": sha(res), "count": data.count(pat)})
    for name, src, pat, rep in REGRESSIONS:
        res = src.replace(pat, rep)
        rec = {"name": name, "pat": pat.hex(), "rep": rep.hex(), "n_in": len(src), "n_out": len(res),
               "sha_out": sha(res)}
        if len(src) <= 256:          # large inputs are regenerated from tests/cases.py
            rec["src"] = src.hex()
            rec["out"] = res.hex()
        out["regressions"].append(rec)
    # seeded applies: out = S[:F] + S[F:].replace(P, R), F = max(0, seed + L)
    seeded = [
        (b"aaaaaaaa", b"aa", b"X", 0), (b"aaaaaaaa", b"aa", b"X", 1), (b"aaaaaaaa", b"aa", b"X", -1),
        (b"abababab", b"aba", b"Z", 0), (b"abababab", b"aba", b"Z", 3), (b"xxabxxab", b"ab", b"QQQ", 2),
        (b"aaaa", b"aa", b"X", -2), (b"aaaa", b"aa", b"X", 5),
    ]
    for src, pat, rep, seed in seeded:
        f = max(0, seed + len(pat))
        tail = src[f:]
        res = src[:f] + tail.replace(pat, rep)
        ks, i = [], tail.find(pat)
        while i != -1:
            ks.append(f + i)
            i = tail.find(pat, i + len(pat))
        out["seeded"].append({"src": src.hex(), "pat": pat.hex(), "rep": rep.hex(), "seed": seed,
                              "out": res.hex(), "last": ks[-1] if ks else -1, "kept": ks})
    # benchmark inputs (seeded): input and output digests
    a = workloads.uniform_abcd(64 << 20, 0)
    r = a.replace(b"ab", b"X")
    out["configs"].append({"name": "uniform_abcd_64mib_ab_X", "n_in": len(a), "sha_in": sha(a),
                           "pat": b"ab".hex(), "rep": b"X".hex(), "n_out": len(r), "sha_out": sha(r)})
    text, frag, rep_e, rep_s = workloads.text_config(64 << 20)
    for tag, rep in (("expand", rep_e), ("shrink", rep_s)):
        r = text.replace(frag, rep)
        out["configs"].append({"name": f"zipf_text_64mib_{tag}", "n_in": len(text), "sha_in": sha(text),
                               "pat": frag.hex(), "rep": rep.hex(), "n_out": len(r), "sha_out": sha(r)})
    with open(os.path.join(HERE, "vectors.json"), "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main()
