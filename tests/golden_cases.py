"""Loads the committed golden fixtures (tests/golden/*.json) as typed cases.

The fixtures are data written by tests/golden/make_golden.py: inputs plus
expected outputs of extend / extend_hint / run_dispatch (extend.rs:293-428)
and of EA-Myers (myers_ea.rs:118-185).
"""
from __future__ import annotations

import json
import os

import numpy as np

from esperanto_amd.records import EditDir, ExtendParams

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _b(s: str) -> bytes:
    return s.encode("latin-1")


def extend_cases():
    with open(os.path.join(HERE, "extend_cases.json")) as f:
        raw = json.load(f)
    for c in raw:
        c["read"], c["window"] = _b(c["read"]), _b(c["window"])
        p = c["params"]
        c["prm"] = ExtendParams(p[0], p[1], p[2], p[3], EditDir(c["dir"]))
    return raw


def myers_cases():
    with open(os.path.join(HERE, "myers_cases.json")) as f:
        raw = json.load(f)
    for c in raw:
        c["read"], c["text"] = _b(c["read"]), _b(c["text"])
    return raw


def groups(cases):
    """Cases grouped by (mode, params, dir): one batch call per group."""
    out = {}
    for c in cases:
        out.setdefault((c["mode"], tuple(c["params"]), c["dir"]), []).append(c)
    return out


def check_batch(cases, ext, touch, cig, cig_off, cig_n, aux, label):
    for k, c in enumerate(cases):
        ctx = (label, k, c["mode"], c["dir"], c["params"], c["offset"], c["half"],
               len(c["read"]), len(c["window"]))
        assert [int(x) for x in ext[k]] == c["ext"], ctx
        a = int(cig_off[k])
        assert [int(x) for x in cig[a:a + int(cig_n[k])]] == c["cigar"], ctx
        if c["mode"] == 2:
            assert int(aux[k]) == c["path"], ctx
        else:
            assert [int(x) for x in touch[k]] == c["touch"], ctx


def batch_of(cases):
    from esperanto_amd.batch import PairBatch
    return PairBatch.from_pairs([(c["read"], c["window"], c["offset"], c["half"]) for c in cases])


def myers_batch_of(cases):
    from esperanto_amd.batch import PairBatch
    return PairBatch.from_pairs([(c["read"], c["text"], 0, 0) for c in cases])


def myers_expected(cases):
    return (np.array([c["dist"] for c in cases]), np.array([c["end"] for c in cases]),
            np.array([c["start"] for c in cases]))
