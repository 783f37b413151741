"""Readers of the reference fixtures at the benchmark configurations' own
sizes (tests/golden/baseline, written by tests/golden/make_baseline.py):
compact encodings decoded into the layout of the small fixtures, so the
same system builders apply."""
import base64
import json
import os
import zlib

import numpy as np

from oracle_py import GOLDEN

BASE = os.path.join(GOLDEN, "baseline")


def dec(e):
    """a base64 array ({dtype, shape, b64}) or packed flags ({bits, b64})"""
    raw = base64.b64decode(e["b64"])
    if e.get("zlib"):
        raw = zlib.decompress(raw)
    if "bits" in e:
        return np.unpackbits(np.frombuffer(raw, dtype=np.uint8))[:e["bits"]].astype(np.int32)
    return np.frombuffer(raw, dtype=np.dtype(e["dtype"])).reshape(e["shape"]).copy()


def mt_text_words(txt):
    """libstdc++ mt19937 text state -> (624 words, position)"""
    v = [int(t) for t in txt.split()]
    return np.array(v[:624], dtype=np.uint32), v[624]


def load(name):
    with open(os.path.join(BASE, name + ".json")) as f:
        fx = json.load(f)
    p = fx["particles"]
    xyz = dec(p["xyz_q16"]).astype(np.float64) * p["unit"]
    ids = dec(p["id"]).astype(np.int64)
    if "radius_steps" in p:
        radius = p["radius_base"] + dec(p["radius_steps"]).astype(np.float64) * p["radius_step"]
    else:
        radius = np.array(p["radius_by_id"])[ids]
    fx["particles"] = {"xyz": xyz.reshape(-1), "charge": np.array(p["charge_by_id"])[ids],
                       "radius": radius, "mw": np.array(p["mw_by_id"])[ids],
                       "id": ids.astype(np.int32)}
    e = fx["energy"]
    for k in ("g2g", "i2i"):
        e[k] = dec(e[k])
    return fx
