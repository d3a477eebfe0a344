"""Shared test helpers: package and oracle imports, state comparison."""
import importlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

pkg = importlib.import_module("godot-avbd_amd")
import oracle  # noqa: E402

scenes = pkg.scenes

# The reference's recorded state digests (tools/baseline/digests.txt:1-3).
DIGEST_STACK_300 = 0xFBAC4F08CB1C977C
DIGEST_INTERLEAVED_300 = 0xB03ADEB37742850A
DIGEST_MIXED_240 = 0xFB3CA7A3162035C2

STATE_KEYS = ("position", "rotation", "velocity", "angular_velocity")


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def state_mismatch(g, o):
    """Names of the state fields that differ bit for bit."""
    return [k for k in STATE_KEYS if not bits_equal(g[k], o[k])]


def dt120():
    return float(np.float32(1.0) / np.float32(120.0))


def has_gpu():
    try:
        return pkg.abi.load_library().avbd_device_count() > 0
    except Exception:
        return False
