"""MI355X-native AVBD rigid-body step: Python handles over the C ABI of libavbd_hip.so.

The package is importable as `importlib.import_module("godot-avbd_amd")`.  The compute
path is the HIP library only; importing the handles loads it and fails loudly when it has
not been built.
"""
from . import abi, scenes  # noqa: F401
from .world import Batch, World, state_digest  # noqa: F401

default_params = abi.default_params
