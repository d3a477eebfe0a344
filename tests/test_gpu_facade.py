"""GPU: the C++ facade (include/avbd_solver.hpp) run by tests/cpp/facade_test.cpp, which
builds the reference's digest scenes with `new Rigid` / `new Joint` / `step()` exactly as the
reference's core tests do and checks the reference's recorded digests, then springs and
collision filters against closed forms, generic joints against the reference's rig checks and
round shapes against the reference's recorded runs."""
import os
import subprocess

import pytest

from helpers import ROOT

pytestmark = pytest.mark.gpu


def test_facade_program():
    exe = os.path.join(ROOT, "godot-avbd_amd", "bin", "facade_test")
    assert os.path.exists(exe), "build the facade test with godot-avbd_amd/build.sh"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "11 checks, 0 failures" in r.stdout
