"""GPU: the C++ facade (include/avbd_solver.hpp) run by tests/cpp/facade_test.cpp, which
builds the reference's digest scenes with `new Rigid` / `new Joint` / `step()` exactly as the
reference's core tests do and checks the reference's recorded digests, then springs and
collision filters against closed forms, generic joints against the reference's rig checks and
round shapes against the reference's recorded runs."""
import os
import subprocess

import numpy as np
import pytest

from helpers import ROOT, bits_equal, oracle, pkg, scenes

pytestmark = pytest.mark.gpu


def test_facade_program():
    exe = os.path.join(ROOT, "godot-avbd_amd", "bin", "facade_test")
    assert os.path.exists(exe), "build the facade test with godot-avbd_amd/build.sh"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "11 checks, 0 failures" in r.stdout


def _read_loops(path):
    """Parse the facade program's server-loop output (floats in C99 hex)."""
    fx = float.fromhex
    lines = open(path).read().split("\n")
    k = 0
    out = {"manifolds": [], "report": {}}
    m = int(lines[k].split()[1]); k += 1
    for _ in range(m):
        t = lines[k].split(); k += 1
        a, b, cnt = int(t[0]), int(t[1]), int(t[2])
        normal = np.array([fx(x) for x in t[3:6]], np.float32)
        contacts = []
        for _ in range(cnt):
            u = lines[k].split(); k += 1
            contacts.append((int(u[0]), np.array([fx(x) for x in u[1:10]], np.float32)))
        out["manifolds"].append((a, b, normal, contacts))
    t = lines[k].split(); k += 1
    out["deep"] = (int(t[1]), int(t[2]), int(t[3]), np.array([fx(x) for x in t[4:13]], np.float32).reshape(3, 3))
    out["fresh"] = []
    for _ in range(out["deep"][2]):
        u = lines[k].split(); k += 1
        out["fresh"].append((int(u[0]), np.array([fx(x) for x in u[1:7]], np.float32)))
    out["penetration"] = fx(lines[k].split()[1]); k += 1
    while lines[k].startswith("R "):
        _, i, cnt = lines[k].split(); k += 1
        rows = []
        for _ in range(int(cnt)):
            u = lines[k].split(); k += 1
            rows.append((np.array([fx(x) for x in u[0:6]], np.float32), int(u[6])))
        out["report"][int(i)] = rows
    out["times"] = [float(x) for x in lines[k].split()[1:]]
    return out


def test_server_loops_through_the_facade(tmp_path):
    """The Godot server's deepest-contact re-measure and contact collection loops
    (avbd_physics_server3d.cpp:1547-1585, 1620-1652), written against the facade as the server
    writes them, on a mixed pile after 60 steps: every manifold accessor value, the re-measured
    contacts of Manifold::collide, and each body's reported contacts equal the oracle's."""
    b, j = scenes.round_pile(160, seed=21, pitch=2.2)
    n = len(b["mass"])
    steps = 60
    rep = (np.arange(n) % 4).astype(np.int32)            # max_contacts_reported, in creation order
    shape = b.get("shape", np.zeros(n, np.uint8))
    scene = tmp_path / "scene.txt"
    with open(scene, "w") as f:
        f.write(f"{n} {steps}\n")
        for c in range(n):                               # creation order = reversed list order
            i = n - 1 - c
            vals = [*b["size"][i], b["mass"][i], *b["moment"][i], b["friction"][i], b["radius"][i],
                    *b["position"][i], *b["rotation"][i], *b["velocity"][i], *b["angular_velocity"][i]]
            f.write(f"{int(shape[i])} " + " ".join(repr(float(np.float32(v))) for v in vals) + f" {rep[c]}\n")
    out = tmp_path / "loops.txt"
    exe = os.path.join(ROOT, "godot-avbd_amd", "bin", "facade_test")
    r = subprocess.run([exe, "--server-loops", str(scene), str(out)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    got = _read_loops(out)

    o = oracle.OracleWorld(pkg.default_params())
    o.upload(b, j)
    o.step(steps)
    om = o.manifolds()
    assert len(got["manifolds"]) == len(om["body_a"]) > 50
    for k, (a, bb, normal, contacts) in enumerate(got["manifolds"]):
        assert (a, bb, len(contacts)) == (om["body_a"][k], om["body_b"][k], om["num_contacts"][k])
        assert bits_equal(normal, om["basis"][k].reshape(3, 3)[0])
        for i, (key, v) in enumerate(contacts):
            assert key == om["feature"][k][i]
            want = np.concatenate([om["c0"][k][3 * i:3 * i + 3], om["r_a"][k][3 * i:3 * i + 3],
                                   om["r_b"][k][3 * i:3 * i + 3]])
            assert bits_equal(v, want), (k, i)
    # deepest stored contact, re-measured at the current pose
    depth = [(np.float32(0.001) - om["c0"][k][3 * i], k) for k in range(len(om["body_a"]))
             for i in range(om["num_contacts"][k])]
    deepest = max((float(d), -k) for d, k in depth if float(d) > 0.0)
    kd = -deepest[1]
    a, bb, cnt, basis = got["deep"]
    assert (a, bb) == (om["body_a"][kd], om["body_b"][kd])
    wn, wkeys, wra, wrb, wbasis = o.collide_pair(a, bb, pkg.default_params().dt)
    assert cnt == wn > 0 and bits_equal(basis, wbasis)
    for i, (key, v) in enumerate(got["fresh"]):
        assert key == wkeys[i] and bits_equal(v, np.concatenate([wra[i], wrb[i]]))
    # each body's reported contacts (the oracle's walk is the server's, pinned separately)
    rep_list = rep[::-1].copy()                          # per list index
    cnt_o, pos_o, nrm_o, oth_o = o.contact_report(rep_list)
    for i in range(n):
        rows = got["report"].get(i, [])
        assert len(rows) == cnt_o[i], i
        for c, (v, other) in enumerate(rows):
            assert other == oth_o[i][c]
            assert np.allclose(v[:3], pos_o[i][c], atol=1e-6) and bits_equal(v[3:], nrm_o[i][c]), (i, c)
    threads, *times = got["times"]
    assert threads == 1 and all(t >= 0.0 for t in times) and times[0] > 0.0
