"""GPU tests of the drop-in pieces around the step: the command-line driver
(the reference's code/treecode.cu flags) and checkpoint / resume through
tipsy snapshots (a run split by a snapshot equals the run in one piece)."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "bonsai2_amd", "treecode")


def test_driver_plummer_steps_and_snapshot(bt, tmp_path):
    out = str(tmp_path / "final.tipsy")
    r = subprocess.run([DRIVER, "-n", "65536", "-l", "16", "-g", "64", "-o", "0.4", "-e", "0.01",
                        "-t", "4", "-d", "0.0078125", "-w", out],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    text = r.stdout + r.stderr
    assert "Using Plummer model with nPtcl= 65536" in text
    m = re.search(r"\|dE/E\|= ([0-9.eE+-]+)", text)
    assert m and float(m.group(1)) < 1e-4, text
    pos, vel = bt.read_tipsy(out)
    assert pos.shape == (65536, 4) and np.isfinite(pos).all() and np.isfinite(vel).all()


def test_driver_rejects_bad_nleaf():
    r = subprocess.run([DRIVER, "-n", "1024", "-l", "17"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode != 0 and "nleaf" in r.stderr


def test_resume_from_snapshot_equals_continuous_run(bt, tmp_path):
    n, dt = 40000, 1.0 / 128
    pos, vel = bt.plummer(n, seed=31)

    def fresh(p, v):
        t = bt.Treecode(0.01, 0.5, 64)
        t.alloc(n)
        t.ptcl_h2d(p, v)
        t.buildTree(16)
        t.computeMultipoles()
        t.makeGroups(5, 64)
        t.computeForces()
        return t

    a = fresh(pos, vel)
    for _ in range(6):
        a.step(dt, 16, 5, 64)
    pa, va = a.ptcl_d2h(bt.ORDER_INPUT)
    a.close()

    b = fresh(pos, vel)
    for _ in range(3):
        b.step(dt, 16, 5, 64)
    snap = str(tmp_path / "mid.tipsy")
    pb, vb = b.ptcl_d2h(bt.ORDER_INPUT)
    bt.write_tipsy(snap, pb, vb, time=3 * dt, eps=0.01)
    b.close()
    p2, v2 = bt.read_tipsy(snap)
    v2[:, 3] = p2[:, 3]
    c = fresh(p2, v2)
    for _ in range(3):
        c.step(dt, 16, 5, 64)
    pc, vc = c.ptcl_d2h(bt.ORDER_INPUT)
    c.close()
    # forces are recomputed from the restored positions, so the split run is
    # the continuous run bit for bit
    assert np.array_equal(pa, pc) and np.array_equal(va[:, :3], vc[:, :3])
