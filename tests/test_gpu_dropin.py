"""The drop-in C++ boundary on the GPU: tests/cpp/ref_caller (built by
__graft_entry__.build()) follows the reference driver's call sequence
(code/treecode.cu:61-165) on include/Treecode.h; its results must equal, bit
for bit, the same sequence through the Python mirror of the same library."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "ref_caller")


def test_reference_call_sequence(bt, tmp_path):
    assert os.path.exists(EXE), "tests/cpp/ref_caller is not built (run __graft_entry__.build())"
    n, nleaf, ngroup, theta, eps = 1 << 16, 16, 64, 0.4, 0.01
    out = str(tmp_path / "ref_caller.bin")
    print("ref_caller n=%d" % n, flush=True)
    r = subprocess.run([EXE, str(n), str(nleaf), str(ngroup), repr(theta), repr(eps), out],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    raw = np.fromfile(out, np.uint8)
    head = raw[:64].view(np.float64)
    acc_c = raw[64:64 + 16 * n].view(np.float32).reshape(n, 4)
    pos_c = raw[64 + 16 * n:].view(np.float32).reshape(n, 4)

    pos, vel = bt.plummer(n, seed=19810614)
    t = bt.Treecode(eps, theta, 64)
    try:
        t.alloc(n)
        t.ptcl_h2d(pos, vel)
        t.buildTree(nleaf)
        t.computeMultipoles()
        t.makeGroups(5, ngroup)
        st = t.computeForces()
        assert tuple(head[:4]) == st
        assert head[4] == t.grav_potential
        # d_ptclAcc is PH order, as the reference's (code/computeForces.cu:611-617)
        assert np.array_equal(acc_c, t.acc(bt.ORDER_PH))
        # moveParticles then computeEnergies: the energies are evaluated at the
        # moved positions with the step's closing half kick applied, which is
        # one full kick-drift-kick step (bt_step) from the same state
        t.moveParticles(1.0 / 64)
        ek, ep = t.computeEnergies()
        assert (head[5], head[6]) == (ek[0], ep[0])
        assert head[7] == t.grav_potential
        p_tree, _ = t.ptcl_d2h()          # tree order, as the reference's ptcl_d2h
        assert np.array_equal(pos_c, p_tree)
        s = bt.Treecode(eps, theta, 64)
        try:
            s.alloc(n)
            s.ptcl_h2d(pos, vel)
            s.buildTree(nleaf)
            s.computeMultipoles()
            s.makeGroups(5, ngroup)
            s.computeForces()
            s.step(1.0 / 64, nleaf, 5, ngroup)
            ek2, ep2 = s.computeEnergies()
            assert (ek2[0], ep2[0]) == (ek[0], ep[0])
            assert np.array_equal(s.ptcl_d2h()[0], p_tree)
        finally:
            s.close()
    finally:
        t.close()


def test_reference_call_sequence_double(bt, tmp_path):
    """The same caller instantiating Treecode<double>: its FP64 results equal
    the Python mirror's FP64 context bit for bit."""
    n, nleaf, ngroup, theta, eps = 1 << 15, 16, 64, 0.4, 0.01
    out = str(tmp_path / "ref_caller64.bin")
    print("ref_caller f64 n=%d" % n, flush=True)
    r = subprocess.run([EXE, str(n), str(nleaf), str(ngroup), repr(theta), repr(eps), out, "f64"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    raw = np.fromfile(out, np.uint8)
    head = raw[:64].view(np.float64)
    acc_c = raw[64:64 + 32 * n].view(np.float64).reshape(n, 4)
    pos_c = raw[64 + 32 * n:].view(np.float64).reshape(n, 4)
    pos, vel = bt.plummer(n, seed=19810614)
    t = bt.Treecode(eps, theta, 64, fp64=True)
    try:
        t.alloc(n)
        t.ptcl_h2d(pos.astype(np.float64), vel.astype(np.float64))
        t.buildTree(nleaf)
        t.computeMultipoles()
        t.makeGroups(5, ngroup)
        st = t.computeForces()
        assert tuple(head[:4]) == st and head[4] == t.grav_potential
        assert np.array_equal(acc_c, t.acc(bt.ORDER_PH))
        t.moveParticles(1.0 / 64)
        ek, ep = t.computeEnergies()
        assert (head[5], head[6]) == (ek[0], ep[0])
        assert np.array_equal(pos_c, t.ptcl_d2h()[0])
    finally:
        t.close()
