"""CPU tests of the product's host side: the C ABI library loads and exports
every declared entry point, host-only entry points (initial conditions,
tipsy I/O) behave, and errors come back as codes, never aborts."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "bonsai2_amd.h")).read()
    return sorted(set(re.findall(r"\b(bt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(bt):
    L = bt.lib()
    decl = declared_symbols()
    assert len(decl) >= 30
    missing = [s for s in decl if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(decl) == sorted(bt.EXPORTED_SYMBOLS)


def test_no_device_is_an_error_code_not_an_abort(bt):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    h = ctypes.c_void_p()
    rc = bt.lib().bt_create(ctypes.byref(h), 0.01, 0.5, 64)
    assert rc == -6 and not h.value  # BT_ERR_NODEVICE


def test_plummer_sampler_statistics(bt, oracle):
    n = 8192
    pos, vel = bt.plummer(n, seed=11, rmax=22.8)
    assert pos.dtype == np.float32 and pos.shape == (n, 4)
    assert np.allclose(pos[:, 3], 1.0 / n) and np.array_equal(pos[:, 3], vel[:, 3])
    # centre of mass and momentum removed (FP64, then cast)
    assert np.abs(pos[:, :3].astype(np.float64).mean(0)).max() < 1e-6
    assert np.abs(vel[:, :3].astype(np.float64).mean(0)).max() < 1e-6
    # cumulative mass profile M(<r) = r^3/(1+r^2)^1.5 in structural units
    r = np.linalg.norm(pos[:, :3], axis=1) / (3 * np.pi / 16)
    for rr in (0.5, 1.0, 2.0, 5.0):
        m_exact = rr ** 3 / (1 + rr ** 2) ** 1.5 / (22.8 ** 3 / (1 + 22.8 ** 2) ** 1.5)
        assert abs((r < rr).mean() - m_exact) < 4 * np.sqrt(m_exact * (1 - m_exact) / n)
    assert r.max() < 22.8
    # Henon units: E = -1/4, virial 2K/|W| = 1 (FP64 direct sum, tiny softening)
    k = 0.5 * (vel[:, :3].astype(np.float64) ** 2).sum() / n
    phi = oracle.direct_fp64(pos, np.arange(n, dtype=np.int32), 1e-4)[:, 3]
    w = 0.5 * (phi + (1.0 / n) / 1e-4).sum() / n
    assert abs(k + w + 0.25) < 0.02 and abs(2 * k / abs(w) - 1) < 0.05


def test_samplers_are_deterministic_and_seeded(bt):
    a = bt.plummer(3000, seed=1)
    b = bt.plummer(3000, seed=1)
    c = bt.plummer(3000, seed=2)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])
    # counter-based: the first particles do not depend on n (before COM removal)
    x = bt.plummer(100, seed=1)[0][:, :3].astype(np.float64)
    y = bt.plummer(200, seed=1)[0][:100, :3].astype(np.float64)
    assert np.allclose(x - x.mean(0), y - y.mean(0), atol=2e-2)


def test_clustered_sampler(bt):
    n = 20000
    pos, vel = bt.clustered(n, nclust=64, seed=3)
    assert np.isclose(pos[:, 3].astype(np.float64).sum(), 1.0, rtol=1e-5)
    assert np.abs(pos[:, :3]).max() < 10.0 + 2.0
    # 64 compact clumps: most particles sit far from the global centre
    assert np.median(np.linalg.norm(pos[:, :3], axis=1)) > 2.0


def test_halo_disk_sampler(bt):
    n = 20000
    pos, vel = bt.halo_disk(n, seed=4, mh=0.8, rd=0.3, z0=0.03)
    disk = np.arange(n) >= int(round(0.8 * n))
    z = np.abs(pos[disk, 2] - pos[:, 2].mean())
    assert np.median(z) < 0.05                          # sech^2 height 0.03
    R = np.linalg.norm(pos[disk, :2] - pos[:, :2].mean(0), axis=1)
    assert 0.3 < np.median(R) < 0.7                     # Gamma(2, 0.3): median ~0.50
    # disk velocities are tangential
    vr = (pos[disk, 0] * vel[disk, 0] + pos[disk, 1] * vel[disk, 1]) / np.maximum(R, 1e-6)
    vt = np.linalg.norm(vel[disk, :2], axis=1)
    assert np.median(np.abs(vr) / vt) < 0.05


def test_ic_bad_arguments_are_errors(bt):
    L = bt.lib()
    z = np.zeros((4, 4), np.float32)
    p = z.ctypes.data_as(ctypes.c_void_p)
    assert L.bt_ic_plummer(0, 1, 22.8, p, p) == -1
    assert L.bt_ic_plummer(4, 1, -1.0, p, p) == -1
    assert L.bt_ic_clustered(4, 0, 1, 10.0, 1e-3, 0.5, p, p) == -1
    assert L.bt_ic_halo_disk(4, 1, 1.5, 0.3, 0.03, p, p) == -1


def test_tipsy_round_trip(bt, tmp_path):
    pos, vel = bt.plummer(1000, seed=7)
    f = str(tmp_path / "snap.tipsy")
    bt.write_tipsy(f, pos, vel, time=0.25, eps=0.01)
    assert os.path.getsize(f) == 32 + 36 * 1000
    p2, v2 = bt.read_tipsy(f)
    assert np.array_equal(p2, pos)
    assert np.array_equal(v2[:, :3], vel[:, :3]) and np.array_equal(v2[:, 3], pos[:, 3])


def test_tipsy_written_file_reads_with_reference_reader(bt, oracle, tmp_path):
    """code/read_tipsy.h (compiled into oracle/_ref) reads our snapshot: pos.w =
    mass, vel.w = eps (:117-124)."""
    R = oracle.ref_host()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    pos, vel = bt.plummer(500, seed=9)
    f = str(tmp_path / "snap.tipsy")
    bt.write_tipsy(f, pos, vel, eps=0.02)
    rp = np.zeros((500, 4), np.float32)
    rv = np.zeros((500, 4), np.float32)
    ids = np.zeros(500, np.int32)
    n = R.ref_read_tipsy(f.encode(), 500, oracle._ptr(rp), oracle._ptr(rv), oracle._ptr(ids))
    assert n == 500
    assert np.array_equal(rp, pos)
    assert np.array_equal(rv[:, :3], vel[:, :3]) and np.allclose(rv[:, 3], 0.02)
    assert np.array_equal(ids, np.arange(500))


def test_tipsy_errors(bt, tmp_path):
    with pytest.raises(bt.BonsaiError):
        bt.read_tipsy(str(tmp_path / "missing.tipsy"))
    pos, vel = bt.plummer(100, seed=1)
    f = str(tmp_path / "t.tipsy")
    bt.write_tipsy(f, pos, vel)
    with open(f, "rb") as fh:
        data = fh.read()
    g = str(tmp_path / "trunc.tipsy")
    with open(g, "wb") as fh:
        fh.write(data[:-10])
    with pytest.raises(bt.BonsaiError):
        bt.read_tipsy(g)


def test_reference_style_caller_compiles_against_treecode_header(tmp_path):
    """tests/cpp/ref_caller.cpp follows code/treecode.cu:61-165 (double4
    statistics, Tree::Particle, h_ptclPos/h_ptclVel, d_ptclAcc); it compiles
    with a plain C++11 host compiler and the flags `hipconfig --cpp_config`
    prints, and links against the library."""
    import subprocess
    from bonsai2_amd import build as bbuild
    exe = str(tmp_path / "ref_caller")
    cmd = (["g++", "-O1", "-std=c++11", "-Wall", "-Werror"] + bbuild.HIP_HOST_FLAGS +
           ["-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp",
                                                               "ref_caller.cpp"),
            "-o", exe, "-L" + os.path.join(ROOT, "bonsai2_amd"), "-lbonsai2_amd"])
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    # the header's double4 is HIP's vector type
    src = '#include "Treecode.h"\nstatic_assert(sizeof(double4) == 32, "");\n' \
          'double4 f(Treecode<float> &t) { return t.computeForces(); }\n'
    p = tmp_path / "t.cpp"
    p.write_text(src)
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only"] + bbuild.HIP_HOST_FLAGS +
                       ["-I" + os.path.join(ROOT, "include"), str(p)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
