"""CPU tests of host-side tools: bench.py's one-GPU contract, the trained
network file, and C-ABI argument errors that need no device."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "single" in r.stderr
    assert r.stdout.strip() == ""


def test_trained_network_file_round_trip(tmp_path):
    from bonsai2_amd import learn
    params, budget, lo, hi = learn.load()
    assert params.shape == (1377,) and np.isfinite(params).all()
    assert 0 < lo < hi and budget > 0
    with open(learn.DEFAULT_PATH) as f:
        meta = json.load(f)["meta"]
    # trained on the two Plummer configurations, with a calibration table
    assert [c["name"] for c in meta["configs"]] == ["plummer_2p16", "plummer_2p22"]
    assert meta["loss_last"] < 0.2 * meta["target_var"]
    assert any(r["budget"] == budget and r["p99"] <= meta["p99_target"]
               for r in meta["calibration"])
    p = str(tmp_path / "net.json")
    learn.save(p, params, budget, lo, hi, {"note": "copy"})
    p2, b2, lo2, hi2 = learn.load(p)
    assert np.array_equal(p2, params) and (b2, lo2, hi2) == (budget, lo, hi)


def test_null_context_and_bad_arguments_are_error_codes(bt):
    import ctypes
    L = bt.lib()
    assert L.bt_set_opening(None, 1, 1e-4, 0.25, 1.0) == -1
    assert L.bt_set_quadrupole(None, 0) == -1
    assert L.bt_step_graph(None, 0.01, 16, 5, 64, 1) == -1
    assert L.bt_mlp_forward(None, 1, None, None) == -1
    assert L.bt_learn_collect(None, 1, None, 0.25, 1.0, 1, None, None, None, None, None,
                              None) == -1
    out = (ctypes.c_double * 6)()
    assert L.bt_get_opening_check(None, out) == -1
