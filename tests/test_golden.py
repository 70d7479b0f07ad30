"""Golden fixtures (tests/golden/mt_np16_elastic.npz, written by tests/golden/make_golden.py).

CPU: the oracle, recomputed from scratch, reproduces the committed vectors.
GPU (test_gpu_parity.py): the library reproduces them through the C-ABI.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402

FIX = os.path.join(HERE, "golden", "mt_np16_elastic.npz")


def load():
    with np.load(FIX) as d:
        return {k: d[k] for k in d.files}


def test_oracle_reproduces_golden_fixture():
    g = load()
    c = make_golden.compute()
    assert float(c["E_bound"]) == float(g["E_bound"])
    assert np.allclose(c["e_SWP_unco"], g["e_SWP_unco"], rtol=1e-13, atol=0)
    assert np.array_equal(c["q_idx"], g["q_idx"]) and np.array_equal(c["deut"], g["deut"])
    assert np.abs(c["G"] - g["G"]).max() <= 1e-13 * np.abs(g["G"]).max()
    assert np.array_equal(c["p123_colptr"], g["p123_colptr"])
    assert np.array_equal(c["p123_row"], g["p123_row"])
    assert np.abs(c["p123_val_sample"] - g["p123_val_sample"]).max() <= 1e-14 * np.abs(g["p123_val_sample"]).max()
    assert abs(float(c["p123_abs_sum"]) - float(g["p123_abs_sum"])) <= 1e-13 * float(g["p123_abs_sum"])
    assert np.abs(c["U"] - g["U"]).max() <= 1e-12 * np.abs(g["U"]).max()
    assert np.abs(c["neumann"] - g["neumann"]).max() <= 1e-12 * np.abs(g["neumann"]).max()
