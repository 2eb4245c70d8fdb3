"""The oracle against the reference's committed golden waveforms (spice21/resources/*.json).

The reference compares at 1e-6 absolute (assert.rs:124-143); the oracle reproduces every sample
bit for bit, which pins the Newton loop, the per-iteration Markowitz re-pivoting, the LU, Mos1/Mos0,
R/C/V stamps, initial-condition forcing and Backward-Euler integration together.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle_py as O
from tests.circuits import GOLDEN_RUNS, ro3

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("fixture", sorted(GOLDEN_RUNS))
def test_oracle_matches_golden_bit_exact(fixture):
    kind, tstep, tstop = GOLDEN_RUNS[fixture]
    g = json.load(open(os.path.join(GOLDEN, fixture)))
    cn = ro3(kind).to_c()
    t, X, names, _ = O.tran(cn, tstep, tstop, [("1", 0.0)])
    assert sorted(names + ["time"]) == sorted(g.keys())
    assert np.array_equal(t, np.array(g["time"]))
    for k, nm in enumerate(names):
        ref = np.array(g[nm])
        assert X.shape[0] == len(ref)
        assert np.max(np.abs(X[:, k] - ref)) <= 1e-6  # the reference's own tolerance
        assert np.array_equal(X[:, k], ref), f"{fixture}:{nm} not bit-exact"


def test_planted_mistake_is_caught():
    """A wrong IC value (a plausible bug) moves the waveform far beyond the 1e-6 tolerance."""
    g = json.load(open(os.path.join(GOLDEN, "test_mos1_cmos_ro_tran.json")))
    cn = ro3("cmos").to_c()
    t, X, names, _ = O.tran(cn, 1e-11, 1e-8, [("1", 0.1)])
    k = names.index("1")
    assert np.max(np.abs(X[:, k] - np.array(g["1"]))) > 1e-3
