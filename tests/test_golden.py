"""The FP64 oracle against its committed golden frame and step logs
(tests/golden/cpu256_a09_64x64.npz, written by make_oracle_golden.py).

The oracle is pinned against the mathematics elsewhere; this test fixes
its numbers, so a later edit of the oracle that the device tracer mirrors
cannot change both sides unseen.  Equality is bitwise: the oracle is
compiled with -ffp-contract=off and explicit fma, so a rebuild on any
x86-64 host gives the same doubles.
"""
import os

import numpy as np

from oracle import oracle as O
from tests.golden.make_oracle_golden import OUT, make


def test_oracle_reproduces_golden_frame_and_logs():
    assert os.path.exists(OUT)
    with np.load(OUT) as z:
        gold = {k: z[k] for k in z.files}
    now = make()
    assert set(now) == set(gold)
    for k in gold:
        np.testing.assert_array_equal(now[k], gold[k], err_msg=k)
    # the fixture's own consistency: logs end where the rays' step counts
    # say, and the frame holds every end class
    assert int(gold["log_len"].sum()) == len(gold["logs"])
    assert np.array_equal(gold["log_stats"][:, 4], gold["log_len"])
    ends = set((gold["cls"].reshape(-1) & 3).tolist())
    assert ends >= {1, 2}
    assert (gold["cls"] & 4).any() and (gold["cls"] & 8).any()
