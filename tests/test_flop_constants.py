"""The FP64 operation counts the benchmark's roofline uses
(tests/golden/flop_constants.json, written by make_flop_constants.py)
against the oracle's operation-counting scalar.

bench.py prices the tracer's work as (RHS evaluations x rhs + attempts x
attempt_overhead + accepted steps x commit) FLOP; an edit of the oracle's
operation sequence that left the fixture behind would misprice every
roofline figure.  The counts are also the same for any state and spin (the
counted routines have no data-dependent branches).  The fixture counts the
exact right-hand side, the form of the measured workloads; the reference's
as-written dr/dy weight costs one multiplication more (its (u + 4a^2) / 2s
halves 1/s with a product).
"""
import json
import os

import numpy as np

from kerrtrace_amd import config as K
from oracle import oracle as O

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "flop_constants.json")


def test_fixture_matches_the_oracle_count():
    with open(FIXTURE) as fh:
        fc = json.load(fh)
    cfg, _ = K.preset("hd1080_a0998")
    rhs, att, commit = O.flop_counts(cfg, [10.0, 1.0, 2.0, 0.5, 0.1, 0.2], -1.0)
    assert (rhs, att, commit) == (fc["rhs"], fc["attempt_overhead"], fc["commit"])


def test_counts_do_not_depend_on_state_or_spin():
    cfg, _ = K.preset("hd1080_a0998")
    assert cfg.rhs_mode == K.KT_RHS_HAMILTONIAN
    ref = O.flop_counts(cfg, [10.0, 1.0, 2.0, 0.5, 0.1, 0.2], -1.0)
    rng = np.random.default_rng(5)
    for rhs_mode, extra in ((K.KT_RHS_HAMILTONIAN, 0), (K.KT_RHS_REFERENCE, 1)):
        for spin in (0.0, 0.5, -0.9, 0.998):
            c = cfg.copy()
            c.spin = spin
            c.rhs_mode = rhs_mode
            for _ in range(5):
                pos = rng.normal(size=3) * 20.0
                mom = rng.normal(size=3)
                got = O.flop_counts(c, np.concatenate([pos, mom]), -1.0)
                assert got == (ref[0] + extra, ref[1], ref[2]), (spin, rhs_mode, got, ref)
