"""Writes flop_constants.json: exact FP64 operation counts of the tracer's
units of work, measured by running the oracle's restated routines on its
operation-counting scalar (FMA = 2 FLOP; add, sub, mul, div, sqrt = 1).

  rhs     one evaluation of the Hamiltonian right-hand side
  attempt one DP5(4) attempt without its right-hand sides: six stage
          inputs, the 5th-order solution, the error norm, the controller
  commit  per committed step: radius + potential-scaled tolerance of the
          next step and the radius test after the step
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from kerrtrace_amd import config as K  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    cfg, _ = K.preset("hd1080_a0998")
    rhs, att, commit = O.flop_counts(cfg, [10.0, 1.0, 2.0, 0.5, 0.1, 0.2], -1.0)
    out = {"rhs": rhs, "attempt_overhead": att, "commit": commit,
           "note": "FP64 operations; FMA counts 2; transcendentals excluded"}
    with open(os.path.join(HERE, "flop_constants.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(out)


if __name__ == "__main__":
    main()
