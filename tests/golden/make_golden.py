"""Writes tests/golden/ldc32_oracle.npz: the 32x32 lid-driven cavity
(tau 0.6, dt 0.01, lid u = 0.1, upwind) advanced by the CPU oracle, cell
pdfs after 1, 10, 100 and 1000 FP32 steps and after 1000 FP64 steps. The
reference ships no golden vectors (its notebooks have no stored outputs), so
these are regression anchors for the oracle and the GPU path, not an
independent pin; the pins are in tests/test_oracle.py.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from cases import ldc_case, oracle_env, oracle_state  # noqa: E402
from oracle import oracle  # noqa: E402


def main():
    m, dyn, cells, faces, nodes = ldc_case()
    out = {}
    env = oracle_env(cells, faces, nodes, "f32")
    st = oracle_state(env, nodes, "f32")
    done = 0
    for n in (1, 10, 100, 1000):
        oracle.run(env, st, dyn.tau, dyn.delta_t, n - done, "f32")
        done = n
        out[f"pdf_f32_{n}"] = st["cell_pdf"].copy()
    env64 = oracle_env(cells, faces, nodes, "f64")
    st64 = oracle_state(env64, nodes, "f64")
    oracle.run(env64, st64, dyn.tau, dyn.delta_t, 1000, "f64")
    out["pdf_f64_1000"] = st64["cell_pdf"].copy()
    np.savez_compressed(os.path.join(HERE, "ldc32_oracle.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
