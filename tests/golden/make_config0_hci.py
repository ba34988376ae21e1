"""Writes config0_hci.json: the HCI variational stage of the n_orb=16, 4a4b
synthetic system (seed 0, K=32 Cholesky vectors) at eps_var = 3e-3, run by
risq_amd/hci_select.py on the CPU oracle (elements, incremental builds and the
reference-flow Davidson of oracle/).  The GPU test compares its own run with it.

    python tests/golden/make_config0_hci.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

from risq_amd import hci_select as hs  # noqa: E402
from risq_amd.fcidump import synthetic_molecular  # noqa: E402
from test_hci_select import OracleSolver  # noqa: E402


def main():
    n_orb, k, seed, nup, ndn, eps_var = 16, 32, 0, 4, 4, 3e-3
    ints = synthetic_molecular(n_orb, k, seed)
    t = time.time()
    r = hs.variational(OracleSolver(ints), ints, nup, ndn, eps_var)
    words = hs._words(r.dets)
    out = {"n_orb": n_orb, "n_chol": k, "seed": seed, "nup": nup, "ndn": ndn, "eps_var": eps_var,
           "energy": r.energy, "n_dets": len(r.dets), "iterations": r.iterations, "converged": r.converged,
           "history": [[e, n, en] for e, n, en in r.history],
           "dets_checksum": int(np.bitwise_xor.reduce(words[:, 0] * np.uint64(0x9E3779B97F4A7C15)
                                                      ^ words[:, 2]))}
    with open(os.path.join(HERE, "config0_hci.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"E={r.energy:.10f} dets={len(r.dets)} ({time.time() - t:.1f}s)")


if __name__ == "__main__":
    main()
