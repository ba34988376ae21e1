"""HCI variational stage on the GPU library: the selection loop of
risq_amd/hci_select.py with elements, diagonals, incremental FULL-storage H
builds and warm-started Davidson (reference control flow) from the HIP library
reproduces the reference's integration-test energies
(tests/integration_test_suite.rs:30-89, tolerance 1e-6 Ha at :13)."""
import json
import os
import time

import pytest

from risq_amd import hci_select as hs
from risq_amd.fcidump import read_fcidump

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = json.load(open(os.path.join(GOLD, "hci_cases.json")))


@pytest.mark.timeout(300)
@pytest.mark.parametrize("name", ["he2", "be", "li2", "be2", "f2"])
def test_reference_energies_on_gpu(gpu_ctx, name):
    c = CASES[name]
    ints = read_fcidump(os.path.join(GOLD, c["fcidump"]), c["norb"])
    solver = hs.GpuSolver(gpu_ctx, ints)
    t = time.time()
    r = hs.variational(solver, ints, c["nup"], c["ndn"], c["eps_var"], n_core=c["norb_core"])
    solver.close()
    print(f"{name}: E={r.energy:.10f} (reference {c['energy']:.10f}, diff {r.energy - c['energy']:.1e}) "
          f"dets={len(r.dets)} iterations={r.iterations} {time.time() - t:.1f}s")
    assert r.converged
    assert abs(r.energy - c["energy"]) < 1e-6
