"""Determinant-list helpers and the deterministic space generator."""
import numpy as np

from risq_amd.dets import energy_truncated_space, from_orbitals, full_space, to_int_pairs


def test_words_layout():
    w = from_orbitals([([0, 2, 70], [1, 127])])
    assert w.tolist() == [[5, 1 << 6, 2, 1 << 63]]
    assert to_int_pairs(w) == [(5 | (1 << 70), 2 | (1 << 127))]


def test_energy_space_exact_size_and_order():
    eps = -1.0 + 2.0 * np.arange(16) / 15.0
    d = energy_truncated_space(eps, 4, 4, 3000)
    assert d.shape == (3000, 4)
    assert len({tuple(r) for r in d.tolist()}) == 3000
    pairs = to_int_pairs(d)
    assert all(bin(u).count("1") == 4 and bin(v).count("1") == 4 for u, v in pairs)
    assert pairs[0] == (0b1111, 0b1111)  # the HF determinant first (src/wf/mod.rs:1697-1698)
    e = [sum(eps[k] for k in range(16) if (u >> k) & 1) + sum(eps[k] for k in range(16) if (v >> k) & 1)
         for u, v in pairs]
    assert np.all(np.diff(e) >= -1e-12)


def test_energy_space_two_words():
    eps = np.linspace(-1, 3, 70)
    d = energy_truncated_space(eps, 3, 3, 500)
    assert d.shape == (500, 4)


def test_full_space_size():
    assert full_space(6, 3, 2).shape == (20 * 15, 4)
