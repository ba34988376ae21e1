"""Integral generator, FCIDUMP writer/reader (reference src/ham/read_ints.rs:73-190)."""
import os

import numpy as np
import pytest

from risq_amd.fcidump import (FcidumpError, SplitMix64, combine_4, hubbard_2d, read_fcidump,
                              synthetic_molecular, table_sizes, write_fcidump)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_splitmix64_published_vector():
    # first outputs of splitmix64 seeded with 1234567 (published reference values)
    g = SplitMix64(1234567)
    assert [g.next_u64() for _ in range(3)] == [6457827717110365317, 3203168211198807973,
                                                9817491932198370423]


def test_table_sizes_match_reference_allocation():
    # read_ints.rs:78-80: combine_2(n+1, n+1), combine_4(n+1, n+1, n+1, n+1)
    assert table_sizes(4) == (15, 120)
    assert table_sizes(40)[1] == 371091


def test_generator_symmetry_and_psd():
    ints = synthetic_molecular(8, 12, 3)
    h, eri = ints.dense()
    assert np.array_equal(h, h.T)
    assert np.allclose(np.diag(h), -1.0 + 2.0 * np.arange(8) / 15.0)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        assert np.array_equal(eri, eri.transpose(perm))
    m = eri.reshape(64, 64)
    assert np.linalg.eigvalsh(m).min() > -1e-12
    again = synthetic_molecular(8, 12, 3)
    assert np.array_equal(again.two_body, ints.two_body)
    assert not np.array_equal(synthetic_molecular(8, 12, 4).two_body, ints.two_body)


def test_hubbard_integrals():
    ints = hubbard_2d(6, 6, 1.0, 4.0, 0.1, 2)
    h, eri = ints.dense()
    assert np.count_nonzero(eri) == 36 and np.all(eri[np.arange(36), np.arange(36), np.arange(36), np.arange(36)] == 4.0)
    off = h - np.diag(np.diag(h))
    assert np.all(np.isin(off, [0.0, -1.0])) and np.all((off != 0).sum(axis=1) == 4)
    assert np.all(np.abs(np.diag(h)) <= 0.1)


def test_write_read_round_trip_is_exact(tmp_path):
    ints = synthetic_molecular(7, 9, 5)
    ints.e_core = 1.2345678901234567
    p = tmp_path / "FCIDUMP"
    write_fcidump(str(p), ints, 6)
    back = read_fcidump(str(p), 7)
    assert np.array_equal(back.one_body, ints.one_body)
    assert np.array_equal(back.two_body, ints.two_body)
    assert back.e_core == ints.e_core
    assert "\n\n" not in p.read_text()


def test_reader_rules(tmp_path):
    hdr = "&FCI NORB=2,NELEC=2,\n&END\n"
    p = tmp_path / "f"
    p.write_text(hdr + "0.5 1 1 1 1\n0.25 2 1 0 0\n-1.5 0 0 0 0\n")
    ints = read_fcidump(str(p), 2)
    assert ints.eri(0, 0, 0, 0) == 0.5 and ints.h1(0, 1) == 0.25 and ints.e_core == -1.5
    for bad, msg in [("0.5 1 1 1\n", "Missing s index"), ("0.5 1 x 1 1\n", "Invalid q index"),
                     ("0.5 1 3 1 1\n", "out of range"), ("\n", "Missing integral value")]:
        p.write_text(hdr + bad)
        with pytest.raises(FcidumpError, match=msg):
            read_fcidump(str(p), 2)


def test_he2_example_reads():
    ints = read_fcidump(os.path.join(GOLDEN, "he2_FCIDUMP"), 4)
    assert ints.eri(0, 0, 0, 0) == 0.7912276603999572
    assert ints.two_body[combine_4(1, 1, 1, 1)] == 0.7912276603999572
