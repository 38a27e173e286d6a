"""Matrix dumps (DumpMatrixEquation, femtools/Petsc_Tools.F90:1487-1504;
read back by tools/petsc_readnsolve.F90:60-83) on the CPU: the oracle's
restatement of PETSc's binary layout against a hand-built byte string, the
library's reader and writer against the oracle byte for byte, the committed
fixture, and the reader's errors.  No GPU call is made."""
import os
import struct

import numpy as np
import pytest

import petsc_binary as PB
from fluidity_amd import readnsolve as R

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "c1_8.matrixdump")

# [[4, -1, 0], [-1, 4, -1], [0, -2, 5]], b = (1, 2, 3), x0 = (0.5, 0, -0.5)
FINDRM = np.array([1, 3, 6, 8], np.int32)
COLM = np.array([1, 2, 1, 2, 3, 2, 3], np.int32)
VAL = np.array([4.0, -1.0, -1.0, 4.0, -1.0, -2.0, 5.0])
B = np.array([1.0, 2.0, 3.0])
X0 = np.array([0.5, 0.0, -0.5])


def hand_bytes(with_x0=True):
    """The published layout, spelled out with struct (big-endian)."""
    out = struct.pack(">4i", 1211216, 3, 3, 7)
    out += struct.pack(">3i", 2, 3, 2)
    out += struct.pack(">7i", 0, 1, 0, 1, 2, 1, 2)
    out += struct.pack(">7d", 4.0, -1.0, -1.0, 4.0, -1.0, -2.0, 5.0)
    out += struct.pack(">2i", 1211214, 3) + struct.pack(">3d", 1.0, 2.0, 3.0)
    if with_x0:
        out += struct.pack(">2i", 1211214, 3) + struct.pack(">3d", 0.5, 0.0, -0.5)
    return out


def test_oracle_layout_matches_hand_built_bytes():
    assert PB.encode(FINDRM, COLM, VAL, B, X0) == hand_bytes()
    f, c, v, b, x0 = PB.decode(hand_bytes())
    assert np.array_equal(f, FINDRM) and np.array_equal(c, COLM) and np.array_equal(v, VAL)
    assert np.array_equal(b, B) and np.array_equal(x0, X0)
    assert PB.decode(hand_bytes(False))[4] is None


@pytest.mark.parametrize("with_x0", [True, False])
def test_library_writer_and_reader_match_oracle(tmp_path, with_x0):
    path = str(tmp_path / "matrixdump")
    R.matrixdump_write(path, FINDRM, COLM, VAL, B, X0 if with_x0 else None)
    assert open(path, "rb").read() == hand_bytes(with_x0)
    f, c, v, b, x0 = R.matrixdump_read(path)
    assert np.array_equal(f, FINDRM) and np.array_equal(c, COLM) and np.array_equal(v, VAL)
    assert np.array_equal(b, B)
    assert (x0 is None) == (not with_x0)
    if with_x0:
        assert np.array_equal(x0, X0)


def test_fixture_reads_the_same_both_ways():
    data = open(FIXTURE, "rb").read()
    ref = PB.decode(data)
    got = R.matrixdump_read(FIXTURE)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    assert ref[0].size - 1 == 81
    assert R.centrm_of(ref[0], ref[1]).min() > 0  # every row holds its diagonal


def test_reader_errors(tmp_path):
    good = hand_bytes()
    cases = {
        "classid": struct.pack(">i", 42) + good[4:],
        "truncated": good[:40],
        "unsorted": good[:16 + 12] + struct.pack(">7i", 1, 0, 0, 1, 2, 1, 2) + good[16 + 12 + 28:],
        "nonsquare": struct.pack(">4i", 1211216, 3, 4, 7) + good[16:],
        # a row length near 2^31 would wrap a 32-bit running offset
        "huge_row": good[:16] + struct.pack(">3i", 2, 2**31 - 2, 3) + good[28:],
        "short_rows": good[:16] + struct.pack(">3i", 2, 3, 1) + good[28:],
    }
    for name, data in cases.items():
        p = str(tmp_path / name)
        open(p, "wb").write(data)
        with pytest.raises(Exception):
            R.matrixdump_read(p)
    with pytest.raises(Exception):
        R.matrixdump_read(str(tmp_path / "missing"))
