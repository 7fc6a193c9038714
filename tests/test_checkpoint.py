"""Dump2 file layout: header parsing on the host (the device round trip is a GPU test)."""
import numpy as np
import pytest

import ffvhc


def test_dump_header_reads_reference_layout(tmp_path):
    p = tmp_path / "pP-00000-0000000010.bin"
    head = np.array([16, 16, 16, 2, 1, 3], np.int32)
    with open(p, "wb") as fh:
        fh.write(head.tobytes())
        fh.write(np.zeros(3 * 20 ** 3).tobytes())
    assert list(ffvhc.dump_header(p)) == [16, 16, 16, 2, 1, 3]


def test_dump_header_errors(tmp_path):
    p = tmp_path / "short.bin"
    p.write_bytes(b"\x00\x01")
    with pytest.raises(ffvhc.FFVError):
        ffvhc.dump_header(p)
    with pytest.raises(ffvhc.FFVError):
        ffvhc.dump_header(tmp_path / "missing.bin")
