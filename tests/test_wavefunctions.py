"""The wavefunction working file (reference ``qepy_pw/pw/buffers.py:109-232``)
on the host: the reference's own tests (``tests/unit/io/test_wavefunction_
buffers.py:55-79``) run against ``qe_python_amd.wavefunctions``, the record
bytes against the oracle's element-by-element layout, and the errors."""

import numpy as np
import pytest

from oracle import wavefunctions as ow
from qe_python_amd.wavefunctions import QEInputError, WavefunctionBuffer


def test_reference_round_trip_of_fixed_binary_records(tmp_path):
    buffer = WavefunctionBuffer(tmp_path, "silicon", [2, 3], 2)
    first = np.array([[1 + 2j, 3 + 4j], [5 + 6j, 7 + 8j]])
    second = np.arange(6).reshape(3, 2) + 1j * np.arange(6, 12).reshape(3, 2)
    buffer.write(0, first)
    buffer.write(1, second)
    assert buffer.path == tmp_path / "silicon.wfc"
    assert buffer.path.stat().st_size == 2 * 3 * 2 * 16
    assert buffer.path.read_bytes()[:8] != b"\x89HDF\r\n\x1a\n"
    np.testing.assert_array_equal(buffer[0], first)
    np.testing.assert_array_equal(buffer[1], second)
    np.testing.assert_array_equal(buffer[-1], second)
    assert [b.shape for b in buffer] == [(2, 2), (3, 2)]
    assert [b.shape for b in buffer[0:2]] == [(2, 2), (3, 2)]
    # the bytes are the band columns written out one after the other, each
    # zero-padded to the largest row count
    raw = buffer.path.read_bytes()
    assert raw == ow.encode_record(first, 3) + ow.encode_record(second, 3)
    col = np.frombuffer(raw[:48], dtype="<c16")
    np.testing.assert_array_equal(col, [1 + 2j, 5 + 6j, 0])
    np.testing.assert_array_equal(ow.decode_record(raw[96:], 3, 2, 3), second)


def test_reference_rank_suffix(tmp_path):
    buffer = WavefunctionBuffer(tmp_path, "pwscf", [1], 1, rank=2, processes=4)
    assert buffer.path.name == "pwscf.wfc3" == ow.file_name("pwscf", 2, 4)
    assert buffer.record_bytes == ow.record_bytes([1], 1)


def test_errors(tmp_path):
    with pytest.raises(ValueError):
        WavefunctionBuffer(tmp_path, "x", [], 2)
    with pytest.raises(ValueError):
        WavefunctionBuffer(tmp_path, "x", [2, -1], 2)
    with pytest.raises(ValueError):
        WavefunctionBuffer(tmp_path, "x", [2], 0)
    buffer = WavefunctionBuffer(tmp_path, "x", [2, 3], 2)
    with pytest.raises(QEInputError) as err:
        buffer[1]
    assert err.value.routine == "get_buffer"
    with pytest.raises(QEInputError) as err:
        buffer.write(0, np.zeros((3, 2)))
    assert err.value.routine == "save_buffer"
    with pytest.raises(IndexError):
        buffer.write(2, np.zeros((2, 2)))
    with pytest.raises(IndexError):
        buffer[5]
    assert not buffer.has_record(0)
    buffer.write(0, np.ones((2, 2)))
    assert buffer.has_record(0) and not buffer.has_record(1)
    with pytest.raises(QEInputError) as err:
        WavefunctionBuffer(tmp_path / "missing" / "dir", "x", [1], 1)
    assert err.value.routine == "open_buffer"
