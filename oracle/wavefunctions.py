"""CPU restatement of the reference's wavefunction working-file record format
(test infrastructure only: the oracle checks the device path, it is never
the measured or shipped code).

``qepy_pw/pw/buffers.py:109-232`` (``WavefunctionBuffer``): one fixed-size
record per k-point at byte offset ``index * record_bytes``, record_bytes =
max(row_counts) * nbands * 16; a record holds the (npw, nbands) complex128
matrix band-major (each band's npw coefficients, then zeros up to the
largest row count), native little-endian, C order.  The file is
``{prefix}.wfc`` for one process and ``{prefix}.wfc{rank + 1}`` otherwise.

Pinned by the reference's own tests (``tests/unit/io/test_wavefunction_
buffers.py:55-79``: file size 2 x 3 x 2 x 16 for row counts [2, 3] and two
bands, no HDF5 signature, round trip, rank suffix) and by the closed form of
the layout: each record's bytes are compared with the band columns written
out element by element.
"""

from __future__ import annotations

import numpy as np


def record_bytes(row_counts, number_of_bands) -> int:
    return max(int(r) for r in row_counts) * int(number_of_bands) * 16


def file_name(prefix: str, rank: int = 0, processes: int = 1) -> str:
    return f"{prefix}.wfc" + ("" if processes == 1 else str(rank + 1))


def encode_record(matrix: np.ndarray, record_rows: int) -> bytes:
    """Band-major, zero-padded bytes of one (npw, nbands) record."""
    m = np.asarray(matrix, dtype=np.complex128)
    out = bytearray()
    pad = np.zeros(record_rows - m.shape[0], dtype="<c16").tobytes()
    for band in range(m.shape[1]):
        out += np.ascontiguousarray(m[:, band], dtype="<c16").tobytes() + pad
    return bytes(out)


def decode_record(payload: bytes, rows: int, number_of_bands: int, record_rows: int) -> np.ndarray:
    stored = np.frombuffer(payload, dtype="<c16").reshape(number_of_bands, record_rows)
    return np.ascontiguousarray(stored[:, :rows].T)
