"""The SCF's wavefunction working file with device eigenvectors
(``WavefunctionBuffer``, reference ``qepy_pw/pw/buffers.py:109-232``).

Same file, same records, same errors as the reference: one fixed-size
record per k-point, each the (npw, nbands) complex128 matrix stored
band-major and zero-padded to the largest row count, in ``{prefix}.wfc``
(``{prefix}.wfc{rank + 1}`` for several processes).  Host arrays are
written and read as the reference does.  Device eigenvectors -- a (npw, nb)
tensor in reference row order, as ``davidson(..., device_output=True)``
returns them, or a (nb, npw) block in a plan's device order -- are packed
into the record layout on the device (``qepw_permute_rows`` into a
zero-padded staging block, so the reordering and padding cost no host
pass) and reach the file through one pinned device-to-host copy;
``read_device`` reverses the path.  The SCF writes every k-point's
eigenvectors after each diagonalisation and reads them back as the next
trial block (``pw/scf.py:3410-3416, 3643-3647``).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from . import _native as nat
from . import device as dv


class QEInputError(ValueError):
    """The reference's I/O error (``qepy_pw/errors.py:16-32``): a ValueError
    carrying the QE routine name and error code."""

    def __init__(self, message: str, *, routine: str = "pw.py", code: int = 1):
        super().__init__(message)
        self.routine = routine
        self.code = code


class WavefunctionBuffer:
    """Fixed-record binary storage of per-k wavefunction matrices
    (reference constructor, ``__len__``, ``__getitem__``, ``__iter__``,
    ``has_record``, ``write``), plus ``write`` of device tensors and
    ``read_device``."""

    def __init__(self, directory, prefix, row_counts, number_of_bands, *, rank=0, processes=1,
                 truncate=True):
        self.row_counts = tuple(int(v) for v in row_counts)
        self.number_of_bands = int(number_of_bands)
        if not self.row_counts or min(self.row_counts) < 0:
            raise ValueError("wavefunction row counts must be nonnegative")
        if self.number_of_bands < 1:
            raise ValueError("number_of_bands must be positive")
        self.record_rows = max(self.row_counts)
        self.record_bytes = self.record_rows * self.number_of_bands * 16
        suffix = "" if processes == 1 else str(rank + 1)
        self.path = Path(directory) / f"{prefix}.wfc{suffix}"
        self._written = [False] * len(self.row_counts)
        self._staging = {}
        if truncate:
            try:
                with self.path.open("w+b") as stream:
                    stream.truncate(self.record_bytes * len(self.row_counts))
            except OSError as exc:
                raise QEInputError(f"cannot open wavefunction buffer {self.path}: {exc}",
                                   routine="open_buffer") from exc

    def __len__(self):
        return len(self.row_counts)

    def has_record(self, index):
        return self._written[index]

    def _check_index(self, index):
        if index < 0:
            index += len(self)
        if not 0 <= index < len(self):
            raise IndexError(index)
        return index

    # -- file access ----------------------------------------------------------
    def _read_payload(self, index, into=None):
        """The record's bytes, read into the writable buffer ``into`` when
        given (no intermediate copy), else returned as bytes."""
        if not self._written[index]:
            raise QEInputError(f"wavefunction record {index + 1} has not been written",
                               routine="get_buffer")
        try:
            with self.path.open("rb") as stream:
                stream.seek(index * self.record_bytes)
                if into is None:
                    payload = stream.read(self.record_bytes)
                    got = len(payload)
                else:
                    payload = into
                    got = stream.readinto(into)
        except OSError as exc:
            raise QEInputError(f"cannot read wavefunction buffer {self.path}: {exc}",
                               routine="get_buffer") from exc
        if got != self.record_bytes:
            raise QEInputError(f"short wavefunction record in {self.path}", routine="get_buffer")
        return payload

    def _write_payload(self, index, payload) -> None:
        try:
            with self.path.open("r+b") as stream:
                stream.seek(index * self.record_bytes)
                stream.write(payload)
                stream.flush()
        except OSError as exc:
            raise QEInputError(f"cannot write wavefunction buffer {self.path}: {exc}",
                               routine="save_buffer") from exc
        self._written[index] = True

    def __getitem__(self, index):
        if isinstance(index, slice):
            return [self[i] for i in range(*index.indices(len(self)))]
        index = self._check_index(index)
        stored = np.frombuffer(self._read_payload(index), dtype=np.complex128).reshape(
            self.number_of_bands, self.record_rows)
        return np.ascontiguousarray(stored[:, :self.row_counts[index]].T)

    def __iter__(self):
        for index in range(len(self)):
            yield self[index]

    # -- device staging ---------------------------------------------------------
    def _stage(self, device):
        """(nbands, record_rows) device staging block and its pinned host
        twin, shared by the records (``write`` clears a shorter record's
        padding columns)."""
        key = str(device)
        st = self._staging.get(key)
        if st is None:
            dev = torch.zeros((self.number_of_bands, self.record_rows), dtype=torch.complex128,
                              device=device)
            host = torch.zeros((self.number_of_bands, self.record_rows), dtype=torch.complex128,
                               pin_memory=True)
            st = self._staging[key] = (dev, host)
        return st

    def write(self, index, coefficients, *, perm=None) -> None:
        """Store record ``index``: a host or device (npw, nbands) matrix in
        reference row order, or -- with ``perm``, the plan's device
        permutation (``DeviceFFTPlan.perm_device``) -- a device (nbands, npw)
        block in device order."""
        if not 0 <= index < len(self):
            raise IndexError(index)
        rows = self.row_counts[index]
        if not isinstance(coefficients, torch.Tensor) or coefficients.device.type == "cpu":
            if perm is not None:
                raise ValueError("perm applies to device-order device blocks only")
            if isinstance(coefficients, torch.Tensor):
                coefficients = coefficients.detach().resolve_conj().resolve_neg().numpy()
            matrix = np.asarray(coefficients, dtype=np.complex128)
            if matrix.shape != (rows, self.number_of_bands):
                raise QEInputError(f"wavefunction record has shape {matrix.shape}, expected "
                                   f"{(rows, self.number_of_bands)}", routine="save_buffer")
            record = np.zeros((self.number_of_bands, self.record_rows), dtype=np.complex128)
            record[:, :rows] = matrix.T
            self._write_payload(index, record.tobytes(order="C"))
            return
        t = coefficients
        if perm is None:
            if tuple(t.shape) != (rows, self.number_of_bands):
                raise QEInputError(f"wavefunction record has shape {tuple(t.shape)}, expected "
                                   f"{(rows, self.number_of_bands)}", routine="save_buffer")
            block = t.T                   # (nbands, npw) view of the reference-order columns
            direction, pdev = 0, None
        else:
            if tuple(t.shape) != (self.number_of_bands, rows) or perm.numel() != rows:
                raise QEInputError(f"device block has shape {tuple(t.shape)}, expected "
                                   f"{(self.number_of_bands, rows)}", routine="save_buffer")
            block, direction, pdev = t, 1, perm
        # records are complex128, as the reference's np.asarray(..., complex128)
        block = block.detach().resolve_conj().resolve_neg().to(torch.complex128)
        if rows > 1 and block.stride(1) != 1:
            block = block.contiguous()
        dev, host = self._stage(t.device)
        if rows < self.record_rows:
            dev[:, rows:].zero_()
        if rows:
            dv.permute_rows(direction, pdev, block, dev, rows)
        host.copy_(dev)
        self._write_payload(index, memoryview(host.numpy()).cast("B"))

    def read_device(self, index, device=None, *, perm=None) -> torch.Tensor:
        """Record ``index`` on the device: a (npw, nbands) reference-order
        tensor, or with ``perm`` the (nbands, npw) device-order block."""
        index = self._check_index(index)
        rows = self.row_counts[index]
        dev_t = dv.require_device(device)
        dev, host = self._stage(dev_t)
        self._read_payload(index, into=memoryview(host.numpy()).cast("B"))
        dev.copy_(host, non_blocking=True)
        out = torch.empty((self.number_of_bands, rows), dtype=torch.complex128, device=dev_t)
        if rows:
            if perm is not None and perm.numel() != rows:
                raise QEInputError("device permutation does not match the record's rows",
                                   routine="get_buffer")
            dv.permute_rows(0, perm, dev, out, rows)
        # the staging block is reused: finish the unpack before the next call
        torch.cuda.current_stream(dev_t).synchronize()
        return out if perm is not None else out.T
