"""CPU restatement of the reference's head-file and restart-file formats.
TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
msg123_write_head / msg123_write_restart / msg123_read_restart.

Both files are Fortran stream-unformatted (open_new_wbin,
Utility_module.f90:174-175): raw little-endian values, no record markers.

* head file, one record per output time (write_out_headf,
  Write_output.f90:336-362): write_header_bin's real(SP) time
  (Write_module.f90:209-225), then write_3dbin's nx*ny*nz real(SP) values
  real(head*len_scal, SP), x fastest, SNOVAL = -9999 where no calculation cell
  lies (Write_module.f90:146-183);
* restart file (write_out_restf, Write_output.f90:415-432): real(time, DP),
  then head(i)*len_scal in DP for i = 1..ncalc;
* restart read (assign_initv, Assign_calculation.f90:416-435): the values
  times len_scal_inv, which the reference holds in SP as 1/len_scal
  (Read_input.f90:1108), and the time rounded to SP.

Pinned by the reference's own head and restart files of a run
(tests/golden/ref_io_box9.npz) written from the FP64 heads it dumped.
"""
from __future__ import annotations

import numpy as np


def head_record(time: float, heads: np.ndarray, len_scal: float) -> bytes:
    """One head-file record of a full box (every grid cell a calculation cell,
    in grid order)."""
    v = (np.asarray(heads, np.float64) * np.float64(np.float32(len_scal))).astype("<f4")
    return np.float32(time).astype("<f4").tobytes() + v.tobytes()


def restart_file(time: float, heads: np.ndarray, len_scal: float) -> bytes:
    v = (np.asarray(heads, np.float64) * np.float64(np.float32(len_scal))).astype("<f8")
    return np.float64(np.float32(time)).astype("<f8").tobytes() + v.tobytes()


def read_restart(data: bytes, ncalc: int, len_scal: float):
    """(heads, time) from the bytes of a restart file."""
    a = np.frombuffer(data, "<f8", count=ncalc + 1)
    inv = np.float64(np.float32(1.0) / np.float32(len_scal))
    return a[1:] * inv, float(np.float32(a[0]))


def read_head_records(data: bytes, ncell: int):
    """[(time, float32 values)] of every record of a head file."""
    rec = 4 * (ncell + 1)
    if len(data) % rec:
        raise ValueError(f"head file of {len(data)} bytes is not a whole number of "
                         f"{rec}-byte records")
    out = []
    for k in range(len(data) // rec):
        a = np.frombuffer(data, "<f4", count=ncell + 1, offset=k * rec)
        out.append((float(a[0]), a[1:].copy()))
    return out
