"""Restart files on the host (no GPU): the writer and reader of ppk_mhd_write_restart /
ppk_mhd_read_restart, which ppk_mhd_save_state / ppk_mhd_load_state use.  The reference's
restart content (IO_HDF5.h:430-484): the fields with their ghost zone, "time step" and
"total time"."""
import numpy as np
import pytest

from ppkmhd_amd import _abi as A
from ppkmhd_amd import default_problem, make_params
from ppkmhd_amd._lib import PPKError
from ppkmhd_amd.solver import initial_state, read_restart, write_restart


@pytest.mark.parametrize("dim", [2, 3])
def test_round_trip_is_exact(tmp_path, dim):
    p = make_params(dim, 12, 10, 8, gamma0=5 / 3)
    U = initial_state(p, default_problem(A.PROBLEM_ORSZAG_TANG, p))
    U[0, ..., 1] = np.nextafter(U[0, ..., 1], 2.0)  # low bits must survive
    f = tmp_path / "r.bin"
    write_restart(f, p, U, total_time=0.123456789, time_step=7, iteration=41)
    V, t, ts, it = read_restart(f, p)
    assert np.array_equal(V, U) and (t, ts, it) == (0.123456789, 7, 41)
    # header (128 bytes) + the doubles of the padded box, nothing else
    assert f.stat().st_size == 128 + 8 * U.size
    raw = f.read_bytes()
    assert raw[:8] == b"PPKMHDRS"
    assert np.array_equal(np.frombuffer(raw[128:], dtype="<f8"), U.ravel())


def test_mismatched_grid_truncated_and_foreign_files_are_refused(tmp_path):
    p = make_params(3, 12, 10, 8)
    U = initial_state(p, default_problem(A.PROBLEM_BLAST, p))
    f = tmp_path / "r.bin"
    write_restart(f, p, U, 1.0, 0, 3)
    q = make_params(3, 12, 10, 9)
    with pytest.raises(PPKError) as e:
        read_restart(f, q)
    assert e.value.code == A.PPK_ERR_INVALID_ARGUMENT and "12x10x8" in str(e.value)
    g = tmp_path / "short.bin"
    g.write_bytes(f.read_bytes()[:-8])
    with pytest.raises(PPKError) as e:
        read_restart(g, p)
    assert "truncated" in str(e.value)
    h = tmp_path / "other.bin"
    h.write_bytes(b"\0" * 256)
    with pytest.raises(PPKError) as e:
        read_restart(h, p)
    assert "not a restart file" in str(e.value)
    with pytest.raises(PPKError):
        read_restart(tmp_path / "missing.bin", p)
