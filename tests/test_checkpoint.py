"""Dump formats of the time loop (wavebem_amd.checkpoint): deal.II's Vector::block_write
and PersistentTriangulation::write_flags layouts byte for byte on hand-built cases, round
trips, the dump schedule of source/free_surface.cc:1003-1010, and a restart of the BDF
integrator from a dump."""
import io
import math
import struct

import numpy as np
import pytest

from wavebem_amd import checkpoint as C
from wavebem_amd import timeloop as TL


def test_block_write_layout():
    f = io.BytesIO()
    C.block_write(f, [1.0, -2.5])
    assert f.getvalue() == b"2\n[" + struct.pack("=2d", 1.0, -2.5) + b"]"
    f = io.BytesIO()
    C.block_write(f, np.zeros(0))
    assert f.getvalue() == b"0\n[]"


def test_block_round_trip_and_errors():
    v = np.random.default_rng(1).standard_normal(1001)
    f = io.BytesIO()
    C.block_write(f, v)
    C.block_write(f, v[:3])
    f.seek(0)
    assert np.array_equal(C.block_read(f), v)
    assert np.array_equal(C.block_read(f), v[:3])
    for bad in (b"3\n(" + bytes(24) + b"]", b"3\n[" + bytes(24) + b")", b"3\n[" + bytes(16),
                b"12345"):
        with pytest.raises(ValueError):
            C.block_read(io.BytesIO(bad))


def test_write_flags_layout():
    f = io.StringIO()
    C.write_flags(f)
    assert f.getvalue() == "41472 0\n41473\n"
    f = io.StringIO()
    refine = [1, 0, 0, 1, 0, 0, 0, 0, 1, 1]
    C.write_flags(f, [refine], [[0, 0, 0]])
    assert f.getvalue() == ("41472 1\n40960 10\n9 3 \n40961\n40976 3\n0 \n40977\n41473\n")
    f.seek(0)
    r, c = C.read_flags(f)
    assert r == [[bool(b) for b in refine]] and c == [[False] * 3]
    with pytest.raises(ValueError):
        C.read_flags(io.StringIO("41472 1\n40961 0\n0 \n40961\n"))


def test_dump_schedule_and_files(tmp_path):
    d = C.Dumper(str(tmp_path / "run"), 0.25)
    names = [d.step(t, np.array([t]), np.array([2 * t]))
             for t in (0.1, 0.2, 0.3, 0.45, 0.5, 0.76)]
    assert [n is not None for n in names] == [False, False, True, False, True, True]
    assert d.written == [str(tmp_path / f"run_dump_{k}") for k in (1, 2, 3)]
    y, yp, (r, c) = C.restore_solution(d.written[1])
    assert y.tolist() == [0.5] and yp.tolist() == [1.0] and r == [] and c == []


class _Osc:
    A = np.array([[0.0, 1.0], [-4.0, 0.0]])

    def residual(self, t, y, yp):
        return yp - self.A @ y

    def jacobian(self, t, y, yp, v, alpha):
        return alpha * v - self.A @ v


def _dumped_run(tmp_path, dt, T=1.0, t_dump=0.5):
    """Runs to t_dump, dumps, restarts a new integrator from the files and runs to T.
    Returns (error of the uninterrupted run, error of the restarted run, restored y equals
    the state at the dump)."""
    steps = int(round(T / dt))
    it = TL.BDFIntegrator(_Osc(), np.array([1.0, 0.0]), dt, rtol=1e-12, atol=1e-14)
    dump = C.Dumper(str(tmp_path / f"osc{steps}"), t_dump)
    y_mid = None
    for s in range(steps):
        it.step()
        if dump.step(it.t + 1e-12, it.y, it.last_yp) and y_mid is None:
            y_mid, yp_mid = it.y.copy(), it.last_yp.copy()
    y, yp, _ = C.restore_solution(dump.written[0])
    it2 = TL.BDFIntegrator(_Osc(), np.zeros(2), dt, rtol=1e-12, atol=1e-14)
    it2.restart(y, t_dump)
    for _ in range(int(round((T - t_dump) / dt))):
        it2.step()
    exact = np.array([math.cos(2 * T), -2 * math.sin(2 * T)])
    return (np.abs(it.y - exact).max(), np.abs(it2.y - exact).max(),
            np.array_equal(y, y_mid) and np.array_equal(yp, yp_mid))


def test_restart_from_dump(tmp_path):
    """A run dumped at t = 0.5 and restarted from the files (one order-1 step, as IDA's
    restart) keeps the global second order: halving dt divides its error by about 4."""
    e1a, e2a, same = _dumped_run(tmp_path, 0.01)
    e1b, e2b, _ = _dumped_run(tmp_path, 0.005)
    assert same
    assert abs(e1a / e1b - 4.0) < 0.6 and abs(e2a / e2b - 4.0) < 0.8, (e1a, e1b, e2a, e2b)
    assert e2a < 5 * e1a
