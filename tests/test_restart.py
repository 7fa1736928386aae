"""Restart files in scallion's format (allium_amd/restart.py; src/restart.cpp).
CPU: the byte layout of RestartMetaData, the rotation of older files, the
errors where the reference aborts.  GPU: a TRT run continued from a restart
file repeats the uninterrupted run bit for bit."""
import os
import struct

import numpy as np
import pytest

from allium_amd import restart as R


def test_layout_and_round_trip(tmp_path):
    x = np.random.default_rng(0).normal(size=37)
    R.write_restart(str(tmp_path), x, output_cycle=3, simulation_cycle=17, time=0.125, time_step=1e-3)
    path = R.restart_file(str(tmp_path), 0, 0)
    assert os.path.basename(path) == "restart_0.000000.dat"   # mfem::MakeParFilename, width 6
    raw = open(path, "rb").read()
    assert len(raw) == 40 + 8 * x.size
    # struct RestartMetaData on x86-64: 3 ints, padding, 2 doubles, int, padding
    assert struct.unpack_from("<iii", raw, 0) == (1, 3, 17)
    assert struct.unpack_from("<dd", raw, 16) == (0.125, 1e-3)
    assert struct.unpack_from("<i", raw, 32) == (37,)
    assert np.array_equal(np.frombuffer(raw[40:], "<f8"), x)
    r = R.load_restart(str(tmp_path), 0, size=37)
    assert np.array_equal(r["x"], x)
    assert (r["output_cycle"], r["simulation_cycle"], r["time"], r["time_step"]) == (3, 17, 0.125, 1e-3)


def test_rotation_keeps_the_newest_first(tmp_path):
    root = str(tmp_path / "rs")
    for k in range(4):
        R.write_restart(root, np.full(3, float(k)), simulation_cycle=k, keep=3)
    got = [R.load_restart(root, n)["simulation_cycle"] for n in range(3)]
    assert got == [3, 2, 1]
    assert not os.path.exists(R.restart_file(root, 3))


def test_errors(tmp_path):
    root = str(tmp_path)
    with pytest.raises(FileNotFoundError):
        R.load_restart(root, 0)
    R.write_restart(root, np.zeros(5), num_procs=2)
    with pytest.raises(ValueError, match="ranks"):
        R.load_restart(root, 0)
    R.write_restart(root, np.zeros(5))
    with pytest.raises(ValueError, match="size"):
        R.load_restart(root, 0, size=6)


@pytest.mark.gpu
def test_trt_continues_from_restart_bit_for_bit(tmp_path):
    """scallion's restart path: 4 steps straight, or 2 steps, a restart file,
    a new solver loaded from it and 2 more steps."""
    print("trt restart", flush=True)
    from allium_amd.trt import TRT
    kw = dict(nel=(6, 6, 6), sn=4, groups=4, dt=1e-3)
    a = TRT(**kw)
    for _ in range(4):
        a.step()
    b = TRT(**kw)
    for _ in range(2):
        b.step()
    b.write_restart(str(tmp_path), output_cycle=1)
    c = TRT(**kw)
    meta = c.load_restart(str(tmp_path), 0)
    assert meta["simulation_cycle"] == 2 and meta["output_cycle"] == 1
    info = None
    for _ in range(2):
        info = c.step()
    xa, cyc_a, ta = a.state()
    xc, cyc_c, tc = c.state()
    assert cyc_a == cyc_c == 4 and info["step"] == 4 and ta == tc
    assert np.array_equal(xa, xc)
    assert np.array_equal(a.scalar_flux(), c.scalar_flux())
    # scallion adopts the file's time step (use_restart_dt, its default): a solver
    # created with another dt continues exactly as the uninterrupted run
    d = TRT(nel=(6, 6, 6), sn=4, groups=4, dt=2e-3)
    d.load_restart(str(tmp_path), 0)
    assert d.dt == 1e-3
    for _ in range(2):
        d.step()
    xd, cyc_d, td = d.state()
    assert cyc_d == 4 and np.array_equal(xa, xd)
    # without it the solver keeps its own time step
    e = TRT(nel=(6, 6, 6), sn=4, groups=4, dt=2e-3)
    e.load_restart(str(tmp_path), 0, use_restart_dt=False)
    e.step()
    assert e.dt == 2e-3 and abs(e.state()[2] - (2e-3 + 2e-3)) < 1e-15
