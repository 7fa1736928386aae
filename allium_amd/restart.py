"""scallion's restart files (src/restart.cpp, exe/scallion.cpp:604-631,1326-1340).

One file per rank, ``<root>/restart_<n>.<rank, 6 digits>.dat``
(mfem::MakeParFilename), holding the RestartMetaData struct as the
reference's compiler lays it out on x86-64 -- int num_procs, int output_cycle,
int simulation_cycle, 4 bytes of padding, double time, double time_step, int
size, 4 bytes of padding (40 bytes) -- followed by `size` doubles of the
state vector.  Writing shifts restart_<n-1> to restart_<n> for n < keep first
(RestartWriter::CycleRestartFiles), so restart_0 is always the newest.  For
the TRT solver the state is the block vector [psi, T] (exe/scallion.cpp:560-580).
"""
import os
import struct

import numpy as np

_META = struct.Struct("<iii4xddi4x")
assert _META.size == 40


def restart_file(root, n=0, rank=0):
    return os.path.join(root, f"restart_{int(n)}.{int(rank):06d}.dat")


def write_restart(root, x, output_cycle=0, simulation_cycle=0, time=0.0, time_step=0.0, keep=2, rank=0,
                  num_procs=1):
    """RestartWriter::Write: rotate the older files, then write restart_0."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    os.makedirs(root, exist_ok=True)
    for lvl in range(int(keep) - 1, 0, -1):
        old, new = restart_file(root, lvl - 1, rank), restart_file(root, lvl, rank)
        if os.path.exists(old):
            os.replace(old, new)
    with open(restart_file(root, 0, rank), "wb") as fh:
        fh.write(_META.pack(int(num_procs), int(output_cycle), int(simulation_cycle), float(time), float(time_step),
                            x.size))
        fh.write(x.astype("<f8").tobytes())


def load_restart(base, restart_num=0, size=None, rank=0, num_procs=1):
    """LoadFromRestart: the state vector and its metadata.  Raises where the
    reference aborts: a missing file, another rank count, a size mismatch."""
    path = restart_file(base, restart_num, rank)
    if not os.path.exists(path):
        raise FileNotFoundError(f"restart file not found: {path}")
    with open(path, "rb") as fh:
        head = fh.read(_META.size)
        if len(head) != _META.size:
            raise ValueError(f"{path}: truncated restart header")
        nprocs, out_cycle, sim_cycle, time, time_step, n = _META.unpack(head)
        if nprocs != num_procs:
            raise ValueError(f"MPI ranks does not match restart data ({nprocs} in the file, {num_procs} here)")
        if size is not None and n != size:
            raise ValueError(f"restart file size wrong ({n} values in the file, {size} expected)")
        x = np.frombuffer(fh.read(8 * n), dtype="<f8")
        if x.size != n:
            raise ValueError(f"{path}: truncated restart data")
    return {"x": x.astype(np.float64), "output_cycle": out_cycle, "simulation_cycle": sim_cycle, "time": time,
            "time_step": time_step}
