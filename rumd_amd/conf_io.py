"""Reader and writer for RUMD configuration files (``.xyz`` / ``.xyz.gz``).

Follows ConfigurationWriterReader (src/ConfigurationWriterReader.cu:20-343) and
ConfigurationMetaData (src/ConfigurationMetaData.cc): line 1 holds N, line 2
``key=value`` tokens (ioformat 1: ``boxLengths=``, ``mass=``, ``numTypes=``,
``Nose-Hoover-Ps=``, ``dt=``; ioformat 2: ``sim_box=RectangularSimulationBox,Lx,Ly,Lz``,
``integrator=Name,params...``), then one line per particle with the columns
named by ``columns=``.  The ``vir`` column is w_i / 2
(src/ConfigurationWriterReader.cu:172-173).
"""
from __future__ import annotations

import gzip
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np


@dataclass
class Configuration:
    n: int
    box: np.ndarray                      # (3,) box lengths
    types: np.ndarray                    # (n,) int32
    positions: np.ndarray                # (n, 3) float32
    masses: np.ndarray                   # (num_types,) float64
    images: np.ndarray | None = None     # (n, 3) float32
    velocities: np.ndarray | None = None
    forces: np.ndarray | None = None
    pe: np.ndarray | None = None         # per-particle u_i
    vir: np.ndarray | None = None        # per-particle w_i / 2 as written
    meta: dict = field(default_factory=dict)

    @property
    def num_types(self) -> int:
        return int(self.masses.size)


def _open(path: Path, mode: str):
    path = Path(path)
    if path.suffix == ".gz":
        return gzip.open(path, mode + "t")
    return open(path, mode)


def _read_frame(fh, path) -> Configuration | None:
    """One frame (count line, comment line, N particle lines); None at the
    end of the file."""
    first = fh.readline()
    if not first.strip():
        return None
    n = int(first.split()[0])
    meta = {}
    for tok in fh.readline().split():
        if "=" in tok:
            k, v = tok.split("=", 1)
            meta[k] = v
    lines = [fh.readline() for _ in range(n)]
    if n and not lines[-1].strip():
        raise ValueError(f"{path}: expected {n} particle lines")
    rows = np.array([ln.split() for ln in lines], dtype=np.float64).reshape(n, -1)
    cols = meta.get("columns", "type,x,y,z").split(",")
    col = {c: k for k, c in enumerate(cols)}
    if "sim_box" in meta:
        parts = meta["sim_box"].split(",")
        if parts[0] != "RectangularSimulationBox":
            raise ValueError(f"{path}: unsupported simulation box {parts[0]}")
        box = np.array([float(x) for x in parts[1:4]])
    else:
        box = np.array([float(x) for x in meta["boxLengths"].split(",")])
    masses = np.array([float(x) for x in meta.get("mass", "1.0").split(",")])
    nt = int(meta.get("numTypes", masses.size))
    if masses.size < nt:
        masses = np.concatenate([masses, np.ones(nt - masses.size)])

    def take(names, dtype=np.float32):
        if all(c in col for c in names):
            return np.ascontiguousarray(rows[:, [col[c] for c in names]], dtype=dtype)
        return None

    return Configuration(
        n=n,
        box=box,
        types=rows[:, col["type"]].astype(np.int32),
        positions=take(["x", "y", "z"]),
        masses=masses,
        images=take(["imx", "imy", "imz"]),
        velocities=take(["vx", "vy", "vz"]),
        forces=take(["fx", "fy", "fz"]),
        pe=rows[:, col["pe"]].copy() if "pe" in col else None,
        vir=rows[:, col["vir"]].copy() if "vir" in col else None,
        meta=meta,
    )


def read_conf(path) -> Configuration:
    """The first frame of a configuration file."""
    with _open(path, "r") as fh:
        conf = _read_frame(fh, path)
    if conf is None:
        raise ValueError(f"{path}: empty configuration file")
    return conf


def read_trajectory(path) -> list[Configuration]:
    """Every frame of a trajectory file (frames appended one after another)."""
    frames = []
    with _open(path, "r") as fh:
        while True:
            c = _read_frame(fh, path)
            if c is None:
                return frames
            frames.append(c)


def write_conf(path, conf: Configuration, precision: int = 9, integrator_info: str | None = None,
               box_info: str | None = None, loglin: tuple | None = None, mode: str = "w") -> None:
    """Writes ioformat 2 (the reference's current format,
    src/ConfigurationWriterReader.cu:90-180).  ``loglin`` = (timeStepIndex,
    block, base, index, maxIndex, maxInterval) adds the trajectory's schedule
    tokens; ``box_info`` replaces the sim_box token (the box's GetInfoString);
    mode "a" appends a frame (a new gzip member for .gz files)."""
    if mode not in ("w", "a"):
        raise ValueError("mode must be 'w' or 'a'")
    cols = ["type", "x", "y", "z"]
    data = [conf.types.astype(np.float64)[:, None], conf.positions.astype(np.float64)]
    if conf.images is not None:
        cols += ["imx", "imy", "imz"]
        data.append(conf.images.astype(np.float64))
    if conf.velocities is not None:
        cols += ["vx", "vy", "vz"]
        data.append(conf.velocities.astype(np.float64))
    if conf.forces is not None:
        cols += ["fx", "fy", "fz"]
        data.append(conf.forces.astype(np.float64))
    if conf.pe is not None:
        cols += ["pe"]
        data.append(np.asarray(conf.pe, dtype=np.float64)[:, None])
    if conf.vir is not None:
        cols += ["vir"]
        data.append(np.asarray(conf.vir, dtype=np.float64)[:, None])
    arr = np.hstack(data)
    header = ["ioformat=2"]
    if loglin is not None:
        header.append("timeStepIndex=%d logLin=%d,%d,%d,%d,%d" % tuple(int(x) for x in loglin))
    header.append(f"numTypes={conf.num_types}")
    if integrator_info:
        header.append(f"integrator={integrator_info}")
    if box_info:
        header.append(f"sim_box={box_info}")
    else:
        header.append("sim_box=RectangularSimulationBox," + ",".join(f"{x:.{precision}g}" for x in conf.box))
    header.append("mass=" + ",".join(f"{m:.{precision}f}" for m in conf.masses))
    header.append("columns=" + ",".join(cols))
    fmt = ["%d"] + [f"%.{precision}f"] * 3
    if conf.images is not None:
        fmt += ["%.0f"] * 3
    fmt += [f"%.{precision}f"] * (arr.shape[1] - len(fmt))
    with _open(path, mode) as fh:
        fh.write(f"{conf.n}\n")
        fh.write(" ".join(header) + "\n")
        np.savetxt(fh, arr, fmt=fmt)
