"""Synthetic start configurations for the benchmark and test workloads.

Lattice builders in the spirit of RUMD's ``rumd_init_conf`` tool
(Tools/rumd_init_conf.cc:132-305 lattices, 328-337 Maxwell velocities):
positions on a simple-cubic or FCC lattice centred on the origin, types drawn
by seeded permutation, velocities Gaussian with zero total momentum.  Every
generator is deterministic in its seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# Kob-Andersen 80:20 binary LJ (W. Kob and H. C. Andersen, PRE 51, 4626 (1995)),
# as set up in the reference's examples (doc/examples/NpTEnsemble/run.py).
KA_SIGMA = {(0, 0): 1.0, (0, 1): 0.8, (1, 0): 0.8, (1, 1): 0.88}
KA_EPSILON = {(0, 0): 1.0, (0, 1): 1.5, (1, 0): 1.5, (1, 1): 0.5}


@dataclass
class System:
    positions: np.ndarray   # (n, 3) float32, inside [-L/2, L/2)
    velocities: np.ndarray  # (n, 3) float32
    types: np.ndarray       # (n,) int32
    box: np.ndarray         # (3,) float64
    masses: np.ndarray      # (num_types,)
    name: str = ""

    @property
    def n(self) -> int:
        return int(self.types.size)

    @property
    def num_types(self) -> int:
        return int(self.masses.size)


def _maxwell(rng: np.random.Generator, n: int, temperature: float, mass_per_particle: np.ndarray) -> np.ndarray:
    v = rng.normal(size=(n, 3)) * np.sqrt(temperature / mass_per_particle)[:, None]
    p = (v * mass_per_particle[:, None]).sum(axis=0) / mass_per_particle.sum()
    v -= p
    return v.astype(np.float32)


def _wrap(x: np.ndarray, L: np.ndarray) -> np.ndarray:
    return (x - L * np.rint(x / L)).astype(np.float32)


def simple_cubic(cells: int, density: float):
    n = cells ** 3
    L = (n / density) ** (1.0 / 3.0)
    a = L / cells
    g = np.arange(cells) * a + 0.5 * a - 0.5 * L
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1), np.array([L, L, L])


def fcc(cells: int, density: float):
    n = 4 * cells ** 3
    L = (n / density) ** (1.0 / 3.0)
    a = L / cells
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) * a
    g = np.arange(cells) * a
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    origins = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    pos = (origins[:, None, :] + basis[None, :, :]).reshape(-1, 3) + 0.25 * a - 0.5 * L
    return pos, np.array([L, L, L])


def lj_liquid(n_side: int = 10, density: float = 0.8442, temperature: float = 0.72, jitter: float = 0.05,
              seed_pos: int = 1001, seed_vel: int = 1002) -> System:
    """Single-component LJ start: n_side^3 simple-cubic sites with seeded uniform
    jitter of +-jitter per coordinate and Maxwell velocities at T."""
    pos, L = simple_cubic(n_side, density)
    rng = np.random.default_rng(seed_pos)
    pos = pos + rng.uniform(-jitter, jitter, size=pos.shape)
    n = pos.shape[0]
    vel = _maxwell(np.random.default_rng(seed_vel), n, temperature, np.ones(n))
    return System(_wrap(pos, L), vel, np.zeros(n, dtype=np.int32), L, np.ones(1), f"LJ_N{n}")


def ka_random_lattice(n: int, density: float = 1.2, temperature: float = 0.45, seed: int = 2001,
                      lattice: str = "sc") -> System:
    """Kob-Andersen 80:20 start: n sites picked by seeded permutation from the
    smallest simple-cubic lattice holding n (or an exact FCC lattice when
    n = 4 k^3 and lattice='fcc'), round(0.8 n) A particles chosen by a second
    permutation, Maxwell velocities at T."""
    rng = np.random.default_rng(seed)
    L = (n / density) ** (1.0 / 3.0)
    if lattice == "fcc":
        k = round((n / 4) ** (1.0 / 3.0))
        if 4 * k ** 3 != n:
            raise ValueError("FCC start needs n = 4 k^3")
        pos, _ = fcc(k, density)
    else:
        # k^3 >= n simple-cubic sites spanning the box of the target density
        k = math.ceil(n ** (1.0 / 3.0) - 1e-9)
        a = L / k
        g = np.arange(k) * a + 0.5 * a - 0.5 * L
        x, y, z = np.meshgrid(g, g, g, indexing="ij")
        sites = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
        pos = sites[np.sort(rng.permutation(sites.shape[0])[:n])]
    n_a = int(round(0.8 * n))
    types = np.ones(n, dtype=np.int32)
    types[rng.permutation(n)[:n_a]] = 0
    vel = _maxwell(np.random.default_rng(seed + 2), n, temperature, np.ones(n))
    box = np.array([L, L, L])
    return System(_wrap(pos, box), vel, types, box, np.ones(2), f"KA_N{n}")


def ka_replica_set(nrep: int = 512, n_min: int = 256, n_max: int = 8192, t_min: float = 0.4, t_max: float = 2.0,
                   density: float = 1.2, seed_n: int = 4001, seed_t: int = 4002, seed_start: int = 4100):
    """Batched-replica workload: nrep Kob-Andersen replicas with N drawn
    seeded-uniform in [n_min, n_max] and T uniform in [t_min, t_max]; replica r
    starts from ka_random_lattice(N_r, density, T_r, seed=seed_start + r).
    Returns (systems, temperatures)."""
    ns = np.random.default_rng(seed_n).integers(n_min, n_max + 1, size=nrep)
    ts = np.random.default_rng(seed_t).uniform(t_min, t_max, size=nrep)
    systems = [ka_random_lattice(int(n), density=density, temperature=float(t), seed=seed_start + r)
               for r, (n, t) in enumerate(zip(ns, ts))]
    return systems, [float(t) for t in ts]


def polymer_melt(n_chains: int = 10000, beads: int = 10, density: float = 1.0, temperature: float = 1.0,
                 seed: int = 5001, grid=(50, 50, 40)):
    """Linear bead-spring chains on a lattice (the polymer workload: 10^4
    chains of 10 beads, rho = 1.0, N = 100,000).  Sites of a grid[0] x grid[1]
    x grid[2] lattice filling the cubic box; each chain takes `beads`
    consecutive sites along x, so bonds start at the x spacing (L / grid[0]).
    Returns (System, bonds (m, 3) as (a, b, 0), angles (m, 4) as (a, b, c, 0)).
    Velocities are Maxwellian at `temperature` (seed) with zero total momentum."""
    gx, gy, gz = grid
    n = n_chains * beads
    if gx * gy * gz != n or gx % beads:
        raise ValueError("the lattice must hold the chains exactly, with whole chains along x")
    L = (n / density) ** (1.0 / 3.0)
    ix, iy, iz = np.meshgrid(np.arange(gx), np.arange(gy), np.arange(gz), indexing="ij")
    # x fastest within a chain: order sites by (z, y, x)
    order = np.lexsort((ix.ravel(), iy.ravel(), iz.ravel()))
    pos = np.stack([(ix.ravel()[order] + 0.5) * L / gx, (iy.ravel()[order] + 0.5) * L / gy,
                    (iz.ravel()[order] + 0.5) * L / gz], axis=1) - L / 2
    rng = np.random.default_rng(seed)
    types = np.zeros(n, np.int32)
    masses = np.ones(1)
    vel = _maxwell(rng, n, temperature, masses[types])
    first = np.arange(n_chains) * beads
    bonds = np.stack([(first[:, None] + np.arange(beads - 1)).ravel(),
                      (first[:, None] + np.arange(1, beads)).ravel(), np.zeros(n_chains * (beads - 1), int)], axis=1)
    angles = np.stack([(first[:, None] + np.arange(beads - 2)).ravel(),
                       (first[:, None] + np.arange(1, beads - 1)).ravel(),
                       (first[:, None] + np.arange(2, beads)).ravel(), np.zeros(n_chains * (beads - 2), int)], axis=1)
    sysm = System(_wrap(pos, np.array([L, L, L])).astype(np.float32), vel.astype(np.float32), types,
                  np.array([L, L, L]), masses, name=f"polymer_{n_chains}x{beads}")
    return sysm, bonds.astype(np.int64), angles.astype(np.int64)


def ka_params(rcut_in_sigma: float = 2.5):
    """(i, j, sigma, epsilon, rcut-in-units-of-sigma) rows for Pot_LJ_12_6.SetParams."""
    return [(i, j, KA_SIGMA[(i, j)], KA_EPSILON[(i, j)], rcut_in_sigma) for i in (0, 1) for j in (0, 1)]


def lj_param_table(rows, num_types: int) -> np.ndarray:
    """RUMD's per-pair parameter table (num_types^2 x 12 float32) for the LJ
    12-6 potential: Pot_LJ_12_6::SetParams (include/rumd/PairPotential.h:202-207)
    gives Rcut = rc*sigma, A12 = 4 eps sigma^12, A6 = -4 eps sigma^6; the shifts
    param[11] = u(rc), param[10] = s(rc) rc (src/PairPotential.cu:82-125) are
    formed here in FP64 from those float values.  Pairs not given stay zero;
    with Newton's third law the (j, i) row mirrors (i, j)."""
    tab = np.zeros((num_types * num_types, 12), dtype=np.float32)
    for (i, j, sigma, eps, rcut) in rows:
        rc = np.float32(np.float32(rcut) * np.float32(sigma))
        a12 = np.float32(4.0 * float(np.float32(eps)) * float(np.float32(sigma)) ** 12)
        a6 = np.float32(-4.0 * float(np.float32(eps)) * float(np.float32(sigma)) ** 6)
        row = np.zeros(12, dtype=np.float32)
        row[0], row[1], row[2] = rc, a12, a6
        if rc > 0:
            r2 = float(rc) * float(rc)
            i6 = 1.0 / r2 ** 3
            u = i6 * (float(a12) * i6 + float(a6))
            s = i6 / r2 * (12.0 * float(a12) * i6 + 6.0 * float(a6))
            row[11] = np.float32(u)
            row[10] = np.float32(s * float(rc))
        tab[i * num_types + j] = row
        tab[j * num_types + i] = row
    return tab


def cut_skin2_table(param_table: np.ndarray, skin: float) -> np.ndarray:
    """(rc_ij + skin)^2 formed in FP32 (src/NeighborList.cu:1239-1240)."""
    rc = param_table[:, 0].astype(np.float32)
    cs = (rc + np.float32(skin)).astype(np.float32)
    return (cs * cs).astype(np.float32)


def box8(L) -> np.ndarray:
    """RUMD's float[8] box (include/rumd/SimulationBox.h:165-178)."""
    L = np.asarray(L, dtype=np.float64).reshape(-1)
    if L.size == 1:
        L = np.repeat(L, 3)
    Lf = L.astype(np.float32)
    cubic = 1.0 if (abs(Lf[0] / Lf[1] - 1.0) < 1e-8 and abs(Lf[1] / Lf[2] - 1.0) < 1e-8) else 0.0
    b = np.zeros(8, dtype=np.float32)
    b[0] = cubic
    b[1:4] = Lf
    b[4:7] = (np.float32(1.0) / Lf).astype(np.float32)
    return b


def pack_pos4(positions: np.ndarray, types: np.ndarray) -> np.ndarray:
    """float4 rows (x, y, z, type bits) as RUMD stores them (ParticleData.h:41)."""
    n = types.size
    out = np.zeros((n, 4), dtype=np.float32)
    out[:, :3] = positions
    out[:, 3] = np.ascontiguousarray(types, dtype=np.int32).view(np.float32)
    return out


def pack_vel4(velocities: np.ndarray, types: np.ndarray, masses: np.ndarray) -> np.ndarray:
    n = types.size
    out = np.zeros((n, 4), dtype=np.float32)
    out[:, :3] = velocities
    out[:, 3] = (1.0 / np.asarray(masses, dtype=np.float64)[types]).astype(np.float32)
    return out
