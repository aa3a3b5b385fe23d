"""ctypes front end of the CPU oracle (rumd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the rumd_amd package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

_lib = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build() -> Path:
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.orc_forces_brute.argtypes = [C.c_int, f32p, f64p, C.c_int, f32p, C.c_int, f64p, f64p]
        L.orc_stress_brute.argtypes = [C.c_int, f32p, f64p, C.c_int, f32p, C.c_int, f64p]
        L.orc_forces_form_brute.argtypes = [C.c_int, f32p, f64p, C.c_int, f32p, C.c_int, C.c_int, f64p, f64p]
        L.orc_bonds.argtypes = [C.c_int, f32p, f64p, C.c_int, i32p, f32p, C.c_int, f64p, f64p]
        L.orc_dihedrals_ryckaert.argtypes = [C.c_int, f32p, f64p, C.c_int, i32p, f32p, f64p]
        L.orc_angles_cossq.argtypes = [C.c_int, f32p, f64p, C.c_int, i32p, f32p, f64p]
        L.orc_angles_sq.argtypes = [C.c_int, f32p, f64p, C.c_int, i32p, f32p, f64p]
        L.orc_dihedrals_periodic.argtypes = [C.c_int, f32p, f64p, C.c_int, i32p, f32p, f64p]
        L.orc_forces_nl.argtypes = [C.c_int, f32p, f64p, C.c_int, f32p, C.c_int, i64p, i32p, f64p, f64p]
        L.orc_forces_cells.argtypes = [C.c_int, f32p, f64p, C.c_int, f32p, C.c_int, f64p, f64p]
        L.orc_forces_cells.restype = C.c_int
        L.orc_nblist_brute_f32.argtypes = [C.c_int, f32p, f32p, C.c_int, f32p, i32p, C.c_void_p, C.c_void_p]
        L.orc_cell_grid.argtypes = [f32p, C.c_float, C.c_float, i32p]
        L.orc_cell_of.argtypes = [C.c_int, f32p, f32p, i32p, i32p]
        L.orc_nblist_cells_f32.argtypes = [C.c_int, f32p, f32p, C.c_int, f32p, i32p, i32p, C.c_void_p, C.c_void_p]
        L.orc_skin_exceeded.argtypes = [C.c_int, f32p, f32p, f32p, C.c_float]
        L.orc_skin_exceeded.restype = C.c_int
        L.orc_nvt_step.argtypes = [C.c_int, f32p, f32p, f32p, f32p, C.c_void_p, f32p, C.c_float, C.c_double,
                                   C.c_uint32, C.POINTER(C.c_double)]
        L.orc_thermostat_update.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double, C.c_double,
                                            C.c_uint32, C.c_float, C.c_float]
        L.orc_kinetic_energy.argtypes = [C.c_int, f32p, f32p, f32p, C.c_float, C.c_double, C.c_uint32]
        L.orc_kinetic_energy.restype = C.c_double
        L.orc_sum_f4.argtypes = [C.c_int, f32p, C.c_int]
        L.orc_sum_f4.restype = C.c_double
        L.orc_md_run.argtypes = [C.c_int, f64p, f64p, i32p, f64p, f64p, C.c_int, f32p, C.c_int, C.c_double,
                                 C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]
        L.orc_md_run.restype = C.c_int
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _box3(L) -> np.ndarray:
    """Box lengths as RUMD holds them: float32 values (SimulationBox stores a
    float[8]), widened to double for the FP64 restatements."""
    L = np.asarray(L, dtype=np.float64).reshape(-1)
    if L.size == 1:
        L = np.repeat(L, 3)
    return np.ascontiguousarray(L[:3].astype(np.float32).astype(np.float64))


def forces_brute(pos4, L, nt, params, cutoff):
    """FP64 forces over all pairs: returns (f4 [n,4] float64, vir [n])."""
    pos4 = _f32(pos4)
    n = pos4.shape[0]
    f4 = np.zeros((n, 4))
    vir = np.zeros(n)
    lib().orc_forces_brute(n, pos4, _box3(L), nt, _f32(params), cutoff, f4, vir)
    return f4, vir


def forces_form_brute(pos4, L, nt, params, form, cutoff):
    """FP64 forces of a pair form (RUMDX_POT_* number) over all pairs:
    (f4 [n,4], vir [n]) as forces_brute."""
    pos4 = _f32(pos4)
    n = pos4.shape[0]
    f4 = np.zeros((n, 4))
    vir = np.zeros(n)
    lib().orc_forces_form_brute(n, pos4, _box3(L), nt, _f32(params), form, cutoff, f4, vir)
    return f4, vir


def bonded(pos4, L, bonds=None, bond_params=None, angles=None, angle_params=None, fene=False, dihedrals=None,
           dihedral_params=None, angle_sq=False, periodic_dihedrals=False):
    """FP64 intramolecular forces: bonds (m, 3) int (a, b, type) with
    bond_params (l0 or R0, k) per type (harmonic, or FENE when fene); angles
    (m, 4) int (a, b, c, type) with angle_params (theta0, k) per type;
    dihedrals (m, 5) int (a, b, c, d, type) with dihedral_params c_0..c_5 per
    type (psi0, k1, k2, k3 with periodic_dihedrals); angle_sq selects the
    squared-angle form for the angles.  Returns
    (f4 [n,4], vir [n])."""
    pos4 = _f32(pos4)
    n = pos4.shape[0]
    f4 = np.zeros((n, 4))
    vir = np.zeros(n)
    if bonds is not None and len(bonds):
        b = np.ascontiguousarray(bonds, dtype=np.int32).reshape(-1, 3)
        lib().orc_bonds(n, pos4, _box3(L), b.shape[0], b, _f32(bond_params).reshape(-1), int(fene), f4, vir)
    if dihedrals is not None and len(dihedrals):
        d = np.ascontiguousarray(dihedrals, dtype=np.int32).reshape(-1, 5)
        dp = np.zeros((len(dihedral_params), 6), np.float32)  # six per type, zero-padded
        for t, row in enumerate(dihedral_params):
            dp[t, :len(row)] = row
        fn = lib().orc_dihedrals_periodic if periodic_dihedrals else lib().orc_dihedrals_ryckaert
        fn(n, pos4, _box3(L), d.shape[0], d, dp.reshape(-1), f4)
    if angles is not None and len(angles):
        a = np.ascontiguousarray(angles, dtype=np.int32).reshape(-1, 4)
        fn = lib().orc_angles_sq if angle_sq else lib().orc_angles_cossq
        fn(n, pos4, _box3(L), a.shape[0], a, _f32(angle_params).reshape(-1), f4)
    return f4, vir


def stress_brute(pos4, L, nt, params, cutoff):
    """FP64 per-particle configurational stress over all pairs: [n, 6] as
    (sxx, syy, szz, syz, sxz, sxy), each -sum_j d_a d_b s_ij."""
    pos4 = _f32(pos4)
    n = pos4.shape[0]
    sts = np.zeros((n, 6))
    lib().orc_stress_brute(n, pos4, _box3(L), nt, _f32(params), cutoff, sts)
    return sts


def forces_nl(pos4, L, nt, params, cutoff, offsets, idx):
    pos4 = _f32(pos4)
    n = pos4.shape[0]
    f4 = np.zeros((n, 4))
    vir = np.zeros(n)
    lib().orc_forces_nl(n, pos4, _box3(L), nt, _f32(params), cutoff,
                        np.ascontiguousarray(offsets, dtype=np.int64), np.ascontiguousarray(idx, dtype=np.int32),
                        f4, vir)
    return f4, vir


def forces_cells(pos4, L, nt, params, cutoff):
    pos4 = _f32(pos4)
    n = pos4.shape[0]
    f4 = np.zeros((n, 4))
    vir = np.zeros(n)
    rc = lib().orc_forces_cells(n, pos4, _box3(L), nt, _f32(params), cutoff, f4, vir)
    if rc != 0:
        raise ValueError("box too small for the cell-list oracle")
    return f4, vir


def _csr_from_counts(counts):
    off = np.zeros(counts.size + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    return off


def nblist_brute_f32(pos4, box8, nt, cut_skin2):
    """RUMD N2 criterion in FP32: returns CSR (offsets, idx) over input order."""
    pos4 = _f32(pos4)
    n = pos4.shape[0]
    counts = np.zeros(n, dtype=np.int32)
    L = lib()
    L.orc_nblist_brute_f32(n, pos4, _f32(box8), nt, _f32(cut_skin2), counts, None, None)
    off = _csr_from_counts(counts)
    idx = np.zeros(max(int(off[-1]), 1), dtype=np.int32)
    L.orc_nblist_brute_f32(n, pos4, _f32(box8), nt, _f32(cut_skin2), counts, off.ctypes.data, idx.ctypes.data)
    return off, idx[: off[-1]]


def cell_grid(box8, max_cutoff, skin):
    nc = np.zeros(3, dtype=np.int32)
    lib().orc_cell_grid(_f32(box8), max_cutoff, skin, nc)
    return nc


def cell_of(pos4, box8, nc):
    pos4 = _f32(pos4)
    out = np.zeros(pos4.shape[0], dtype=np.int32)
    lib().orc_cell_of(pos4.shape[0], pos4, _f32(box8), np.ascontiguousarray(nc, dtype=np.int32), out)
    return out


def nblist_cells_f32(pos4, box8, nt, cut_skin2, nc):
    """RUMD sorted-cell criterion in FP32: returns CSR (offsets, idx) over input order."""
    pos4 = _f32(pos4)
    n = pos4.shape[0]
    nc = np.ascontiguousarray(nc, dtype=np.int32)
    counts = np.zeros(n, dtype=np.int32)
    L = lib()
    L.orc_nblist_cells_f32(n, pos4, _f32(box8), nt, _f32(cut_skin2), nc, counts, None, None)
    off = _csr_from_counts(counts)
    idx = np.zeros(max(int(off[-1]), 1), dtype=np.int32)
    L.orc_nblist_cells_f32(n, pos4, _f32(box8), nt, _f32(cut_skin2), nc, counts, off.ctypes.data, idx.ctypes.data)
    return off, idx[: off[-1]]


def skin_exceeded(pos4, last4, box8, skin) -> bool:
    pos4 = _f32(pos4)
    return bool(lib().orc_skin_exceeded(pos4.shape[0], pos4, _f32(last4), _f32(box8), skin))


def nvt_step(pos4, vel4, f4, im4, box8, dt, ps, on_mask, rlo4=None):
    """One FP32-exact leapfrog/NH step in place; returns the FP64 KE2 sum."""
    ke2 = C.c_double(0.0)
    rlo_ptr = None if rlo4 is None else rlo4.ctypes.data
    lib().orc_nvt_step(pos4.shape[0], pos4, vel4, _f32(f4), im4, rlo_ptr, _f32(box8), dt, ps, on_mask,
                       C.byref(ke2))
    return ke2.value


def thermostat_update(ps, target_T, dT, ke2, dof, dt, omega2):
    p = C.c_double(ps)
    t = C.c_double(target_T)
    lib().orc_thermostat_update(C.byref(p), C.byref(t), dT, ke2, dof, dt, omega2)
    return p.value, t.value


def kinetic_energy(pos4, vel4, f4, dt, ps, on_mask) -> float:
    pos4 = _f32(pos4)
    return lib().orc_kinetic_energy(pos4.shape[0], pos4, _f32(vel4), _f32(f4), dt, ps, on_mask)


def sum_f4(a4, comp) -> float:
    a4 = _f32(a4)
    return lib().orc_sum_f4(a4.shape[0], a4, comp)


def md_run(pos, vel, types, inv_mass, L, nt, params, cutoff, skin, dt, nsteps, thermostat, target_T, omega2, ps):
    """FP64 CPU MD (the CPU baseline).  Arrays are modified in place.
    Returns (builds, pe, ps)."""
    ps_c = C.c_double(ps)
    pe_c = C.c_double(0.0)
    b = lib().orc_md_run(pos.shape[0], pos, vel, np.ascontiguousarray(types, dtype=np.int32),
                         np.ascontiguousarray(inv_mass, dtype=np.float64), _box3(L), nt, _f32(params), cutoff,
                         skin, dt, nsteps, int(thermostat), target_T, omega2, C.byref(ps_c), C.byref(pe_c))
    if b < 0:
        raise RuntimeError(f"orc_md_run failed ({b})")
    return b, pe_c.value, ps_c.value


def num_threads() -> int:
    return lib().orc_num_threads()


def set_num_threads(k: int) -> None:
    lib().orc_set_num_threads(k)
