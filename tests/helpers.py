"""Shared helpers for the parity tests."""
from __future__ import annotations

import numpy as np

from rumd_amd import configs, conf_io

# Per-particle force tolerance against the FP64 oracle, guarded metric
# |dF_i| / max(|F_i|, F_rms) (FP32 pair arithmetic, FP64 displacement).
FORCE_TOL_GUARDED = 1e-5


def force_errors(f_test: np.ndarray, f_ref: np.ndarray) -> dict:
    """Per-particle relative force errors: unguarded |dF|/|F| and guarded
    |dF|/max(|F|, F_rms)."""
    d = np.linalg.norm(np.asarray(f_test, dtype=np.float64) - f_ref, axis=1)
    mag = np.linalg.norm(f_ref, axis=1)
    frms = np.sqrt(np.mean(mag ** 2))
    unguarded = d / np.maximum(mag, 1e-300)
    guarded = d / np.maximum(mag, frms)
    return {
        "guarded_max": float(guarded.max()),
        "unguarded_max": float(unguarded.max()),
        "unguarded_p99": float(np.percentile(unguarded, 99)),
        "unguarded_median": float(np.median(unguarded)),
    }


def load_golden(path):
    c = conf_io.read_conf(path)
    return c


def ka_table(rcut=2.5):
    return configs.lj_param_table(configs.ka_params(rcut), 2)


def lj_table(rcut=2.5):
    return configs.lj_param_table([(0, 0, 1.0, 1.0, rcut)], 1)


def csr_to_sets(off, idx):
    return [set(idx[off[i]:off[i + 1]].tolist()) for i in range(off.size - 1)]


def gpu_sets_original_ids(cnt, nbl, uidx):
    """Neighbour sets in original particle ids from a slot-ordered GPU list."""
    n = cnt.size
    sets = [None] * n
    for s in range(n):
        js = nbl[: cnt[s], s]
        sets[uidx[s]] = set(uidx[js].tolist())
    return sets


def unsort(a_slot: np.ndarray, uidx: np.ndarray) -> np.ndarray:
    out = np.empty_like(a_slot)
    out[uidx] = a_slot
    return out


def sorted_rows(off, idx) -> np.ndarray:
    """The rows of a CSR list (offsets[n+1], ids) as a dense (n, longest) int32
    array, each row sorted ascending and padded with INT32_MAX: two lists hold
    the same neighbour sets exactly when these arrays are equal (vectorised,
    for lists of a million particles)."""
    off = np.asarray(off, dtype=np.int64)
    idx = np.asarray(idx)
    counts = np.diff(off)
    n = counts.size
    width = int(counts.max()) if n else 0
    dense = np.full((n, max(width, 1)), np.iinfo(np.int32).max, dtype=np.int32)
    rows = np.repeat(np.arange(n, dtype=np.int64), counts)
    cols = np.arange(idx.size, dtype=np.int64) - np.repeat(off[:-1], counts)
    dense[rows, cols] = idx
    dense.sort(axis=1)
    return dense
