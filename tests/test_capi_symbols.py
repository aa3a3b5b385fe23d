"""The C-ABI library loads on a machine without a GPU and exports every entry
point include/rumdx.h declares.  No compute call is made here."""
from __future__ import annotations

import ctypes as C
import subprocess

import numpy as np

import capi


def test_header_declares_entry_points():
    names = capi.declared_symbols()
    assert "rumdx_pair_lj" in names and "rumdx_nvt_step" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", str(capi.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [s for s in capi.declared_symbols() if s not in exported]
    assert not missing, missing
    lib = capi.lib()
    for s in capi.declared_symbols():
        assert getattr(lib, s) is not None


def test_host_side_helpers_without_gpu():
    """Entry points that only do host arithmetic: version, sizes, cell grid,
    Morton rank table; argument validation reports errors."""
    lib = capi.lib()
    assert lib.rumdx_version() == 1
    assert lib.rumdx_workspace_bytes(1 << 20) >= 256 + 8 * 4 * 4096
    nc = np.zeros(3, dtype=np.int32)
    from rumd_amd import configs
    assert lib.rumdx_cell_grid(configs.box8(95.6033), 2.5, 0.3, nc) == 0
    assert list(nc) == [68, 68, 68]
    rank = np.zeros(int(nc.prod()), dtype=np.uint32)
    assert lib.rumdx_cell_rank_table(nc, rank) == 0
    assert np.array_equal(np.sort(rank), np.arange(rank.size, dtype=np.uint32))
    # Morton order: the first 8 ranks are the 2x2x2 corner block
    first = np.nonzero(rank < 8)[0]
    assert sorted(first.tolist()) == [0, 1, 68, 69, 68 * 68, 68 * 68 + 1, 68 * 68 + 68, 68 * 68 + 69]
    assert lib.rumdx_pair_lj(None, None, None, 2, configs.box8(10.0), 1, 1, 0, None) == -1
    assert lib.rumdx_pair_lj_fp64(None, None, None, 2, configs.box8(10.0), 1, 1, 0, None) == -1
    assert b"null" in lib.rumdx_last_error()
    # lanes per particle must be a power of two <= 8: rejected before any launch
    p = capi.Particles()
    p.r_dev = p.f_dev = p.w_dev = 16
    p.n, p.nvp = 10, 32
    nb = capi.NbList()
    nb.nbl_dev = nb.num_nbrs_dev = 16
    nb.max_nbrs, nb.lanes_per_particle = 8, 3
    assert lib.rumdx_pair_lj(C.byref(p), C.byref(nb), 16, 2, configs.box8(10.0), 1, 1, 0, None) == -1
    assert b"lanes_per_particle" in lib.rumdx_last_error()
    bad = np.zeros(3, dtype=np.int32)
    assert lib.rumdx_cell_rank_table(bad, rank) == -1


def test_tile_table_bricks_are_contiguous_rank_ranges():
    """Aligned 4^3 bricks of Morton-ranked cells are contiguous rank ranges,
    ordered by rank; edge bricks of grids that are not multiples of 4 too."""
    lib = capi.lib()
    for nc_list in ([68, 68, 68], [27, 27, 27], [6, 7, 5], [9, 12, 10]):
        nc = np.array(nc_list, dtype=np.int32)
        ncells = int(nc.prod())
        rank = np.zeros(ncells, dtype=np.uint32)
        assert lib.rumdx_cell_rank_table(nc, rank) == 0
        nt = lib.rumdx_tile_table(nc, None, None, None)
        assert nt == int(np.prod((nc + 3) // 4))
        first = np.zeros(nt + 1, dtype=np.uint32)
        origin = np.zeros(nt, dtype=np.uint32)
        assert lib.rumdx_tile_table(nc, rank.ctypes.data, first.ctypes.data, origin.ctypes.data) == nt
        assert first[0] == 0 and first[-1] == ncells and np.all(np.diff(first.astype(np.int64)) > 0)
        inv = np.argsort(rank)  # rank -> linear cell
        for t in range(nt):
            cells = inv[first[t]:first[t + 1]]
            x, y, z = cells % nc[0], (cells // nc[0]) % nc[1], cells // (nc[0] * nc[1])
            ox, oy, oz = origin[t] & 1023, (origin[t] >> 10) & 1023, (origin[t] >> 20) & 1023
            assert x.min() == ox and y.min() == oy and z.min() == oz
            bw = [min(4, nc[a] - o) for a, o in enumerate((ox, oy, oz))]
            assert cells.size == bw[0] * bw[1] * bw[2]
            assert x.max() < ox + 4 and y.max() < oy + 4 and z.max() < oz + 4
    cap = lib.rumdx_halo_cap_estimate(np.array([68, 68, 68], np.int32), 1 << 20)
    assert cap % 64 == 0 and 512 * (1 << 20) / 68 ** 3 < cap <= 3072


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the ABI structs have the C compiler's sizes and offsets."""
    src = tmp_path / "layout.c"
    fields = {
        "rumdx_particles": capi.Particles,
        "rumdx_nblist": capi.NbList,
        "rumdx_nh_state": capi.NhState,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "rumdx.h"', "int main(void) {"]
    for cname, py in fields.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", f"-I{capi.ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    got = dict(ln.split() for ln in subprocess.run([str(exe)], capture_output=True, text=True,
                                                   check=True).stdout.splitlines())
    for cname, py in fields.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)
