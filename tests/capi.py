"""ctypes bindings of the rumdx C ABI (include/rumdx.h) and of the few HIP
runtime calls the kernel-level tests need for device memory.

The structures below mirror the header field by field; test_capi_symbols.py
checks their sizes against the header's layout rules.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
LIB_PATH = ROOT / "rumd_amd" / "librumdx.so"
HEADER = ROOT / "include" / "rumdx.h"

RUMDX_FLAG_REBUILD, RUMDX_FLAG_OVERFLOW, RUMDX_FLAG_BUILDS, RUMDX_FLAG_MAXLEN, RUMDX_FLAG_DONE = range(5)


class Particles(C.Structure):
    _fields_ = [
        ("r_dev", C.c_void_p),
        ("v_dev", C.c_void_p),
        ("f_dev", C.c_void_p),
        ("w_dev", C.c_void_p),
        ("im_dev", C.c_void_p),
        ("sts_dev", C.c_void_p),
        ("rlo_dev", C.c_void_p),
        ("unsorted_index_dev", C.c_void_p),
        ("n", C.c_uint32),
        ("nvp", C.c_uint32),
    ]


class NbList(C.Structure):
    _fields_ = [
        ("nbl_dev", C.c_void_p),
        ("num_nbrs_dev", C.c_void_p),
        ("last_r_dev", C.c_void_p),
        ("flags_dev", C.c_void_p),
        ("cut_skin2_dev", C.c_void_p),
        ("max_nbrs", C.c_uint32),
        ("ntypes", C.c_uint32),
        ("skin", C.c_float),
        ("nc", C.c_int32 * 3),
        ("cell_rank_dev", C.c_void_p),
        ("cell_start_dev", C.c_void_p),
        ("cell_count_dev", C.c_void_p),
        ("slot_cell_dev", C.c_void_p),
        ("sort_scratch_dev", C.c_void_p),
        ("nbl16_dev", C.c_void_p),
        ("ntiles", C.c_uint32),
        ("tile_first_rank_dev", C.c_void_p),
        ("tile_origin_dev", C.c_void_p),
        ("halo_cap", C.c_uint32),
        ("lane_order_dev", C.c_void_p),
        ("tile_halo_dev", C.c_void_p),
        ("tile_info_dev", C.c_void_p),
        ("lanes_per_particle", C.c_uint32),
        ("excl_start_dev", C.c_void_p),
        ("excl_ids_dev", C.c_void_p),
        ("nbl16_in_dev", C.c_void_p),
        ("num_in_dev", C.c_void_p),
        ("prune_r_dev", C.c_void_p),
        ("cut_in2_dev", C.c_void_p),
        ("inner_skin", C.c_float),
        ("tile_waves_per_eu", C.c_uint32),
    ]


class NhState(C.Structure):
    _fields_ = [("ps", C.c_double), ("target_T", C.c_double), ("dT_per_step", C.c_double), ("ke2_sum", C.c_double)]


_lib = None


def declared_symbols() -> list[str]:
    """Function names declared in include/rumdx.h."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char\*)\s+(rumdx_\w+)\s*\(", text, flags=re.M)))


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(str(LIB_PATH))
    vp, u32, f32, i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
    PP, NP, NHP = C.POINTER(Particles), C.POINTER(NbList), C.POINTER(NhState)
    f32a = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    i32a = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    u32a = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
    L.rumdx_version.restype = C.c_int
    L.rumdx_last_error.restype = C.c_char_p
    L.rumdx_workspace_bytes.argtypes = [u32]
    L.rumdx_workspace_bytes.restype = C.c_size_t
    L.rumdx_sort_scratch_bytes.argtypes = [u32, u32]
    L.rumdx_sort_scratch_bytes.restype = C.c_size_t
    L.rumdx_cell_grid.argtypes = [f32a, f32, f32, i32a]
    L.rumdx_cell_rank_table.argtypes = [i32a, u32a]
    L.rumdx_nblist_check.argtypes = [PP, NP, f32a, vp]
    L.rumdx_nblist_request_rebuild.argtypes = [NP, vp]
    L.rumdx_nblist_build_n2.argtypes = [PP, NP, f32a, vp]
    L.rumdx_nblist_build_sorted.argtypes = [PP, NP, f32a, vp]
    L.rumdx_nblist_rebuild_sorted.argtypes = [PP, NP, f32a, vp]
    L.rumdx_nblist_build_stats.argtypes = [PP, NP, vp]
    L.rumdx_nblist_prune.argtypes = [PP, NP, f32a, vp]
    L.rumdx_pair_lj.argtypes = [PP, NP, vp, u32, f32a, i32, i32, i32, vp]
    L.rumdx_pair_lj_fp64.argtypes = [PP, NP, vp, u32, f32a, i32, i32, i32, vp]
    L.rumdx_pair_form.argtypes = [PP, NP, vp, u32, f32a, i32, i32, i32, i32, vp]
    L.rumdx_nvt_step.argtypes = [PP, NHP, f32, u32, f32, u32, f32a, NP, vp, vp]
    L.rumdx_pair_lj_nvt.argtypes = [PP, NP, vp, u32, f32a, i32, vp, NHP, f32, u32, f32, u32, vp, vp]
    L.rumdx_zero_momentum.argtypes = [PP, vp, vp]
    L.rumdx_kinetic_energy.argtypes = [PP, NHP, f32, u32, vp, vp, vp]
    L.rumdx_reduce_f4.argtypes = [vp, u32, u32, vp, vp, vp]
    L.rumdx_tile_table.argtypes = [i32a, vp, vp, vp]
    L.rumdx_halo_cap_estimate.argtypes = [i32a, u32]
    L.rumdx_nblist_export.argtypes = [PP, NP, vp, vp]
    for name in declared_symbols():
        if name not in ("rumdx_version", "rumdx_last_error", "rumdx_workspace_bytes", "rumdx_sort_scratch_bytes",
                        "rumdx_halo_cap_estimate"):
            getattr(L, name).restype = C.c_int
    L.rumdx_halo_cap_estimate.restype = u32
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(f"rumdx error {rc}: {lib().rumdx_last_error().decode()}")


# ---------------------------------------------------------------- HIP runtime (device memory only)
_hip = None


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        H = C.CDLL("libamdhip64.so")
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipDeviceSynchronize.argtypes = []
        H.hipGetErrorString.restype = C.c_char_p
        H.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        _hip = H
    return _hip


def hip_check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


def gpu_available() -> bool:
    try:
        n = C.c_int(0)
        return hip().hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


class DeviceBuffer:
    """A hipMalloc'd buffer with numpy upload/download."""

    def __init__(self, nbytes: int):
        self.nbytes = max(int(nbytes), 16)
        self.ptr = C.c_void_p()
        hip_check(hip().hipMalloc(C.byref(self.ptr), self.nbytes), "hipMalloc")
        self.zero()

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        hip_check(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        hip_check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def zero(self) -> None:
        z = np.zeros(self.nbytes, dtype=np.uint8)
        hip_check(hip().hipMemcpy(self.ptr, z.ctypes.data, self.nbytes, 1), "hipMemcpy H2D")

    @property
    def value(self) -> int:
        return self.ptr.value

    def __del__(self):
        try:
            if self.ptr and _hip is not None:
                _hip.hipFree(self.ptr)
        except Exception:
            pass


def sync() -> None:
    hip_check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class GpuSystem:
    """Device-resident particle arrays plus neighbour-list state, driven
    directly through the C ABI (no host C++ layer)."""

    def __init__(self, pos4, vel4, box8, nt, param_table, skin=0.3, max_nbrs=None, method="sorted", rlo=False,
                 with_sts=False, pad=64, halo_cap=None):
        self.halo_cap_request = halo_cap
        self.n = int(pos4.shape[0])
        self.nvp = ((self.n + pad - 1) // pad) * pad
        self.nt = nt
        self.box8 = np.ascontiguousarray(box8, dtype=np.float32)
        self.params = np.ascontiguousarray(param_table, dtype=np.float32)
        nvp = self.nvp

        def f4buf(a=None):
            b = DeviceBuffer(16 * nvp)
            if a is not None:
                b.upload(np.ascontiguousarray(a, dtype=np.float32))
            return b

        self.r = f4buf(pos4)
        self.v = f4buf(vel4)
        self.f = f4buf()
        self.w = f4buf()
        self.im = f4buf()
        self.sts = f4buf() if with_sts else None
        self.rlo = f4buf() if rlo else None
        self.idx = DeviceBuffer.from_array(np.arange(nvp, dtype=np.uint32))
        self.p = Particles(self.r.value, self.v.value, self.f.value, self.w.value, self.im.value,
                           self.sts.value if self.sts else None, self.rlo.value if self.rlo else None,
                           self.idx.value, self.n, nvp)
        self.d_params = DeviceBuffer.from_array(self.params)
        self.ws = DeviceBuffer(lib().rumdx_workspace_bytes(self.n))
        self.nh = DeviceBuffer(C.sizeof(NhState))
        self.out = DeviceBuffer(64)
        self.method = method
        self.nb = None
        if method in ("sorted", "n2"):
            self._make_nblist(skin, max_nbrs)

    def _make_nblist(self, skin, max_nbrs):
        rc = self.params[:, 0]
        cut2 = ((rc + np.float32(skin)) * (rc + np.float32(skin))).astype(np.float32)
        if max_nbrs is None:
            max_nbrs = min(self.n, 512)
        self.max_nbrs = max_nbrs
        self.d_cut2 = DeviceBuffer.from_array(cut2)
        self.nbl = DeviceBuffer(4 * self.nvp * max_nbrs) if self.method == "n2" else None
        # tile-format entries come in 16-byte groups of 8 per slot
        self.nbl16 = DeviceBuffer(2 * self.nvp * ((max_nbrs + 7) // 8 * 8)) if self.method == "sorted" else None
        self.export_buf = DeviceBuffer(4 * self.nvp * max_nbrs)
        self.num_nbrs = DeviceBuffer(4 * self.nvp)
        self.last_r = DeviceBuffer(16 * self.nvp)
        self.flags = DeviceBuffer(4 * 16)  # RUMDX_NB_FLAGS
        nb = NbList()
        nb.nbl_dev = self.nbl.value if self.nbl else None
        nb.num_nbrs_dev, nb.last_r_dev = self.num_nbrs.value, self.last_r.value
        nb.flags_dev, nb.cut_skin2_dev = self.flags.value, self.d_cut2.value
        nb.max_nbrs, nb.ntypes, nb.skin = max_nbrs, self.nt, skin
        if self.method == "sorted":
            nc = np.zeros(3, dtype=np.int32)
            check(lib().rumdx_cell_grid(self.box8, float(rc.max()), skin, nc))
            ncells = int(nc.prod())
            rank = np.zeros(ncells, dtype=np.uint32)
            check(lib().rumdx_cell_rank_table(nc, rank))
            self.nc = nc
            self.cell_rank = DeviceBuffer.from_array(rank)
            self.cell_start = DeviceBuffer(4 * (ncells + 1))
            self.cell_count = DeviceBuffer(4 * ncells)
            self.slot_cell = DeviceBuffer(4 * self.nvp)
            self.lane_order = DeviceBuffer(2 * self.nvp)
            self.sort_scratch = DeviceBuffer(lib().rumdx_sort_scratch_bytes(self.nvp, ncells))
            nb.nc[:] = [int(x) for x in nc]
            nb.cell_rank_dev, nb.cell_start_dev = self.cell_rank.value, self.cell_start.value
            nb.cell_count_dev, nb.slot_cell_dev = self.cell_count.value, self.slot_cell.value
            nb.sort_scratch_dev = self.sort_scratch.value
            nb.lane_order_dev = self.lane_order.value
            ntiles = lib().rumdx_tile_table(nc, None, None, None)
            assert ntiles > 0
            first = np.zeros(ntiles + 1, dtype=np.uint32)
            origin = np.zeros(ntiles, dtype=np.uint32)
            check(0 if lib().rumdx_tile_table(nc, rank.ctypes.data, first.ctypes.data, origin.ctypes.data) == ntiles
                  else -1)
            self.tile_first = DeviceBuffer.from_array(first)
            self.tile_origin = DeviceBuffer.from_array(origin)
            self.ntiles = ntiles
            cap = self.halo_cap_request
            if cap is None:
                cap = lib().rumdx_halo_cap_estimate(nc, self.n)
            nb.nbl16_dev = self.nbl16.value
            self.tile_halo = DeviceBuffer(4 * ntiles * max(cap, 1))
            self.tile_info = DeviceBuffer(16 * ntiles)
            nb.tile_halo_dev, nb.tile_info_dev = self.tile_halo.value, self.tile_info.value
            nb.ntiles = ntiles
            nb.tile_first_rank_dev, nb.tile_origin_dev = self.tile_first.value, self.tile_origin.value
            nb.halo_cap = cap
        self.nb = nb
        check(lib().rumdx_nblist_request_rebuild(C.byref(self.nb), None))

    def enable_inner(self, inner_skin):
        """Inner lists (sorted method): the entries within rc + inner_skin,
        pruned after builds and when the skin check asks."""
        assert self.method == "sorted"
        rc = self.params[:, 0]
        ci2 = ((rc + np.float32(inner_skin)) * (rc + np.float32(inner_skin))).astype(np.float32)
        self.d_cut_in2 = DeviceBuffer.from_array(ci2)
        self.nbl16_in = DeviceBuffer(self.nbl16.nbytes)
        self.num_in = DeviceBuffer(4 * self.nvp)
        self.prune_r = DeviceBuffer(16 * self.nvp)
        nb = self.nb
        nb.nbl16_in_dev, nb.num_in_dev = self.nbl16_in.value, self.num_in.value
        nb.prune_r_dev, nb.cut_in2_dev, nb.inner_skin = self.prune_r.value, self.d_cut_in2.value, inner_skin

    # ------------------------------------------------------------ operations
    def update_nblist(self, resort=False):
        """The skin check and the list build; with resort the sorted build is
        the tile re-sort (rumdx_nblist_rebuild_sorted), which needs a previous
        sorted build of these particles.  With inner lists, then the prune."""
        L = lib()
        check(L.rumdx_nblist_check(C.byref(self.p), C.byref(self.nb), self.box8, None))
        if self.method == "sorted":
            build = L.rumdx_nblist_rebuild_sorted if resort else L.rumdx_nblist_build_sorted
            check(build(C.byref(self.p), C.byref(self.nb), self.box8, None))
            if self.nb.nbl16_in_dev:
                check(L.rumdx_nblist_prune(C.byref(self.p), C.byref(self.nb), self.box8, None))
        else:
            check(L.rumdx_nblist_build_n2(C.byref(self.p), C.byref(self.nb), self.box8, None))

    def calc_f(self, cutoff, initialize=True, stress=False, fp64=False, form=0):
        if self.nb is not None:
            self.update_nblist()
        nbp = C.byref(self.nb) if self.nb is not None else None
        if form:
            check(lib().rumdx_pair_form(C.byref(self.p), nbp, self.d_params.value, self.nt, self.box8, form, cutoff,
                                        int(initialize), int(stress), None))
            return
        fn = lib().rumdx_pair_lj_fp64 if fp64 else lib().rumdx_pair_lj
        check(fn(C.byref(self.p), nbp, self.d_params.value, self.nt, self.box8, cutoff, int(initialize), int(stress),
                 None))

    def set_nh(self, ps=0.0, target_T=1.0, dT=0.0):
        self.nh.upload(np.frombuffer(bytes(NhState(ps, target_T, dT, 0.0)), dtype=np.uint8))

    def get_nh(self) -> NhState:
        raw = self.nh.download(np.uint8, C.sizeof(NhState))
        return NhState.from_buffer_copy(raw.tobytes())

    def nvt_step(self, dt, on_mask=0, omega2=0.0, dof=None, fuse_skin=False):
        dof = 3 * self.n - 3 if dof is None else dof
        nbp = C.byref(self.nb) if (fuse_skin and self.nb is not None) else None
        check(lib().rumdx_nvt_step(C.byref(self.p), C.cast(self.nh.value, C.POINTER(NhState)), dt, on_mask,
                                   omega2, dof, self.box8, nbp, self.ws.value, None))

    def fused_step(self, cutoff, dt, on_mask=0, omega2=0.0, dof=None, r_out=None):
        """rumdx_pair_lj_nvt: the list update, then forces and the NVT step in
        one kernel with the new positions in r_out (a buffer of nvp x 4
        floats); returns the entry point's code."""
        dof = 3 * self.n - 3 if dof is None else dof
        self.update_nblist()
        return lib().rumdx_pair_lj_nvt(C.byref(self.p), C.byref(self.nb), self.d_params.value, self.nt, self.box8,
                                       cutoff, r_out.value if r_out is not None else None,
                                       C.cast(self.nh.value, C.POINTER(NhState)), dt, on_mask, omega2, dof,
                                       self.ws.value, None)

    def reduce_f4(self, buf, comp_mask=15):
        check(lib().rumdx_reduce_f4(buf.value, self.n, comp_mask, self.out.value, self.ws.value, None))
        return self.out.download(np.float64, 4)

    # ------------------------------------------------------------ downloads (slot order)
    def get(self, name):
        b = getattr(self, name)
        return b.download(np.float32, (self.nvp, 4))[: self.n]

    def unsorted_index(self):
        return self.idx.download(np.uint32, self.nvp)[: self.n]

    def flags_host(self):
        """The list flags, after rumdx_nblist_build_stats has gathered the last
        sorted build's statistics (as the host does before it grows a list)."""
        if self.nb is not None and self.method == "sorted":
            check(lib().rumdx_nblist_build_stats(C.byref(self.p), C.byref(self.nb), None))
        return self.flags.download(np.uint32, 16)

    def neighbor_lists(self, inner=False):
        """(counts, slots) in RUMD's column-major u32 format, via
        rumdx_nblist_export; with inner, the inner lists (exported through a
        copy of the descriptor that names them as the tile lists)."""
        nb = self.nb
        counts = self.num_nbrs
        if inner:
            nb = NbList.from_buffer_copy(bytes(self.nb))
            nb.nbl16_dev, nb.num_nbrs_dev = self.nbl16_in.value, self.num_in.value
            counts = self.num_in
        check(lib().rumdx_nblist_export(C.byref(self.p), C.byref(nb), self.export_buf.value, None))
        cnt = counts.download(np.uint32, self.nvp)[: self.n]
        nbl = self.export_buf.download(np.uint32, (self.max_nbrs, self.nvp))
        return cnt, nbl
