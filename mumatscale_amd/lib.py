"""Python binding of the CA step library (libmmca.so, C ABI in include/mmca.h).

`Subblock` mirrors the reference's per-subblock operator API: one method per
function that `doiteration()` calls (calculate.c:193-352), plus `step()` for
`loop()`'s timestep++/doiteration pair (loop.c:222-227).  Errors that make the
reference call error() (debug.h:54-58) raise `MmcaError`.

There is no CPU fallback: without the built library or without a GPU the
constructor raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmmca.so")
_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    # PyTorch's ROCm wheels carry their own HIP runtime with the same soname
    # (libamdhip64.so.7) as /opt/rocm's.  The first one loaded serves the
    # whole process; torch fails to find a GPU on a runtime other than its
    # own, so when torch is installed it is loaded first and this library
    # shares its runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("%s is not built: run __graft_entry__.build() (make -C mumatscale_amd/csrc)"
                           % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    i32p = C.POINTER(C.c_int32)
    L.mmca_abi_version.restype = C.c_int
    L.mmca_status_string.restype = C.c_char_p
    L.mmca_status_string.argtypes = [C.c_int]
    L.mmca_device_count.argtypes = [i32p]
    L.mmca_eval_exp.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64]
    L.mmca_bandwidth_probe.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
    L.mmca_create.argtypes = [C.POINTER(abi.Params), C.c_int, C.POINTER(vp)]
    L.mmca_destroy.argtypes = [vp]
    L.mmca_init_nucleation.argtypes = [vp, C.POINTER(abi.NucleationCfg)]
    L.mmca_step.argtypes = [vp, C.c_int, C.c_int, C.POINTER(abi.Report)]
    L.mmca_set_timestep.argtypes = [vp, C.c_uint64]
    L.mmca_get_timestep.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.mmca_cell_nucleation.argtypes = [vp, i32p]
    L.mmca_grow_cell_reduction.argtypes = [vp, i32p]
    for n in ("mmca_activate_new_grains", "mmca_sb_diffuse_alloy_decentered",
              "mmca_fs_change_diffuse", "mmca_grow_octahedra",
              "mmca_capture_octahedra_diffuse", "mmca_doiteration"):
        getattr(L, n).argtypes = [vp]
    L.mmca_solid_volume.argtypes = [vp, C.POINTER(C.c_double)]
    L.mmca_upload_field.argtypes = [vp, C.c_int, vp]
    L.mmca_download_field.argtypes = [vp, C.c_int, vp]
    L.mmca_get_grains.argtypes = [vp, C.POINTER(abi.Grain), C.c_int32, i32p]
    L.mmca_get_interface_list.argtypes = [vp, i32p, C.c_int32, i32p]
    L.mmca_set_grains.argtypes = [vp, C.POINTER(abi.Grain), C.c_int32]
    L.mmca_io_fields.argtypes = [vp, vp, vp, vp, vp]
    L.mmca_rand.argtypes = [vp, i32p, C.c_int32]
    L.mmca_save_state.argtypes = [vp, vp, C.c_int64, C.POINTER(C.c_int64)]
    L.mmca_save_state_ex.argtypes = [vp, C.c_int, vp, C.c_int64, C.POINTER(C.c_int64)]
    L.mmca_get_list_slots.argtypes = [vp, C.POINTER(C.c_int32), C.c_int64]
    L.mmca_set_list_slots.argtypes = [vp, C.POINTER(C.c_int32), C.c_int64]
    L.mmca_get_nucleation_points.argtypes = [vp, vp, C.c_int32, C.POINTER(C.c_int32)]
    L.mmca_load_state.argtypes = [vp, vp, C.c_int64]
    L.mmca_padded_cells.argtypes = [vp, C.POINTER(C.c_int64)]
    L.mmca_set_kernel_timing.argtypes = [vp, C.c_int]
    L.mmca_kernel_time_ms.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    L.mmca_set_tuning.argtypes = [vp, C.c_int, C.c_int]
    L.mmca_get_tuning.argtypes = [vp, C.c_int, i32p]
    L.mmca_batch_create.argtypes = [C.POINTER(abi.Params), C.POINTER(abi.NucleationCfg), C.c_int, C.c_int,
                                    C.POINTER(vp)]
    L.mmca_batch_destroy.argtypes = [vp]
    L.mmca_batch_size.argtypes = [vp, i32p]
    L.mmca_batch_domain.argtypes = [vp, C.c_int, C.POINTER(vp)]
    L.mmca_batch_step.argtypes = [vp, C.c_int, C.c_int, C.POINTER(abi.Report)]
    L.mmca_batch_set_tuning.argtypes = [vp, C.c_int, C.c_int]
    L.mmca_batch_couple.argtypes = [vp, i32p]
    L.mmca_batch_save_state.argtypes = [vp, vp, C.c_int64, C.POINTER(C.c_int64)]
    L.mmca_batch_save_state_ex.argtypes = [vp, C.c_int, vp, C.c_int64, C.POINTER(C.c_int64)]
    L.mmca_batch_load_state.argtypes = [vp, vp, C.c_int64]
    L.mmca_batch_set_kernel_timing.argtypes = [vp, C.c_int]
    L.mmca_batch_kernel_time_ms.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    L.mmca_batch_autotune.argtypes = [C.POINTER(abi.Params), C.POINTER(abi.NucleationCfg), C.c_int, C.c_int,
                                      C.c_int, C.c_int, i32p, i32p, C.POINTER(C.c_double)]
    L.mmca_nucleation_sites.argtypes = [C.POINTER(abi.Params), C.POINTER(abi.NucleationCfg),
                                        i32p, C.POINTER(C.c_float), i32p, C.c_int64,
                                        C.POINTER(C.c_int64)]
    L.mmca_orientation_draws.argtypes = [C.POINTER(abi.Params), C.POINTER(abi.NucleationCfg),
                                         C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.mmca_set_record.argtypes = [vp, vp, C.c_int64]
    dp = C.c_void_p
    L.mmca_replay_step.argtypes = [C.POINTER(abi.Params), C.c_uint64, vp, dp, dp, dp, dp, dp, dp, vp]
    L.mmca_adjoint_workspace_bytes.argtypes = [C.POINTER(abi.Params), C.POINTER(C.c_int64)]
    L.mmca_record_bytes.argtypes = [C.POINTER(abi.Params), C.POINTER(C.c_int64)]
    L.mmca_adjoint_step.argtypes = [C.POINTER(abi.Params), C.c_uint64, vp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp,
                                    vp, vp]
    L.mmca_adjoint_step_noalloc.argtypes = [C.POINTER(abi.Params), C.c_uint64, vp, dp, dp, dp, dp, dp, dp, dp,
                                            dp, dp, dp, vp]
    if L.mmca_abi_version() != abi.ABI_VERSION:
        raise RuntimeError("libmmca ABI version mismatch")
    _lib = L
    return L


def nucleation_sites(params, cfg):
    """Sites the set-up creates (padded index, threshold, matching fixed entry);
    runs on the host only."""
    n = C.c_int64(0)
    _check(lib().mmca_nucleation_sites(C.byref(params), C.byref(cfg), None, None, None, 0,
                                       C.byref(n)), "mmca_nucleation_sites")
    idx = np.zeros(n.value, np.int32)
    thr = np.zeros(n.value, np.float32)
    fixed = np.zeros(n.value, np.int32)
    _check(lib().mmca_nucleation_sites(C.byref(params), C.byref(cfg),
                                       idx.ctypes.data_as(C.POINTER(C.c_int32)),
                                       thr.ctypes.data_as(C.POINTER(C.c_float)),
                                       fixed.ctypes.data_as(C.POINTER(C.c_int32)), n.value,
                                       C.byref(n)), "mmca_nucleation_sites")
    return idx, thr, fixed


def orientation_draws(params, cfg, n):
    """First n random orientations after the set-up: (angles[n,3], rotmat[n,3,3])."""
    ang = np.zeros((n, 3))
    rot = np.zeros((n, 3, 3))
    _check(lib().mmca_orientation_draws(C.byref(params), C.byref(cfg), n,
                                        ang.ctypes.data_as(C.POINTER(C.c_double)),
                                        rot.ctypes.data_as(C.POINTER(C.c_double))),
           "mmca_orientation_draws")
    return ang, rot


def device_count():
    n = C.c_int32(0)
    lib().mmca_device_count(C.byref(n))
    return n.value


def eval_exp(x, device=0):
    """exp of every element of x on the GPU, as the Gaussian-beam temperature
    evaluates it (bit-identical with the host C library's exp)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    f64p = C.POINTER(C.c_double)
    _check(lib().mmca_eval_exp(device, x.ctypes.data_as(f64p), y.ctypes.data_as(f64p), x.size),
           "mmca_eval_exp")
    return y


_FIELD_DTYPES = {
    abi.FIELD_CE: (np.float64, 1), abi.FIELD_CL: (np.float64, 1),
    abi.FIELD_FS: (np.float64, 1), abi.FIELD_GR: (np.int32, 1),
    abi.FIELD_DC: (np.float64, 3), abi.FIELD_D: (np.float64, 1),
    abi.FIELD_TEMP: (np.float64, 1), abi.FIELD_THRESHOLD: (np.float32, 1),
    abi.FIELD_MOLD: (np.int8, 1), abi.FIELD_OCE: (np.float64, 1),
}


def _check(st, where):
    if st != abi.OK:
        raise abi.MmcaError(st, where)


class Subblock:
    """One subblock whose state lives on one GPU."""

    def __init__(self, params, nucleation=None, device=0):
        self.params = params
        self.shape = (params.nz + 2, params.ny + 2, params.nx + 2)
        h = C.c_void_p()
        st = lib().mmca_create(C.byref(params), device, C.byref(h))
        self._h = h
        if st != abi.OK:
            self.close()
            raise abi.MmcaError(st, "mmca_create")
        if nucleation is not None:
            self.init_nucleation(nucleation)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().mmca_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- set-up -----------------------------------------------------------
    def init_nucleation(self, cfg):
        """init_nucleation_sites + init_sb_nucleation (grain.c:603-713, 931-1010)."""
        _check(lib().mmca_init_nucleation(self._h, C.byref(cfg)), "mmca_init_nucleation")

    # ---- stepping ---------------------------------------------------------
    def step(self, n=1, want_sums=False, raise_on_error=True):
        """n x (timestep++, doiteration()); returns the mmca_report."""
        rep = abi.Report()
        st = lib().mmca_step(self._h, n, 1 if want_sums else 0, C.byref(rep))
        if raise_on_error:
            _check(st, "mmca_step")
        return rep

    def doiteration(self):
        _check(lib().mmca_doiteration(self._h), "mmca_doiteration")

    def cell_nucleation(self):
        n = C.c_int32(0)
        _check(lib().mmca_cell_nucleation(self._h, C.byref(n)), "mmca_cell_nucleation")
        return n.value

    def activate_new_grains(self):
        _check(lib().mmca_activate_new_grains(self._h), "mmca_activate_new_grains")

    def sb_diffuse_alloy_decentered(self):
        _check(lib().mmca_sb_diffuse_alloy_decentered(self._h), "mmca_sb_diffuse_alloy_decentered")

    def fs_change_diffuse(self):
        _check(lib().mmca_fs_change_diffuse(self._h), "mmca_fs_change_diffuse")

    def grow_octahedra(self):
        _check(lib().mmca_grow_octahedra(self._h), "mmca_grow_octahedra")

    def grow_cell_reduction(self):
        g = C.c_int32(0)
        _check(lib().mmca_grow_cell_reduction(self._h, C.byref(g)), "mmca_grow_cell_reduction")
        return g.value

    def capture_octahedra_diffuse(self):
        _check(lib().mmca_capture_octahedra_diffuse(self._h), "mmca_capture_octahedra_diffuse")

    def solid_volume(self):
        v = C.c_double(0)
        _check(lib().mmca_solid_volume(self._h, C.byref(v)), "mmca_solid_volume")
        return v.value

    # ---- state ------------------------------------------------------------
    @property
    def timestep(self):
        t = C.c_uint64(0)
        _check(lib().mmca_get_timestep(self._h, C.byref(t)), "mmca_get_timestep")
        return t.value

    @timestep.setter
    def timestep(self, t):
        _check(lib().mmca_set_timestep(self._h, t), "mmca_set_timestep")

    def download(self, field):
        dt, comps = _FIELD_DTYPES[field]
        shp = self.shape + ((3,) if comps == 3 else ())
        out = np.empty(shp, dtype=dt)
        _check(lib().mmca_download_field(self._h, field, out.ctypes.data), "mmca_download_field")
        return out

    def upload(self, field, arr):
        dt, comps = _FIELD_DTYPES[field]
        a = np.ascontiguousarray(arr, dtype=dt)
        shp = self.shape + ((3,) if comps == 3 else ())
        if a.shape != shp:
            raise ValueError("field shape %s, expected %s" % (a.shape, shp))
        _check(lib().mmca_upload_field(self._h, field, a.ctypes.data), "mmca_upload_field")

    @property
    def num_grains(self):
        n = C.c_int32(0)
        _check(lib().mmca_get_grains(self._h, None, 0, C.byref(n)), "mmca_get_grains")
        return n.value

    def grains(self):
        n = self.num_grains
        arr = (abi.Grain * max(1, n))()
        _check(lib().mmca_get_grains(self._h, arr, n, None), "mmca_get_grains")
        return [arr[i] for i in range(n)]

    def save_state(self):
        """Restart image of this block (bytes)."""
        n = C.c_int64(0)
        _check(lib().mmca_save_state(self._h, None, 0, C.byref(n)), "mmca_save_state")
        buf = C.create_string_buffer(n.value)
        _check(lib().mmca_save_state(self._h, buf, n.value, C.byref(n)), "mmca_save_state")
        return buf.raw[:n.value]

    def load_state(self, image):
        buf = C.create_string_buffer(image, len(image))
        _check(lib().mmca_load_state(self._h, buf, len(image)), "mmca_load_state")

    def rand(self, n):
        """The next n rand() values of this block's random stream."""
        out = np.zeros(n, np.int32)
        _check(lib().mmca_rand(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), n), "mmca_rand")
        return out

    def io_fields(self):
        """cache_io_data (loop.c:74-113): FP32 fs, ce, temperature and gr,
        padded, as the output writers read them."""
        fs = np.empty(self.shape, np.float32)
        ce = np.empty(self.shape, np.float32)
        t = np.empty(self.shape, np.float32)
        gr = np.empty(self.shape, np.int32)
        _check(lib().mmca_io_fields(self._h, fs.ctypes.data, ce.ctypes.data, t.ctypes.data, gr.ctypes.data),
               "mmca_io_fields")
        return {"fs": fs, "ce": ce, "temperature": t, "gr": gr}

    def set_grains(self, grains):
        """Replace the grain table (list of abi.Grain, record 0 = liquid)."""
        arr = (abi.Grain * len(grains))(*grains)
        _check(lib().mmca_set_grains(self._h, arr, len(grains)), "mmca_set_grains")

    def interface_list(self):
        g = C.c_int32(0)
        _check(lib().mmca_get_interface_list(self._h, None, 0, C.byref(g)), "mmca_get_interface_list")
        out = np.zeros(g.value, dtype=np.int32)
        if g.value:
            _check(lib().mmca_get_interface_list(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)),
                                                 g.value, None), "mmca_get_interface_list")
        return out

    @property
    def padded_cells(self):
        n = C.c_int64(0)
        _check(lib().mmca_padded_cells(self._h, C.byref(n)), "mmca_padded_cells")
        return n.value

    def set_tuning(self, key, value):
        _check(lib().mmca_set_tuning(self._h, key, value), "mmca_set_tuning")

    def set_record(self, codes_ptr, n_steps):
        """Record the partitioning branch of every cell for the next n_steps
        steps into device memory at codes_ptr (n_steps x padded cells bytes);
        (None, 0) stops recording."""
        _check(lib().mmca_set_record(self._h, codes_ptr, n_steps), "mmca_set_record")

    def get_tuning(self, key):
        v = C.c_int32(0)
        _check(lib().mmca_get_tuning(self._h, key, C.byref(v)), "mmca_get_tuning")
        return v.value

    def set_kernel_timing(self, enable=True):
        _check(lib().mmca_set_kernel_timing(self._h, 1 if enable else 0), "mmca_set_kernel_timing")

    def kernel_time_ms(self):
        ms = C.c_double(0)
        n = C.c_int64(0)
        _check(lib().mmca_kernel_time_ms(self._h, C.byref(ms), C.byref(n)), "mmca_kernel_time_ms")
        return ms.value, n.value


class _Member(Subblock):
    """A domain owned by a Batch (not destroyed on its own)."""

    def __init__(self, handle, params):
        self.params = params
        self.shape = (params.nz + 2, params.ny + 2, params.nx + 2)
        self._h = handle

    def close(self):
        self._h = C.c_void_p()


class Batch:
    """Independent subblocks of any sizes advanced together (one set of
    launches per step for all of them)."""

    def __init__(self, params_list, nucleation_list=None, device=0):
        n = len(params_list)
        P = (abi.Params * n)(*params_list)
        N = None
        if nucleation_list is not None:
            N = (abi.NucleationCfg * n)(*nucleation_list)
            self._keep = [getattr(c, "_keep", None) for c in nucleation_list]
        h = C.c_void_p()
        st = lib().mmca_batch_create(P, N, n, device, C.byref(h))
        self._h = h
        if st != abi.OK:
            self.close()
            raise abi.MmcaError(st, "mmca_batch_create")
        self.params = list(params_list)
        self.domains = []
        for i in range(n):
            hd = C.c_void_p()
            _check(lib().mmca_batch_domain(self._h, i, C.byref(hd)), "mmca_batch_domain")
            self.domains.append(_Member(hd, self.params[i]))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().mmca_batch_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, n=1, want_sums=False, raise_on_error=True):
        reps = (abi.Report * len(self.domains))()
        st = lib().mmca_batch_step(self._h, n, 1 if want_sums else 0, reps)
        if raise_on_error:
            _check(st, "mmca_batch_step")
        return list(reps)

    def set_tuning(self, key, value):
        _check(lib().mmca_batch_set_tuning(self._h, key, value), "mmca_batch_set_tuning")

    def couple(self, nsub):
        """Make the domains the subblocks (ranks) of one global domain split
        nsub[0] x nsub[1] x nsub[2] (mmca_batch_couple)."""
        arr = (C.c_int32 * 3)(*[int(v) for v in nsub])
        _check(lib().mmca_batch_couple(self._h, arr), "mmca_batch_couple")

    def save_state(self):
        """Restart image of the whole batch (bytes), coupled ranks included."""
        n = C.c_int64(0)
        _check(lib().mmca_batch_save_state(self._h, None, 0, C.byref(n)), "mmca_batch_save_state")
        buf = C.create_string_buffer(n.value)
        _check(lib().mmca_batch_save_state(self._h, buf, n.value, C.byref(n)), "mmca_batch_save_state")
        return buf.raw[:n.value]

    def load_state(self, image):
        buf = C.create_string_buffer(image, len(image))
        _check(lib().mmca_batch_load_state(self._h, buf, len(image)), "mmca_batch_load_state")

    def set_kernel_timing(self, enable=True):
        _check(lib().mmca_batch_set_kernel_timing(self._h, 1 if enable else 0), "mmca_batch_set_kernel_timing")

    def kernel_time_ms(self):
        ms = C.c_double(0)
        n = C.c_int64(0)
        _check(lib().mmca_batch_kernel_time_ms(self._h, C.byref(ms), C.byref(n)), "mmca_batch_kernel_time_ms")
        return ms.value, n.value


def autotune_batch(params_list, nucleation_list=None, device=0, warmup=20, trial_steps=30):
    """Fastest solute-kernel shape (planes, rows, kernel ms) for these domains."""
    n = len(params_list)
    P = (abi.Params * n)(*params_list)
    N = (abi.NucleationCfg * n)(*nucleation_list) if nucleation_list is not None else None
    zp = C.c_int32(0)
    zr = C.c_int32(0)
    ms = C.c_double(0)
    _check(lib().mmca_batch_autotune(P, N, n, device, warmup, trial_steps, C.byref(zp), C.byref(zr),
                                     C.byref(ms)), "mmca_batch_autotune")
    return zp.value, zr.value, ms.value
