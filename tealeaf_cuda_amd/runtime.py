"""ctypes bindings: the kernel-layer C-ABI and the host driver.

``Engine`` binds the 29 Fortran-ABI entry points of include/tea_leaf_engine.h
(every argument by reference, exactly as the TeaLeaf Fortran driver passes them)
plus the engine extensions.  ``run_deck`` runs a whole tea.in deck through the
host driver (libtea_driver.so) over a chosen kernel library.

The HIP library is the product; loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .build import DRIVER_SO, ENGINE_SO

TEA_MAX_STEPS = 4096
TEA_MAX_SUMMARIES = 1024

# data.f90:37-75
CHUNK_LEFT, CHUNK_RIGHT, CHUNK_BOTTOM, CHUNK_TOP = 1, 2, 3, 4
EXTERNAL_FACE = -1
FIELD_DENSITY, FIELD_ENERGY0, FIELD_ENERGY1, FIELD_U, FIELD_P, FIELD_SD, FIELD_R = range(1, 8)
NUM_FIELDS = 7
XFIELD_U0, XFIELD_W, XFIELD_Z, XFIELD_MI, XFIELD_KX, XFIELD_KY = 8, 9, 10, 11, 12, 13
XFIELD_RTEMP, XFIELD_UTEMP, XFIELD_RSTORE, XFIELD_CP, XFIELD_BFP, XFIELD_VOLUME = 14, 15, 16, 17, 18, 19
G_RECT, G_CIRC, G_POINT = 1, 2, 3
CONDUCTIVITY, RECIP_CONDUCTIVITY = 1, 2
PREC_NONE, PREC_JAC_DIAG, PREC_JAC_BLOCK = 1, 2, 3
SOLVER_JACOBI, SOLVER_CG, SOLVER_CHEBYSHEV, SOLVER_PPCG = 0, 1, 2, 3


class TeaError(RuntimeError):
    """A fatal error reported by the kernel layer or the driver."""


def engine_library() -> Path:
    """Path of the built HIP kernel library; raises if it has not been built."""
    if not ENGINE_SO.exists():
        raise TeaError(f"HIP engine library {ENGINE_SO} is missing: run tealeaf_cuda_amd.build.build_all()")
    return ENGINE_SO


def driver_library() -> Path:
    if not DRIVER_SO.exists():
        raise TeaError(f"driver library {DRIVER_SO} is missing: run tealeaf_cuda_amd.build.build_all()")
    return DRIVER_SO


# ----------------------------------------------------------------------------- ABI

_I = ctypes.POINTER(ctypes.c_int)
_D = ctypes.POINTER(ctypes.c_double)

_ABI = {
    "initialise_cuda_": [_I, _I, _I, _I],
    "initialise_chunk_kernel_cuda_": [_D, _D, _D, _D],
    "generate_chunk_kernel_cuda_": [_I, _D, _D, _D, _D, _D, _D, _D, _I, _I, _I, _I],
    "set_field_kernel_cuda_": [],
    "field_summary_kernel_cuda_": [_D, _D, _D, _D],
    "update_halo_kernel_cuda_": [_I, _I, _I],
    "cuda_pack_buffers_": [_I, _I, _I, _I, _D],
    "cuda_unpack_buffers_": [_I, _I, _I, _I, _D],
    "tea_leaf_common_init_kernel_cuda_": [_I, _D, _D, _D, _I, _I],
    "tea_leaf_calc_residual_cuda_": [],
    "tea_leaf_common_finalise_kernel_cuda_": [],
    "tea_leaf_calc_2norm_kernel_cuda_": [_I, _D],
    "tea_leaf_cg_init_kernel_cuda_": [_D],
    "tea_leaf_cg_calc_w_kernel_cuda_": [_D],
    "tea_leaf_cg_calc_ur_kernel_cuda_": [_D, _D],
    "tea_leaf_cg_calc_p_kernel_cuda_": [_D],
    "tea_leaf_cheby_init_kernel_cuda_": [_D, _D, _I, _D],
    "tea_leaf_cheby_iterate_kernel_cuda_": [_I],
    "tea_leaf_jacobi_solve_kernel_cuda_": [_D],
    "tea_leaf_ppcg_init_constants_cuda_": [_D, _D, _I],
    "tea_leaf_ppcg_init_kernel_cuda_": [_I, _D],
    "tea_leaf_ppcg_init_sd_kernel_cuda_": [_D],
    "tea_leaf_ppcg_init_sd_new_kernel_cuda_": [_D],
    "tea_leaf_ppcg_inner_kernel_cuda_": [_I, _I, _I],
    "tea_leaf_ppcg_store_r_kernel_cuda_": [],
    "tea_leaf_ppcg_update_z_kernel_cuda_": [],
    "tea_leaf_ppcg_calc_rrn_kernel_cuda_": [_D],
    "tea_leaf_ppcg_calc_2norm_kernel_cuda_": [_D],
    "tea_leaf_ppcg_calc_p_kernel_cuda_": [_D],
}

_EXT = {
    "tea_engine_set_deck_path": ([ctypes.c_char_p], None),
    "tea_engine_set_error_mode": ([ctypes.c_int], None),
    "tea_engine_last_error": ([ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
    "tea_engine_clear_error": ([], None),
    "tea_engine_info": ([_I, _I, _I, _I, _I], None),
    "tea_engine_get_field": ([ctypes.c_int, _D], ctypes.c_int),
    "tea_engine_set_field": ([ctypes.c_int, _D], ctypes.c_int),
    "tea_engine_profile_enable": ([ctypes.c_int], None),
    "tea_engine_profile_only": ([ctypes.c_char_p], None),
    "tea_engine_profile_sample": ([ctypes.c_int], None),
    "tea_engine_profile_reset": ([], None),
    "tea_engine_profile_get": ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_long), _D], ctypes.c_int),
    "tea_engine_profile_names": ([ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
    "tea_engine_stream": ([], ctypes.c_void_p),
    "tea_engine_synchronize": ([], None),
    "tea_engine_finalize": ([], None),
    "tea_engine_ppcg_begin": ([ctypes.c_int, ctypes.c_double, _D], None),
    "tea_engine_ppcg_ur": ([ctypes.c_double, ctypes.c_double, _D], None),
    "tea_engine_ppcg_end": ([_D], None),
    "tea_engine_ppcg_inner2": ([ctypes.c_int], None),
    "tea_engine_ppcg_inner_n": ([ctypes.c_int, ctypes.c_int], None),
    "tea_engine_ppcg_begin_ur": ([ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double)], None),
    "tea_engine_ppcg_end_begin_ur": ([ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)], None),
    "tea_engine_ppcg_end_begin_ur_inner": ([ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double)], None),
    "tea_engine_ppcg_iteration": ([ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_double)], None),
    "tea_engine_cheby_iterate_n": ([ctypes.c_int, ctypes.c_int], None),
    "tea_engine_ppcg_inner_plan": ([ctypes.c_int], ctypes.c_int),
    "tea_engine_cg_begin_ur": ([ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double)], None),
    "tea_engine_set_precision": ([ctypes.c_int], ctypes.c_int),
    "tea_engine_mp_leave": ([], ctypes.c_int),
    "tea_engine_mp_setup": ([], None),
    "tea_engine_mp_begin": ([_D], None),
    "tea_engine_mp_end": ([], None),
    "tea_engine_mp_replace": ([ctypes.c_double, _D, _D], None),
    "tea_engine_get_field32": ([ctypes.c_int, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "tea_engine_set_field32": ([ctypes.c_int, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
}

# Entry points on caller-owned device memory: exported by the HIP library only.
_DEVICE = {
    "tea_engine_calc_2norm_noalloc": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
}

ABI_SYMBOLS = tuple(_ABI)
EXTENSION_SYMBOLS = tuple(_EXT)
DEVICE_SYMBOLS = tuple(_DEVICE)


def _ip(v: int):
    return ctypes.byref(ctypes.c_int(int(v)))


def _iarr(vals) -> ctypes.Array:
    return (ctypes.c_int * len(vals))(*[int(v) for v in vals])


def _darr(vals) -> ctypes.Array:
    return (ctypes.c_double * len(vals))(*[float(v) for v in vals])


class Engine:
    """Direct access to one kernel library through its Fortran-style C-ABI.

    Fatal errors of the library are turned into ``TeaError`` (the library runs
    in sticky-error mode), instead of the reference's abort.
    """

    def __init__(self, lib_path: str | os.PathLike | None = None):
        path = Path(lib_path) if lib_path is not None else engine_library()
        self.path = path
        self.lib = ctypes.CDLL(str(path), mode=os.RTLD_NOW | os.RTLD_LOCAL)
        for name, args in _ABI.items():
            fn = getattr(self.lib, name)
            fn.argtypes = args
            fn.restype = None
        for name, (args, ret) in _EXT.items():
            fn = getattr(self.lib, name)
            fn.argtypes = args
            fn.restype = ret
        for name, (args, ret) in _DEVICE.items():
            fn = getattr(self.lib, name, None)
            if fn is not None:
                fn.argtypes = args
                fn.restype = ret
        self.lib.tea_engine_set_error_mode(1)
        self.refresh()

    def refresh(self):
        """Reads the chunk geometry of the library's current chunk (if any)."""
        nx_, ny_, h_, s_, p_ = (ctypes.c_int() for _ in range(5))
        self.lib.tea_engine_info(*(ctypes.byref(v) for v in (nx_, ny_, h_, s_, p_)))
        self.nx, self.ny, self.h, self.solver, self.prec = nx_.value, ny_.value, h_.value, s_.value, p_.value

    # -- error handling
    def _check(self):
        buf = ctypes.create_string_buffer(1024)
        if self.lib.tea_engine_last_error(buf, 1024):
            self.lib.tea_engine_clear_error()
            raise TeaError(buf.value.decode())

    def _call(self, name, *args):
        getattr(self.lib, name)(*args)
        self._check()

    # -- set-up
    def initialise(self, deck_path, nx: int, ny: int):
        self.lib.tea_engine_set_deck_path(str(deck_path).encode())
        self._call("initialise_cuda_", _ip(1), _ip(nx), _ip(1), _ip(ny))
        self.refresh()

    def initialise_chunk(self, xmin, ymin, dx, dy):
        d = [ctypes.c_double(v) for v in (xmin, ymin, dx, dy)]
        self._call("initialise_chunk_kernel_cuda_", *(ctypes.byref(v) for v in d))

    def generate_chunk(self, density, energy, xmin, xmax, ymin, ymax, radius, geometry):
        n = len(density)
        self._call("generate_chunk_kernel_cuda_", _ip(n), _darr(density), _darr(energy), _darr(xmin),
                   _darr(xmax), _darr(ymin), _darr(ymax), _darr(radius), _iarr(geometry), _ip(G_RECT),
                   _ip(G_CIRC), _ip(G_POINT))

    def set_field_kernel(self):
        self._call("set_field_kernel_cuda_")

    def field_summary(self):
        v = [ctypes.c_double() for _ in range(4)]
        self._call("field_summary_kernel_cuda_", *(ctypes.byref(x) for x in v))
        return tuple(x.value for x in v)

    def update_halo(self, fields, depth, neighbours=(EXTERNAL_FACE,) * 4):
        self._call("update_halo_kernel_cuda_", _iarr(neighbours), _iarr(fields), _ip(depth))

    def pack(self, fields, offsets, depth, face, nelems):
        buf = np.zeros(nelems, dtype=np.float64)
        self._call("cuda_pack_buffers_", _iarr(fields), _iarr(offsets), _ip(depth), _ip(face),
                   buf.ctypes.data_as(_D))
        return buf

    def unpack(self, fields, offsets, depth, face, buf):
        buf = np.ascontiguousarray(buf, dtype=np.float64)
        self._call("cuda_unpack_buffers_", _iarr(fields), _iarr(offsets), _ip(depth), _ip(face),
                   buf.ctypes.data_as(_D))

    # -- solver: common
    def common_init(self, coefficient, dt, zero_boundary=(1, 1, 1, 1), reflective=0):
        rx, ry = ctypes.c_double(), ctypes.c_double()
        self._call("tea_leaf_common_init_kernel_cuda_", _ip(coefficient), ctypes.byref(ctypes.c_double(dt)),
                   ctypes.byref(rx), ctypes.byref(ry), _iarr(zero_boundary), _ip(reflective))
        return rx.value, ry.value

    def _scalar(self, name, *pre):
        out = ctypes.c_double()
        self._call(name, *pre, ctypes.byref(out))
        return out.value

    def calc_residual(self):
        self._call("tea_leaf_calc_residual_cuda_")

    def finalise(self):
        self._call("tea_leaf_common_finalise_kernel_cuda_")

    def calc_2norm(self, norm_array):
        return self._scalar("tea_leaf_calc_2norm_kernel_cuda_", _ip(norm_array))

    # -- CG
    def cg_init(self):
        return self._scalar("tea_leaf_cg_init_kernel_cuda_")

    def cg_calc_w(self):
        return self._scalar("tea_leaf_cg_calc_w_kernel_cuda_")

    def cg_calc_ur(self, alpha):
        return self._scalar("tea_leaf_cg_calc_ur_kernel_cuda_", ctypes.byref(ctypes.c_double(alpha)))

    def cg_calc_p(self, beta):
        self._call("tea_leaf_cg_calc_p_kernel_cuda_", ctypes.byref(ctypes.c_double(beta)))

    # -- Chebyshev
    def cheby_init(self, alphas, betas, theta):
        self._call("tea_leaf_cheby_init_kernel_cuda_", _darr(alphas), _darr(betas), _ip(len(alphas)),
                   ctypes.byref(ctypes.c_double(theta)))

    def cheby_iterate(self, step):
        self._call("tea_leaf_cheby_iterate_kernel_cuda_", _ip(step))

    def cheby_iterate_n(self, step: int, nsteps: int):
        self._call("tea_engine_cheby_iterate_n", int(step), int(nsteps))

    # -- Jacobi
    def jacobi_solve(self):
        return self._scalar("tea_leaf_jacobi_solve_kernel_cuda_")

    # -- PPCG
    def ppcg_init_constants(self, alphas, betas):
        self._call("tea_leaf_ppcg_init_constants_cuda_", _darr(alphas), _darr(betas), _ip(len(alphas)))

    def ppcg_init(self, step):
        return self._scalar("tea_leaf_ppcg_init_kernel_cuda_", _ip(step))

    def ppcg_init_sd(self, theta):
        self._call("tea_leaf_ppcg_init_sd_kernel_cuda_", ctypes.byref(ctypes.c_double(theta)))

    def ppcg_init_sd_new(self, theta):
        self._call("tea_leaf_ppcg_init_sd_new_kernel_cuda_", ctypes.byref(ctypes.c_double(theta)))

    def ppcg_inner(self, step, bounds_extra=0, neighbours=(EXTERNAL_FACE,) * 4):
        self._call("tea_leaf_ppcg_inner_kernel_cuda_", _ip(step), _ip(bounds_extra), _iarr(neighbours))

    def ppcg_store_r(self):
        self._call("tea_leaf_ppcg_store_r_kernel_cuda_")

    def ppcg_update_z(self):
        self._call("tea_leaf_ppcg_update_z_kernel_cuda_")

    def ppcg_calc_rrn(self):
        return self._scalar("tea_leaf_ppcg_calc_rrn_kernel_cuda_")

    def ppcg_calc_2norm(self):
        return self._scalar("tea_leaf_ppcg_calc_2norm_kernel_cuda_")

    def ppcg_calc_p(self, beta):
        self._call("tea_leaf_ppcg_calc_p_kernel_cuda_", ctypes.byref(ctypes.c_double(beta)))

    # -- fields
    def shape(self):
        return (self.ny + 2 * self.h, self.nx + 2 * self.h)

    def get_field(self, field_id) -> np.ndarray:
        self.refresh()  # the driver may have (re)initialised the chunk since
        a = np.empty(self.shape(), dtype=np.float64)
        if self.lib.tea_engine_get_field(int(field_id), a.ctypes.data_as(_D)) != 0:
            self._check()
            raise TeaError(f"get_field({field_id}) failed")
        return a

    def set_field(self, field_id, a) -> None:
        self.refresh()
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.shape():
            raise ValueError(f"field shape {a.shape} != {self.shape()}")
        if self.lib.tea_engine_set_field(int(field_id), a.ctypes.data_as(_D)) != 0:
            self._check()
            raise TeaError(f"set_field({field_id}) failed")

    # -- profiler
    def profile_enable(self, on=True):
        self.lib.tea_engine_profile_enable(1 if on else 0)

    def profile_only(self, prefix: str = ""):
        self.lib.tea_engine_profile_only(prefix.encode())

    def profile_sample(self, every: int = 1):
        self.lib.tea_engine_profile_sample(int(every))

    def profile_reset(self):
        self.lib.tea_engine_profile_reset()

    def profile(self) -> dict:
        n = self.lib.tea_engine_profile_names(None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        self.lib.tea_engine_profile_names(buf, n + 1)
        out = {}
        for name in buf.value.decode().split():
            c, ms = ctypes.c_long(), ctypes.c_double()
            self.lib.tea_engine_profile_get(name.encode(), ctypes.byref(c), ctypes.byref(ms))
            out[name] = (c.value, ms.value)
        return out

    # -- fused PPCG outer iteration
    def ppcg_begin(self, with_p: bool, beta: float):
        out = (ctypes.c_double * 2)()
        self._call("tea_engine_ppcg_begin", int(bool(with_p)), float(beta), out)
        return out[0], out[1]

    def ppcg_begin_ur(self, with_p: bool, beta: float, theta: float):
        two, rrn = (ctypes.c_double * 2)(), ctypes.c_double()
        self._call("tea_engine_ppcg_begin_ur", int(bool(with_p)), float(beta), float(theta), two, ctypes.byref(rrn))
        return two[0], two[1], rrn.value

    def ppcg_end_begin_ur(self, rro_prev: float, theta: float):
        rrn_end, two, rrn = ctypes.c_double(), (ctypes.c_double * 2)(), ctypes.c_double()
        self._call("tea_engine_ppcg_end_begin_ur", float(rro_prev), float(theta), ctypes.byref(rrn_end), two,
                   ctypes.byref(rrn))
        return rrn_end.value, two[0], two[1], rrn.value

    def ppcg_end_begin_ur_inner(self, rro_prev: float, theta: float, inner_steps: int):
        rrn_end, two, rrn = ctypes.c_double(), (ctypes.c_double * 2)(), ctypes.c_double()
        self._call("tea_engine_ppcg_end_begin_ur_inner", float(rro_prev), float(theta), int(inner_steps),
                   ctypes.byref(rrn_end), two, ctypes.byref(rrn))
        return rrn_end.value, two[0], two[1], rrn.value

    def ppcg_iteration(self, with_p: bool, beta: float, theta: float, inner_steps: int):
        """Returns (pw, rro, rrn_ur, rrn_end)."""
        two, rrn_ur, rrn_end = (ctypes.c_double * 2)(), ctypes.c_double(), ctypes.c_double()
        self._call("tea_engine_ppcg_iteration", 1 if with_p else 0, float(beta), float(theta), int(inner_steps), two,
                   ctypes.byref(rrn_ur), ctypes.byref(rrn_end))
        return two[0], two[1], rrn_ur.value, rrn_end.value

    def ppcg_ur(self, alpha: float, theta: float) -> float:
        out = ctypes.c_double()
        self._call("tea_engine_ppcg_ur", float(alpha), float(theta), ctypes.byref(out))
        return out.value

    def ppcg_inner2(self, step: int):
        self._call("tea_engine_ppcg_inner2", int(step))

    def ppcg_inner_n(self, step: int, nsteps: int):
        self._call("tea_engine_ppcg_inner_n", int(step), int(nsteps))

    def ppcg_inner_plan(self, remaining: int) -> int:
        n = self.lib.tea_engine_ppcg_inner_plan(int(remaining))
        self._check()
        return int(n)

    def cg_begin_ur(self, with_p: bool, beta: float, rro: float):
        pw, rrn = ctypes.c_double(), ctypes.c_double()
        self._call("tea_engine_cg_begin_ur", int(bool(with_p)), float(beta), float(rro), ctypes.byref(pw),
                   ctypes.byref(rrn))
        return pw.value, rrn.value

    def ppcg_end(self) -> float:
        out = ctypes.c_double()
        self._call("tea_engine_ppcg_end", ctypes.byref(out))
        return out.value

    # -- mixed precision
    def set_precision(self, bits: int):
        if self.lib.tea_engine_set_precision(int(bits)) != 0:
            self._check()
            raise TeaError(f"set_precision({bits}) failed")

    def mp_leave(self):
        if self.lib.tea_engine_mp_leave() != 0:
            self._check()
            raise TeaError("mp_leave failed")

    def mp_setup(self):
        self._call("tea_engine_mp_setup")

    def mp_begin(self) -> float:
        out = ctypes.c_double()
        self._call("tea_engine_mp_begin", ctypes.byref(out))
        return out.value

    def mp_end(self):
        self._call("tea_engine_mp_end")

    def mp_replace(self, theta: float):
        rr64, rr32 = ctypes.c_double(), ctypes.c_double()
        self._call("tea_engine_mp_replace", float(theta), ctypes.byref(rr64), ctypes.byref(rr32))
        return rr64.value, rr32.value

    def get_field32(self, field_id) -> np.ndarray:
        self.refresh()  # the driver may have (re)initialised the chunk since
        a = np.empty(self.shape(), dtype=np.float32)
        if self.lib.tea_engine_get_field32(int(field_id), a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) != 0:
            raise TeaError(f"get_field32({field_id}) failed")
        return a

    def set_field32(self, field_id, a) -> None:
        self.refresh()
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.shape != self.shape():
            raise ValueError(f"field shape {a.shape} != {self.shape()}")
        if self.lib.tea_engine_set_field32(int(field_id), a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) != 0:
            raise TeaError(f"set_field32({field_id}) failed")

    def calc_2norm_noalloc(self, a_ptr: int, b_ptr: int, nx: int, ny: int, halo: int, pitch: int, out_ptr: int,
                           stream: int | None = None) -> None:
        """sum a*b over the interior of caller-owned device arrays into device
        memory at out_ptr, queued on `stream`; allocates nothing (HIP library
        only)."""
        if self.lib.tea_engine_calc_2norm_noalloc(a_ptr, b_ptr, nx, ny, halo, pitch, out_ptr, stream) != 0:
            self._check()
            raise TeaError("calc_2norm_noalloc failed")

    def synchronize(self):
        self.lib.tea_engine_synchronize()

    def finalize(self):
        self.lib.tea_engine_finalize()


# ----------------------------------------------------------------------------- driver

class _RunResult(ctypes.Structure):
    _fields_ = [
        ("n_steps", ctypes.c_int),
        ("iterations", ctypes.c_int * TEA_MAX_STEPS),
        ("ppcg_inner", ctypes.c_int * TEA_MAX_STEPS),
        ("conduction_error", ctypes.c_double * TEA_MAX_STEPS),
        ("exact_error", ctypes.c_double * TEA_MAX_STEPS),
        ("initial_residual", ctypes.c_double * TEA_MAX_STEPS),
        ("step_seconds", ctypes.c_double * TEA_MAX_STEPS),
        ("refinements", ctypes.c_int * TEA_MAX_STEPS),
        ("warm_start", ctypes.c_int * TEA_MAX_STEPS),
        ("n_summaries", ctypes.c_int),
        ("summary_step", ctypes.c_int * TEA_MAX_SUMMARIES),
        ("summary", (ctypes.c_double * 6) * TEA_MAX_SUMMARIES),
        ("eigmin", ctypes.c_double),
        ("eigmax", ctypes.c_double),
        ("total_seconds", ctypes.c_double),
        ("error_code", ctypes.c_int),
        ("error_msg", ctypes.c_char * 512),
    ]


@dataclass
class Summary:
    step: int
    time: float
    volume: float
    mass: float
    density: float
    energy: float
    u: float

    def values(self):
        return (self.volume, self.mass, self.density, self.energy, self.u)


@dataclass
class RunResult:
    iterations: list
    ppcg_inner: list
    conduction_error: list
    exact_error: list
    initial_residual: list
    step_seconds: list
    refinements: list = field(default_factory=list)
    warm_start: list = field(default_factory=list)
    summaries: list = field(default_factory=list)
    eigmin: float = 0.0
    eigmax: float = 0.0
    total_seconds: float = 0.0

    @property
    def final(self) -> Summary:
        return self.summaries[-1]


_STEP_CB = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_void_p)
_driver = None


def _driver_lib():
    global _driver
    if _driver is None:
        _driver = ctypes.CDLL(str(driver_library()))
        _driver.tea_driver_run.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                           ctypes.POINTER(_RunResult)]
        _driver.tea_driver_run.restype = ctypes.c_int
        _driver.tea_driver_set_step_callback.argtypes = [_STEP_CB, ctypes.c_void_p]
        _driver.tea_driver_set_step_callback.restype = None
    return _driver


def run_deck(deck_path, backend: str | os.PathLike | None = None, out_path=None, on_step=None) -> RunResult:
    """Runs a tea.in deck through the host driver over the kernel library
    ``backend`` (default: the HIP engine) and returns per-step results.
    ``on_step(step)`` is called after each timestep has completed on the device."""
    lib = Path(backend) if backend is not None else engine_library()
    res = _RunResult()
    drv = _driver_lib()
    cb = _STEP_CB(lambda step, _user: on_step(step)) if on_step is not None else _STEP_CB()
    drv.tea_driver_set_step_callback(cb, None)
    try:
        rc = drv.tea_driver_run(str(deck_path).encode(), str(lib).encode(),
                                str(out_path).encode() if out_path else None, ctypes.byref(res))
    finally:
        drv.tea_driver_set_step_callback(_STEP_CB(), None)
    if rc != 0:
        raise TeaError(res.error_msg.decode())
    n = res.n_steps
    out = RunResult(
        iterations=list(res.iterations[:n]),
        ppcg_inner=list(res.ppcg_inner[:n]),
        conduction_error=list(res.conduction_error[:n]),
        exact_error=list(res.exact_error[:n]),
        initial_residual=list(res.initial_residual[:n]),
        step_seconds=list(res.step_seconds[:n]),
        refinements=list(res.refinements[:n]),
        warm_start=list(res.warm_start[:n]),
        eigmin=res.eigmin,
        eigmax=res.eigmax,
        total_seconds=res.total_seconds,
    )
    for k in range(res.n_summaries):
        s = res.summary[k]
        out.summaries.append(Summary(res.summary_step[k], s[0], s[1], s[2], s[3], s[4], s[5]))
    return out
