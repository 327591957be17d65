"""ctypes binding of ``liblbm_hip.so`` (declared in ``include/lbm_hip.h``).

The library is loaded once per process. There is no fallback: if the shared
object is missing or does not load, every HIP backend raises
``BackendUnavailable``; nothing runs the lattice update on the CPU instead.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_int64, c_uint32, c_void_p
from functools import cache
from pathlib import Path

LIBRARY_PATH = Path(__file__).resolve().parent.parent / "liblbm_hip.so"

LATTICE_CODES = {"D2Q9": 0, "D3Q19": 1}
DTYPE_CODES = {"float32": 0, "float64": 1}
FORCE_UNIFORM, FORCE_FIELD, FORCE_DENSE = 0, 1, 2

_i, _d, _p, _i64 = c_int, c_double, c_void_p, c_int64

# name -> argument types; every entry point returns int (0 = success).
SIGNATURES: dict[str, list] = {
    "lbm_lattice_tables": [_i, POINTER(c_int), POINTER(c_int), POINTER(c_int), POINTER(c_double)],
    "lbm_lattice_tables_device": [_i, _i, _p, _p, _p, _p],
    "lbm_pull_step": [_i, _i, _p, _p, _p, _p, _p, _p, _i, _i, _i, _d, _d, _d, _d, _d, _i, _i, _i, _i, _i, _p],
    "lbm_pull_step_links": [_i, _i, _p, _p, _p, _p, _p, _p, _i, _i, _i, _d, _d, _d, _d, _d, _i,
                            _p, _p, _i, _p, _p, _p, _i64, _i, _p, _p, _p, _p, _i64, _i, _p, _p, _i, _i, _i, _p],
    "lbm_force_totals": [_p, _i64, _i, _i, _p, _p],
    "lbm_affine_force_groups": [_i64, POINTER(c_int)],
    "lbm_aa_even": [_i, _i, _p, _p, _p, _p, _i, _i, _i, _d, _d, _d, _d, _d, _i, _i, _i, _i, _i, _p],
    "lbm_aa_odd": [_i, _i, _p, _p, _p, _p, _p, _i, _i, _i, _d, _d, _d, _d, _d, _i, _i, _i, _i, _i, _p],
    "lbm_lds_step": [_i, _i, _p, _p, _p, _p, _p, _p, _i, _i, _i, _d, _d, _d, _d, _d, _i, _i, _i, _i, _i, _p],
    "lbm_init_equilibrium": [_i, _i, _p, _p, _p, _p, _p, _i64, _i, _p],
    "lbm_macroscopic": [_i, _i, _p, _p, _p, _i64, _i, _d, _d, _d, _i, _p, _p, _p],
    "lbm_health": [_i, _i, _p, _p, _i64, _i, _p, _p],
    "lbm_rules_apply": [_i, _p, _p, _p, _p, _p, _i64, _i, _p, _p],
    "lbm_rules_apply_direct": [_i, _p, _p, _p, _p, _p, _p, _i64, _i, _p],
    "lbm_rules_apply_noalloc": [_i, _p, _p, _p, _p, _p, _i64, _i, _p],
    "lbm_affine_values": [_i, _p, _p, _p, _p, _i64, _i, _p, _p],
    "lbm_affine_force": [_i, _p, _p, _p, _p, _p, _i64, _i, _i, _p, _p, _p],
    "lbm_affine_force_scratch_size": [_i64, POINTER(c_int64)],
    "lbm_affine_force_rules": [_i, _p, _p, _p, _p, _p, _i64, _i, _i, _p, _p, _p, _p, _p, _i64, _i, _p],
    "lbm_kernel_resources": [_i, _i, POINTER(c_int), _i, c_char_p, _i, POINTER(c_int)],
    "lbm_device_copy": [_p, _p, _i64, _p],
    "lbm_rules_apply_widths": [_i, _p, _p, _p, _p, _p, _p, _i64, _i, _p, _p],
    "lbm_batched_desc_size": [_i, POINTER(c_int64)],
    "lbm_batched_describe": [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p],
    "lbm_batched_pull_step": [_i, _i, _p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _p],
    "lbm_adj_pull_step": [_i, _i, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _d, _d, _d, _d, _d, _i, _i, _p, _p, _i, _i, _p],
    "lbm_adj_pull_step_rules": [_i, _i, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _d, _d, _d, _d, _d, _i, _i, _p, _p,
                                _i, _i, _p, _p, _p, _p, _p, _i64, _i, _i, _p],
    "lbm_adj_rules": [_i, _p, _p, _i, _p, _p, _p, _p, _i, _p, _p],
    "lbm_ibm_adjoint_scratch_size": [_i, _i, _i, _i, _i, POINTER(c_int64)],
    "lbm_ibm_adjoint": [_i, _i, _p, _p, _p, _p, _i, _p, _i, _p, _p, _p, _i, _p, _p, _i, _d, _d, _d, _i64, _p, _p,
                        _p, _p, _p, _p, _p, _i, _p, _p, _p, _i, _p, _p, _p, _p, _i, _p, _i, _p, _p],
    "lbm_ibm_adjoint_pull": [_i, _i, _p, _p, _p, _p, _i, _p, _i, _p, _p, _p, _i, _p, _p, _i, _d, _d, _d, _i64, _p,
                             _p, _p, _p, _p, _p, _p, _i, _p, _p, _p, _i, _p, _p, _p, _p, _i, _p, _i, _p, _p, _p, _p,
                             _p, _p, _p, _i, _i, _i, _d, _d, _i, _i, _p],
    "lbm_adj_fused_supported": [_i, _i, _i, _i, _i, _i, _i, _i, POINTER(c_int)],
    "lbm_adj_fused_step": [_i, _i, _p, _p, _p, _p, _i, _p, _i, _p, _p, _p, _i, _p, _p, _i, _d, _d, _d, _i64, _p, _p,
                           _p, _p, _p, _p, _p, _i, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _d, _d, _p, _p, _i, _i, _p,
                           _p, _i, _i, _i, _p, _p, _p, _p, _p, _i, _p, _p, _p, _i64, _i, _p, _p, _p, _p],
    "lbm_ibm_scratch_size": [_i, _i, _i, _i, POINTER(c_int64)],
    "lbm_ibm_pull_step_supported": [_i, _i, _i, _i, _i, _i, _i, _i, POINTER(c_int)],
    "lbm_ibm_pull_step": [_i, _i, _p, _p, _p, _p, _p, _p, _i, _i, _i, _d, _d, _d, _d, _d, _i, _p, _p, _p, _i, _p, _i,
                          _p, _p, _p, _i, _p, _p, _i, _p, _p, _p, _p, _i, _p, _p, _p, _i64, _i, _i, _i, _i, _p],
    "lbm_ibm_force": [_i, _i, _p, _p, _p, _p, _i, _p, _i, _p, _p, _p, _i, _p, _p, _i, _d, _d, _d, _i64, _p, _p, _p, _p,
                      _p],
    "lbm_ibm_record_size": [_i, _i, _i, _i, _i, POINTER(c_int64)],
    "lbm_ibm_force_multi": [_i, _i, _p, _p, _p, _p, _i, _p, _i, _p, _p, _p, _i, _p, _p, _i, _d, _d, _d, _i64, _p, _p,
                            _p, _p],
}


class NativeError(RuntimeError):
    """A library entry point returned an error code."""


class Library:
    def __init__(self, path: Path = LIBRARY_PATH):
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build() or python -m lbm._build")
        self.path = path
        self._lib = ctypes.CDLL(str(path))
        self._lib.lbm_last_error.restype = c_char_p
        self._lib.lbm_last_error.argtypes = []
        self._lib.lbm_version.restype = c_char_p
        self._lib.lbm_version.argtypes = []
        self._functions = {}
        for name, argtypes in SIGNATURES.items():
            fn = getattr(self._lib, name)
            fn.argtypes = argtypes
            fn.restype = c_int
            self._functions[name] = fn

    def has(self, name: str) -> bool:
        return hasattr(self._lib, name)

    def raw(self, name: str):
        return getattr(self._lib, name)

    @property
    def version(self) -> str:
        return self._lib.lbm_version().decode()

    def call(self, name: str, *args) -> None:
        code = self._functions[name](*args)
        if code != 0:
            self._raise(code)

    def call_all(self, calls) -> None:
        """Run a prepared launch list [(entry point name, argument tuple), ...] in order."""
        functions = self._functions
        for name, args in calls:
            code = functions[name](*args)
            if code != 0:
                self._raise(code)

    def _raise(self, code: int) -> None:
        message = self._lib.lbm_last_error().decode()
        if code == -1:
            raise ValueError(message)
        raise NativeError(f"{message} (code {code})")


@cache
def library() -> Library:
    return Library()


def pointer(tensor) -> int | None:
    """Device address of a torch tensor, or None for a missing optional argument."""
    return None if tensor is None else tensor.data_ptr()


def uint32_array(values) -> ctypes.Array:
    return (c_uint32 * len(values))(*values)
