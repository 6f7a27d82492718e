"""Zero-copy tensor views of device buffers through DLPack.

The reference exports its CUDA-imported shared buffers to PyTorch as DLPack capsules
(``Runtime.shared_buffer_to_dlpack``, python/bindings.cpp:282-340 and :3303, with the tensor
descriptor of :500-600).  Here the resident buffers are HIP allocations already, so no interop
import is needed: ``to_dlpack`` wraps a device pointer, a shape and a dtype in a ``DLManagedTensor``
(the DLPack ABI: dlpack.h v0.8 structs) inside a ``"dltensor"`` capsule that
``torch.from_dlpack`` consumes without copying.  The capsule keeps ``owner`` alive until the
consumer's tensor is released (the lease the reference takes on its shared buffer).
"""
from __future__ import annotations

import ctypes as C

KDL_CPU = 1
KDL_ROCM = 10

DTYPES = {  # name -> (DLDataTypeCode, bits)
    "float32": (2, 32),
    "int32": (0, 32),
    "uint32": (1, 32),
    "uint16": (1, 16),
}


class DLDevice(C.Structure):
    _fields_ = [("device_type", C.c_int32), ("device_id", C.c_int32)]


class DLDataType(C.Structure):
    _fields_ = [("code", C.c_uint8), ("bits", C.c_uint8), ("lanes", C.c_uint16)]


class DLTensor(C.Structure):
    _fields_ = [("data", C.c_void_p), ("device", DLDevice), ("ndim", C.c_int32), ("dtype", DLDataType),
                ("shape", C.POINTER(C.c_int64)), ("strides", C.POINTER(C.c_int64)), ("byte_offset", C.c_uint64)]


class DLManagedTensor(C.Structure):
    pass


_DELETER = C.CFUNCTYPE(None, C.POINTER(DLManagedTensor))
DLManagedTensor._fields_ = [("dl_tensor", DLTensor), ("manager_ctx", C.c_void_p), ("deleter", _DELETER)]

# the destructor receives the capsule while it is being freed: a raw pointer, never a Python reference
_CAPSULE_DESTRUCTOR = C.CFUNCTYPE(None, C.c_void_p)
_PyCapsule_New = C.pythonapi.PyCapsule_New
_PyCapsule_New.restype = C.py_object
_PyCapsule_New.argtypes = [C.c_void_p, C.c_char_p, _CAPSULE_DESTRUCTOR]
_PyCapsule_IsValid = C.pythonapi.PyCapsule_IsValid
_PyCapsule_IsValid.restype = C.c_int
_PyCapsule_IsValid.argtypes = [C.c_void_p, C.c_char_p]
_PyCapsule_GetPointer = C.pythonapi.PyCapsule_GetPointer
_PyCapsule_GetPointer.restype = C.c_void_p
_PyCapsule_GetPointer.argtypes = [C.c_void_p, C.c_char_p]

# managed-tensor address -> (struct, shape, strides, owner): alive until the consumer's deleter runs
_live: dict = {}


@_DELETER
def _deleter(p):
    entry = _live.pop(C.addressof(p.contents), None)
    if entry is not None and hasattr(entry[3], "_release_view"):
        entry[3]._release_view()


@_CAPSULE_DESTRUCTOR
def _capsule_destructor(cap):
    # an unconsumed capsule still owns its tensor; a consumed one was renamed "used_dltensor"
    if _PyCapsule_IsValid(cap, b"dltensor"):
        ptr = _PyCapsule_GetPointer(cap, b"dltensor")
        if ptr:
            _deleter(C.cast(ptr, C.POINTER(DLManagedTensor)))


def to_dlpack(ptr: int, shape, dtype: str, device_type: int, device_id: int, owner=None, strides=None,
              byte_offset: int = 0, size_bytes: int | None = None):
    """A ``"dltensor"`` capsule over ``ptr`` (compact row-major unless ``strides`` in elements are
    given).  Raises ValueError on a non-positive shape or stride, an unknown dtype or, when
    ``size_bytes`` is given, a view that would reach past the buffer (the reference's checks)."""
    if not ptr:
        raise ValueError("null buffer")
    shape = [int(d) for d in shape]
    if not shape or any(d <= 0 for d in shape):
        raise ValueError("tensor shape dimensions must be positive")
    if dtype not in DTYPES:
        raise ValueError(f"unsupported dtype {dtype}")
    code, bits = DTYPES[dtype]
    if strides is None:
        strides = [1] * len(shape)
        for i in range(len(shape) - 1, 0, -1):
            strides[i - 1] = strides[i] * shape[i]
    strides = [int(s) for s in strides]
    if len(strides) != len(shape) or any(s <= 0 for s in strides):
        raise ValueError("tensor strides must be positive and match the shape's rank")
    last = sum(s * (d - 1) for s, d in zip(strides, shape))
    if size_bytes is not None and byte_offset + (last + 1) * (bits // 8) > size_bytes:
        raise ValueError("tensor view exceeds the buffer")
    cshape = (C.c_int64 * len(shape))(*shape)
    cstrides = (C.c_int64 * len(shape))(*strides)
    m = DLManagedTensor()
    m.dl_tensor = DLTensor(C.c_void_p(ptr), DLDevice(device_type, device_id), len(shape), DLDataType(code, bits, 1),
                           cshape, cstrides, byte_offset)
    m.manager_ctx = None
    m.deleter = _deleter
    _live[C.addressof(m)] = (m, cshape, cstrides, owner)
    if owner is not None and hasattr(owner, "_acquire_view"):
        owner._acquire_view()
    return _PyCapsule_New(C.addressof(m), b"dltensor", _capsule_destructor)
