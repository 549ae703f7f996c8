"""Exceptions of the GPU path.

``Unsupported`` plays the role of the reference's ``gpu.xc.Unsupported``
(gpu/xc.py:30-32): input outside the GPU path's scope; the hooks fall back
to the original PySCF callable.  ``GpuOOM`` triggers the halve-and-retry
policy (gpu/install.py:99-109) before falling back.
"""


class Unsupported(RuntimeError):
    pass


class GpuError(RuntimeError):
    pass


class GpuOOM(GpuError, MemoryError):
    pass


def is_oom(exc):
    """Reference policy (gpu/device.py:62-70): MemoryError-like errors."""
    return isinstance(exc, MemoryError) or 'out of memory' in str(exc).lower()
