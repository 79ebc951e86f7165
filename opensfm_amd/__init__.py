"""MI355X-native PatchMatch multi-view depth estimation.

The compute path is the HIP library ``libopensfm_pm.so`` (C ABI in
``include/opensfm_pm.h``) and the pybind11 module ``pydense`` over it, both
built in-tree by ``python -m opensfm_amd.build``.  Importing this package
does not load them; ``opensfm_amd.native()`` does, and raises if they are
missing.
"""
import os

__all__ = ["native", "library_path"]

_HERE = os.path.dirname(os.path.abspath(__file__))


def library_path() -> str:
    return os.path.join(_HERE, "libopensfm_pm.so")


def _share_torch_runtime() -> None:
    """Load PyTorch's HIP runtime before the library's.

    torch ships its own libamdhip64.so.7; the dynamic loader keeps the first
    library of that name for the whole process.  Loading torch's first
    makes the library and torch share one runtime, so torch tensors and the
    library can be used together (the autograd wrapper); the other order
    leaves torch without a device.
    """
    try:
        import torch  # noqa: F401
    except ImportError:  # pragma: no cover - torch is optional for the library itself
        pass


def native():
    """Import the compiled ``pydense`` module (RuntimeError if it was not built)."""
    _share_torch_runtime()
    try:
        from . import pydense  # noqa: F401
    except ImportError as e:  # pragma: no cover - depends on the build
        raise RuntimeError(
            "opensfm_amd native extension is not built; run `python -m opensfm_amd.build`") from e
    return pydense
