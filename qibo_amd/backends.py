"""Backend registry: the plugin entry point and the global-backend helpers.

qibo loads a backend plugin with ``import_module(name).MetaBackend.load(**kw)``
(``src/qibo/backends/__init__.py:325-350``) and lists plugins with
``MetaBackend().list_available()`` (:312-322). ``MetaBackend`` here is that
entry point; ``qibo.set_backend("qibo_amd")`` therefore returns an
``MI355XBackend``. The module also carries qibo's global-backend helpers
(``_Global``, ``set_backend``, ``get_backend``, ``_check_backend``,
``set_dtype``, ``set_seed``; :81-298) for the circuits of this package, which
run without qibo installed.
"""

from .config import raise_error


class MetaBackend:
    """Plugin loader (``MetaBackend.load`` / ``list_available``)."""

    @staticmethod
    def load(platform=None, **kwargs):
        from .backend import MI355XBackend

        if platform not in (None, "mi355x", "hip"):
            raise_error(ValueError, f"Unsupported platform {platform} for qibo_amd.")
        return MI355XBackend(**kwargs)

    def list_available(self) -> dict:
        try:
            from . import _lib

            _lib.load()
            import torch

            ok = bool(torch.cuda.is_available())
        except Exception:  # noqa: BLE001 - availability probe
            ok = False
        return {"mi355x": ok}


class _Global:
    _backend = None
    _dtype = "complex128"

    @classmethod
    def backend(cls):
        if cls._backend is None:
            cls._backend = MetaBackend.load(dtype=cls._dtype)
        return cls._backend

    @classmethod
    def set_backend(cls, backend, **kwargs):
        cls._backend = construct_backend(backend, **kwargs)


def construct_backend(backend, **kwargs):
    if backend in ("qibo_amd", "qibo-amd", "mi355x"):
        return MetaBackend.load(**kwargs)
    raise_error(ValueError, f"Backend {backend} is not available in qibo_amd.")


def set_backend(backend="qibo_amd", **kwargs):
    _Global.set_backend(backend, **kwargs)


def get_backend():
    return _Global.backend()


def _check_backend(backend=None):
    return _Global.backend() if backend is None else backend


def set_dtype(dtype):
    _Global._dtype = dtype
    if _Global._backend is not None:
        _Global._backend.set_dtype(dtype)


def get_dtype():
    return _Global.backend().dtype


def set_seed(seed):
    _Global.backend().set_seed(seed)


def list_available_backends() -> dict:
    return {"qibo_amd": MetaBackend().list_available()}
