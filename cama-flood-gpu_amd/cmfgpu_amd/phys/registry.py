"""Per-op backend registry (the reference's hydroforge BackendRegistry surface).

``BackendRegistry({"hip": factory}, name=..., spec=...).selected`` returns a
zero-argument callable, as at cmfgpu/phys/outflow.py:51-60: calling it runs
the op on the fields of the model that is currently bound (a CaMaFlood calls
its ops inside ``bind``).  Only the "hip" backend exists; selecting any other
name raises, and there is no CPU fallback.
"""
from __future__ import annotations

import contextlib
import threading

from .specs import KernelSpec

_BOUND = threading.local()


@contextlib.contextmanager
def bind(model):
    """Bind `model` as the target of zero-argument op calls on this thread."""
    prev = getattr(_BOUND, "model", None)
    _BOUND.model = model
    try:
        yield model
    finally:
        _BOUND.model = prev


def bound_model():
    model = getattr(_BOUND, "model", None)
    if model is None:
        raise RuntimeError("no CaMaFlood model is bound; call ops inside `with bind(model):`")
    return model


class SelectedOp:
    def __init__(self, name: str, spec: KernelSpec, impl):
        self.name = name
        self.spec = spec
        self._impl = impl

    def __call__(self, model=None):
        model = bound_model() if model is None else model
        model.check_spec(self.spec)
        return self._impl(model)

    def __repr__(self):
        return f"<op {self.name} (hip)>"


class BackendRegistry:
    BACKENDS = ("hip",)

    def __init__(self, factories: dict, *, name: str, spec: KernelSpec, backend: str = "hip"):
        if spec.name != name:
            raise ValueError(f"spec {spec.name!r} does not match op name {name!r}")
        unknown = set(factories) - set(self.BACKENDS)
        if unknown:
            raise ValueError(f"{name}: unsupported backends {sorted(unknown)}; this build provides 'hip' only")
        if backend not in factories:
            raise ValueError(f"{name}: backend {backend!r} is not registered")
        self.name = name
        self.spec = spec
        self.factories = dict(factories)
        self.backend = backend
        self._selected = None

    @property
    def selected(self) -> SelectedOp:
        if self._selected is None:
            self._selected = SelectedOp(self.name, self.spec, self.factories[self.backend]())
        return self._selected
