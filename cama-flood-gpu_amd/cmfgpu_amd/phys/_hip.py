"""HIP backend factories: each returns a callable that runs one C-ABI entry
point (include/cmf_hip.h) on a bound model's field table and network."""
from __future__ import annotations

from .. import _lib


def hip_op(op_name: str):
    symbol = _lib.OP_SYMBOLS[op_name]

    def factory():
        fn = getattr(_lib.lib(), symbol)

        def run(model):
            _lib.check(fn(model.network_handle, model.field_table_ref(), _lib.current_stream()), op_name)

        return run

    return factory
