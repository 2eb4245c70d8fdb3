"""MI355X-native engine for Spice21's Newton inner loop.

* ``spice21_amd.api``      -- the reference's Python analysis API (dcop / tran / ac over protobuf)
* ``spice21_amd.batch``    -- batched analyses of many circuit instances on one GPU
* ``spice21_amd.netlist``  -- flattened-netlist builder for the batch API
* ``spice21_amd.elab``     -- protobuf circuits through the engine's own decoder and elaborator
* ``spice21_amd.protos``   -- protobuf message classes of the wire schema
"""
from ._lib import S21Error  # noqa: F401


def __getattr__(name):
    if name in ("dcop", "tran", "ac", "add", "circuit", "health"):
        from . import api

        return getattr(api, name)
    raise AttributeError(name)
