"""Spice21's Python analysis API (spice21py/spice21py/__init__.py:16-161), served by the MI355X engine.

Each call encodes its protobuf argument, runs the engine's byte ABI (the counterpart of
``CallableProto::call_bytes``, spice21/src/proto.rs:32-45) and decodes the protobuf result into a dict.
Engine errors raise :class:`spice21_amd.S21Error`, a ``RuntimeError`` carrying the reference's error text,
as the PyO3 binding does (spice21py/src/lib.rs:20-25).
"""
import ctypes as C
from typing import Any, Dict, List, Optional, Union

from . import protos
from ._lib import S21Error, check, lib
from .protos import Capacitor, Diode, Isrc, Mos, Resistor, Vsrc


def health() -> str:
    lib()
    return "alive"


def _call(fn_name: str, enc: bytes) -> bytes:
    L = lib()
    fn = getattr(L, fn_name)
    fn.restype = C.c_int
    fn.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t)]
    L.s21_free_bytes.argtypes = [C.POINTER(C.c_uint8)]
    out = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    check(fn(enc, len(enc), C.byref(out), C.byref(n)))
    try:
        return C.string_at(out, n.value)
    finally:
        L.s21_free_bytes(out)


def _arg(arg, kind, ckt, opts, args=None):
    if arg is None:
        return kind(ckt=ckt, opts=opts, args=args) if args is not None or kind is not protos.Op else kind(ckt=ckt, opts=opts)
    if isinstance(arg, protos.Circuit):
        if kind is protos.Op:
            return kind(ckt=arg, opts=opts)
        return kind(ckt=arg, opts=opts, args=args)
    if isinstance(arg, kind):
        return arg
    raise TypeError


def dcop(arg: Optional[Union["protos.Circuit", "protos.Op"]] = None, *, ckt=None, opts=None) -> Dict[str, float]:
    """DC operating point."""
    op = _arg(arg, protos.Op, ckt, opts)
    rv = protos.OpResult()
    rv.ParseFromString(_call("s21_op_bytes", op.SerializeToString()))
    return dict(rv.vals)


def tran(arg=None, *, ckt=None, opts=None, args=None) -> Dict[str, List[float]]:
    """Transient analysis."""
    tr = _arg(arg, protos.Tran, ckt, opts, args)
    rv = protos.TranResult()
    rv.ParseFromString(_call("s21_tran_bytes", tr.SerializeToString()))
    return {name: list(arr.vals) for name, arr in dict(rv.vals).items()}


def ac(arg=None, *, ckt=None, opts=None, args=None) -> Dict[str, List[complex]]:
    """AC analysis."""
    a = _arg(arg, protos.Ac, ckt, opts, args)
    rv = protos.AcResult()
    rv.ParseFromString(_call("s21_ac_bytes", a.SerializeToString()))
    return {name: [complex(c.re, c.im) for c in arr.vals] for name, arr in dict(rv.vals).items()}


def add(ckt: "protos.Circuit", comp: Any) -> None:
    """Append an instance or definition to a circuit, wrapping it in Instance / Def."""
    if isinstance(comp, protos.Instance):
        return ckt.comps.append(comp)
    for cls, key in ((Diode, "d"), (Capacitor, "c"), (Resistor, "r"), (Mos, "m"), (Isrc, "i"), (Vsrc, "v"),
                     (protos.ModuleInstance, "x")):
        if isinstance(comp, cls):
            return ckt.comps.append(protos.Instance(**{key: comp}))
    for cls, key in ((protos.Bsim4Model, "bsim4model"), (protos.Bsim4InstParams, "bsim4inst"),
                     (protos.Mos1Model, "mos1model"), (protos.Mos1InstParams, "mos1inst"),
                     (protos.DiodeModel, "diodemodel"), (protos.DiodeInstParams, "diodeinst"),
                     (protos.Module, "module")):
        if isinstance(comp, cls):
            return ckt.defs.append(protos.Def(**{key: comp}))
    raise TypeError(f"Invalid Circuit Component {type(comp)}")


def circuit(*args) -> "protos.Circuit":
    """Circuit from a component list."""
    items = args[0] if len(args) == 1 and isinstance(args[0], (list, tuple)) else args
    c = protos.Circuit()
    for a in items:
        add(c, a)
    return c


__all__ = ["health", "dcop", "tran", "ac", "add", "circuit", "S21Error"]
