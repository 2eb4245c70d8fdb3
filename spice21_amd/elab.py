"""Protobuf circuits through the engine's own decoder and elaborator (circuit.rs, elab.rs restated in
``csrc/circuit.cpp``).  ``elaborate`` returns the flattened netlist the batch API takes."""
import ctypes as C

from . import _abi as A
from ._lib import check, lib


class Elaborated:
    """Flattened netlist owned by the engine library (``s21_elab``); usable wherever a CNetlist is."""

    def __init__(self, ckt, mos0=None):
        L = lib()
        L.s21_ckt_decode.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_void_p)]
        L.s21_ckt_add_mos0.argtypes = [C.c_void_p, C.c_char_p, C.c_int32]
        L.s21_ckt_free.argtypes = [C.c_void_p]
        L.s21_elaborate.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.s21_elab_netlist.argtypes = [C.c_void_p, C.POINTER(A.s21_netlist)]
        L.s21_elab_free.argtypes = [C.c_void_p]
        buf = ckt if isinstance(ckt, (bytes, bytearray)) else ckt.SerializeToString()
        h = C.c_void_p()
        check(L.s21_ckt_decode(bytes(buf), len(buf), C.byref(h)))
        try:
            for name, typ in (mos0 or {}).items():
                check(L.s21_ckt_add_mos0(h, name.encode(), int(typ)))
            e = C.c_void_p()
            check(L.s21_elaborate(h, C.byref(e)))
        finally:
            L.s21_ckt_free(h)
        self._h = e
        self.net = A.s21_netlist()
        check(L.s21_elab_netlist(e, C.byref(self.net)))
        self.n_vars = self.net.n_vars
        self.var_names = [self.net.var_name[k].decode() for k in range(self.n_vars)]
        self.var_kinds = [self.net.var_kind[k] for k in range(self.n_vars)]
        self.comps = self.net.comps

    @property
    def ptr(self):
        return C.byref(self.net)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().s21_elab_free(self._h)
            self._h = None


def elaborate(ckt, mos0=None):
    return Elaborated(ckt, mos0)
