"""ctypes mirror of include/spice21_mi355x.h (the engine's C ABI).

Only plain data layouts live here; loading the shared library is done by ``spice21_amd._lib``.
"""
import ctypes as C

S21_GND = -1
S21_PARAM_SLOTS = 40

# status codes
S21_OK, S21_ERR, S21_CONVERGENCE, S21_SINGULAR, S21_PIVOT, S21_DECODE, S21_INVALID, S21_DEVICE, S21_NONFINITE = range(9)
STATUS_NAMES = {
    S21_OK: "ok",
    S21_ERR: "error",
    S21_CONVERGENCE: "Convergence Failed",
    S21_SINGULAR: "Singular Matrix",
    S21_PIVOT: "Pivot Search Fail",
    S21_DECODE: "Decode Error",
    S21_INVALID: "invalid argument",
    S21_DEVICE: "HIP device error",
    S21_NONFINITE: "non-finite Newton update (reserved)",
}

# variable and component kinds (analysis.rs:33-38, comps/mod.rs:60-72)
S21_VAR_V, S21_VAR_I = 0, 1
S21_COMP_R, S21_COMP_C, S21_COMP_I, S21_COMP_V, S21_COMP_MOS1, S21_COMP_MOS0, S21_COMP_DIODE, S21_COMP_BSIM4 = range(8)
S21_OVR_VALUE, S21_OVR_MODEL, S21_OVR_INST = 0, 1, 2
S21_FP64, S21_MIXED = 0, 1

# proto field numbers of the parameter messages (protos/mos.proto:22-76, protos/spice21.proto:43-68)
MOS1_MODEL_FIELDS = dict(
    mos_type=2, vt0=3, kp=4, gamma=5, phi=6, lambda_=7, rd=8, rs=9, cbd=10, cbs=11, is_=12, pb=13, cgso=14,
    cgdo=15, cgbo=16, rsh=17, cj=18, mj=19, cjsw=20, mjsw=21, js=22, tox=23, ld=24, u0=25, fc=26, nsub=27,
    tpg=28, nss=29, tnom=30, kf=31, af=32,
)
MOS1_INST_FIELDS = dict(m=2, l=3, w=4, a_d=5, a_s=6, pd=7, ps=8, nrd=9, nrs=10, temp=11)
DIODE_MODEL_FIELDS = dict(
    tnom=2, is_=3, n=4, tt=5, vj=6, m=7, eg=8, xti=9, kf=10, af=11, fc=12, bv=13, ibv=14, rs=15, cj0=16,
)
DIODE_INST_FIELDS = dict(area=4, temp=5)
# protos/bsim4.proto:44-46 (Bsim4Model) and :6-41 (Bsim4InstParams)
BSIM4_MODEL_FIELDS = dict(mos_type=1)
BSIM4_INST_FIELDS = dict(
    l=1, w=2, nf=3, sa=4, sb=5, sd=6, sca=7, scb=8, scc=9, sc=10, ad=11, as_=12, pd=13, ps=14, nrd=15, nrs=16,
    min=17, rgeomod=18, rbdb=19, rbsb=20, rbpb=21, rbps=22, rbpd=23, delvto=24, xgw=25, ngcon=26, trnqsmod=27,
    acnqsmod=28, rbodymod=29, rgatemod=30, geomod=31)


def _field(table, name):
    key = name if name in table else name + "_"
    if key not in table:
        raise KeyError(f"unknown parameter {name!r}")
    return table[key]


class s21_params(C.Structure):
    _fields_ = [("v", C.c_double * S21_PARAM_SLOTS), ("set", C.c_uint64)]

    @classmethod
    def make(cls, table, **kw):
        p = cls()
        for name, val in kw.items():
            if val is None:
                continue
            f = _field(table, name)
            p.v[f] = float(val)
            p.set |= 1 << f
        return p


class s21_comp(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("node", C.c_int32 * 6),
        ("model", C.c_int32),
        ("inst", C.c_int32),
        ("value", C.c_double),
        ("acm", C.c_double),
    ]


class s21_netlist(C.Structure):
    _fields_ = [
        ("n_vars", C.c_int32),
        ("var_kind", C.POINTER(C.c_int32)),
        ("var_name", C.POINTER(C.c_char_p)),
        ("n_comps", C.c_int32),
        ("comps", C.POINTER(s21_comp)),
        ("n_mos1_models", C.c_int32),
        ("mos1_models", C.POINTER(s21_params)),
        ("n_mos1_insts", C.c_int32),
        ("mos1_insts", C.POINTER(s21_params)),
        ("n_diode_models", C.c_int32),
        ("diode_models", C.POINTER(s21_params)),
        ("n_diode_insts", C.c_int32),
        ("diode_insts", C.POINTER(s21_params)),
        ("n_bsim4_models", C.c_int32),
        ("bsim4_models", C.POINTER(s21_params)),
        ("n_bsim4_insts", C.c_int32),
        ("bsim4_insts", C.POINTER(s21_params)),
    ]


class s21_options(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("temp", "tnom", "gmin", "iabstol", "reltol")]

    @classmethod
    def default(cls, **kw):
        o = cls(300.15, 300.15, 1e-12, 1e-12, 1e-3)  # analysis.rs:665-693
        for k, v in kw.items():
            setattr(o, k, v)
        return o


class s21_tran_options(C.Structure):
    _fields_ = [
        ("tstep", C.c_double),
        ("tstop", C.c_double),
        ("n_ic", C.c_int32),
        ("ic_var", C.POINTER(C.c_int32)),
        ("ic_val", C.POINTER(C.c_double)),
        ("ic_name", C.POINTER(C.c_char_p)),
    ]


class s21_ac_options(C.Structure):
    _fields_ = [("fstart", C.c_uint64), ("fstop", C.c_uint64), ("npts", C.c_uint64)]


class s21_override(C.Structure):
    _fields_ = [
        ("comp", C.c_int32),
        ("target", C.c_int32),
        ("field", C.c_int32),
        ("vals", C.POINTER(C.c_double)),
    ]


class s21_group(C.Structure):
    _fields_ = [
        ("net", C.POINTER(s21_netlist)),
        ("n_inst", C.c_int32),
        ("n_over", C.c_int32),
        ("over", C.POINTER(s21_override)),
    ]


class s21_trap_options(C.Structure):
    _fields_ = [("lte", C.c_int32), ("reltol", C.c_double), ("abstol", C.c_double), ("trtol", C.c_double),
                ("dt_min", C.c_double)]
