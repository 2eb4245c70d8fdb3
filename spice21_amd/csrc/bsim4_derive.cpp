// BSIM4 parameter derivation on the host: the model values at their defaults (model/vals.rs::resolve), the
// temperature-dependent model values (bsim4derive.rs::derive) and the size-dependent and instance values
// (bsim4inst.rs::from), for one polarity and one set of instance parameters.  Only what the operating
// point, its stamps and its charges read at the default mode selectors is kept (bsim4.h).
#include <cmath>
#include <string>

#include "bsim4.h"

namespace s21 {
namespace b4 {

namespace {
// The model values of model/vals.rs::resolve that derivation reads beyond B4M.
struct ModelExtra {
    double dtox, nsub, phig, epsrgate, easub, epsrsub, ni0sub, bg0sub, tbgasub, tbgbsub, vbm, xt, kt1, kt1l, kt2, k3,
        k3b, w0, lpe0, lpeb, dvtp0, dvtp1, dvtp2, dvtp3, dvtp4, dvtp5, dvt0, dvt1, dvt2, dvt0w, dvt1w, dvt2w, drout,
        dsub, vth0, vfb, eu, ucs, ua, ua1, ub, ub1, uc, uc1, ud, ud1, up, lp, u0, ute, ucste, voff, minv, minvcv,
        fprout, pdits, pditsd, delta, rdsw, rdw, rsw, prwg, prwb, prt, eta0, etab, pclm, pdiblc1, pdiblc2, pdiblcb,
        pscbe1, pscbe2, pvag, wr, dwg, dwb, b0, b1, alpha0, alpha1, beta0, agidl, bgidl, cgidl, egidl, rgidl, kgidl,
        fgidl, lambda, vtl, xn, lc, vfbsdoff, tvfbsdoff, tvoff, tnfactor, teta0, tvoffcv, cdsc, cdscb, cdscd, cit,
        nfactor, vsat, at, a0, ags, a1, a2, keta, clc, cle, vfbcv, acde, moin, noff, voffcv, rdswmin, rdwmin, rswmin,
        jss, jsd, jsws, jswd, jswgs, jswgd, njs, njd, xtis, xtid, tcj, tcjsw, tcjswg, tpb, tpbsw, tpbswg, njts,
        njtssw, njtsswg, njtsd, njtsswd, njtsswgd, tnjts, tnjtssw, tnjtsswg, tnjtsd, tnjtsswd, tnjtsswgd, jtss, jtsd,
        jtssws, jtsswd, jtsswgs, jtsswgd, jtweff, xtss, xtsd, xtssws, xtsswd, xtsswgs, xtsswgd, lint, wint, ll, lw,
        lwl, llc, lwc, lwlc, lln, lwn, wl, ww, wwl, wlc, wwc, wwlc, wln, wwn, dwc, dwj, xl, xw, saref, sbref, wlod,
        ku0, kvsat, kvth0, tku0, llodku0, wlodku0, llodvth, wlodvth, lku0, wku0, pku0, lkvth0, wkvth0, pkvth0, stk2,
        lodk2, steta0, lodeta0, dmcg, dmci, dmdg, dmcgt, rsh, bdos, ados, pditsl, toxref;
};

void defaults(bool pmos, B4M& m, ModelExtra& x) {  // model/vals.rs::resolve, no parameter given
    m = B4M{};
    x = ModelExtra{};
    m.p = pmos ? -1.0 : 1.0;
    m.tnom = 300.15;
    x.toxref = 30.0e-10;
    m.toxe = 30.0e-10;
    m.toxp = m.toxe;
    m.toxm = m.toxe;
    x.dtox = 0.0;
    m.epsrox = 3.9;
    x.cdsc = 2.4e-4;
    x.cdscb = 0.0;
    x.cdscd = 0.0;
    x.cit = 0.0;
    x.nfactor = 1.0;
    m.xj = 0.15e-6;
    x.vsat = 8.0e4;
    x.at = 3.3e4;
    x.a0 = 1.0;
    x.ags = 0.0;
    x.a1 = 0.0;
    x.a2 = 1.0;
    x.keta = -0.047;
    x.nsub = 6.0e16;
    x.phig = 4.05;
    x.epsrgate = 11.7;
    x.easub = 4.05;
    x.epsrsub = 11.7;
    x.ni0sub = 1.45e10;
    x.bg0sub = 1.16;
    x.tbgasub = 7.02e-4;
    x.tbgbsub = 1108.0;
    m.ndep = 1.7e17;
    m.nsd = 1.0e20;
    m.phin = 0.0;
    m.ngate = 0.0;
    x.vbm = -3.0;
    x.xt = 1.55e-7;
    x.kt1 = -0.11;
    x.kt1l = 0.0;
    x.kt2 = 0.022;
    x.k3 = 80.0;
    x.k3b = 0.0;
    x.w0 = 2.5e-6;
    x.lpe0 = 1.74e-7;
    x.lpeb = 0.0;
    x.dvtp0 = x.dvtp1 = x.dvtp2 = x.dvtp3 = x.dvtp4 = x.dvtp5 = 0.0;
    x.dvt0 = 2.2;
    x.dvt1 = 0.53;
    x.dvt2 = -0.032;
    x.dvt0w = 0.0;
    x.dvt1w = 5.3e6;
    x.dvt2w = -0.032;
    x.drout = 0.56;
    x.dsub = x.drout;
    x.vth0 = pmos ? -0.7 : 0.7;
    x.vfb = -1.0;
    x.eu = pmos ? 1.0 : 1.67;
    x.ucs = pmos ? 1.0 : 1.67;
    x.ua = 1.0e-9;  // mobmod 0
    x.ua1 = 1.0e-9;
    x.ub = 1.0e-19;
    x.ub1 = -1.0e-18;
    x.uc = -0.0465e-9;
    x.uc1 = -0.056e-9;
    x.ud = 0.0;
    x.ud1 = 0.0;
    x.up = 0.0;
    x.lp = 1.0e-8;
    x.u0 = pmos ? 0.025 : 0.067;
    x.ute = -1.5;
    x.ucste = -4.775e-3;
    x.voff = -0.08;
    m.voffl = 0.0;
    m.voffcvl = 0.0;
    x.minv = 0.0;
    x.minvcv = 0.0;
    x.fprout = 0.0;
    x.pdits = 0.0;
    x.pditsd = 0.0;
    m.pditsl = 0.0;
    x.delta = 0.01;
    x.rdswmin = 0.0;
    x.rdwmin = 0.0;
    x.rswmin = 0.0;
    x.rdsw = 200.0;
    x.rdw = 100.0;
    x.rsw = 100.0;
    x.prwg = 1.0;
    x.prwb = 0.0;
    x.prt = 0.0;
    x.eta0 = 0.08;
    x.etab = -0.07;
    x.pclm = 1.3;
    x.pdiblc1 = 0.39;
    x.pdiblc2 = 0.0086;
    x.pdiblcb = 0.0;
    x.pscbe1 = 4.24e8;
    x.pscbe2 = 1.0e-5;
    x.pvag = 0.0;
    x.wr = 1.0;
    x.dwg = 0.0;
    x.dwb = 0.0;
    x.b0 = 0.0;
    x.b1 = 0.0;
    x.alpha0 = 0.0;
    x.alpha1 = 0.0;
    x.beta0 = 0.0;
    x.agidl = 0.0;
    x.bgidl = 2.3e9;
    x.cgidl = 0.5;
    x.egidl = 0.8;
    x.rgidl = 1.0;
    x.kgidl = 0.0;
    x.fgidl = 1.0;
    m.ijthsfwd = 0.1;
    m.ijthdfwd = m.ijthsfwd;
    x.lambda = 0.0;
    x.vtl = 2.0e5;
    x.xn = 3.0;
    x.lc = 5.0e-9;
    x.vfbsdoff = 0.0;
    x.tvfbsdoff = 0.0;
    x.tvoff = 0.0;
    x.tnfactor = 0.0;
    x.teta0 = 0.0;
    x.tvoffcv = 0.0;
    m.xjbvs = 1.0;
    m.xjbvd = m.xjbvs;
    m.bvs = 10.0;
    m.bvd = m.bvs;
    m.cgsl = 0.0;
    m.cgdl = 0.0;
    m.ckappas = 0.6;
    m.ckappad = m.ckappas;
    x.clc = 0.1e-6;
    x.cle = 0.6;
    x.vfbcv = -1.0;
    x.acde = 1.0;
    x.moin = 15.0;
    x.noff = 1.0;
    x.voffcv = 0.0;
    x.dmcg = 0.0;
    x.dmci = x.dmcg;
    x.dmdg = 0.0;
    x.dmcgt = 0.0;
    x.lint = 0.0;
    x.ll = 0.0;
    x.llc = x.ll;
    x.lln = 1.0;
    x.lw = 0.0;
    x.lwc = x.lw;
    x.lwn = 1.0;
    x.lwl = 0.0;
    x.lwlc = x.lwl;
    x.wint = 0.0;
    x.wl = 0.0;
    x.wlc = x.wl;
    x.wln = 1.0;
    x.ww = 0.0;
    x.wwc = x.ww;
    x.wwn = 1.0;
    x.wwl = 0.0;
    x.wwlc = x.wwl;
    x.dwc = x.wint;
    m.dlc = x.lint;
    x.xl = 0.0;
    x.xw = 0.0;
    x.dwj = x.dwc;
    m.cf = 2.0 * m.epsrox * EPS0 / PI * std::log(1.0 + 0.4e-6 / m.toxe);
    m.xpart = 0.0;
    x.rsh = 0.0;
    m.cjs = 5.0E-4;
    m.cjd = m.cjs;
    m.cjsws = 5.0E-10;
    m.cjswd = m.cjsws;
    m.cjswgs = m.cjsws;
    m.cjswgd = m.cjswgs;
    x.jss = 1.0E-4;
    x.jsd = x.jss;
    x.jsws = 0.0;
    x.jswd = x.jsws;
    x.jswgs = 0.0;
    x.jswgd = x.jswgs;
    m.pbs = 1.0;
    m.pbd = m.pbs;
    m.pbsws = 1.0;
    m.pbswd = m.pbsws;
    m.pbswgs = m.pbsws;
    m.pbswgd = m.pbswgs;
    m.mjs = 0.5;
    m.mjd = m.mjs;
    m.mjsws = 0.33;
    m.mjswd = m.mjsws;
    m.mjswgs = m.mjsws;
    m.mjswgd = m.mjswgs;
    x.njs = 1.0;
    x.njd = x.njs;
    x.xtis = 3.0;
    x.xtid = x.xtis;
    x.jtss = 0.0;
    x.jtsd = x.jtss;
    x.jtssws = 0.0;
    x.jtsswd = x.jtssws;
    x.jtsswgs = 0.0;
    x.jtsswgd = x.jtsswgs;
    x.jtweff = 0.0;
    x.njts = 20.0;
    x.njtssw = 20.0;
    x.njtsswg = 20.0;
    x.njtsd = x.njts;
    x.njtsswd = x.njtssw;
    x.njtsswgd = x.njtsswg;
    x.xtss = 0.02;
    x.xtsd = x.xtss;
    x.xtssws = 0.02;
    x.xtsswd = x.xtssws;
    x.xtsswgs = 0.02;
    x.xtsswgd = x.xtsswgs;
    x.tnjts = 0.0;
    x.tnjtssw = 0.0;
    x.tnjtsswg = 0.0;
    x.tnjtsd = x.tnjts;
    x.tnjtsswd = x.tnjtssw;
    x.tnjtsswgd = x.tnjtsswg;
    m.vtss = 10.0;
    m.vtsd = m.vtss;
    m.vtssws = 10.0;
    m.vtsswd = m.vtssws;
    m.vtsswgs = 10.0;
    m.vtsswgd = m.vtsswgs;
    x.saref = 1e-6;
    x.sbref = 1e-6;
    x.wlod = 0.0;
    x.ku0 = 0.0;
    x.kvsat = 0.0;
    x.kvth0 = 0.0;
    x.tku0 = 0.0;
    x.llodku0 = 0.0;
    x.wlodku0 = 0.0;
    x.llodvth = 0.0;
    x.wlodvth = 0.0;
    x.lku0 = 0.0;
    x.wku0 = 0.0;
    x.pku0 = 0.0;
    x.lkvth0 = 0.0;
    x.wkvth0 = 0.0;
    x.pkvth0 = 0.0;
    x.stk2 = 0.0;
    x.lodk2 = 1.0;
    x.steta0 = 0.0;
    x.lodeta0 = 1.0;
    m.bdos = 1.0;
    m.ados = 1.0;
    m.delta = x.delta;
    const double coxe = m.epsrox * EPS0 / m.toxe;
    m.cgso = 0.6 * m.xj * coxe;  // dlc not given
    m.cgdo = 0.6 * m.xj * coxe;
    m.cgbo = 2.0 * x.dwc * coxe;
}
}  // namespace


// bsim4inst.rs:1387-1509: perimeters and areas of the layout at geomod 0 (isolated ends, shared interiors).
static void pa_eff_geo(double nf, int minSD, double Weffcj, double DMCG, double DMCI, double DMDG, double& Ps,
                       double& Pd, double& As, double& Ad) {
    double nuIntD = 0.0, nuEndD = 0.0, nuIntS = 0.0, nuEndS = 0.0;
    {
        const long NF = (long)nf;
        if ((NF % 2) != 0) {
            const double nint = 2.0 * std::fmax((nf - 1.0) / 2.0, 0.0);
            nuIntD = nint;
            nuEndD = 1.0;
            nuIntS = nint;
            nuEndS = 1.0;
        } else if (minSD == 1) {
            nuEndD = 2.0;
            nuIntD = 2.0 * std::fmax((nf / 2.0 - 1.0), 0.0);
            nuEndS = 0.0;
            nuIntS = nf;
        } else {
            nuEndD = 0.0;
            nuIntD = nf;
            nuEndS = 2.0;
            nuIntS = 2.0 * std::fmax((nf / 2.0 - 1.0), 0.0);
        }
    }
    const double tmp = DMCG + DMCI;
    const double p_iso = tmp + tmp + Weffcj;
    const double p_sha = DMCG + DMCG;
    const double a_iso = tmp * Weffcj;
    const double a_sha = DMCG * Weffcj;
    (void)DMDG;
    Ps = nuEndS * p_iso + nuIntS * p_sha;
    Pd = nuEndD * p_iso + nuIntD * p_sha;
    As = nuEndS * a_iso + nuIntS * a_sha;
    Ad = nuEndD * a_iso + nuIntD * a_sha;
}
static double dio_ijth_vjm(double Nvtm, double Ijth, double Isb, double XExpBV) {  // bsim4inst.rs:1380-1385
    const double Tc = XExpBV;
    const double Tb = 1.0 + Ijth / Isb - Tc;
    const double EVjmovNv = 0.5 * (Tb + std::sqrt(Tb * Tb + 4.0 * Tc));
    return Nvtm * std::log(EVjmovNv);
}

B4Par derive(bool pmos, const B4Inst& in) {
    B4Par P{};
    ModelExtra x;
    defaults(pmos, P.m, x);
    const B4M& model = P.m;
    B4D& md = P.d;
    B4S& sp = P.s;
    B4I& ip = P.i;

    // bsim4derive.rs::derive (mtrlmod 0)
    md.epssub = EPSSI;
    const double epssub = md.epssub;
    md.coxp = model.epsrox * EPS0 / model.toxp;
    md.coxe = model.epsrox * EPS0 / model.toxe;
    md.vcrit = VT_REF * std::log(VT_REF / (SQRT2 * 1.0e-14));
    md.factor1 = std::sqrt(epssub / (model.epsrox * EPS0) * model.toxe);
    const double Temp = 300.15;
    const double Tnom = model.tnom;
    md.TempRatio = Temp / Tnom;
    const double Vtm0 = KB_OVER_Q * Tnom;
    md.vtm0 = Vtm0;
    const double Eg0 = 1.16 - 7.02e-4 * Tnom * Tnom / (Tnom + 1108.0);
    const double ni = 1.45e10 * (Tnom / 300.15) * std::sqrt(Tnom / 300.15) * std::exp(21.5565981 - Eg0 / (2.0 * Vtm0));
    md.Eg0 = Eg0;
    md.vtm = KB_OVER_Q * Temp;
    md.ni = ni;
    {
        const double Eg = 1.16 - 7.02e-4 * Temp * Temp / (Temp + 1108.0);
        if (Temp != Tnom) {
            const double T0 = Eg0 / Vtm0 - Eg / md.vtm;
            const double T1 = std::log(Temp / Tnom);
            double T2 = T0 + x.xtis * T1;
            double T3 = std::exp(T2 / x.njs);
            md.SjctTempSatCurDensity = x.jss * T3;
            md.SjctSidewallTempSatCurDensity = x.jsws * T3;
            md.SjctGateSidewallTempSatCurDensity = x.jswgs * T3;
            T2 = T0 + x.xtid * T1;
            T3 = std::exp(T2 / x.njd);
            md.DjctTempSatCurDensity = x.jsd * T3;
            md.DjctSidewallTempSatCurDensity = x.jswd * T3;
            md.DjctGateSidewallTempSatCurDensity = x.jswgd * T3;
        } else {
            md.SjctTempSatCurDensity = x.jss;
            md.SjctSidewallTempSatCurDensity = x.jsws;
            md.SjctGateSidewallTempSatCurDensity = x.jswgs;
            md.DjctTempSatCurDensity = x.jsd;
            md.DjctSidewallTempSatCurDensity = x.jswd;
            md.DjctGateSidewallTempSatCurDensity = x.jswgd;
        }
    }
    if (md.SjctTempSatCurDensity < 0.0) md.SjctTempSatCurDensity = 0.0;
    if (md.SjctSidewallTempSatCurDensity < 0.0) md.SjctSidewallTempSatCurDensity = 0.0;
    if (md.SjctGateSidewallTempSatCurDensity < 0.0) md.SjctGateSidewallTempSatCurDensity = 0.0;
    if (md.DjctTempSatCurDensity < 0.0) md.DjctTempSatCurDensity = 0.0;
    if (md.DjctSidewallTempSatCurDensity < 0.0) md.DjctSidewallTempSatCurDensity = 0.0;
    if (md.DjctGateSidewallTempSatCurDensity < 0.0) md.DjctGateSidewallTempSatCurDensity = 0.0;
    const double delTemp = Temp - model.tnom;
    {
        double T0 = x.tcj * delTemp;
        if (T0 >= -1.0) {
            md.SunitAreaTempJctCap = model.cjs * (1.0 + T0);
            md.DunitAreaTempJctCap = model.cjd * (1.0 + T0);
        } else {
            if (model.cjs > 0.0) md.SunitAreaTempJctCap = 0.0;
            if (model.cjd > 0.0) md.DunitAreaTempJctCap = 0.0;
        }
        T0 = x.tcjsw * delTemp;
        if (T0 >= -1.0) {
            md.SunitLengthSidewallTempJctCap = model.cjsws * (1.0 + T0);
            md.DunitLengthSidewallTempJctCap = model.cjswd * (1.0 + T0);
        } else {
            if (model.cjsws > 0.0) md.SunitLengthSidewallTempJctCap = 0.0;
            if (model.cjswd > 0.0) md.DunitLengthSidewallTempJctCap = 0.0;
        }
        T0 = x.tcjswg * delTemp;
        if (T0 >= -1.0) {
            md.SunitLengthGateSidewallTempJctCap = model.cjswgs * (1.0 + T0);
            md.DunitLengthGateSidewallTempJctCap = model.cjswgd * (1.0 + T0);
        } else {
            if (model.cjswgs > 0.0) md.SunitLengthGateSidewallTempJctCap = 0.0;
            if (model.cjswgd > 0.0) md.DunitLengthGateSidewallTempJctCap = 0.0;
        }
    }
    md.PhiBS = model.pbs - x.tpb * delTemp;
    if (md.PhiBS < 0.01) md.PhiBS = 0.01;
    md.PhiBD = model.pbd - x.tpb * delTemp;
    if (md.PhiBD < 0.01) md.PhiBD = 0.01;
    md.PhiBSWS = model.pbsws - x.tpbsw * delTemp;
    if (md.PhiBSWS <= 0.01) md.PhiBSWS = 0.01;
    md.PhiBSWD = model.pbswd - x.tpbsw * delTemp;
    if (md.PhiBSWD <= 0.01) md.PhiBSWD = 0.01;
    md.PhiBSWGS = model.pbswgs - x.tpbswg * delTemp;
    if (md.PhiBSWGS <= 0.01) md.PhiBSWGS = 0.01;
    md.PhiBSWGD = model.pbswgd - x.tpbswg * delTemp;
    if (md.PhiBSWGD <= 0.01) md.PhiBSWGD = 0.01;
    {
        const double T0 = md.TempRatio - 1.0;
        const double njtsstemp = x.njts * (1.0 + x.tnjts * T0);
        const double njtsswstemp = x.njtssw * (1.0 + x.tnjtssw * T0);
        const double njtsswgstemp = x.njtsswg * (1.0 + x.tnjtsswg * T0);
        const double njtsdtemp = x.njtsd * (1.0 + x.tnjtsd * T0);
        const double njtsswdtemp = x.njtsswd * (1.0 + x.tnjtsswd * T0);
        const double njtsswgdtemp = x.njtsswgd * (1.0 + x.tnjtsswgd * T0);
        md.Nvtms = md.vtm * x.njs;
        md.Nvtmd = md.vtm * x.njd;
        md.Nvtmrssws = md.vtm0 * njtsswstemp;
        md.Nvtmrsswgs = md.vtm0 * njtsswgstemp;
        md.Nvtmrss = md.vtm0 * njtsstemp;
        md.Nvtmrsswd = md.vtm0 * njtsswdtemp;
        md.Nvtmrsswgd = md.vtm0 * njtsswgdtemp;
        md.Nvtmrsd = md.vtm0 * njtsdtemp;
    }

    // bsim4inst.rs::from
    ip.l = in.given(B4I_L) ? in.v[B4I_L] : 5.0e-6;
    ip.w = in.given(B4I_W) ? in.v[B4I_W] : 5.0e-6;
    ip.nf = in.given(B4I_NF) ? in.v[B4I_NF] : 1.0;
    const double sa = in.given(B4I_SA) ? in.v[B4I_SA] : 0.0;
    const double sb = in.given(B4I_SB) ? in.v[B4I_SB] : 0.0;
    const double sd = in.given(B4I_SD) ? in.v[B4I_SD] : 2.0 * x.dmcg;
    const double delvto = in.given(B4I_DELVTO) ? in.v[B4I_DELVTO] : 0.0;
    const int minSD = in.given(B4I_MIN) ? (int)in.v[B4I_MIN] : 0;
    const double Ldrn = ip.l;
    const double Lnew = ip.l + x.xl;
    const double Wnew = ip.w / ip.nf + x.xw;
    double T0, T1, T2, T3, T4, T5, T8, T9, tmp, tmp1, tmp2;
    T0 = std::pow(Lnew, x.lln);
    T1 = std::pow(Wnew, x.lwn);
    tmp1 = x.ll / T0 + x.lw / T1 + x.lwl / (T0 * T1);
    sp.dl = x.lint + tmp1;
    tmp2 = x.llc / T0 + x.lwc / T1 + x.lwlc / (T0 * T1);
    sp.dlc = model.dlc + tmp2;
    T2 = std::pow(Lnew, x.wln);
    T3 = std::pow(Wnew, x.wwn);
    tmp1 = x.wl / T2 + x.ww / T3 + x.wwl / (T2 * T3);
    sp.dw = x.wint + tmp1;
    tmp2 = x.wlc / T2 + x.wwc / T3 + x.wwlc / (T2 * T3);
    sp.dwc = x.dwc + tmp2;
    sp.dwj = x.dwj + tmp2;
    sp.leff = Lnew - 2.0 * sp.dl;
    if (sp.leff <= 0.0) throw std::domain_error("BSIM4: Effective channel length <= 0");
    sp.weff = Wnew - 2.0 * sp.dw;
    if (sp.weff <= 0.0) throw std::domain_error("BSIM4: Effective channel width <= 0");
    sp.leffCV = Lnew - 2.0 * sp.dlc;
    if (sp.leffCV <= 0.0) throw std::domain_error("BSIM4: Effective channel length for C-V <= 0");
    sp.weffCV = Wnew - 2.0 * sp.dwc;
    if (sp.weffCV <= 0.0) throw std::domain_error("BSIM4: Effective channel width for C-V <= 0");
    sp.weffCJ = Wnew - 2.0 * sp.dwj;
    if (sp.weffCJ <= 0.0) throw std::domain_error("BSIM4: Effective channel width for S/D junctions <= 0");
    // binning: every length, width and cross coefficient is zero, so each value is its model value
    sp.cdsc = x.cdsc;
    sp.cdscb = x.cdscb;
    sp.cdscd = x.cdscd;
    sp.cit = x.cit;
    sp.nfactor = x.nfactor;
    const double tnfactor = x.tnfactor;
    sp.xj = model.xj;
    sp.vsat = x.vsat;
    sp.at = x.at;
    sp.a0 = x.a0;
    sp.ags = x.ags;
    sp.a1 = x.a1;
    sp.a2 = x.a2;
    sp.keta = x.keta;
    sp.nsub = x.nsub;
    sp.ndep = model.ndep;
    sp.nsd = model.nsd;
    sp.phin = model.phin;
    sp.ngate = model.ngate;
    sp.gamma1 = 0.0;
    sp.gamma2 = 0.0;
    sp.vbx = 0.0;
    sp.vbm = x.vbm;
    sp.xt = x.xt;
    sp.vfb = x.vfb;
    sp.kt1 = x.kt1;
    sp.kt1l = x.kt1l;
    sp.kt2 = x.kt2;
    sp.k3 = x.k3;
    sp.k3b = x.k3b;
    sp.w0 = x.w0;
    sp.lpe0 = x.lpe0;
    sp.lpeb = x.lpeb;
    sp.dvtp0 = x.dvtp0;
    sp.dvtp1 = x.dvtp1;
    sp.dvtp2 = x.dvtp2;
    sp.dvtp3 = x.dvtp3;
    sp.dvtp4 = x.dvtp4;
    sp.dvtp5 = x.dvtp5;
    sp.dvt0 = x.dvt0;
    sp.dvt1 = x.dvt1;
    sp.dvt2 = x.dvt2;
    sp.dvt0w = x.dvt0w;
    sp.dvt1w = x.dvt1w;
    sp.dvt2w = x.dvt2w;
    sp.drout = x.drout;
    sp.dsub = x.dsub;
    sp.vth0 = x.vth0;
    sp.ua = x.ua;
    sp.ua1 = x.ua1;
    sp.ub = x.ub;
    sp.ub1 = x.ub1;
    sp.uc = x.uc;
    sp.uc1 = x.uc1;
    sp.ud = x.ud;
    sp.ud1 = x.ud1;
    sp.up = x.up;
    sp.lp = x.lp;
    sp.eu = x.eu;
    sp.u0 = x.u0;
    sp.ute = x.ute;
    sp.ucs = x.ucs;
    sp.ucste = x.ucste;
    sp.voff = x.voff;
    sp.tvoff = x.tvoff;
    sp.minv = x.minv;
    sp.minvcv = x.minvcv;
    sp.fprout = x.fprout;
    sp.pdits = x.pdits;
    sp.pditsd = x.pditsd;
    sp.delta = x.delta;
    sp.rdsw = x.rdsw;
    sp.rdw = x.rdw;
    sp.rsw = x.rsw;
    sp.prwg = x.prwg;
    sp.prwb = x.prwb;
    sp.prt = x.prt;
    sp.eta0 = x.eta0;
    sp.teta0 = x.teta0;
    sp.tvoffcv = x.tvoffcv;
    sp.etab = x.etab;
    sp.pclm = x.pclm;
    sp.pdibl1 = x.pdiblc1;
    sp.pdibl2 = x.pdiblc2;
    sp.pdiblb = x.pdiblcb;
    sp.pscbe1 = x.pscbe1;
    sp.pscbe2 = x.pscbe2;
    sp.pvag = x.pvag;
    sp.wr = x.wr;
    sp.dwg = x.dwg;
    sp.dwb = x.dwb;
    sp.b0 = x.b0;
    sp.b1 = x.b1;
    sp.alpha0 = x.alpha0;
    sp.alpha1 = x.alpha1;
    sp.beta0 = x.beta0;
    sp.agidl = x.agidl;
    sp.bgidl = x.bgidl;
    sp.cgidl = x.cgidl;
    sp.egidl = x.egidl;
    sp.rgidl = x.rgidl;
    sp.kgidl = x.kgidl;
    sp.fgidl = x.fgidl;
    sp.agisl = x.agidl;
    sp.bgisl = x.bgidl;
    sp.cgisl = x.cgidl;
    sp.egisl = x.egidl;
    sp.rgisl = x.rgidl;
    sp.kgisl = x.kgidl;
    sp.fgisl = x.fgidl;
    sp.lambda = x.lambda;
    sp.vtl = x.vtl;
    sp.xn = x.xn;
    sp.vfbsdoff = x.vfbsdoff;
    sp.tvfbsdoff = x.tvfbsdoff;
    sp.cgsl = model.cgsl;
    sp.cgdl = model.cgdl;
    sp.ckappas = model.ckappas;
    sp.ckappad = model.ckappad;
    sp.cf = model.cf;
    sp.clc = x.clc;
    sp.cle = x.cle;
    sp.vfbcv = x.vfbcv;
    sp.acde = x.acde;
    sp.moin = x.moin;
    sp.noff = x.noff;
    sp.voffcv = x.voffcv;
    sp.abulkCVfactor = 1.0 + std::pow((sp.clc / sp.leffCV), sp.cle);
    T0 = md.TempRatio - 1.0;
    const double PowWeffWr = std::pow(sp.weffCJ * 1.0e6, sp.wr) * ip.nf;
    T1 = T2 = T3 = T4 = 0.0;
    sp.ucs = sp.ucs * std::pow(md.TempRatio, sp.ucste);
    // tempmod 0
    sp.ua = sp.ua + sp.ua1 * T0;
    sp.ub = sp.ub + sp.ub1 * T0;
    sp.uc = sp.uc + sp.uc1 * T0;
    sp.ud = sp.ud + sp.ud1 * T0;
    sp.vsattemp = sp.vsat - sp.at * T0;
    {
        const double T10 = sp.prt * T0;
        sp.rds0 = (sp.rdsw + T10) * ip.nf / PowWeffWr;
        sp.rdswmin = (x.rdswmin + T10) * ip.nf / PowWeffWr;
    }
    sp.rd0 = T1 / PowWeffWr;
    sp.rdwmin = T2 / PowWeffWr;
    sp.rs0 = T3 / PowWeffWr;
    sp.rswmin = T4 / PowWeffWr;
    if (sp.u0 > 1.0) sp.u0 = sp.u0 / 1.0e4;
    T5 = 1.0 - sp.up * std::exp(-sp.leff / sp.lp);
    sp.u0temp = sp.u0 * T5 * std::pow(md.TempRatio, sp.ute);
    if (sp.eu < 0.0) sp.eu = 0.0;
    if (sp.ucs < 0.0) sp.ucs = 0.0;
    sp.vfbsdoff = sp.vfbsdoff * (1.0 + sp.tvfbsdoff * delTemp);
    sp.voff = sp.voff * (1.0 + sp.tvoff * delTemp);
    sp.nfactor = sp.nfactor + tnfactor * delTemp / model.tnom;
    sp.voffcv = sp.voffcv * (1.0 + sp.tvoffcv * delTemp);
    sp.eta0 = sp.eta0 + sp.teta0 * delTemp / model.tnom;
    // vtl is not given: no source-end velocity limit factor
    sp.cgdo = (model.cgdo + sp.cf) * sp.weffCV;
    sp.cgso = (model.cgso + sp.cf) * sp.weffCV;
    sp.cgbo = model.cgbo * sp.leffCV * ip.nf;
    sp.phi = md.vtm0 * std::log(sp.ndep / md.ni) + sp.phin + 0.4;
    sp.sqrtPhi = std::sqrt(sp.phi);
    sp.phis3 = sp.sqrtPhi * sp.phi;
    sp.Xdep0 = std::sqrt(2.0 * md.epssub / (Q * sp.ndep * 1.0e6)) * sp.sqrtPhi;
    sp.sqrtXdep0 = std::sqrt(sp.Xdep0);
    sp.litl = std::sqrt(3.0 * 3.9 / model.epsrox * sp.xj * model.toxe);
    sp.vbi = md.vtm0 * std::log(sp.nsd * sp.ndep / (md.ni * md.ni));
    sp.vfbsd = (sp.ngate > 0.0) ? md.vtm0 * std::log(sp.ngate / sp.nsd) : 0.0;
    sp.cdep0 = std::sqrt(Q * md.epssub * sp.ndep * 1.0e6 / 2.0 / sp.phi);
    sp.mstar = 0.5 + std::atan(sp.minv) / PI;
    sp.mstarcv = 0.5 + std::atan(sp.minvcv) / PI;
    sp.voffcbn = sp.voff + model.voffl / sp.leff;
    sp.voffcbncv = sp.voffcv + model.voffcvl / sp.leff;
    sp.ldeb = std::sqrt(md.epssub * md.vtm0 / (Q * sp.ndep * 1.0e6)) / 3.0;
    sp.acde *= std::pow((sp.ndep / 2.0e16), -0.25);
    // k1 and k2 not given (:579-624)
    sp.k1 = 0.53;
    sp.k2 = -0.0186;
    sp.vbx = sp.phi - 7.7348e-4 * sp.ndep * sp.xt * sp.xt;
    if (sp.vbx > 0.0) sp.vbx = -sp.vbx;
    if (sp.vbm > 0.0) sp.vbm = -sp.vbm;
    sp.gamma1 = 5.753e-12 * std::sqrt(sp.ndep) / md.coxe;
    sp.gamma2 = 5.753e-12 * std::sqrt(sp.nsub) / md.coxe;
    T0 = sp.gamma1 - sp.gamma2;
    T1 = std::sqrt(sp.phi - sp.vbx) - sp.sqrtPhi;
    T2 = std::sqrt(sp.phi * (sp.phi - sp.vbm)) - sp.phi;
    sp.k2 = T0 * T1 / (2.0 * T2 + sp.vbm);
    sp.k1 = sp.gamma2 - 2.0 * sp.k2 * std::sqrt(sp.phi - sp.vbm);
    // vfb not given, vth0 not given, mtrlmod 0 (:626-645)
    sp.vfb = -1.0;
    sp.vth0 = model.p * (sp.vfb + sp.phi + sp.k1 * sp.sqrtPhi);
    sp.k1ox = sp.k1 * model.toxe / model.toxm;
    tmp = std::sqrt(md.epssub / (model.epsrox * EPS0) * model.toxe * sp.Xdep0);
    auto theta = [](double t0) {
        if (t0 < EXP_THRESHOLD) {
            const double t1 = std::exp(t0);
            const double t2 = t1 - 1.0;
            const double t3 = t2 * t2;
            const double t4 = t3 + 2.0 * t1 * MIN_EXP;
            return t1 / t4;
        }
        return 1.0 / (MAX_EXP - 2.0);
    };
    T0 = sp.dsub * sp.leff / tmp;
    sp.theta0vb0 = theta(T0);
    T0 = sp.drout * sp.leff / tmp;
    T5 = theta(T0);
    sp.thetaRout = sp.pdibl1 * T5 + sp.pdibl2;
    tmp = std::sqrt(sp.Xdep0);
    tmp1 = sp.vbi - sp.phi;
    tmp2 = md.factor1 * tmp;
    T0 = sp.dvt1w * sp.weff * sp.leff / tmp2;
    T8 = theta(T0);
    T0 = sp.dvt0w * T8;
    T8 = T0 * tmp1;
    T0 = sp.dvt1 * sp.leff / tmp2;
    T9 = theta(T0);
    T9 = sp.dvt0 * T9 * tmp1;
    T4 = model.toxe * sp.phi / (sp.weff + sp.w0);
    T0 = std::sqrt(1.0 + sp.lpe0 / sp.leff);
    T3 = (sp.kt1 + sp.kt1l / sp.leff) * (md.TempRatio - 1.0);
    T5 = sp.k1ox * (T0 - 1.0) * sp.sqrtPhi + T3;
    sp.vfbzbfactor = -T8 - T9 + sp.k3 * T4 + T5 - sp.phi - sp.k1 * sp.sqrtPhi;
    // stress effect (:709-737)
    T0 = std::pow(Lnew, x.llodku0);
    const double W_tmp = Wnew + x.wlod;
    T1 = std::pow(W_tmp, x.wlodku0);
    tmp1 = x.lku0 / T0 + x.wku0 / T1 + x.pku0 / (T0 * T1);
    sp.ku0 = 1.0 + tmp1;
    T0 = std::pow(Lnew, x.llodvth);
    T1 = std::pow(W_tmp, x.wlodvth);
    tmp1 = x.lkvth0 / T0 + x.wkvth0 / T1 + x.pkvth0 / (T0 * T1);
    sp.kvth0 = 1.0 + tmp1;
    sp.kvth0 = std::sqrt(sp.kvth0 * sp.kvth0 + 1e-9);
    T0 = (md.TempRatio - 1.0);
    sp.ku0temp = sp.ku0 * (1.0 + x.tku0 * T0) + 1e-9;
    const double Inv_saref = 1.0 / (x.saref + 0.5 * Ldrn);
    const double Inv_sbref = 1.0 / (x.sbref + 0.5 * Ldrn);
    sp.inv_od_ref = Inv_saref + Inv_sbref;
    sp.rho_ref = x.ku0 / sp.ku0temp * sp.inv_od_ref;
    // DITS (:766-769)
    T0 = -sp.dvtp3 * std::log(sp.leff);
    T1 = dexpb(T0);
    sp.dvtp2factor = sp.dvtp5 + sp.dvtp2 * T1;

    // instance (:772-808)
    if ((sa > 0.0) && (sb > 0.0) && ((ip.nf == 1.0) || ((ip.nf > 1.0) && (sd > 0.0)))) {
        double Inv_sa = 0.0, Inv_sb = 0.0;
        double kvsat = x.kvsat;
        if (x.kvsat < -1.0) kvsat = -1.0;
        if (x.kvsat > 1.0) kvsat = 1.0;
        const long nfi = (long)ip.nf;
        for (long i = 0; i < nfi; i++) {
            const double t0 = 1.0 / ip.nf / (sa + 0.5 * Ldrn + (double)i * (sd + Ldrn));
            const double t1 = 1.0 / ip.nf / (sb + 0.5 * Ldrn + (double)i * (sd + Ldrn));
            Inv_sa += t0;
            Inv_sb += t1;
        }
        const double Inv_ODeff = Inv_sa + Inv_sb;
        const double rho = x.ku0 / sp.ku0temp * Inv_ODeff;
        T0 = (1.0 + rho) / (1.0 + sp.rho_ref);
        ip.u0temp = sp.u0temp * T0;
        T1 = (1.0 + kvsat * rho) / (1.0 + kvsat * sp.rho_ref);
        ip.vsattemp = sp.vsattemp * T1;
        const double OD_offset = Inv_ODeff - sp.inv_od_ref;
        const double dvth0_lod = x.kvth0 / sp.kvth0 * OD_offset;
        const double dk2_lod = x.stk2 / std::pow(sp.kvth0, x.lodk2) * OD_offset;
        const double deta0_lod = x.steta0 / std::pow(sp.kvth0, x.lodeta0) * OD_offset;
        ip.vth0 = sp.vth0 + dvth0_lod;
        ip.eta0 = sp.eta0 + deta0_lod;
        ip.k2 = sp.k2 + dk2_lod;
    } else {
        ip.u0temp = sp.u0temp;
        ip.vth0 = sp.vth0;
        ip.vsattemp = sp.vsattemp;
        ip.eta0 = sp.eta0;
        ip.k2 = sp.k2;
    }
    // wpemod 0: no well-proximity effect
    ip.vth0 += delvto;
    ip.vfb = sp.vfb + model.p * delvto;
    T3 = model.p * ip.vth0 - ip.vfb - sp.phi;
    T4 = T3 + T3;
    T5 = 2.5 * T3;
    ip.vtfbphi1 = (model.p > 0.0) ? T4 : T5;
    if (ip.vtfbphi1 < 0.0) ip.vtfbphi1 = 0.0;
    ip.vtfbphi2 = 4.0 * T3;
    if (ip.vtfbphi2 < 0.0) ip.vtfbphi2 = 0.0;
    if (ip.k2 < 0.0) {
        T0 = 0.5 * sp.k1 / ip.k2;
        ip.vbsc = 0.9 * (sp.phi - T0 * T0);
        if (ip.vbsc > -3.0)
            ip.vbsc = -3.0;
        else if (ip.vbsc < -30.0)
            ip.vbsc = -30.0;
    } else {
        ip.vbsc = -30.0;
    }
    if (ip.vbsc > sp.vbm) ip.vbsc = sp.vbm;
    ip.k2ox = ip.k2 * model.toxe / model.toxm;
    ip.vfbzb = sp.vfbzbfactor + model.p * ip.vth0;
    // rbodymod 0: no body resistances; gate resistance unused at rgatemod 0
    // junction geometry (:1022-1070), permod 1
    {
        const double DMCGeff = x.dmcg - x.dmcgt;
        const double DMCIeff = x.dmci;
        const double DMDGeff = x.dmdg - x.dmcgt;
        double ps_calc, pd_calc, as_calc, ad_calc;
        pa_eff_geo(ip.nf, minSD, sp.weffCJ, DMCGeff, DMCIeff, DMDGeff, ps_calc, pd_calc, as_calc, ad_calc);
        if (in.given(B4I_PS)) {
            const double val = in.v[B4I_PS];
            ip.Pseff = val < 0.0 ? 0.0 : val - sp.weffCJ * ip.nf;
        } else {
            ip.Pseff = ps_calc;
        }
        if (ip.Pseff < 0.0) ip.Pseff = 0.0;
        if (in.given(B4I_PD)) {
            const double val = in.v[B4I_PD];
            ip.Pdeff = val < 0.0 ? 0.0 : val - sp.weffCJ * ip.nf;
        } else {
            ip.Pdeff = pd_calc;
        }
        if (ip.Pdeff < 0.0) ip.Pdeff = 0.0;
        ip.Aseff = in.given(B4I_AS) ? in.v[B4I_AS] : as_calc;
        if (ip.Aseff < 0.0) ip.Aseff = 0.0;
        ip.Adeff = in.given(B4I_AD) ? in.v[B4I_AD] : ad_calc;
        if (ip.Adeff < 0.0) ip.Adeff = 0.0;
    }
    // S/D conductances (rsh 0: the reference's 1e3 S floor; used only with rdsmod != 0)
    ip.sourceConductance = 1.0e3;
    ip.drainConductance = 1.0e3;
    // junction saturation currents, diomod 1 (:1118-1205, 1343-1361)
    const double Nvtms = md.vtm * x.njs;
    double SourceSatCurrent;
    if ((ip.Aseff <= 0.0) && (ip.Pseff <= 0.0))
        SourceSatCurrent = 0.0;
    else
        SourceSatCurrent = ip.Aseff * md.SjctTempSatCurDensity + ip.Pseff * md.SjctSidewallTempSatCurDensity +
                           sp.weffCJ * ip.nf * md.SjctGateSidewallTempSatCurDensity;
    if (SourceSatCurrent > 0.0) {
        ip.vjsmFwd = dio_ijth_vjm(Nvtms, model.ijthsfwd, SourceSatCurrent, 0.0);
        ip.IVjsmFwd = SourceSatCurrent * std::exp(ip.vjsmFwd / Nvtms);
    }
    const double Nvtmd = md.vtm * x.njd;
    double DrainSatCurrent;
    if ((ip.Adeff <= 0.0) && (ip.Pdeff <= 0.0))
        DrainSatCurrent = 0.0;
    else
        DrainSatCurrent = ip.Adeff * md.DjctTempSatCurDensity + ip.Pdeff * md.DjctSidewallTempSatCurDensity +
                          sp.weffCJ * ip.nf * md.DjctGateSidewallTempSatCurDensity;
    if (DrainSatCurrent > 0.0) {
        ip.vjdmFwd = dio_ijth_vjm(Nvtmd, model.ijthdfwd, DrainSatCurrent, 0.0);
        ip.IVjdmFwd = DrainSatCurrent * std::exp(ip.vjdmFwd / Nvtmd);
    }
    // trap-assisted tunnelling saturation currents (:1207-1227)
    {
        T0 = md.TempRatio - 1.0;
        const double T7 = md.Eg0 / md.vtm * T0;
        const double Ta = dexpb(x.xtss * T7), Tb = dexpb(x.xtsd * T7), Tc = dexpb(x.xtssws * T7),
                     Td = dexpb(x.xtsswd * T7), Te = dexpb(x.xtsswgs * T7), Tf = dexpb(x.xtsswgd * T7);
        const double T11 = std::sqrt(x.jtweff / sp.weffCJ) + 1.0;
        const double T10 = sp.weffCJ * ip.nf;
        ip.SjctTempRevSatCur = Ta * ip.Aseff * x.jtss;
        ip.DjctTempRevSatCur = Tb * ip.Adeff * x.jtsd;
        ip.SswTempRevSatCur = Tc * ip.Pseff * x.jtssws;
        ip.DswTempRevSatCur = Td * ip.Pdeff * x.jtsswd;
        ip.SswgTempRevSatCur = Te * T10 * T11 * x.jtsswgs;
        ip.DswgTempRevSatCur = Tf * T10 * T11 * x.jtsswgd;
    }
    ip.toxp = model.toxp;
    ip.coxp = md.coxp;
    ip.SourceSatCurrent = SourceSatCurrent;
    ip.DrainSatCurrent = DrainSatCurrent;
    return P;
}

B4Par derive_params(const s21_params& model, const s21_params& inst) {
    const bool pmos = ((model.set >> 1) & 1ull) && model.v[1] == 1.0;  // Bsim4ModelSpecs::from, model/mod.rs:21-30
    static const int field[B4I_N] = {1, 2, 3, 4, 5, 6, 11, 12, 13, 14, 15, 16, 17, 24};
    B4Inst in{};
    for (int k = 0; k < B4I_N; k++) {
        const int f = field[k];
        if ((inst.set >> f) & 1ull) {
            in.v[k] = inst.v[f];
            in.set |= 1u << k;
        }
    }
    return derive(pmos, in);
}

}  // namespace b4
}  // namespace s21
