// BSIM4 MOSFET (the reference's comps/bsim4, itself a restatement of BSIM 4.8): parameters, device state and
// the operating-point evaluation with its matrix stamps and Backward-Euler charge companions, compiled for
// the host (parameter derivation, the CPU check) and for gfx950 (the large-instance path's device kernel).
//
// Scope: the model as the reference's API reaches it.  A Bsim4Model carries only its polarity
// (protos/bsim4.proto:45-46; Bsim4ModelSpecs::from, comps/bsim4/model/mod.rs:21-30), so every model
// parameter takes its default: every mode selector at its default (mobmod 0, capmod 2, diomod 1, rdsmod 0,
// rbodymod 0, rgatemod 0, trnqsmod 0, acnqsmod 0, igcmod 0, igbmod 0, tempmod 0, mtrlmod 0, permod 1,
// geomod 0, gidlmod 0, cvchargemod 0, wpemod 0) and every binning coefficient zero, so that each
// size-dependent parameter equals its model value exactly.  The branches of the other modes are not
// compiled.  Instance parameters (L, W, NF, AS/AD/PS/PD, SA/SB/SD, DELVTO, ...) are general.
//
// The arithmetic follows the reference statement by statement (bsim4solver.rs:145-3754, stamp.rs, tran.rs,
// bsim4derive.rs, bsim4inst.rs), so that its results are reproduced to the last bits.
#pragma once

#if defined(__HIPCC_RTC__)
// run-time compilation provides the HIP declarations and the math functions itself
#elif defined(__HIPCC__)
#include <hip/hip_runtime.h>
#else  // a plain host build (the CPU check compiles these functions with g++)
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#endif

#ifndef __HIPCC_RTC__
#include <cmath>
#include <stdexcept>
#endif

#include "spice21_mi355x.h"

namespace s21 {
namespace b4 {

constexpr double EXP_THRESHOLD = 34.0;  // comps/bsim4/mod.rs:34-38
constexpr double MAX_EXP = 5.834617425e14;
constexpr double MIN_EXP = 1.713908431e-15;
constexpr double DELTA_1 = 0.02, DELTA_3 = 0.02, DELTA_4 = 0.02;
constexpr double KB = 1.3806226e-23, Q = 1.6021918e-19;  // comps/mod.rs:24-37
constexpr double KB_OVER_Q = KB / Q;
constexpr double TEMP_REF = 273.15 + 27.0;
constexpr double VT_REF = KB * TEMP_REF / Q;
constexpr double SQRT2 = 1.4142135624;
constexpr double EPS0 = 8.85418e-12;
constexpr double EPSSI = 1.03594e-10;
constexpr double PI = 3.14159265358979323846;

__host__ __device__ inline double dexpb(double A) {  // mod.rs:45-53
    if (A > EXP_THRESHOLD) return MAX_EXP * (1.0 + (A)-EXP_THRESHOLD);
    if (A < -EXP_THRESHOLD) return MIN_EXP;
    return exp(A);
}
__host__ __device__ inline double dexpc(double A) {  // mod.rs:54-62
    if (A > EXP_THRESHOLD) return MAX_EXP;
    if (A < -EXP_THRESHOLD) return 0.0;
    return exp(A);
}

// ---- parameters ------------------------------------------------------------------------------------
// Model values (model/vals.rs::resolve at its defaults, polarity p = +1 NMOS / -1 PMOS).
struct B4M {
    double p;
    double toxe, toxp, toxm, epsrox, xj, ndep, nsd, ngate, phin;
    double vtss, vtsd, vtssws, vtsswd, vtsswgs, vtsswgd;
    double xjbvs, xjbvd, bvs, bvd, ijthsfwd, ijthdfwd;
    double cf, cgso, cgdo, cgbo, ckappas, ckappad, cgsl, cgdl, dlc;
    double tnom, xpart, voffl, voffcvl;
    double cjs, cjd, cjsws, cjswd, cjswgs, cjswgd, mjs, mjd, mjsws, mjswd, mjswgs, mjswgd;
    double pbs, pbd, pbsws, pbswd, pbswgs, pbswgd;
    double delta, bdos, ados, pditsl;
};
// Derived model values (bsim4derive.rs).
struct B4D {
    double epssub, coxp, coxe, vcrit, factor1, TempRatio, vtm0, vtm, Eg0, ni;
    double SjctTempSatCurDensity, SjctSidewallTempSatCurDensity, SjctGateSidewallTempSatCurDensity;
    double DjctTempSatCurDensity, DjctSidewallTempSatCurDensity, DjctGateSidewallTempSatCurDensity;
    double SunitAreaTempJctCap, DunitAreaTempJctCap, SunitLengthSidewallTempJctCap, DunitLengthSidewallTempJctCap;
    double SunitLengthGateSidewallTempJctCap, DunitLengthGateSidewallTempJctCap;
    double PhiBS, PhiBD, PhiBSWS, PhiBSWD, PhiBSWGS, PhiBSWGD;
    double Nvtms, Nvtmd, Nvtmrss, Nvtmrssws, Nvtmrsswgs, Nvtmrsd, Nvtmrsswd, Nvtmrsswgd;
};
// Size-dependent values (bsim4inst.rs::from, the first part).
struct B4S {
    double leff, weff, leffCV, weffCV, weffCJ, dl, dw, dlc, dwc, dwj;
    double cdsc, cdscb, cdscd, cit, nfactor, xj, vsat, at, a0, ags, a1, a2, keta, nsub, ndep, nsd, phin, ngate;
    double gamma1, gamma2, vbx, vbm, xt, vfb, kt1, kt1l, kt2, k3, k3b, w0, lpe0, lpeb;
    double dvtp0, dvtp1, dvtp2, dvtp3, dvtp4, dvtp5, dvt0, dvt1, dvt2, dvt0w, dvt1w, dvt2w, drout, dsub, vth0;
    double ua, ua1, ub, ub1, uc, uc1, ud, ud1, up, lp, eu, u0, ute, ucs, ucste, voff, tvoff, minv, minvcv, fprout;
    double pdits, pditsd, delta, rdsw, rdw, rsw, prwg, prwb, prt, eta0, teta0, tvoffcv, etab, pclm, pdibl1, pdibl2;
    double pdiblb, pscbe1, pscbe2, pvag, wr, dwg, dwb, b0, b1, alpha0, alpha1, beta0;
    double agidl, bgidl, cgidl, egidl, rgidl, kgidl, fgidl, agisl, bgisl, cgisl, egisl, rgisl, kgisl, fgisl;
    double lambda, vtl, xn, lc, tfactor, vfbsdoff, tvfbsdoff, cgsl, cgdl, ckappas, ckappad, cf, clc, cle, vfbcv;
    double acde, moin, noff, voffcv, abulkCVfactor, ucs_t, vsattemp, rds0, rdswmin, rd0, rdwmin, rs0, rswmin;
    double u0temp, cgdo, cgso, cgbo, phi, sqrtPhi, phis3, Xdep0, sqrtXdep0, litl, vbi, vfbsd, cdep0;
    double mstar, mstarcv, voffcbn, voffcbncv, ldeb, k1, k2, k1ox, theta0vb0, thetaRout, vfbzbfactor;
    double ku0, kvth0, ku0temp, inv_od_ref, rho_ref, dvtp2factor;
};
// Instance values (bsim4inst.rs::from, the second part).
struct B4I {
    double l, w, nf, u0temp, vsattemp, vth0, eta0, k2, vfb, vtfbphi1, vtfbphi2, vbsc, k2ox, vfbzb;
    double Pseff, Pdeff, Aseff, Adeff, sourceConductance, drainConductance;
    double SourceSatCurrent, DrainSatCurrent, vjsmFwd, IVjsmFwd, vjdmFwd, IVjdmFwd;
    double SjctTempRevSatCur, DjctTempRevSatCur, SswTempRevSatCur, DswTempRevSatCur, SswgTempRevSatCur,
        DswgTempRevSatCur;
    double toxp, coxp, grgeltd;
};
struct B4Par {
    B4M m;
    B4D d;
    B4S s;
    B4I i;
    // field K among the structure's doubles (the parameter-view interface of the evaluation functions)
    template <int K>
    __host__ __device__ double get() const {
        return reinterpret_cast<const double*>(this)[K];
    }
};

// Instance parameters a Bsim4InstParams message can set (protos/bsim4.proto:6-41) that reach the model at
// the default mode selectors: slot -> value, with a bit of `set` per given slot.
enum { B4I_L = 0, B4I_W, B4I_NF, B4I_SA, B4I_SB, B4I_SD, B4I_AD, B4I_AS, B4I_PD, B4I_PS, B4I_NRD, B4I_NRS, B4I_MIN,
       B4I_DELVTO, B4I_N };
struct B4Inst {
    double v[B4I_N];
    unsigned set;
    bool given(int k) const { return (set >> k) & 1u; }
};
// Host derivation of one device's parameters (bsim4_derive.cpp).  Throws std::domain_error with the
// reference's message where the reference panics (an effective length or width <= 0,
// bsim4inst.rs:200-223).
B4Par derive(bool pmos, const B4Inst& inst);
// The same from the netlist's parameter sets: Bsim4Model (slot 1 = mos_type, 1 PMOS) and Bsim4InstParams
// keyed by field number (protos/bsim4.proto:6-46).  Instance fields that reach nothing at the default mode
// selectors (sca/scb/scc/sc: wpemod; nrd/nrs, rgeomod: rdsmod; rb*: rbodymod; xgw/ngcon: rgatemod; the
// instance mode fields, which the reference takes from the model, bsim4inst.rs:134-146) are ignored.
B4Par derive_params(const s21_params& model, const s21_params& inst);

// ---- operating point -----------------------------------------------------------------------------------
// The fields of Bsim4OpPoint (mod.rs:202-...) that the next load, the stamps and the charge integration
// read.  `guess` is the last load's point (limiting starts from it); `op` the committed one (charges).
struct B4Op {
    int mode;
    double vds, vgs, vbs, vbd, vges, vgms, vdbs, vdbd, vsbs, vses, vdes, vgb, vgmb, vbs_jct, vbd_jct, von;
    double gbs, cbs, gbd, cbd, thetavth, nstar, vgs_eff, vgd_eff, dvgs_eff_dvg, dvgd_eff_dvg;
    double Vgsteff, Abulk, vdsat, Vdseff, cd, gds, gm, gmbs, csub, gbbs, gbgs, gbds, IdovVds;
    double vgd, Igidl, ggidld, ggidlg, ggidlb, ggidls, Igisl, ggisls, ggislg, ggislb, ggisld;
    double qinv, Coxeff, qgate, qbulk, qdrn, qsrc, qdef;
    double cggb, cgsb, cgdb, cdgb, cdsb, cddb, cbgb, cbsb, cbdb, csgb, cssb, csdb;
    double cgdo, qgdo, cgso, qgso, capbs, capbd, qbs, qbd;
    double qg, qd, qs, qb, cqg, cqd, cqb, cqs;
};

}  // namespace b4
}  // namespace s21
