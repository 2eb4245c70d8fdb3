// BSIM4 operating point (bsim4solver.rs:145-3754) at the default mode selectors (see bsim4.h), for one device
// whose internal nodes all coincide with its terminals (rdsmod, rgatemod and rbodymod 0, no NQS charge node).
// `G` is the point of the device's previous load (the limiting starts from it, Bsim4::load_dc_tr keeps it
// as `guess`); the result is written to `N`.
#pragma once

#include "bsim4.h"

// A parameter read in the functions below: field f of part m (model values), d (derived), s (size-dependent)
// or i (instance) of the device's B4Par, through the parameter view P -- a B4Par itself, or the batched
// kernel's view that gathers each field from where it is stored -- as P.get<K>(), K the field's index among
// the structure's doubles.
#define B4_FIELD(part, type, f)                                                                          \
    P.template get<(int)((__builtin_offsetof(::s21::b4::B4Par, part) + __builtin_offsetof(::s21::b4::type, f)) / \
                         sizeof(double))>()
#define B4_M(f) B4_FIELD(m, B4M, f)
#define B4_D(f) B4_FIELD(d, B4D, f)
#define B4_S(f) B4_FIELD(s, B4S, f)
#define B4_I(f) B4_FIELD(i, B4I, f)

namespace s21 {
namespace b4 {

__host__ __device__ inline void poly_depletion(double phi, double ngate, double epsgate, double coxe, double Vgs,
                                               double& v, double& dv) {  // bsim4solver.rs:3771-3788
    if ((ngate > 1.0e18) && (ngate < 1.0e25) && (Vgs > phi) && (epsgate != 0.0)) {
        const double T1 = 1.0e6 * Q * epsgate * ngate / (coxe * coxe);
        const double T8 = Vgs - phi;
        const double T4 = sqrt(1.0 + 2.0 * T8 / T1);
        const double T2 = 2.0 * T8 / (T4 + 1.0);
        const double T3 = 0.5 * T2 * T2 / T1;
        const double T7 = 1.12 - T3 - 0.05;
        const double T6 = sqrt(T7 * T7 + 0.224);
        const double T5 = 1.12 - 0.5 * (T7 + T6);
        v = Vgs - T5;
        dv = 1.0 - (0.5 - 0.5 / T4) * (1.0 + T7 / T6);
    } else {
        v = Vgs;
        dv = 1.0;
    }
}
__host__ __device__ inline double limvds(double vnew, double vold) {  // :3790-3806
    if (vold >= 3.5) {
        if (vnew > vold) return fmin(vnew, (3.0 * vold) + 2.0);
        if (vnew < 3.5) return fmax(vnew, 2.0);
    } else {
        if (vnew > vold) return fmin(vnew, 4.0);
        return fmax(vnew, -0.5);
    }
    return vnew;
}
__host__ __device__ inline double fetlim(double vnew, double vold, double vto) {  // :3808-3859
    const double vtsthi = fabs(2.0 * (vold - vto)) + 2.0;
    const double vtstlo = vtsthi / 2.0 + 2.0;
    const double vtox = vto + 3.5;
    const double delv = vnew - vold;
    if (vold >= vto) {
        if (vold >= vtox) {
            if (delv <= 0.0) {
                if (vnew >= vtox) {
                    if (-delv > vtstlo) return vold - vtstlo;
                } else {
                    return fmax(vnew, vto + 2.0);
                }
            } else {
                if (delv >= vtsthi) return vold + vtsthi;
            }
        } else {
            if (delv <= 0.0) return fmax(vnew, vto - 0.5);
            return fmin(vnew, vto + 4.0);
        }
    } else {
        if (delv <= 0.0) {
            if (-delv > vtsthi) return vold - vtsthi;
        } else {
            const double vtemp = vto + 0.5;
            if (vnew <= vtemp) {
                if (delv > vtstlo) return vold + vtstlo;
            } else {
                return vtemp;
            }
        }
    }
    return vnew;
}
__host__ __device__ inline double pnjlim(double vnew, double vold, double vt, double vcrit) {  // :3861-3873
    if (vnew > vcrit && fabs(vnew - vold) > (vt + vt)) {
        if (vold > 0.0) {
            const double arg = 1.0 + (vnew - vold) / vt;
            if (arg > 0.0) return vold + vt * log(arg);
            return vcrit;
        }
        return vt * log(vnew / vt);
    }
    return vnew;
}

// Trap-assisted tunnelling factor of one junction component (bsim4solver.rs:682-766): T and dT/dVb.
__host__ __device__ inline void tat(double vts, double Nvtmr, double v, double& T, double& dT) {
    double T9, T0, T10;
    if ((vts - v) < (vts * 1e-3)) {
        T9 = 1.0e3;
        T0 = -v / Nvtmr * T9;
        T = dexpb(T0);
        T10 = dexpc(T0);
        dT = T10 / Nvtmr * T9;
    } else {
        T9 = 1.0 / (vts - v);
        T0 = -v / Nvtmr * vts * T9;
        const double dT0_dVb = vts / Nvtmr * (T9 + v * T9 * T9);
        T = dexpb(T0);
        T10 = dexpc(T0);
        dT = T10 * dT0_dVb;
    }
}

// Junction diode DC model, diomod 1 (:528-552, 592-612).
__host__ __device__ inline void junction(double Isat, double Nvtm, double vjmFwd, double IVjmFwd, double v, double gmin,
                                         double& g, double& c) {
    if (Isat <= 0.0) {
        g = gmin;
        c = g * v;
        return;
    }
    const double T2 = v / Nvtm;
    if (T2 < -EXP_THRESHOLD) {
        g = gmin;
        c = Isat * (MIN_EXP - 1.0) + gmin * v;
    } else if (v <= vjmFwd) {
        const double ev = exp(T2);
        g = Isat * ev / Nvtm + gmin;
        c = Isat * (ev - 1.0) + gmin * v;
    } else {
        const double T0 = IVjmFwd / Nvtm;
        g = T0 + gmin;
        c = IVjmFwd - Isat + T0 * (v - vjmFwd) + gmin * v;
    }
}

// The DC part of the operating point and the intrinsic charges (bsim4solver.rs:145-3754, default modes).
template <class PV>
__host__ __device__ inline void op_eval(const PV& P, const B4Op& G, double vd_, double vg_, double vs_, double vb_,
                                        double gmin, B4Op& N) {
    const double p = B4_M(p);
    double T0, T1, T2, T3, T4, T5, T6, T7, T8, T9, T10, T11, T12, T13, T14, tmp, tmp1, tmp2, tmp3, tmp4;
    double dT0_dVg, dT0_dVd, dT0_dVb, dT1_dVg, dT1_dVd, dT1_dVb, dT2_dVg, dT2_dVd, dT2_dVb;
    double dT3_dVg, dT3_dVd, dT3_dVb, dT4_dVd, dT5_dVb, dT1b, dT2b, dT3b, dT4b, dT5b, dT6b;
    dT0_dVg = dT0_dVd = dT0_dVb = dT1_dVg = dT1_dVd = dT1_dVb = dT2_dVg = dT2_dVd = dT2_dVb = 0.0;
    dT3_dVg = dT3_dVd = dT3_dVb = dT4_dVd = dT5_dVb = 0.0;
    N = B4Op{};

    // terminal voltages: every internal node is its terminal (bsim4ports.rs:37-118), no charge node
    double vds = p * (vd_ - vs_);
    double vgs = p * (vg_ - vs_);
    double vbs = p * (vb_ - vs_);
    double vges = p * (vg_ - vs_);
    double vgms = p * (vg_ - vs_);
    double vdbs = p * (vb_ - vs_);
    double vsbs = p * (vb_ - vs_);
    double vses = p * (vs_ - vs_);
    double vdes = p * (vd_ - vs_);
    const double vgdo = G.vgs - G.vds;
    double vbd = vbs - vds;
    double vdbd = vdbs - vds;
    double vgd = vgs - vds;
    const double von = G.von;
    // limiting (:459-526)
    if (G.vds >= 0.0) {
        vgs = fetlim(vgs, G.vgs, von);
        vds = vgs - vgd;
        vds = limvds(vds, G.vds);
        vgd = vgs - vds;
    } else {
        vgd = fetlim(vgd, vgdo, von);
        vds = vgs - vgd;
        vds = -limvds(-vds, -G.vds);
        vgs = vgd + vds;
    }
    if (vds >= 0.0) {
        vbs = pnjlim(vbs, G.vbs, VT_REF, B4_D(vcrit));
        vbd = vbs - vds;
    } else {
        vbd = pnjlim(vbd, G.vbd, VT_REF, B4_D(vcrit));
        vbs = vbd + vds;
    }
    vbd = vbs - vds;
    vgd = vgs - vds;
    const double vgb = vgs - vbs;
    const double vgmb = vgms - vbs;
    vdbd = vdbs - vds;
    const double vbs_jct = vsbs;
    const double vbd_jct = vdbd;

    // source/drain junction diodes (diomod 1) and trap-assisted tunnelling (:528-771)
    junction(B4_I(SourceSatCurrent), B4_D(Nvtms), B4_I(vjsmFwd), B4_I(IVjsmFwd), vbs_jct, gmin, N.gbs, N.cbs);
    junction(B4_I(DrainSatCurrent), B4_D(Nvtmd), B4_I(vjdmFwd), B4_I(IVjdmFwd), vbd_jct, gmin, N.gbd, N.cbd);
    double Ts1, Ts2, Ts3, Ts4, Ts5, Ts6;
    tat(B4_M(vtss), B4_D(Nvtmrss), vbs_jct, Ts1, dT1b);
    tat(B4_M(vtsd), B4_D(Nvtmrsd), vbd_jct, Ts2, dT2b);
    tat(B4_M(vtssws), B4_D(Nvtmrssws), vbs_jct, Ts3, dT3b);
    tat(B4_M(vtsswd), B4_D(Nvtmrsswd), vbd_jct, Ts4, dT4b);
    tat(B4_M(vtsswgs), B4_D(Nvtmrsswgs), vbs_jct, Ts5, dT5b);
    tat(B4_M(vtsswgd), B4_D(Nvtmrsswgd), vbd_jct, Ts6, dT6b);
    N.gbs += B4_I(SjctTempRevSatCur) * dT1b + B4_I(SswTempRevSatCur) * dT3b + B4_I(SswgTempRevSatCur) * dT5b;
    N.cbs -= B4_I(SjctTempRevSatCur) * (Ts1 - 1.0) + B4_I(SswTempRevSatCur) * (Ts3 - 1.0) + B4_I(SswgTempRevSatCur) * (Ts5 - 1.0);
    N.gbd += B4_I(DjctTempRevSatCur) * dT2b + B4_I(DswTempRevSatCur) * dT4b + B4_I(DswgTempRevSatCur) * dT6b;
    N.cbd -= B4_I(DjctTempRevSatCur) * (Ts2 - 1.0) + B4_I(DswTempRevSatCur) * (Ts4 - 1.0) + B4_I(DswgTempRevSatCur) * (Ts6 - 1.0);

    double Vds, Vgs, Vbs, Vdb;
    if (vds >= 0.0) {
        N.mode = 1;
        Vds = vds;
        Vgs = vgs;
        Vbs = vbs;
        Vdb = vds - vbs;
    } else {
        N.mode = -1;
        Vds = -vds;
        Vgs = vgd;
        Vbs = vbd;
        Vdb = -vbs;
    }
    (void)Vgs;
    (void)Vdb;
    const double toxe = B4_M(toxe);
    const double epssub = B4_D(epssub);

    // Vbseff (:790-810)
    double Vbseff, dVbseff_dVb;
    T0 = Vbs - B4_I(vbsc) - 0.001;
    T1 = sqrt(T0 * T0 - 0.004 * B4_I(vbsc));
    if (T0 >= 0.0) {
        Vbseff = B4_I(vbsc) + 0.5 * (T0 + T1);
        dVbseff_dVb = 0.5 * (1.0 + T0 / T1);
    } else {
        T2 = -0.002 / (T1 - T0);
        Vbseff = B4_I(vbsc) * (1.0 + T2);
        dVbseff_dVb = T2 * B4_I(vbsc) / T1;
    }
    T9 = 0.95 * B4_S(phi);
    T0 = T9 - Vbseff - 0.001;
    T1 = sqrt(T0 * T0 + 0.004 * T9);
    Vbseff = T9 - 0.5 * (T0 + T1);
    dVbseff_dVb *= 0.5 * (1.0 + T0 / T1);
    const double Phis = B4_S(phi) - Vbseff;
    const double dPhis_dVb = -1.0;
    const double sqrtPhis = sqrt(Phis);
    const double dsqrtPhis_dVb = -0.5 / sqrtPhis;
    const double Xdep = B4_S(Xdep0) * sqrtPhis / B4_S(sqrtPhi);
    const double dXdep_dVb = (B4_S(Xdep0) / B4_S(sqrtPhi)) * dsqrtPhis_dVb;
    const double Leff = B4_S(leff);
    const double Vtm = B4_D(vtm);

    // Vth (:826-905)
    T3 = sqrt(Xdep);
    const double V0 = B4_S(vbi) - B4_S(phi);
    T0 = B4_S(dvt2) * Vbseff;
    if (T0 >= -0.5) {
        T1 = 1.0 + T0;
        T2 = B4_S(dvt2);
    } else {
        T4 = 1.0 / (3.0 + 8.0 * T0);
        T1 = (1.0 + 3.0 * T0) * T4;
        T2 = B4_S(dvt2) * T4 * T4;
    }
    const double lt1 = B4_D(factor1) * T3 * T1;
    const double dlt1_dVb = B4_D(factor1) * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
    T0 = B4_S(dvt2w) * Vbseff;
    if (T0 >= -0.5) {
        T1 = 1.0 + T0;
        T2 = B4_S(dvt2w);
    } else {
        T4 = 1.0 / (3.0 + 8.0 * T0);
        T1 = (1.0 + 3.0 * T0) * T4;
        T2 = B4_S(dvt2w) * T4 * T4;
    }
    const double ltw = B4_D(factor1) * T3 * T1;
    const double dltw_dVb = B4_D(factor1) * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
    double Theta0, dTheta0_dVb;
    T0 = B4_S(dvt1) * Leff / lt1;
    if (T0 < EXP_THRESHOLD) {
        T1 = exp(T0);
        T2 = T1 - 1.0;
        T3 = T2 * T2;
        T4 = T3 + 2.0 * T1 * MIN_EXP;
        Theta0 = T1 / T4;
        dT1_dVb = -T0 * T1 * dlt1_dVb / lt1;
        dTheta0_dVb = dT1_dVb * (T4 - 2.0 * T1 * (T2 + MIN_EXP)) / T4 / T4;
    } else {
        Theta0 = 1.0 / (MAX_EXP - 2.0);
        dTheta0_dVb = 0.0;
    }
    N.thetavth = B4_S(dvt0) * Theta0;
    const double Delt_vth = N.thetavth * V0;
    const double dDelt_vth_dVb = B4_S(dvt0) * dTheta0_dVb * V0;
    T0 = B4_S(dvt1w) * B4_S(weff) * Leff / ltw;
    if (T0 < EXP_THRESHOLD) {
        T1 = exp(T0);
        T2 = T1 - 1.0;
        T3 = T2 * T2;
        T4 = T3 + 2.0 * T1 * MIN_EXP;
        T5 = T1 / T4;
        dT1_dVb = -T0 * T1 * dltw_dVb / ltw;
        dT5_dVb = dT1_dVb * (T4 - 2.0 * T1 * (T2 + MIN_EXP)) / T4 / T4;
    } else {
        T5 = 1.0 / (MAX_EXP - 2.0);
        dT5_dVb = 0.0;
    }
    T0 = B4_S(dvt0w) * T5;
    T2 = T0 * V0;
    dT2_dVb = B4_S(dvt0w) * dT5_dVb * V0;
    T0 = sqrt(1.0 + B4_S(lpe0) / Leff);
    T1 = B4_S(k1ox) * (T0 - 1.0) * B4_S(sqrtPhi) + (B4_S(kt1) + B4_S(kt1l) / Leff + B4_S(kt2) * Vbseff) * (B4_D(TempRatio) - 1.0);
    const double Vth_NarrowW = toxe * B4_S(phi) / (B4_S(weff) + B4_S(w0));
    T3 = B4_I(eta0) + B4_S(etab) * Vbseff;
    if (T3 < 1.0e-4) {
        T9 = 1.0 / (3.0 - 2.0e4 * T3);
        T3 = (2.0e-4 - T3) * T9;
        T4 = T9 * T9;
    } else {
        T4 = 1.0;
    }
    const double dDIBL_Sft_dVd = T3 * B4_S(theta0vb0);
    const double DIBL_Sft = dDIBL_Sft_dVd * Vds;
    const double Lpe_Vb = sqrt(1.0 + B4_S(lpeb) / Leff);
    double Vth = p * B4_I(vth0) + (B4_S(k1ox) * sqrtPhis - B4_S(k1) * B4_S(sqrtPhi)) * Lpe_Vb - B4_I(k2ox) * Vbseff - Delt_vth - T2 +
                 (B4_S(k3) + B4_S(k3b) * Vbseff) * Vth_NarrowW + T1 - DIBL_Sft;
    double dVth_dVb = Lpe_Vb * B4_S(k1ox) * dsqrtPhis_dVb - B4_I(k2ox) - dDelt_vth_dVb - dT2_dVb + B4_S(k3b) * Vth_NarrowW -
                      B4_S(etab) * Vds * B4_S(theta0vb0) * T4 + B4_S(kt2) * (B4_D(TempRatio) - 1.0);
    double dVth_dVd = -dDIBL_Sft_dVd;

    // n (:907-926)
    double n, dn_dVb, dn_dVd;
    tmp1 = epssub / Xdep;
    N.nstar = B4_D(vtm) / Q * (B4_D(coxe) + tmp1 + B4_S(cit));
    tmp2 = B4_S(nfactor) * tmp1;
    tmp3 = B4_S(cdsc) + B4_S(cdscb) * Vbseff + B4_S(cdscd) * Vds;
    tmp4 = (tmp2 + tmp3 * Theta0 + B4_S(cit)) / B4_D(coxe);
    if (tmp4 >= -0.5) {
        n = 1.0 + tmp4;
        dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + B4_S(cdscb) * Theta0) / B4_D(coxe);
        dn_dVd = B4_S(cdscd) * Theta0 / B4_D(coxe);
    } else {
        T0 = 1.0 / (3.0 + 8.0 * tmp4);
        n = (1.0 + 3.0 * tmp4) * T0;
        T0 *= T0;
        dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + B4_S(cdscb) * Theta0) / B4_D(coxe) * T0;
        dn_dVd = B4_S(cdscd) * Theta0 / B4_D(coxe) * T0;
    }
    // pocket implant (:928-963)
    if (B4_S(dvtp0) > 0.0) {
        T0 = -B4_S(dvtp1) * Vds;
        double dT2_dVd_;
        if (T0 < -EXP_THRESHOLD) {
            T2 = MIN_EXP;
            dT2_dVd_ = 0.0;
        } else {
            T2 = exp(T0);
            dT2_dVd_ = -B4_S(dvtp1) * T2;
        }
        T3 = Leff + B4_S(dvtp0) * (1.0 + T2);
        const double dT3_dVd_ = B4_S(dvtp0) * dT2_dVd_;
        T4 = Vtm * log(Leff / T3);
        dT4_dVd = -Vtm * dT3_dVd_ / T3;
        const double dDITS_Sft_dVd = dn_dVd * T4 + n * dT4_dVd;
        const double dDITS_Sft_dVb = T4 * dn_dVb;
        Vth -= n * T4;
        dVth_dVd -= dDITS_Sft_dVd;
        dVth_dVb -= dDITS_Sft_dVb;
    }
    // DITS_SFT2 (:965-974)
    if (!((B4_S(dvtp4) == 0.0) || (B4_S(dvtp2factor) == 0.0))) {
        T1 = 2.0 * B4_S(dvtp4) * Vds;
        T0 = dexpb(T1);
        T10 = dexpc(T1);
        const double DITS_Sft2 = B4_S(dvtp2factor) * (T0 - 1.0) / (T0 + 1.0);
        const double dDITS_Sft2_dVd = B4_S(dvtp2factor) * B4_S(dvtp4) * 4.0 * T10 / ((T0 + 1.0) * (T0 + 1.0));
        Vth -= DITS_Sft2;
        dVth_dVd -= dDITS_Sft2_dVd;
    }
    N.von = Vth;

    // poly depletion (:978-1004)
    T0 = B4_I(vfb) + B4_S(phi);
    T1 = EPSSI;
    double vgs_eff, dvgs_eff_dvg, vgd_eff, dvgd_eff_dvg;
    poly_depletion(T0, B4_S(ngate), T1, B4_D(coxe), vgs, vgs_eff, dvgs_eff_dvg);
    poly_depletion(T0, B4_S(ngate), T1, B4_D(coxe), vgd, vgd_eff, dvgd_eff_dvg);
    double Vgs_eff, dVgs_eff_dVg;
    if (N.mode > 0) {
        Vgs_eff = vgs_eff;
        dVgs_eff_dVg = dvgs_eff_dvg;
    } else {
        Vgs_eff = vgd_eff;
        dVgs_eff_dVg = dvgd_eff_dvg;
    }
    N.vgs_eff = vgs_eff;
    N.vgd_eff = vgd_eff;
    N.dvgs_eff_dvg = dvgs_eff_dvg;
    N.dvgd_eff_dvg = dvgd_eff_dvg;
    const double Vgst = Vgs_eff - Vth;

    // Vgsteff (:1008-1062)
    double dT10_dVg, dT10_dVd, dT10_dVb, dT9_dVg, dT9_dVd, dT9_dVb;
    T0 = n * Vtm;
    T1 = B4_S(mstar) * Vgst;
    T2 = T1 / T0;
    if (T2 > EXP_THRESHOLD) {
        T10 = T1;
        dT10_dVg = B4_S(mstar) * dVgs_eff_dVg;
        dT10_dVd = -dVth_dVd * B4_S(mstar);
        dT10_dVb = -dVth_dVb * B4_S(mstar);
    } else if (T2 < -EXP_THRESHOLD) {
        T10 = Vtm * log(1.0 + MIN_EXP);
        dT10_dVg = 0.0;
        dT10_dVd = T10 * dn_dVd;
        dT10_dVb = T10 * dn_dVb;
        T10 *= n;
    } else {
        const double ExpVgst = exp(T2);
        T3 = Vtm * log(1.0 + ExpVgst);
        T10 = n * T3;
        dT10_dVg = B4_S(mstar) * ExpVgst / (1.0 + ExpVgst);
        dT10_dVb = T3 * dn_dVb - dT10_dVg * (dVth_dVb + Vgst * dn_dVb / n);
        dT10_dVd = T3 * dn_dVd - dT10_dVg * (dVth_dVd + Vgst * dn_dVd / n);
        dT10_dVg *= dVgs_eff_dVg;
    }
    T1 = B4_S(voffcbn) - (1.0 - B4_S(mstar)) * Vgst;
    T2 = T1 / T0;
    if (T2 < -EXP_THRESHOLD) {
        T3 = B4_D(coxe) * MIN_EXP / B4_S(cdep0);
        T9 = B4_S(mstar) + T3 * n;
        dT9_dVg = 0.0;
        dT9_dVd = dn_dVd * T3;
        dT9_dVb = dn_dVb * T3;
    } else if (T2 > EXP_THRESHOLD) {
        T3 = B4_D(coxe) * MAX_EXP / B4_S(cdep0);
        T9 = B4_S(mstar) + T3 * n;
        dT9_dVg = 0.0;
        dT9_dVd = dn_dVd * T3;
        dT9_dVb = dn_dVb * T3;
    } else {
        const double ExpVgst = exp(T2);
        T3 = B4_D(coxe) / B4_S(cdep0);
        T4 = T3 * ExpVgst;
        T5 = T1 * T4 / T0;
        T9 = B4_S(mstar) + n * T4;
        dT9_dVg = T3 * (B4_S(mstar) - 1.0) * ExpVgst / Vtm;
        dT9_dVb = T4 * dn_dVb - dT9_dVg * dVth_dVb - T5 * dn_dVb;
        dT9_dVd = T4 * dn_dVd - dT9_dVg * dVth_dVd - T5 * dn_dVd;
        dT9_dVg *= dVgs_eff_dVg;
    }
    N.Vgsteff = T10 / T9;
    const double Vgsteff = N.Vgsteff;
    T11 = T9 * T9;
    const double dVgsteff_dVg = (T9 * dT10_dVg - T10 * dT9_dVg) / T11;
    const double dVgsteff_dVd = (T9 * dT10_dVd - T10 * dT9_dVd) / T11;
    const double dVgsteff_dVb = (T9 * dT10_dVb - T10 * dT9_dVb) / T11;

    // effective width (:1064-1078); Rds(V) is computed only for rdsmod > 1 in the reference: zero here
    T9 = sqrtPhis - B4_S(sqrtPhi);
    double Weff = B4_S(weff) - 2.0 * (B4_S(dwg) * Vgsteff + B4_S(dwb) * T9);
    double dWeff_dVg = -2.0 * B4_S(dwg);
    double dWeff_dVb = -2.0 * B4_S(dwb) * dsqrtPhis_dVb;
    if (Weff < 2.0e-8) {
        T0 = 1.0 / (6.0e-8 - 2.0 * Weff);
        Weff = 2.0e-8 * (4.0e-8 - Weff) * T0;
        T0 *= T0 * 4.0e-16;
        dWeff_dVg *= T0;
        dWeff_dVb *= T0;
    }
    const double Rds = 0.0, dRds_dVg = 0.0, dRds_dVb = 0.0;

    // Abulk (:1105-1163)
    T9 = 0.5 * B4_S(k1ox) * Lpe_Vb / sqrtPhis;
    T1 = T9 + B4_I(k2ox) - B4_S(k3b) * Vth_NarrowW;
    dT1_dVb = -T9 / sqrtPhis * dsqrtPhis_dVb;
    T9 = sqrt(B4_S(xj) * Xdep);
    tmp1 = Leff + 2.0 * T9;
    T5 = Leff / tmp1;
    tmp2 = B4_S(a0) * T5;
    tmp3 = B4_S(weff) + B4_S(b1);
    tmp4 = B4_S(b0) / tmp3;
    T2 = tmp2 + tmp4;
    dT2_dVb = -T9 / tmp1 / Xdep * dXdep_dVb;
    T6 = T5 * T5;
    T7 = T5 * T6;
    double Abulk0 = 1.0 + T1 * T2;
    double dAbulk0_dVb = T1 * tmp2 * dT2_dVb + T2 * dT1_dVb;
    T8 = B4_S(ags) * B4_S(a0) * T7;
    double dAbulk_dVg = -T1 * T8;
    double Abulk = Abulk0 + dAbulk_dVg * Vgsteff;
    double dAbulk_dVb = dAbulk0_dVb - T8 * Vgsteff * (dT1_dVb + 3.0 * T1 * dT2_dVb);
    if (Abulk0 < 0.1) {
        T9 = 1.0 / (3.0 - 20.0 * Abulk0);
        Abulk0 = (0.2 - Abulk0) * T9;
        dAbulk0_dVb *= T9 * T9;
    }
    if (Abulk < 0.1) {
        T9 = 1.0 / (3.0 - 20.0 * Abulk);
        Abulk = (0.2 - Abulk) * T9;
        T10 = T9 * T9;
        dAbulk_dVb *= T10;
        dAbulk_dVg *= T10;
    }
    N.Abulk = Abulk;
    T2 = B4_S(keta) * Vbseff;
    if (T2 >= -0.9) {
        T0 = 1.0 / (1.0 + T2);
        dT0_dVb = -B4_S(keta) * T0 * T0;
    } else {
        T1 = 1.0 / (0.8 + T2);
        T0 = (17.0 + 20.0 * T2) * T1;
        dT0_dVb = -B4_S(keta) * T1 * T1;
    }
    dAbulk_dVg *= T0;
    dAbulk_dVb = dAbulk_dVb * T0 + Abulk * dT0_dVb;
    dAbulk0_dVb = dAbulk0_dVb * T0 + Abulk0 * dT0_dVb;
    Abulk *= T0;
    Abulk0 *= T0;

    // mobility, mobmod 0 (:1165-1190, 1297-1315)
    double Denomi, dDenomi_dVg, dDenomi_dVd, dDenomi_dVb;
    T14 = 0.0;
    T0 = Vgsteff + Vth + Vth - T14;
    T2 = B4_S(ua) + B4_S(uc) * Vbseff;
    T3 = T0 / toxe;
    T12 = sqrt(Vth * Vth + 0.0001);
    T9 = 1.0 / (Vgsteff + 2.0 * T12);
    T10 = T9 * toxe;
    T8 = B4_S(ud) * T10 * T10 * Vth;
    T6 = T8 * Vth;
    T5 = T3 * (T2 + B4_S(ub) * T3) + T6;
    T7 = -2.0 * T6 * T9;
    T11 = T7 * Vth / T12;
    dDenomi_dVg = (T2 + 2.0 * B4_S(ub) * T3) / toxe;
    T13 = 2.0 * (dDenomi_dVg + T11 + T8);
    dDenomi_dVd = T13 * dVth_dVd;
    dDenomi_dVb = T13 * dVth_dVb + B4_S(uc) * T3;
    dDenomi_dVg += T7;
    if (T5 >= -0.8) {
        Denomi = 1.0 + T5;
    } else {
        T9 = 1.0 / (7.0 + 10.0 * T5);
        Denomi = (0.6 + T5) * T9;
        T9 *= T9;
        dDenomi_dVg *= T9;
        dDenomi_dVd *= T9;
        dDenomi_dVb *= T9;
    }
    const double ueff = B4_I(u0temp) / Denomi;
    T9 = -ueff / Denomi;
    const double dueff_dVg = T9 * dDenomi_dVg;
    const double dueff_dVd = T9 * dDenomi_dVd;
    const double dueff_dVb = T9 * dDenomi_dVb;

    // Vdsat (:1317-1396)
    const double WVCox = Weff * B4_I(vsattemp) * B4_D(coxe);
    const double WVCoxRds = WVCox * Rds;
    double Esat = 2.0 * B4_I(vsattemp) / ueff;
    double EsatL = Esat * Leff;
    T0 = -EsatL / ueff;
    double dEsatL_dVg = T0 * dueff_dVg;
    double dEsatL_dVd = T0 * dueff_dVd;
    double dEsatL_dVb = T0 * dueff_dVb;
    double Lambda, dLambda_dVg;
    if (B4_S(a1) == 0.0) {
        Lambda = B4_S(a2);
        dLambda_dVg = 0.0;
    } else if (B4_S(a1) > 0.0) {
        T0 = 1.0 - B4_S(a2);
        T1 = T0 - B4_S(a1) * Vgsteff - 0.0001;
        T2 = sqrt(T1 * T1 + 0.0004 * T0);
        Lambda = B4_S(a2) + T0 - 0.5 * (T1 + T2);
        dLambda_dVg = 0.5 * B4_S(a1) * (1.0 + T1 / T2);
    } else {
        T1 = B4_S(a2) + B4_S(a1) * Vgsteff - 0.0001;
        T2 = sqrt(T1 * T1 + 0.0004 * B4_S(a2));
        Lambda = 0.5 * (T1 + T2);
        dLambda_dVg = 0.5 * B4_S(a1) * (1.0 + T1 / T2);
    }
    const double Vgst2Vtm = Vgsteff + 2.0 * Vtm;
    if (Rds > 0.0) {
        tmp2 = dRds_dVg / Rds + dWeff_dVg / Weff;
        tmp3 = dRds_dVb / Rds + dWeff_dVb / Weff;
    } else {
        tmp2 = dWeff_dVg / Weff;
        tmp3 = dWeff_dVb / Weff;
    }
    double Vdsat, dVdsat_dVg, dVdsat_dVd, dVdsat_dVb;
    if ((Rds == 0.0) && (Lambda == 1.0)) {
        T0 = 1.0 / (Abulk * EsatL + Vgst2Vtm);
        tmp1 = 0.0;
        T1 = T0 * T0;
        T2 = Vgst2Vtm * T0;
        T3 = EsatL * Vgst2Vtm;
        Vdsat = T3 * T0;
        dT0_dVg = -(Abulk * dEsatL_dVg + EsatL * dAbulk_dVg + 1.0) * T1;
        dT0_dVd = -(Abulk * dEsatL_dVd) * T1;
        dT0_dVb = -(Abulk * dEsatL_dVb + dAbulk_dVb * EsatL) * T1;
        dVdsat_dVg = T3 * dT0_dVg + T2 * dEsatL_dVg + EsatL * T0;
        dVdsat_dVd = T3 * dT0_dVd + T2 * dEsatL_dVd;
        dVdsat_dVb = T3 * dT0_dVb + T2 * dEsatL_dVb;
    } else {
        tmp1 = dLambda_dVg / (Lambda * Lambda);
        T9 = Abulk * WVCoxRds;
        T8 = Abulk * T9;
        T7 = Vgst2Vtm * T9;
        T6 = Vgst2Vtm * WVCoxRds;
        T0 = 2.0 * Abulk * (T9 - 1.0 + 1.0 / Lambda);
        dT0_dVg = 2.0 * (T8 * tmp2 - Abulk * tmp1 + (2.0 * T9 + 1.0 / Lambda - 1.0) * dAbulk_dVg);
        dT0_dVb = 2.0 * (T8 * (2.0 / Abulk * dAbulk_dVb + tmp3) + (1.0 / Lambda - 1.0) * dAbulk_dVb);
        dT0_dVd = 0.0;
        T1 = Vgst2Vtm * (2.0 / Lambda - 1.0) + Abulk * EsatL + 3.0 * T7;
        dT1_dVg = (2.0 / Lambda - 1.0) - 2.0 * Vgst2Vtm * tmp1 + Abulk * dEsatL_dVg + EsatL * dAbulk_dVg +
                  3.0 * (T9 + T7 * tmp2 + T6 * dAbulk_dVg);
        dT1_dVb = Abulk * dEsatL_dVb + EsatL * dAbulk_dVb + 3.0 * (T6 * dAbulk_dVb + T7 * tmp3);
        dT1_dVd = Abulk * dEsatL_dVd;
        T2 = Vgst2Vtm * (EsatL + 2.0 * T6);
        dT2_dVg = EsatL + Vgst2Vtm * dEsatL_dVg + T6 * (4.0 + 2.0 * Vgst2Vtm * tmp2);
        dT2_dVb = Vgst2Vtm * (dEsatL_dVb + 2.0 * T6 * tmp3);
        dT2_dVd = Vgst2Vtm * dEsatL_dVd;
        T3 = sqrt(T1 * T1 - 2.0 * T0 * T2);
        Vdsat = (T1 - T3) / T0;
        dVdsat_dVg = (dT1_dVg - (T1 * dT1_dVg - dT0_dVg * T2 - T0 * dT2_dVg) / T3 - Vdsat * dT0_dVg) / T0;
        dVdsat_dVb = (dT1_dVb - (T1 * dT1_dVb - dT0_dVb * T2 - T0 * dT2_dVb) / T3 - Vdsat * dT0_dVb) / T0;
        dVdsat_dVd = (dT1_dVd - (T1 * dT1_dVd - T0 * dT2_dVd) / T3) / T0;
    }
    N.vdsat = Vdsat;

    // Vdseff (:1398-1438)
    double Vdseff, dVdseff_dVg, dVdseff_dVd, dVdseff_dVb;
    T1 = Vdsat - Vds - B4_S(delta);
    dT1_dVg = dVdsat_dVg;
    dT1_dVd = dVdsat_dVd - 1.0;
    dT1_dVb = dVdsat_dVb;
    T2 = sqrt(T1 * T1 + 4.0 * B4_S(delta) * Vdsat);
    T0 = T1 / T2;
    T9 = 2.0 * B4_S(delta);
    T3 = T9 / T2;
    dT2_dVg = T0 * dT1_dVg + T3 * dVdsat_dVg;
    dT2_dVd = T0 * dT1_dVd + T3 * dVdsat_dVd;
    dT2_dVb = T0 * dT1_dVb + T3 * dVdsat_dVb;
    if (T1 >= 0.0) {
        Vdseff = Vdsat - 0.5 * (T1 + T2);
        dVdseff_dVg = dVdsat_dVg - 0.5 * (dT1_dVg + dT2_dVg);
        dVdseff_dVd = dVdsat_dVd - 0.5 * (dT1_dVd + dT2_dVd);
        dVdseff_dVb = dVdsat_dVb - 0.5 * (dT1_dVb + dT2_dVb);
    } else {
        T4 = T9 / (T2 - T1);
        T5 = 1.0 - T4;
        T6 = Vdsat * T4 / (T2 - T1);
        Vdseff = Vdsat * T5;
        dVdseff_dVg = dVdsat_dVg * T5 + T6 * (dT2_dVg - dT1_dVg);
        dVdseff_dVd = dVdsat_dVd * T5 + T6 * (dT2_dVd - dT1_dVd);
        dVdseff_dVb = dVdsat_dVb * T5 + T6 * (dT2_dVb - dT1_dVb);
    }
    if (Vds == 0.0) {
        Vdseff = 0.0;
        dVdseff_dVg = 0.0;
        dVdseff_dVb = 0.0;
    }
    if (Vdseff > Vds) Vdseff = Vds;
    const double diffVds = Vds - Vdseff;
    N.Vdseff = Vdseff;
    // velocity overshoot: model lambda is 0 (:1441)

    // Vasat (:1486-1510)
    tmp4 = 1.0 - 0.5 * Abulk * Vdsat / Vgst2Vtm;
    T9 = WVCoxRds * Vgsteff;
    T8 = T9 / Vgst2Vtm;
    T0 = EsatL + Vdsat + 2.0 * T9 * tmp4;
    T7 = 2.0 * WVCoxRds * tmp4;
    dT0_dVg = dEsatL_dVg + dVdsat_dVg + T7 * (1.0 + tmp2 * Vgsteff) -
              T8 * (Abulk * dVdsat_dVg - Abulk * Vdsat / Vgst2Vtm + Vdsat * dAbulk_dVg);
    dT0_dVb = dEsatL_dVb + dVdsat_dVb + T7 * tmp3 * Vgsteff - T8 * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
    dT0_dVd = dEsatL_dVd + dVdsat_dVd - T8 * Abulk * dVdsat_dVd;
    T9 = WVCoxRds * Abulk;
    T1 = 2.0 / Lambda - 1.0 + T9;
    dT1_dVg = -2.0 * tmp1 + WVCoxRds * (Abulk * tmp2 + dAbulk_dVg);
    dT1_dVb = dAbulk_dVb * WVCoxRds + T9 * tmp3;
    const double Vasat = T0 / T1;
    const double dVasat_dVg = (dT0_dVg - Vasat * dT1_dVg) / T1;
    const double dVasat_dVb = (dT0_dVb - Vasat * dT1_dVb) / T1;
    const double dVasat_dVd = dT0_dVd / T1;

    // Idl (:1512-1565)
    tmp1 = B4_I(vtfbphi2);
    tmp2 = 2.0e8 * B4_I(toxp);
    dT0_dVg = 1.0 / tmp2;
    T0 = (Vgsteff + tmp1) * dT0_dVg;
    tmp3 = exp(B4_M(bdos) * 0.7 * log(T0));
    T1 = 1.0 + tmp3;
    T2 = B4_M(bdos) * 0.7 * tmp3 / T0;
    const double Tcen = B4_M(ados) * 1.9e-9 / T1;
    const double dTcen_dVg = -Tcen * T2 * dT0_dVg / T1;
    const double Coxeff = epssub * B4_I(coxp) / (epssub + B4_I(coxp) * Tcen);
    const double dCoxeff_dVg = -Coxeff * Coxeff * dTcen_dVg / epssub;
    const double CoxeffWovL = Coxeff * Weff / Leff;
    const double beta = ueff * CoxeffWovL;
    T3 = ueff / Leff;
    const double dbeta_dVg = CoxeffWovL * dueff_dVg + T3 * (Weff * dCoxeff_dVg + Coxeff * dWeff_dVg);
    const double dbeta_dVd = CoxeffWovL * dueff_dVd;
    const double dbeta_dVb = CoxeffWovL * dueff_dVb + T3 * Coxeff * dWeff_dVb;
    const double AbovVgst2Vtm = Abulk / Vgst2Vtm;
    T0 = 1.0 - 0.5 * Vdseff * AbovVgst2Vtm;
    dT0_dVg = -0.5 * (Abulk * dVdseff_dVg - Abulk * Vdseff / Vgst2Vtm + Vdseff * dAbulk_dVg) / Vgst2Vtm;
    dT0_dVd = -0.5 * Abulk * dVdseff_dVd / Vgst2Vtm;
    dT0_dVb = -0.5 * (Abulk * dVdseff_dVb + dAbulk_dVb * Vdseff) / Vgst2Vtm;
    const double fgche1 = Vgsteff * T0;
    const double dfgche1_dVg = Vgsteff * dT0_dVg + T0;
    const double dfgche1_dVd = Vgsteff * dT0_dVd;
    const double dfgche1_dVb = Vgsteff * dT0_dVb;
    T9 = Vdseff / EsatL;
    const double fgche2 = 1.0 + T9;
    const double dfgche2_dVg = (dVdseff_dVg - T9 * dEsatL_dVg) / EsatL;
    const double dfgche2_dVd = (dVdseff_dVd - T9 * dEsatL_dVd) / EsatL;
    const double dfgche2_dVb = (dVdseff_dVb - T9 * dEsatL_dVb) / EsatL;
    const double gche = beta * fgche1 / fgche2;
    const double dgche_dVg = (beta * dfgche1_dVg + fgche1 * dbeta_dVg - gche * dfgche2_dVg) / fgche2;
    const double dgche_dVd = (beta * dfgche1_dVd + fgche1 * dbeta_dVd - gche * dfgche2_dVd) / fgche2;
    const double dgche_dVb = (beta * dfgche1_dVb + fgche1 * dbeta_dVb - gche * dfgche2_dVb) / fgche2;
    T0 = 1.0 + gche * Rds;
    const double Idl = gche / T0;
    T1 = (1.0 - Idl * Rds) / T0;
    T2 = Idl * Idl;
    const double dIdl_dVg = T1 * dgche_dVg - T2 * dRds_dVg;
    const double dIdl_dVd = T1 * dgche_dVd;
    const double dIdl_dVb = T1 * dgche_dVb - T2 * dRds_dVb;

    // pocket-implant degradation factor (:1567-1576)
    double FP, dFP_dVg;
    if (B4_S(fprout) <= 0.0) {
        FP = 1.0;
        dFP_dVg = 0.0;
    } else {
        T9 = B4_S(fprout) * sqrt(Leff) / Vgst2Vtm;
        FP = 1.0 / (1.0 + T9);
        dFP_dVg = FP * FP * T9 / Vgst2Vtm;
    }
    // VACLM (:1578-1645)
    double PvagTerm, dPvagTerm_dVg, dPvagTerm_dVd, dPvagTerm_dVb;
    T8 = B4_S(pvag) / EsatL;
    T9 = T8 * Vgsteff;
    if (T9 > -0.9) {
        PvagTerm = 1.0 + T9;
        dPvagTerm_dVg = T8 * (1.0 - Vgsteff * dEsatL_dVg / EsatL);
        dPvagTerm_dVb = -T9 * dEsatL_dVb / EsatL;
        dPvagTerm_dVd = -T9 * dEsatL_dVd / EsatL;
    } else {
        T4 = 1.0 / (17.0 + 20.0 * T9);
        PvagTerm = (0.8 + T9) * T4;
        T4 *= T4;
        dPvagTerm_dVg = T8 * (1.0 - Vgsteff * dEsatL_dVg / EsatL) * T4;
        T9 *= T4 / EsatL;
        dPvagTerm_dVb = -T9 * dEsatL_dVb;
        dPvagTerm_dVd = -T9 * dEsatL_dVd;
    }
    double Cclm, dCclm_dVg, dCclm_dVd, dCclm_dVb, VACLM, dVACLM_dVg, dVACLM_dVd, dVACLM_dVb;
    if ((B4_S(pclm) > MIN_EXP) && (diffVds > 1.0e-10)) {
        T0 = 1.0 + Rds * Idl;
        dT0_dVg = dRds_dVg * Idl + Rds * dIdl_dVg;
        dT0_dVd = Rds * dIdl_dVd;
        dT0_dVb = dRds_dVb * Idl + Rds * dIdl_dVb;
        T2 = Vdsat / Esat;
        T1 = Leff + T2;
        dT1_dVg = (dVdsat_dVg - T2 * dEsatL_dVg / Leff) / Esat;
        dT1_dVd = (dVdsat_dVd - T2 * dEsatL_dVd / Leff) / Esat;
        dT1_dVb = (dVdsat_dVb - T2 * dEsatL_dVb / Leff) / Esat;
        Cclm = FP * PvagTerm * T0 * T1 / (B4_S(pclm) * B4_S(litl));
        dCclm_dVg = Cclm * (dFP_dVg / FP + dPvagTerm_dVg / PvagTerm + dT0_dVg / T0 + dT1_dVg / T1);
        dCclm_dVb = Cclm * (dPvagTerm_dVb / PvagTerm + dT0_dVb / T0 + dT1_dVb / T1);
        dCclm_dVd = Cclm * (dPvagTerm_dVd / PvagTerm + dT0_dVd / T0 + dT1_dVd / T1);
        VACLM = Cclm * diffVds;
        dVACLM_dVg = dCclm_dVg * diffVds - dVdseff_dVg * Cclm;
        dVACLM_dVb = dCclm_dVb * diffVds - dVdseff_dVb * Cclm;
        dVACLM_dVd = dCclm_dVd * diffVds + (1.0 - dVdseff_dVd) * Cclm;
    } else {
        VACLM = MAX_EXP;
        Cclm = MAX_EXP;
        dVACLM_dVd = dVACLM_dVg = dVACLM_dVb = 0.0;
        dCclm_dVd = dCclm_dVg = dCclm_dVb = 0.0;
    }
    // VADIBL (:1647-1694)
    double VADIBL, dVADIBL_dVg, dVADIBL_dVd, dVADIBL_dVb;
    if (B4_S(thetaRout) > MIN_EXP) {
        T8 = Abulk * Vdsat;
        T0 = Vgst2Vtm * T8;
        dT0_dVg = Vgst2Vtm * Abulk * dVdsat_dVg + T8 + Vgst2Vtm * Vdsat * dAbulk_dVg;
        dT0_dVb = Vgst2Vtm * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
        dT0_dVd = Vgst2Vtm * Abulk * dVdsat_dVd;
        T1 = Vgst2Vtm + T8;
        dT1_dVg = 1.0 + Abulk * dVdsat_dVg + Vdsat * dAbulk_dVg;
        dT1_dVb = Abulk * dVdsat_dVb + dAbulk_dVb * Vdsat;
        dT1_dVd = Abulk * dVdsat_dVd;
        T9 = T1 * T1;
        T2 = B4_S(thetaRout);
        VADIBL = (Vgst2Vtm - T0 / T1) / T2;
        dVADIBL_dVg = (1.0 - dT0_dVg / T1 + T0 * dT1_dVg / T9) / T2;
        dVADIBL_dVb = (-dT0_dVb / T1 + T0 * dT1_dVb / T9) / T2;
        dVADIBL_dVd = (-dT0_dVd / T1 + T0 * dT1_dVd / T9) / T2;
        T7 = B4_S(pdiblb) * Vbseff;
        if (T7 >= -0.9) {
            T3 = 1.0 / (1.0 + T7);
            VADIBL *= T3;
            dVADIBL_dVg *= T3;
            dVADIBL_dVb = (dVADIBL_dVb - VADIBL * B4_S(pdiblb)) * T3;
            dVADIBL_dVd *= T3;
        } else {
            T4 = 1.0 / (0.8 + T7);
            T3 = (17.0 + 20.0 * T7) * T4;
            dVADIBL_dVg *= T3;
            dVADIBL_dVb = dVADIBL_dVb * T3 - VADIBL * B4_S(pdiblb) * T4 * T4;
            dVADIBL_dVd *= T3;
            VADIBL *= T3;
        }
        dVADIBL_dVg = dVADIBL_dVg * PvagTerm + VADIBL * dPvagTerm_dVg;
        dVADIBL_dVb = dVADIBL_dVb * PvagTerm + VADIBL * dPvagTerm_dVb;
        dVADIBL_dVd = dVADIBL_dVd * PvagTerm + VADIBL * dPvagTerm_dVd;
        VADIBL *= PvagTerm;
    } else {
        VADIBL = MAX_EXP;
        dVADIBL_dVd = dVADIBL_dVg = dVADIBL_dVb = 0.0;
    }
    // Va (:1696-1700)
    const double Va = Vasat + VACLM;
    const double dVa_dVg = dVasat_dVg + dVACLM_dVg;
    const double dVa_dVb = dVasat_dVb + dVACLM_dVb;
    const double dVa_dVd = dVasat_dVd + dVACLM_dVd;
    // VADITS (:1702-1723)
    double VADITS, dVADITS_dVg, dVADITS_dVd;
    T0 = B4_S(pditsd) * Vds;
    if (T0 > EXP_THRESHOLD) {
        T1 = MAX_EXP;
        dT1_dVd = 0.0;
    } else {
        T1 = exp(T0);
        dT1_dVd = T1 * B4_S(pditsd);
    }
    if (B4_S(pdits) > MIN_EXP) {
        T2 = 1.0 + B4_M(pditsl) * Leff;
        VADITS = (1.0 + T2 * T1) / B4_S(pdits);
        dVADITS_dVg = VADITS * dFP_dVg;
        dVADITS_dVd = FP * T2 * dT1_dVd / B4_S(pdits);
        VADITS *= FP;
    } else {
        VADITS = MAX_EXP;
        dVADITS_dVg = 0.0;
        dVADITS_dVd = 0.0;
    }
    // VASCBE (:1725-1744)
    double VASCBE = MAX_EXP, dVASCBE_dVg = 0.0, dVASCBE_dVd = 0.0, dVASCBE_dVb = 0.0;
    if (B4_S(pscbe2) > 0.0 && B4_S(pscbe1) >= 0.0) {
        if (diffVds > B4_S(pscbe1) * B4_S(litl) / EXP_THRESHOLD) {
            T0 = B4_S(pscbe1) * B4_S(litl) / diffVds;
            VASCBE = Leff * exp(T0) / B4_S(pscbe2);
            T1 = T0 * VASCBE / diffVds;
            dVASCBE_dVg = T1 * dVdseff_dVg;
            dVASCBE_dVd = -T1 * (1.0 - dVdseff_dVd);
            dVASCBE_dVb = T1 * dVdseff_dVb;
        } else {
            VASCBE = MAX_EXP * Leff / B4_S(pscbe2);
            dVASCBE_dVg = dVASCBE_dVd = dVASCBE_dVb = 0.0;
        }
    }
    // DIBL, DITS, CLM (:1746-1780)
    double Idsa, dIdsa_dVg, dIdsa_dVd, dIdsa_dVb;
    T9 = diffVds / VADIBL;
    T0 = 1.0 + T9;
    Idsa = Idl * T0;
    dIdsa_dVg = T0 * dIdl_dVg - Idl * (dVdseff_dVg + T9 * dVADIBL_dVg) / VADIBL;
    dIdsa_dVd = T0 * dIdl_dVd + Idl * (1.0 - dVdseff_dVd - T9 * dVADIBL_dVd) / VADIBL;
    dIdsa_dVb = T0 * dIdl_dVb - Idl * (dVdseff_dVb + T9 * dVADIBL_dVb) / VADIBL;
    T9 = diffVds / VADITS;
    T0 = 1.0 + T9;
    dIdsa_dVg = T0 * dIdsa_dVg - Idsa * (dVdseff_dVg + T9 * dVADITS_dVg) / VADITS;
    dIdsa_dVd = T0 * dIdsa_dVd + Idsa * (1.0 - dVdseff_dVd - T9 * dVADITS_dVd) / VADITS;
    dIdsa_dVb = T0 * dIdsa_dVb - Idsa * dVdseff_dVb / VADITS;
    Idsa *= T0;
    T0 = log(Va / Vasat);
    dT0_dVg = dVa_dVg / Va - dVasat_dVg / Vasat;
    dT0_dVb = dVa_dVb / Va - dVasat_dVb / Vasat;
    dT0_dVd = dVa_dVd / Va - dVasat_dVd / Vasat;
    T1 = T0 / Cclm;
    T9 = 1.0 + T1;
    dT9_dVg = (dT0_dVg - T1 * dCclm_dVg) / Cclm;
    dT9_dVb = (dT0_dVb - T1 * dCclm_dVb) / Cclm;
    dT9_dVd = (dT0_dVd - T1 * dCclm_dVd) / Cclm;
    dIdsa_dVg = dIdsa_dVg * T9 + Idsa * dT9_dVg;
    dIdsa_dVb = dIdsa_dVb * T9 + Idsa * dT9_dVb;
    dIdsa_dVd = dIdsa_dVd * T9 + Idsa * dT9_dVd;
    Idsa *= T9;
    // substrate current (:1782-1820)
    double Isub, Gbd, Gbb, Gbg;
    tmp = B4_S(alpha0) + B4_S(alpha1) * Leff;
    if ((tmp <= 0.0) || (B4_S(beta0) <= 0.0)) {
        Isub = Gbd = Gbb = Gbg = 0.0;
    } else {
        T2 = tmp / Leff;
        if (diffVds > B4_S(beta0) / EXP_THRESHOLD) {
            T0 = -B4_S(beta0) / diffVds;
            T1 = T2 * diffVds * exp(T0);
            T3 = T1 / diffVds * (T0 - 1.0);
            dT1_dVg = T3 * dVdseff_dVg;
            dT1_dVd = T3 * (dVdseff_dVd - 1.0);
            dT1_dVb = T3 * dVdseff_dVb;
        } else {
            T3 = T2 * MIN_EXP;
            T1 = T3 * diffVds;
            dT1_dVg = -T3 * dVdseff_dVg;
            dT1_dVd = T3 * (1.0 - dVdseff_dVd);
            dT1_dVb = -T3 * dVdseff_dVb;
        }
        T4 = Idsa * Vdseff;
        Isub = T1 * T4;
        Gbg = T1 * (dIdsa_dVg * Vdseff + Idsa * dVdseff_dVg) + T4 * dT1_dVg;
        Gbd = T1 * (dIdsa_dVd * Vdseff + Idsa * dVdseff_dVd) + T4 * dT1_dVd;
        Gbb = T1 * (dIdsa_dVb * Vdseff + Idsa * dVdseff_dVb) + T4 * dT1_dVb;
        Gbd += Gbg * dVgsteff_dVd;
        Gbb += Gbg * dVgsteff_dVb;
        Gbg *= dVgsteff_dVg;
        Gbb *= dVbseff_dVb;
    }
    N.csub = Isub;
    N.gbbs = Gbb;
    N.gbgs = Gbg;
    N.gbds = Gbd;
    // SCBE (:1822-1840)
    T9 = diffVds / VASCBE;
    T0 = 1.0 + T9;
    const double Ids = Idsa * T0;
    double Gm = T0 * dIdsa_dVg - Idsa * (dVdseff_dVg + T9 * dVASCBE_dVg) / VASCBE;
    double Gds = T0 * dIdsa_dVd + Idsa * (1.0 - dVdseff_dVd - T9 * dVASCBE_dVd) / VASCBE;
    double Gmb = T0 * dIdsa_dVb - Idsa * (dVdseff_dVb + T9 * dVASCBE_dVb) / VASCBE;
    tmp1 = Gds + Gm * dVgsteff_dVd;
    tmp2 = Gmb + Gm * dVgsteff_dVb;
    tmp3 = Gm;
    Gm = (Ids * dVdseff_dVg + Vdseff * tmp3) * dVgsteff_dVg;
    Gds = Ids * (dVdseff_dVd + dVdseff_dVg * dVgsteff_dVd) + Vdseff * tmp1;
    Gmb = (Ids * (dVdseff_dVb + dVdseff_dVg * dVgsteff_dVb) + Vdseff * tmp2) * dVbseff_dVb;
    const double cdrain = Ids * Vdseff;
    // source-end velocity limit: vtl is not given (:1845)
    N.gds = Gds;
    N.gm = Gm;
    N.gmbs = Gmb;
    N.IdovVds = Ids;
    if (N.IdovVds <= 1.0e-9) N.IdovVds = 1.0e-9;
    // Rg (rgatemod, NQS) and bias-dependent S/D resistance (rdsmod) are off
    // GIDL / GISL, gidlmod 0 (:1985-2100)
    double Igidl, Ggidld, Ggidlg, Ggidlb, Igisl, Ggisls, Ggislg, Ggislb;
    {
        const double T0g = 3.0 * toxe;
        vgs_eff = N.vgs_eff;
        dvgs_eff_dvg = N.dvgs_eff_dvg;
        vgd_eff = N.vgd_eff;
        dvgd_eff_dvg = N.dvgd_eff_dvg;
        T1 = (vds - vgs_eff - B4_S(egidl)) / T0g;
        if ((B4_S(agidl) <= 0.0) || (B4_S(bgidl) <= 0.0) || (T1 <= 0.0) || (B4_S(cgidl) <= 0.0) || (vbd > 0.0)) {
            Igidl = Ggidld = Ggidlg = Ggidlb = 0.0;
        } else {
            dT1_dVd = 1.0 / T0g;
            dT1_dVg = -dvgs_eff_dvg * dT1_dVd;
            T2 = B4_S(bgidl) / T1;
            if (T2 < 100.0) {
                Igidl = B4_S(agidl) * B4_S(weffCJ) * T1 * exp(-T2);
                T3 = Igidl * (1.0 + T2) / T1;
                Ggidld = T3 * dT1_dVd;
                Ggidlg = T3 * dT1_dVg;
            } else {
                Igidl = B4_S(agidl) * B4_S(weffCJ) * 3.720075976e-44;
                Ggidld = Igidl * dT1_dVd;
                Ggidlg = Igidl * dT1_dVg;
                Igidl *= T1;
            }
            T4 = vbd * vbd;
            T5 = -vbd * T4;
            T6 = B4_S(cgidl) + T5;
            T7 = T5 / T6;
            T8 = 3.0 * B4_S(cgidl) * T4 / T6 / T6;
            Ggidld = Ggidld * T7 + Igidl * T8;
            Ggidlg = Ggidlg * T7;
            Ggidlb = -Igidl * T8;
            Igidl *= T7;
        }
        T1 = (-vds - vgd_eff - B4_S(egisl)) / T0g;
        if ((B4_S(agisl) <= 0.0) || (B4_S(bgisl) <= 0.0) || (T1 <= 0.0) || (B4_S(cgisl) <= 0.0) || (vbs > 0.0)) {
            Igisl = Ggisls = Ggislg = Ggislb = 0.0;
        } else {
            dT1_dVd = 1.0 / T0g;
            dT1_dVg = -dvgd_eff_dvg * dT1_dVd;
            T2 = B4_S(bgisl) / T1;
            if (T2 < 100.0) {
                Igisl = B4_S(agisl) * B4_S(weffCJ) * T1 * exp(-T2);
                T3 = Igisl * (1.0 + T2) / T1;
                Ggisls = T3 * dT1_dVd;
                Ggislg = T3 * dT1_dVg;
            } else {
                Igisl = B4_S(agisl) * B4_S(weffCJ) * 3.720075976e-44;
                Ggisls = Igisl * dT1_dVd;
                Ggislg = Igisl * dT1_dVg;
                Igisl *= T1;
            }
            T4 = vbs * vbs;
            T5 = -vbs * T4;
            T6 = B4_S(cgisl) + T5;
            T7 = T5 / T6;
            T8 = 3.0 * B4_S(cgisl) * T4 / T6 / T6;
            Ggisls = Ggisls * T7 + Igisl * T8;
            Ggislg = Ggislg * T7;
            Ggislb = -Igisl * T8;
            Igisl *= T7;
        }
    }
    N.Igidl = Igidl;
    N.ggidld = Ggidld;
    N.ggidlg = Ggidlg;
    N.ggidlb = Ggidlb;
    N.Igisl = Igisl;
    N.ggisls = Ggisls;
    N.ggislg = Ggislg;
    N.ggislb = Ggislb;
    // gate tunnelling: igcmod and igbmod are 0
    double cd = cdrain;
    if (B4_I(nf) != 1.0) {  // :2591-2635
        cd *= B4_I(nf);
        N.gds *= B4_I(nf);
        N.gm *= B4_I(nf);
        N.gmbs *= B4_I(nf);
        N.IdovVds *= B4_I(nf);
        N.gbbs *= B4_I(nf);
        N.gbgs *= B4_I(nf);
        N.gbds *= B4_I(nf);
        N.csub *= B4_I(nf);
        N.Igidl *= B4_I(nf);
        N.ggidld *= B4_I(nf);
        N.ggidlg *= B4_I(nf);
        N.ggidlb *= B4_I(nf);
        N.Igisl *= B4_I(nf);
        N.ggisls *= B4_I(nf);
        N.ggislg *= B4_I(nf);
        N.ggislb *= B4_I(nf);
    }
    N.ggidls = -(N.ggidld + N.ggidlg + N.ggidlb);
    N.ggisld = -(N.ggisls + N.ggislg + N.ggislb);
    N.cd = cd;

    // noise-analysis inversion charge, tnoimod 0 (:2643-2667): reassigns Abulk, Vdsat, Vdseff
    Abulk = Abulk0 * B4_S(abulkCVfactor);
    Vdsat = Vgsteff / Abulk;
    T0 = Vdsat - Vds - DELTA_4;
    T1 = sqrt(T0 * T0 + 4.0 * DELTA_4 * Vdsat);
    if (T0 >= 0.0) {
        Vdseff = Vdsat - 0.5 * (T0 + T1);
    } else {
        T3 = (DELTA_4 + DELTA_4) / (T1 - T0);
        T4 = 1.0 - T3;
        Vdseff = Vdsat * T4;
    }
    if (Vds == 0.0) Vdseff = 0.0;
    T0 = Abulk * Vdseff;
    T1 = 12.0 * (Vgsteff - 0.5 * T0 + 1.0e-20);
    T2 = Vdseff / T1;
    T3 = T0 * T2;
    N.qinv = Coxeff * B4_S(weffCV) * B4_I(nf) * B4_S(leffCV) * (Vgsteff - 0.5 * T0 + Abulk * T3);

    // C-V, capmod 2 with cvchargemod 0 and the 40/60 partition of xpart 0 (:2672-3476)
    double qgate, qbulk, qdrn, qsrc;
    {
        double VbseffCV, dVbseffCV_dVb;
        if (Vbseff < 0.0) {
            VbseffCV = Vbseff;
            dVbseffCV_dVb = 1.0;
        } else {
            VbseffCV = B4_S(phi) - Phis;
            dVbseffCV_dVb = -dPhis_dVb;
        }
        const double CoxWL = B4_D(coxe) * B4_S(weffCV) * B4_S(leffCV) * B4_I(nf);
        double VgsteffC, dVgsteffC_dVg, dVgsteffC_dVd, dVgsteffC_dVb;  // VgsteffCV (:2974-3005)
        {
            const double noff = n * B4_S(noff);
            const double dnoff_dVd = B4_S(noff) * dn_dVd;
            const double dnoff_dVb = B4_S(noff) * dn_dVb;
            T0 = Vtm * noff;
            const double voffcv = B4_S(voffcv);
            const double VgstNVt = (Vgst - voffcv) / T0;
            if (VgstNVt > EXP_THRESHOLD) {
                VgsteffC = Vgst - voffcv;
                dVgsteffC_dVg = dVgs_eff_dVg;
                dVgsteffC_dVd = -dVth_dVd;
                dVgsteffC_dVb = -dVth_dVb;
            } else if (VgstNVt < -EXP_THRESHOLD) {
                VgsteffC = T0 * log(1.0 + MIN_EXP);
                dVgsteffC_dVg = 0.0;
                dVgsteffC_dVd = VgsteffC / noff;
                dVgsteffC_dVb = dVgsteffC_dVd * dnoff_dVb;
                dVgsteffC_dVd *= dnoff_dVd;
            } else {
                const double ExpVgst = exp(VgstNVt);
                VgsteffC = T0 * log(1.0 + ExpVgst);
                dVgsteffC_dVg = ExpVgst / (1.0 + ExpVgst);
                dVgsteffC_dVd = -dVgsteffC_dVg * (dVth_dVd + (Vgst - voffcv) / noff * dnoff_dVd) + VgsteffC / noff * dnoff_dVd;
                dVgsteffC_dVb = -dVgsteffC_dVg * (dVth_dVb + (Vgst - voffcv) / noff * dnoff_dVb) + VgsteffC / noff * dnoff_dVb;
                dVgsteffC_dVg *= dVgs_eff_dVg;
            }
        }
        // charge-thickness model (:3215-3474)
        double V3 = B4_I(vfbzb) - Vgs_eff + VbseffCV - DELTA_3;
        if (B4_I(vfbzb) <= 0.0)
            T0 = sqrt(V3 * V3 - 4.0 * DELTA_3 * B4_I(vfbzb));
        else
            T0 = sqrt(V3 * V3 + 4.0 * DELTA_3 * B4_I(vfbzb));
        T1 = 0.5 * (1.0 + V3 / T0);
        const double Vfbeff = B4_I(vfbzb) - 0.5 * (V3 + T0);
        const double dVfbeff_dVg = T1 * dVgs_eff_dVg;
        const double dVfbeff_dVb = -T1 * dVbseffCV_dVb;
        const double Cox = B4_I(coxp);
        double Tox = 1.0e8 * B4_I(toxp);
        T0 = (Vgs_eff - VbseffCV - B4_I(vfbzb)) / Tox;
        dT0_dVg = dVgs_eff_dVg / Tox;
        dT0_dVb = -dVbseffCV_dVb / Tox;
        tmp = T0 * B4_S(acde);
        double Tcen1, dTcen_dVb, dTcen_dVg;
        if ((-EXP_THRESHOLD < tmp) && (tmp < EXP_THRESHOLD)) {
            Tcen1 = B4_S(ldeb) * exp(tmp);
            dTcen_dVg = B4_S(acde) * Tcen1;
            dTcen_dVb = dTcen_dVg * dT0_dVb;
            dTcen_dVg *= dT0_dVg;
        } else if (tmp <= -EXP_THRESHOLD) {
            Tcen1 = B4_S(ldeb) * MIN_EXP;
            dTcen_dVg = 0.0;
            dTcen_dVb = 0.0;
        } else {
            Tcen1 = B4_S(ldeb) * MAX_EXP;
            dTcen_dVg = 0.0;
            dTcen_dVb = 0.0;
        }
        const double LINK = 1.0e-3 * B4_I(toxp);
        V3 = B4_S(ldeb) - Tcen1 - LINK;
        const double V4 = sqrt(V3 * V3 + 4.0 * LINK * B4_S(ldeb));
        const double Tcen2 = B4_S(ldeb) - 0.5 * (V3 + V4);
        T1 = 0.5 * (1.0 + V3 / V4);
        dTcen_dVg *= T1;
        dTcen_dVb *= T1;
        double Ccen = epssub / Tcen2;
        T2 = Cox / (Cox + Ccen);
        double CoxeffC = T2 * Ccen;
        T3 = -Ccen / Tcen2;
        double dCoxeff_dVg_ = T2 * T2 * T3;
        double dCoxeff_dVbC = dCoxeff_dVg_ * dTcen_dVb;
        double dCoxeff_dVgC = dCoxeff_dVg_ * dTcen_dVg;
        double CoxWLcen = CoxWL * CoxeffC / B4_D(coxe);
        const double Qac0 = CoxWLcen * (Vfbeff - B4_I(vfbzb));
        double QovCox = Qac0 / CoxeffC;
        const double dQac0_dVg = CoxWLcen * dVfbeff_dVg + QovCox * dCoxeff_dVgC;
        const double dQac0_dVb = CoxWLcen * dVfbeff_dVb + QovCox * dCoxeff_dVbC;
        T0 = 0.5 * B4_S(k1ox);
        T3 = Vgs_eff - Vfbeff - VbseffCV - VgsteffC;
        if (B4_S(k1ox) == 0.0) {
            T1 = 0.0;
            T2 = 0.0;
        } else if (T3 < 0.0) {
            T1 = T0 + T3 / B4_S(k1ox);
            T2 = CoxWLcen;
        } else {
            T1 = sqrt(T0 * T0 + T3);
            T2 = CoxWLcen * T0 / T1;
        }
        const double Qsub0 = CoxWLcen * B4_S(k1ox) * (T1 - T0);
        QovCox = Qsub0 / CoxeffC;
        const double dQsub0_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteffC_dVg) + QovCox * dCoxeff_dVgC;
        const double dQsub0_dVd = -T2 * dVgsteffC_dVd;
        const double dQsub0_dVb = -T2 * (dVfbeff_dVb + dVbseffCV_dVb + dVgsteffC_dVb) + QovCox * dCoxeff_dVbC;
        double DenomiC;
        if (B4_S(k1ox) <= 0.0) {
            DenomiC = 0.25 * B4_S(moin) * Vtm;
            T0 = 0.5 * B4_S(sqrtPhi);
        } else {
            DenomiC = B4_S(moin) * Vtm * B4_S(k1ox) * B4_S(k1ox);
            T0 = B4_S(k1ox) * B4_S(sqrtPhi);
        }
        T1 = 2.0 * T0 + VgsteffC;
        const double DeltaPhi = Vtm * log(1.0 + T1 * VgsteffC / DenomiC);
        const double dDeltaPhi_dVg = 2.0 * Vtm * (T1 - T0) / (DenomiC + T1 * VgsteffC);
        T0 = VgsteffC - DeltaPhi - 0.001;
        dT0_dVg = 1.0 - dDeltaPhi_dVg;
        T1 = sqrt(T0 * T0 + VgsteffC * 0.004);
        const double VgDP = 0.5 * (T0 + T1);
        const double dVgDP_dVg = 0.5 * (dT0_dVg + (T0 * dT0_dVg + 0.002) / T1);
        Tox += Tox;
        T0 = (VgsteffC + B4_I(vtfbphi2)) / Tox;
        tmp = exp(B4_M(bdos) * 0.7 * log(T0));
        T1 = 1.0 + tmp;
        T2 = B4_M(bdos) * 0.7 * tmp / (T0 * Tox);
        const double Tcen3 = B4_M(ados) * 1.9e-9 / T1;
        double dTcen3_dVg = -Tcen3 * T2 / T1;
        const double dTcen3_dVd = dTcen3_dVg * dVgsteffC_dVd;
        const double dTcen3_dVb = dTcen3_dVg * dVgsteffC_dVb;
        dTcen3_dVg *= dVgsteffC_dVg;
        Ccen = epssub / Tcen3;
        T0 = Cox / (Cox + Ccen);
        CoxeffC = T0 * Ccen;
        T1 = -Ccen / Tcen3;
        dCoxeff_dVg_ = T0 * T0 * T1;
        const double dCoxeff_dVd = dCoxeff_dVg_ * dTcen3_dVd;
        dCoxeff_dVbC = dCoxeff_dVg_ * dTcen3_dVb;
        dCoxeff_dVgC = dCoxeff_dVg_ * dTcen3_dVg;
        CoxWLcen = CoxWL * CoxeffC / B4_D(coxe);
        const double AbulkCV = Abulk0 * B4_S(abulkCVfactor);
        const double dAbulkCV_dVb = B4_S(abulkCVfactor) * dAbulk0_dVb;
        const double VdsatCV = VgDP / AbulkCV;
        T0 = VdsatCV - Vds - DELTA_4;
        dT0_dVg = dVgDP_dVg / AbulkCV;
        dT0_dVb = -VdsatCV * dAbulkCV_dVb / AbulkCV;
        T1 = sqrt(T0 * T0 + 4.0 * DELTA_4 * VdsatCV);
        dT1_dVg = (T0 + DELTA_4 + DELTA_4) / T1;
        dT1_dVd = -T0 / T1;
        dT1_dVb = dT1_dVg * dT0_dVb;
        dT1_dVg *= dT0_dVg;
        double VdseffCV, dVdseffCV_dVg, dVdseffCV_dVd, dVdseffCV_dVb;
        if (T0 >= 0.0) {
            VdseffCV = VdsatCV - 0.5 * (T0 + T1);
            dVdseffCV_dVg = 0.5 * (dT0_dVg - dT1_dVg);
            dVdseffCV_dVd = 0.5 * (1.0 - dT1_dVd);
            dVdseffCV_dVb = 0.5 * (dT0_dVb - dT1_dVb);
        } else {
            T3 = (DELTA_4 + DELTA_4) / (T1 - T0);
            T4 = 1.0 - T3;
            T5 = VdsatCV * T3 / (T1 - T0);
            VdseffCV = VdsatCV * T4;
            dVdseffCV_dVg = dT0_dVg * T4 + T5 * (dT1_dVg - dT0_dVg);
            dVdseffCV_dVd = T5 * (dT1_dVd + 1.0);
            dVdseffCV_dVb = dT0_dVb * (T4 - T5) + T5 * dT1_dVb;
        }
        if (Vds == 0.0) {
            VdseffCV = 0.0;
            dVdseffCV_dVg = 0.0;
            dVdseffCV_dVb = 0.0;
        }
        T0 = AbulkCV * VdseffCV;
        T1 = VgDP;
        T2 = 12.0 * (T1 - 0.5 * T0 + 1.0e-20);
        T3 = T0 / T2;
        T4 = 1.0 - 12.0 * T3 * T3;
        T5 = AbulkCV * (6.0 * T0 * (4.0 * T1 - T0) / (T2 * T2) - 0.5);
        T6 = T5 * VdseffCV / AbulkCV;
        qgate = CoxWLcen * (T1 - T0 * (0.5 - T3));
        QovCox = qgate / CoxeffC;
        double Cgg1 = CoxWLcen * (T4 * dVgDP_dVg + T5 * dVdseffCV_dVg);
        const double Cgd1 = CoxWLcen * T5 * dVdseffCV_dVd + Cgg1 * dVgsteffC_dVd + QovCox * dCoxeff_dVd;
        const double Cgb1 =
            CoxWLcen * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Cgg1 * dVgsteffC_dVb + QovCox * dCoxeff_dVbC;
        Cgg1 = Cgg1 * dVgsteffC_dVg + QovCox * dCoxeff_dVgC;
        T7 = 1.0 - AbulkCV;
        T8 = T2 * T2;
        T9 = 12.0 * T7 * T0 * T0 / (T8 * AbulkCV);
        T10 = T9 * dVgDP_dVg;
        T11 = -T7 * T5 / AbulkCV;
        T12 = -(T9 * T1 / AbulkCV + VdseffCV * (0.5 - T0 / T2));
        qbulk = CoxWLcen * T7 * (0.5 * VdseffCV - T0 * VdseffCV / T2);
        QovCox = qbulk / CoxeffC;
        double Cbg1 = CoxWLcen * (T10 + T11 * dVdseffCV_dVg);
        const double Cbd1 = CoxWLcen * T11 * dVdseffCV_dVd + Cbg1 * dVgsteffC_dVd + QovCox * dCoxeff_dVd;
        const double Cbb1 =
            CoxWLcen * (T11 * dVdseffCV_dVb + T12 * dAbulkCV_dVb) + Cbg1 * dVgsteffC_dVb + QovCox * dCoxeff_dVbC;
        Cbg1 = Cbg1 * dVgsteffC_dVg + QovCox * dCoxeff_dVgC;
        // 40/60 partition
        T2 = T2 / 12.0;
        T3 = 0.5 * CoxWLcen / (T2 * T2);
        T4 = T1 * (2.0 * T0 * T0 / 3.0 + T1 * (T1 - 4.0 * T0 / 3.0)) - 2.0 * T0 * T0 * T0 / 15.0;
        qsrc = -T3 * T4;
        QovCox = qsrc / CoxeffC;
        T8 = 4.0 / 3.0 * T1 * (T1 - T0) + 0.4 * T0 * T0;
        T5 = -2.0 * qsrc / T2 - T3 * (T1 * (3.0 * T1 - 8.0 * T0 / 3.0) + 2.0 * T0 * T0 / 3.0);
        T6 = AbulkCV * (qsrc / T2 + T3 * T8);
        T7 = T6 * VdseffCV / AbulkCV;
        double Csg = T5 * dVgDP_dVg + T6 * dVdseffCV_dVg;
        const double Csd = Csg * dVgsteffC_dVd + T6 * dVdseffCV_dVd + QovCox * dCoxeff_dVd;
        double Csb = Csg * dVgsteffC_dVb + T6 * dVdseffCV_dVb + T7 * dAbulkCV_dVb + QovCox * dCoxeff_dVbC;
        Csg = Csg * dVgsteffC_dVg + QovCox * dCoxeff_dVgC;
        qgate += Qac0 + Qsub0 - qbulk;
        qbulk -= (Qac0 + Qsub0);
        qdrn = -(qgate + qbulk + qsrc);
        const double Cbg = Cbg1 - dQac0_dVg - dQsub0_dVg;
        const double Cbd = Cbd1 - dQsub0_dVd;
        double Cbb = Cbb1 - dQac0_dVb - dQsub0_dVb;
        const double Cgg = Cgg1 - Cbg;
        const double Cgd = Cgd1 - Cbd;
        double Cgb = Cgb1 - Cbb;
        Cgb *= dVbseff_dVb;
        Cbb *= dVbseff_dVb;
        Csb *= dVbseff_dVb;
        N.cggb = Cgg;
        N.cgsb = -(Cgg + Cgd + Cgb);
        N.cgdb = Cgd;
        N.cdgb = -(Cgg + Cbg + Csg);
        N.cdsb = (Cgg + Cgd + Cgb + Cbg + Cbd + Cbb + Csg + Csd + Csb);
        N.cddb = -(Cgd + Cbd + Csd);
        N.cbgb = Cbg;
        N.cbsb = -(Cbg + Cbd + Cbb);
        N.cbdb = Cbd;
    }
    N.csgb = -N.cggb - N.cdgb - N.cbgb;
    N.csdb = -N.cgdb - N.cddb - N.cbdb;
    N.cssb = -N.cgsb - N.cdsb - N.cbsb;
    N.qgate = qgate;
    N.qbulk = qbulk;
    N.qdrn = qdrn;
    N.qsrc = -(qgate + qbulk + qdrn);

    // junction capacitances (:3499-3614)
    {
        const double czbd = B4_D(DunitAreaTempJctCap) * B4_I(Adeff);
        const double czbs = B4_D(SunitAreaTempJctCap) * B4_I(Aseff);
        const double czbdsw = B4_D(DunitLengthSidewallTempJctCap) * B4_I(Pdeff);
        const double czbdswg = B4_D(DunitLengthGateSidewallTempJctCap) * B4_S(weffCJ) * B4_I(nf);
        const double czbssw = B4_D(SunitLengthSidewallTempJctCap) * B4_I(Pseff);
        const double czbsswg = B4_D(SunitLengthGateSidewallTempJctCap) * B4_S(weffCJ) * B4_I(nf);
        auto side = [](double v, double cz, double Phi, double MJ, double& q, double& c) {
            const double arg = 1.0 - v / Phi;
            const double sarg = (MJ == 0.5) ? 1.0 / sqrt(arg) : exp(-MJ * log(arg));
            q += Phi * cz * (1.0 - arg * sarg) / (1.0 - MJ);
            c += cz * sarg;
        };
        if (vbs_jct == 0.0) {
            N.qbs = 0.0;
            N.capbs = czbs + czbssw + czbsswg;
        } else if (vbs_jct < 0.0) {
            N.qbs = 0.0;
            N.capbs = 0.0;
            if (czbs > 0.0) side(vbs_jct, czbs, B4_D(PhiBS), B4_M(mjs), N.qbs, N.capbs);
            if (czbssw > 0.0) side(vbs_jct, czbssw, B4_D(PhiBSWS), B4_M(mjsws), N.qbs, N.capbs);
            if (czbsswg > 0.0) side(vbs_jct, czbsswg, B4_D(PhiBSWGS), B4_M(mjswgs), N.qbs, N.capbs);
        } else {
            T0 = czbs + czbssw + czbsswg;
            T1 = vbs_jct * (czbs * B4_M(mjs) / B4_D(PhiBS) + czbssw * B4_M(mjsws) / B4_D(PhiBSWS) + czbsswg * B4_M(mjswgs) / B4_D(PhiBSWGS));
            N.qbs = vbs_jct * (T0 + 0.5 * T1);
            N.capbs = T0 + T1;
        }
        if (vbd_jct == 0.0) {
            N.qbd = 0.0;
            N.capbd = czbd + czbdsw + czbdswg;
        } else if (vbd_jct < 0.0) {
            N.qbd = 0.0;
            N.capbd = 0.0;
            if (czbd > 0.0) side(vbd_jct, czbd, B4_D(PhiBD), B4_M(mjd), N.qbd, N.capbd);
            if (czbdsw > 0.0) side(vbd_jct, czbdsw, B4_D(PhiBSWD), B4_M(mjswd), N.qbd, N.capbd);
            if (czbdswg > 0.0) side(vbd_jct, czbdswg, B4_D(PhiBSWGD), B4_M(mjswgd), N.qbd, N.capbd);
        } else {
            T0 = czbd + czbdsw + czbdswg;
            T1 = vbd_jct * (czbd * B4_M(mjd) / B4_D(PhiBD) + czbdsw * B4_M(mjswd) / B4_D(PhiBSWD) + czbdswg * B4_M(mjswgd) / B4_D(PhiBSWGD));
            N.qbd = vbd_jct * (T0 + 0.5 * T1);
            N.capbd = T0 + T1;
        }
    }
    N.vds = vds;
    N.vgs = vgs;
    N.vbs = vbs;
    N.vbd = vbd;
    N.vgd = vgd;
    N.vges = vges;
    N.vgms = vgms;
    N.vdbs = vdbs;
    N.vdbd = vdbd;
    N.vsbs = vsbs;
    N.vses = vses;
    N.vdes = vdes;
    N.qdef = p * 0.0;  // no charge node

    // overlap capacitances, capmod != 0 (:3646-3678)
    double cgdo, qgdo, cgso, qgso;
    {
        const double vgdx = vgd, vgsx = vgs;
        T0 = vgdx + DELTA_1;
        T1 = sqrt(T0 * T0 + 4.0 * DELTA_1);
        T2 = 0.5 * (T0 - T1);
        T3 = B4_S(weffCV) * B4_S(cgdl);
        T4 = sqrt(1.0 - 4.0 * T2 / B4_S(ckappad));
        cgdo = B4_S(cgdo) + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
        qgdo = (B4_S(cgdo) + T3) * vgdx - T3 * (T2 + 0.5 * B4_S(ckappad) * (T4 - 1.0));
        T0 = vgsx + DELTA_1;
        T1 = sqrt(T0 * T0 + 4.0 * DELTA_1);
        T2 = 0.5 * (T0 - T1);
        T3 = B4_S(weffCV) * B4_S(cgsl);
        T4 = sqrt(1.0 - 4.0 * T2 / B4_S(ckappas));
        cgso = B4_S(cgso) + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
        qgso = (B4_S(cgso) + T3) * vgsx - T3 * (T2 + 0.5 * B4_S(ckappas) * (T4 - 1.0));
    }
    if (B4_I(nf) != 1.0) {
        cgdo *= B4_I(nf);
        cgso *= B4_I(nf);
        qgdo *= B4_I(nf);
        qgso *= B4_I(nf);
    }
    N.cgdo = cgdo;
    N.qgdo = qgdo;
    N.cgso = cgso;
    N.qgso = qgso;
    N.vgb = vgb;
    N.vgmb = vgmb;
    N.vbs_jct = vbs_jct;
    N.vbd_jct = vbd_jct;
    // charge edits (:3680-3747): in forward mode the reference moves only the gate-drain overlap charge
    if (N.mode > 0) {
        qdrn -= qgdo;
    } else {
        qsrc = qdrn - qgso;
        const double qgb = B4_S(cgbo) * vgb;
        qgate += qgdo + qgso + qgb;
        qbulk -= qgb;
        qdrn = -(qgate + qbulk + qsrc);
    }
    N.qg = qgate;
    N.qd = qdrn - N.qbd;
    N.qs = qsrc - N.qbs;
    N.qb = qbulk;
}

// ---- charge companions (tran.rs::tran_op, trnqsmod/rgatemod/rbodymod 0) ------------------------------
struct B4Tran {
    double gcggb, gcgdb, gcgsb, gcgbb, gcdgb, gcddb, gcdsb, gcdbb, gcsgb, gcsdb, gcssb, gcsbb;
    double gcbgb, gcbdb, gcbsb, gcbbb;
    double ceqqg, ceqqd, ceqqb;
    double dxpart, sxpart;
};
// `O` is the committed operating point (its charges and their currents); the charge currents of `N` are
// filled in (cqb, cqg, cqd: Backward Euler, i = dq / dt, analysis.rs:420-428).
template <class PV>
__host__ __device__ inline void tran_eval(const PV& P, const B4Op& O, B4Op& N, double dt, B4Tran& T) {
    const double ag0 = 1.0 / dt;
    const double cgdo = N.cgdo, cgso = N.cgso;
    T = B4Tran{};
    if (N.mode > 0) {
        T.gcggb = (N.cggb + cgdo + cgso + B4_S(cgbo)) * ag0;
        T.gcgdb = (N.cgdb - cgdo) * ag0;
        T.gcgsb = (N.cgsb - cgso) * ag0;
        T.gcgbb = -(T.gcggb + T.gcgdb + T.gcgsb);
        T.gcdgb = (N.cdgb - cgdo) * ag0;
        T.gcsgb = -(N.cggb + N.cbgb + N.cdgb + cgso) * ag0;
        T.gcbgb = (N.cbgb - B4_S(cgbo)) * ag0;
        T.gcddb = (N.cddb + N.capbd + cgdo) * ag0;
        T.gcdsb = N.cdsb * ag0;
        T.gcsdb = -(N.cgdb + N.cbdb + N.cddb) * ag0;
        T.gcssb = (N.capbs + cgso - (N.cgsb + N.cbsb + N.cdsb)) * ag0;
        T.gcdbb = -(T.gcdgb + T.gcddb + T.gcdsb + 0.0);
        T.gcsbb = -(T.gcsgb + T.gcsdb + T.gcssb + 0.0);
        T.gcbdb = (N.cbdb - N.capbd) * ag0;
        T.gcbsb = (N.cbsb - N.capbs) * ag0;
        T.gcbbb = -(T.gcbdb + T.gcbgb + T.gcbsb + 0.0);
        T.sxpart = 0.6;
        T.dxpart = 0.4;
    } else {
        T.gcggb = (N.cggb + cgdo + cgso + B4_S(cgbo)) * ag0;
        T.gcgdb = (N.cgsb - cgdo) * ag0;
        T.gcgsb = (N.cgdb - cgso) * ag0;
        T.gcgbb = -(T.gcggb + T.gcgdb + T.gcgsb);
        T.gcdgb = -(N.cggb + N.cbgb + N.cdgb + cgdo) * ag0;
        T.gcsgb = (N.cdgb - cgso) * ag0;
        T.gcbgb = (N.cbgb - B4_S(cgbo)) * ag0;
        T.gcddb = (N.capbd + cgdo - (N.cgsb + N.cbsb + N.cdsb)) * ag0;
        T.gcdsb = -(N.cgdb + N.cbdb + N.cddb) * ag0;
        T.gcsdb = N.cdsb * ag0;
        T.gcssb = (N.cddb + N.capbs + cgso) * ag0;
        T.gcdbb = -(T.gcdgb + T.gcddb + T.gcdsb + 0.0);
        T.gcsbb = -(T.gcsgb + T.gcsdb + T.gcssb + 0.0);
        T.gcbdb = (N.cbsb - N.capbd) * ag0;
        T.gcbsb = (N.cbdb - N.capbs) * ag0;
        T.gcbbb = -(T.gcbgb + T.gcbdb + T.gcbsb + 0.0);
        T.sxpart = 0.4;
        T.dxpart = 0.6;
    }
    N.cqb = (N.qb - O.qb) / dt;
    N.cqg = (N.qg - O.qg) / dt;
    N.cqd = (N.qd - O.qd) / dt;
    const double vgb = N.vgb, vbd = N.vbd, vbs = N.vbs, vgmb = N.vgmb, vbd_jct = N.vbd_jct;
    const double gcdbdb = 0.0, gcdgmb = 0.0, gcbgmb = 0.0;
    T.ceqqg = N.cqg - T.gcggb * vgb + T.gcgdb * vbd + T.gcgsb * vbs;
    T.ceqqd = N.cqd - T.gcdgb * vgb - gcdgmb * vgmb + (T.gcddb + gcdbdb) * vbd - gcdbdb * vbd_jct + T.gcdsb * vbs;
    T.ceqqb = N.cqb - T.gcbgb * vgb - gcbgmb * vgmb + T.gcbdb * vbd + T.gcbsb * vbs;
}

// ---- stamps (stamp.rs) -----------------------------------------------------------------------------------
// The 38 matrix stamps then 4 right-hand-side stamps, in the reference's order, with internal nodes
// coincident with the terminals (rdsmod, rgatemod, rbodymod 0).  Stamp k goes to the element
// (B4_ROW[k], B4_COL[k]) of terminal roles D, G, S, B (B4_COL -1: right-hand side of row B4_ROW[k]).
enum { B4D_ = 0, B4G_ = 1, B4S_ = 2, B4B_ = 3 };
constexpr int B4_NJ = 38, B4_NB = 4, B4_NS = B4_NJ + B4_NB;
__host__ __device__ constexpr int b4_row(int k) {
    constexpr int r[B4_NS] = {1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 2, 2, 2, 2, 2, 2, 2, 3, 3, 3,
                              3, 0, 0, 0, 0, 3, 3, 3, 3, 2, 2, 2, 2, 3, 3, 3, 3, 0, 1, 3, 2};
    return r[k];
}
__host__ __device__ constexpr int b4_col(int k) {
    constexpr int c[B4_NS] = {1, 0, 2, 3, 0, 0, 1, 2, 3, 0, 0, 0, 1, 2, 2, 3, 2, 2, 0, 1, 2,
                              3, 0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 2, 3, -1, -1, -1, -1};
    return c[k];
}

// `N` is the load's operating point; `T` its charge companions (all zero in an operating point).
template <class PV, class Put>
__host__ __device__ inline void stamp(const PV& P, const B4Op& N, const B4Tran& T, Put&& put) {
    const double p = B4_M(p);
    double ceqdrn, ceqbd, ceqbs, Gm, Gmbs, FwdSum, RevSum;
    double gbbdp, gbbsp, gbdpg, gbdpdp, gbdpb, gbdpsp, gbspg, gbspdp, gbspb, gbspsp;
    if (N.mode >= 0) {
        Gm = N.gm;
        Gmbs = N.gmbs;
        FwdSum = Gm + Gmbs;
        RevSum = 0.0;
        ceqdrn = p * (N.cd - N.gds * N.vds - Gm * N.vgs - Gmbs * N.vbs);
        ceqbd = p * (N.csub + N.Igidl - (N.gbds + N.ggidld) * N.vds - (N.gbgs + N.ggidlg) * N.vgs -
                     (N.gbbs + N.ggidlb) * N.vbs);
        ceqbs = p * (N.Igisl + N.ggisls * N.vds - N.ggislg * N.vgd - N.ggislb * N.vbd);
        gbbdp = -(N.gbds);
        gbbsp = N.gbds + N.gbgs + N.gbbs;
        gbdpg = N.gbgs;
        gbdpdp = N.gbds;
        gbdpb = N.gbbs;
        gbdpsp = -(gbdpg + gbdpdp + gbdpb);
        gbspg = gbspdp = gbspb = gbspsp = 0.0;
    } else {
        Gm = -N.gm;
        Gmbs = -N.gmbs;
        FwdSum = 0.0;
        RevSum = -(Gm + Gmbs);
        ceqdrn = -p * (N.cd + N.gds * N.vds + Gm * N.vgd + Gmbs * N.vbd);
        ceqbs = p * (N.csub + N.Igisl + (N.gbds + N.ggisls) * N.vds - (N.gbgs + N.ggislg) * N.vgd -
                     (N.gbbs + N.ggislb) * N.vbd);
        ceqbd = p * (N.Igidl - N.ggidld * N.vds - N.ggidlg * N.vgs - N.ggidlb * N.vbs);
        gbbsp = -(N.gbds);
        gbbdp = N.gbds + N.gbgs + N.gbbs;
        gbdpg = gbdpsp = gbdpb = gbdpdp = 0.0;
        gbspg = N.gbgs;
        gbspsp = N.gbds;
        gbspb = N.gbbs;
        gbspdp = -(gbspg + gbspsp + gbspb);
    }
    // gate tunnelling, gate resistance and bias-dependent S/D resistance are off: their terms are 0
    const double gIstotg = 0.0, gIstotd = 0.0, gIstots = 0.0, gIstotb = 0.0, Istoteq = 0.0;
    const double gIdtotg = 0.0, gIdtotd = 0.0, gIdtots = 0.0, gIdtotb = 0.0, Idtoteq = 0.0;
    const double gIbtotg = 0.0, gIbtotd = 0.0, gIbtots = 0.0, gIbtotb = 0.0, Ibtoteq = 0.0;
    const double gIgtotg = 0.0, gIgtotd = 0.0, gIgtots = 0.0, gIgtotb = 0.0, Igtoteq = 0.0;
    const double ceqgcrg = 0.0;
    const double gstot = 0.0, gstotd = 0.0, gstotg = 0.0, gstots = 0.0, gstotb = 0.0, ceqgstot = 0.0;
    const double gdtot = 0.0, gdtotd = 0.0, gdtotg = 0.0, gdtots = 0.0, gdtotb = 0.0, ceqgdtot = 0.0;
    double ceqjs, ceqjd;
    if (p > 0.0) {
        ceqjs = N.cbs - N.gbs * N.vbs_jct;
        ceqjd = N.cbd - N.gbd * N.vbd_jct;
    } else {
        ceqjs = -(N.cbs - N.gbs * N.vbs_jct);
        ceqjd = -(N.cbd - N.gbd * N.vbd_jct);
    }
    double ceqqg = T.ceqqg, ceqqd = T.ceqqd, ceqqb = T.ceqqb;
    const double ceqqgmid = 0.0;
    double ceqgcrg2 = ceqgcrg;
    if (p < 0.0) {
        ceqqg = -ceqqg;
        ceqqd = -ceqqd;
        ceqqb = -ceqqb;
        ceqgcrg2 = -ceqgcrg2;
    }
    const double gjbd = 0.0, gjbs = 0.0;  // the reference passes the junction conductances only with rbodymod
    const double gdpr = 0.0, gspr = 0.0;
    const double ggtg = 0.0, ggtd = 0.0, ggts = 0.0, ggtb = 0.0;
    const double tmp1 = N.qdef * 0.0;  // qdef * gtau, gtau = 0 without NQS
    const double dxpart = T.dxpart, sxpart = T.sxpart;
    const double dd = 0.0;  // ddxpart_*, dsxpart_*: zero without NQS
    put(0, T.gcggb - ggtg + gIgtotg);
    put(1, T.gcgdb - ggtd + gIgtotd);
    put(2, T.gcgsb - ggts + gIgtots);
    put(3, T.gcgbb - ggtb + gIgtotb);
    put(4, gdpr + N.gds + N.gbd + tmp1 * dd - gdtotd + RevSum + T.gcddb + gbdpdp + dxpart * ggtd - gIdtotd);
    put(5, -(gdpr + gdtot));
    put(6, Gm + T.gcdgb - gdtotg + gbdpg - gIdtotg + dxpart * ggtg + tmp1 * dd);
    put(7, -(N.gds + gdtots - dxpart * ggts + gIdtots - tmp1 * dd + FwdSum - T.gcdsb - gbdpsp));
    put(8, -(gjbd + gdtotb - Gmbs - T.gcdbb - gbdpb + gIdtotb - tmp1 * dd - dxpart * ggtb));
    put(9, -(gdpr - gdtotd));
    put(10, gdpr + gdtot);
    put(11, -(N.gds + gstotd + RevSum - T.gcsdb - gbspdp - tmp1 * dd - sxpart * ggtd + gIstotd));
    put(12, T.gcsgb - Gm - gstotg + gbspg + sxpart * ggtg + tmp1 * dd - gIstotg);
    put(13, gspr + N.gds + N.gbs + tmp1 * dd - gstots + FwdSum + T.gcssb + gbspsp + sxpart * ggts - gIstots);
    put(14, -(gspr + gstot));
    put(15, -(gjbs + gstotb + Gmbs - T.gcsbb - gbspb - sxpart * ggtb - tmp1 * dd + gIstotb));
    put(16, -(gspr - gstots));
    put(17, gspr + gstot);
    put(18, T.gcbdb - gjbd + gbbdp - gIbtotd);
    put(19, T.gcbgb - N.gbgs - gIbtotg);
    put(20, T.gcbsb - gjbs + gbbsp - gIbtots);
    put(21, gjbd + gjbs + T.gcbbb - N.gbbs - gIbtotb);
    put(22, N.ggidld);
    put(23, N.ggidlg);
    put(24, -(N.ggidlg + N.ggidld + N.ggidlb));
    put(25, N.ggidlb);
    put(26, -(N.ggidld));
    put(27, -(N.ggidlg));
    put(28, (N.ggidlg + N.ggidld + N.ggidlb));
    put(29, -(N.ggidlb));
    put(30, -(N.ggisls + N.ggislg + N.ggislb));
    put(31, N.ggislg);
    put(32, N.ggisls);
    put(33, N.ggislb);
    put(34, (N.ggislg + N.ggisls + N.ggislb));
    put(35, -(N.ggislg));
    put(36, -(N.ggisls));
    put(37, -(N.ggislb));
    put(38, (ceqjd - ceqbd + ceqgdtot - ceqdrn - ceqqd + Idtoteq));
    put(39, -(ceqqg - ceqgcrg2 + Igtoteq));
    put(40, (ceqbd + ceqbs - ceqjd - ceqjs - ceqqb + Ibtoteq));
    put(41, (ceqdrn - ceqbs + ceqjs + ceqqg + ceqqb + ceqqd + ceqqgmid - ceqgstot + Istoteq));
}

}  // namespace b4
}  // namespace s21
