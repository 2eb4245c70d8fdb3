// Host derivation of per-device parameters from the proto-level model / instance parameter sets.
//   Mos1Model::resolve               comps/mos.rs:140-237
//   Mos1InstanceParams::resolve      comps/mos.rs:271-290
//   Mos1InternalParams::derive       comps/mos.rs:320-476
//   DiodeModel::from                 comps/diode.rs:52-70
//   DiodeIntParams::derive           comps/diode.rs:146-212
// The arithmetic is kept in the reference's order so the derived values equal the reference's.
#include <cmath>

#include "internal.h"

namespace s21 {
namespace {

constexpr double KB = 1.3806226e-23;
constexpr double Q = 1.6021918e-19;
constexpr double KB_OVER_Q = KB / Q;
constexpr double KELVIN_TO_C = 273.15;
constexpr double TEMP_REF = KELVIN_TO_C + 27.0;
constexpr double SIO2_PERMITTIVITY = 3.9 * 8.854214871e-12;
constexpr double SQRT2 = 1.4142135624;

inline bool has(const s21_params& p, int f) { return (p.set >> f) & 1ull; }
inline double get(const s21_params& p, int f, double dflt) { return has(p, f) ? p.v[f] : dflt; }

// proto field numbers, protos/mos.proto:43-76
enum {
    M_TYPE = 2, M_VT0, M_KP, M_GAMMA, M_PHI, M_LAMBDA, M_RD, M_RS, M_CBD, M_CBS, M_IS, M_PB, M_CGSO, M_CGDO, M_CGBO,
    M_RSH, M_CJ, M_MJ, M_CJSW, M_MJSW, M_JS, M_TOX, M_LD, M_U0, M_FC, M_NSUB, M_TPG, M_NSS, M_TNOM, M_KF, M_AF
};
// protos/mos.proto:22-41
enum { I_M = 2, I_L, I_W, I_AD, I_AS, I_PD, I_PS, I_NRD, I_NRS, I_TEMP };

struct Mos1ModelVals {
    double p, vt0, kp, gamma, cox_per_area, phi, lambda, cbd, cbs, is, pb, cj, mj, cjsw, mjsw, js, ld, fc, tnom;
    double cgso, cgdo, cgbo;
    bool has_rd, has_rs, has_rsh;
    double rd, rs, rsh;
};

Mos1ModelVals resolve_mos1(const s21_params& s) {
    Mos1ModelVals m{};
    m.p = (has(s, M_TYPE) && s.v[M_TYPE] == 1.0) ? -1.0 : 1.0;
    const double tnom = has(s, M_TNOM) ? s.v[M_TNOM] + KELVIN_TO_C : TEMP_REF;
    const double vtnom = tnom * KB_OVER_Q;
    const double egfet1 = 1.16 - (7.02e-4 * (tnom * tnom)) / (tnom + 1108.0);
    double cox_per_area = 0.0;
    double vt0 = get(s, M_VT0, 0.0), kp = get(s, M_KP, 2.0e-5), phi = get(s, M_PHI, 0.6), gamma = get(s, M_GAMMA, 0.0);
    if (has(s, M_TOX)) {
        cox_per_area = SIO2_PERMITTIVITY / s.v[M_TOX];
        if (!has(s, M_KP)) kp = get(s, M_U0, 600.0) * cox_per_area * 1e-4;
        if (has(s, M_NSUB)) {
            const double nsub = s.v[M_NSUB];
            if (nsub * 1e6 <= 1.45e16) fail(S21_ERR, "Invalid Mos1 Substrate Doping nsub < ni (1.45e16)");
            if (!has(s, M_PHI)) phi = fmax(2.0 * vtnom * log(nsub * 1e6 / 1.45e16), 0.1);
            const double fermis = m.p * 0.5 * phi;
            double wkfng = 3.2;
            double gate_type = 1.0;
            if (has(s, M_TPG)) {
                gate_type = s.v[M_TPG];
                if (gate_type > 1 || gate_type < -1) fail(S21_ERR, "Invalid Mos1 tps");
            }
            if (gate_type != 0.0) {
                const double fermig = m.p * gate_type * 0.5 * egfet1;
                wkfng = 3.25 + 0.5 * egfet1 - fermig;
            }
            if (!has(s, M_GAMMA)) gamma = sqrt(2.0 * 11.70 * 8.854214871e-12 * Q * nsub * 1e6) / cox_per_area;
            if (!has(s, M_VT0)) {
                const double nss = get(s, M_NSS, 0.0);
                const double wkfngs = wkfng - (3.25 + 0.5 * egfet1 + fermis);
                const double vfb = wkfngs - nss * 1e4 * Q / cox_per_area;
                vt0 = vfb + m.p * (gamma * sqrt(phi) + phi);
            }
        }
    }
    m.vt0 = vt0;
    m.kp = kp;
    m.gamma = gamma;
    m.phi = phi;
    m.cox_per_area = cox_per_area;
    m.tnom = tnom;
    m.lambda = get(s, M_LAMBDA, 0.0);
    m.pb = get(s, M_PB, 0.8);
    m.cbd = get(s, M_CBD, 0.0);
    m.cbs = get(s, M_CBS, 0.0);
    m.cgso = get(s, M_CGSO, 0.0);
    m.cgdo = get(s, M_CGDO, 0.0);
    m.cgbo = get(s, M_CGBO, 0.0);
    m.cj = get(s, M_CJ, 0.0);
    m.cjsw = get(s, M_CJSW, 0.0);
    m.mj = get(s, M_MJ, 0.5);
    m.mjsw = get(s, M_MJSW, 0.5);
    m.is = get(s, M_IS, 1.0e-14);
    m.js = get(s, M_JS, 1.0e-8);
    m.ld = get(s, M_LD, 0.0);
    m.fc = get(s, M_FC, 0.5);
    m.has_rd = has(s, M_RD);
    m.rd = get(s, M_RD, 0.0);
    m.has_rs = has(s, M_RS);
    m.rs = get(s, M_RS, 0.0);
    m.has_rsh = has(s, M_RSH);
    m.rsh = get(s, M_RSH, 0.0);
    return m;
}

}  // namespace

Mos1Derived derive_mos1(const s21_params& model_p, const s21_params& inst_p, const s21_options& o) {
    const Mos1ModelVals model = resolve_mos1(model_p);
    if (has(inst_p, I_TEMP)) fail(S21_ERR, "Mos1 Instance Temperatures Are Not Supported");
    const double l = get(inst_p, I_L, 1e-6), w = get(inst_p, I_W, 1e-6), a_d = get(inst_p, I_AD, 1e-12),
                 a_s = get(inst_p, I_AS, 1e-12), nrd = get(inst_p, I_NRD, 1.0), nrs = get(inst_p, I_NRS, 1.0);
    const double temp = o.temp;
    const double fact1 = model.tnom / TEMP_REF;
    const double vtnom = model.tnom * KB_OVER_Q;
    const double kt1 = KB * model.tnom;
    const double egfet1 = 1.16 - (7.02e-4 * (model.tnom * model.tnom)) / (model.tnom + 1108.0);
    const double arg1 = -egfet1 / 2.0 / kt1 + 1.1150877 / (KB * 2.0 * TEMP_REF);
    const double pbfact1 = -2.0 * vtnom * (1.5 * log(fact1) + Q * arg1);
    const double kt = temp * KB;
    const double vtherm = temp * KB_OVER_Q;
    const double temp_ratio = temp / model.tnom;
    const double fact2 = temp / TEMP_REF;
    const double egfet = 1.16 - (7.02e-4 * (temp * temp)) / (temp + 1108.0);
    const double arg = -egfet / 2.0 / kt + 1.1150877 / (KB * 2.0 * TEMP_REF);
    const double pbfact = -2.0 * vtherm * (1.5 * log(fact2) + Q * arg);
    const double leff = l - 2.0 * model.ld;
    if (leff < 0.0) fail(S21_ERR, "Mos1 Effective Length < 0");
    const double phio = (model.phi - pbfact1) / fact1;
    const double phi_t = fact2 * phio + pbfact;
    const double vbi_t = model.vt0 - model.p * (model.gamma * sqrt(model.phi)) + 0.5 * (egfet1 - egfet) +
                         model.p * 0.5 * (phi_t - model.phi);
    const double vt0_t = vbi_t + model.p * model.gamma * sqrt(phi_t);
    const double isat_t = model.is * exp(-egfet / vtherm + egfet1 / vtnom);
    const double jsat_t = model.js * exp(-egfet / vtherm + egfet1 / vtnom);
    const bool use_default_isat = jsat_t == 0.0 || a_d == 0.0 || a_s == 0.0;
    // Junction saturation currents; the depletion-charge terms of the junctions never reach a stamp
    // (MosJunction::qc's results are only stored, comps/mos.rs:770-771, 885-889).
    const double isat_d = use_default_isat ? isat_t : jsat_t * a_d;
    const double isat_s = use_default_isat ? isat_t : jsat_t * a_s;
    double grs = 0.0, grd = 0.0;
    if (model.has_rs)
        grs = model.rs <= 0.0 ? 0.0 : 1.0 / model.rs;
    else if (model.has_rsh)
        grs = model.rsh <= 0.0 ? 0.0 : 1.0 / model.rsh / nrs;
    if (model.has_rd)
        grd = model.rd <= 0.0 ? 0.0 : 1.0 / model.rd;
    else if (model.has_rsh)
        grd = model.rsh <= 0.0 ? 0.0 : 1.0 / model.rsh / nrd;
    const double kp_t = model.kp / temp_ratio * sqrt(temp_ratio);

    Mos1Derived d{};
    d.core.p = model.p;
    d.core.vt0_t = vt0_t;
    d.core.beta = kp_t * w / leff;
    d.core.phi_t = phi_t;
    d.core.gamma = model.gamma;
    d.core.lambda = model.lambda;
    d.core.vtherm = vtherm;
    d.core.isat_d = isat_d;
    d.core.isat_s = isat_s;
    d.core.grd = grd;
    d.core.grs = grs;
    d.caps.cox = model.cox_per_area * leff * w;
    d.caps.cgs_ov = w * model.cgso;
    d.caps.cgd_ov = w * model.cgdo;
    d.caps.cgb_ov = leff * model.cgbo;
    d.has_caps = d.caps.cox != 0.0 || d.caps.cgs_ov != 0.0 || d.caps.cgd_ov != 0.0 || d.caps.cgb_ov != 0.0;
    d.has_rd = model.has_rd || model.has_rsh;
    d.has_rs = model.has_rs || model.has_rsh;
    return d;
}

// protos/spice21.proto:44-68
enum {
    D_TNOM = 2, D_IS, D_N, D_TT, D_VJ, D_M, D_EG, D_XTI, D_KF, D_AF, D_FC, D_BV, D_IBV, D_RS, D_CJ0
};
enum { DI_AREA = 4, DI_TEMP = 5 };

bool diode_model_has_rs(const s21_params& m) { return get(m, D_RS, 0.0) != 0.0; }

DiodeP derive_diode(const s21_params& mp, const s21_params& ip, const s21_options& o) {
    const double tnom = get(mp, D_TNOM, 300.15), is = get(mp, D_IS, 1e-14), n = get(mp, D_N, 1.0),
                 tt = get(mp, D_TT, 0.0), vj = get(mp, D_VJ, 1.0), m = get(mp, D_M, 0.5), eg = get(mp, D_EG, 1.11),
                 xti = get(mp, D_XTI, 3.0), fc = get(mp, D_FC, 0.5), bv0 = get(mp, D_BV, 0.0),
                 ibv = get(mp, D_IBV, 1e-3), rs = get(mp, D_RS, 0.0), cj0 = get(mp, D_CJ0, 0.0);
    const double temp = has(ip, DI_TEMP) ? ip.v[DI_TEMP] : o.temp;
    const double area = has(ip, DI_AREA) ? ip.v[DI_AREA] : 1.0;
    const double gs = rs != 0.0 ? 1.0 / rs : 0.0;
    const double vt = KB_OVER_Q * temp;
    const double vtnom = KB_OVER_Q * tnom;
    const double fact2 = temp / TEMP_REF;
    const double egfet = 1.16 - (7.02e-4 * temp * temp) / (temp + 1108.0);
    const double arg = -egfet / (2.0 * KB * temp) + 1.1150877 / (2.0 * KB * TEMP_REF);
    const double pbfact = -2.0 * vt * (1.5 * log(fact2) + Q * arg);
    const double egfet1 = 1.16 - (7.02e-4 * tnom) / (tnom + 1108.0);  // tnom, not tnom^2: diode.rs:162
    const double arg1 = -egfet1 / (KB * 2.0 * tnom) + 1.1150877 / (2.0 * KB * TEMP_REF);
    const double fact1 = tnom / TEMP_REF;
    const double pbfact1 = -2.0 * vtnom * (1.5 * log(fact1) + Q * arg1);
    const double pbo = (vj - pbfact1) / fact1;
    const double vjunc = pbfact + fact2 * pbo;
    const double isat = is * exp(((temp / tnom) - 1.0) * eg / n * vt + xti / n * log(temp / tnom));
    const double xfc = 1.0 - log(fc);
    const double f1 = vjunc * (1.0 - exp(1.0 - m * xfc)) / (1.0 - m);
    const double vte = n * vt;
    const double vcrit = vte * (vte / sqrt(2.0) / isat);  // no log: diode.rs:179
    double bv = bv0;
    if (bv0 != 0.0)
        for (int i = 0; i < 25; i++) bv = bv0 - vt * log(ibv / isat + 1.0 - bv / vt);
    const double f2 = exp(xfc * (1.0 + m));
    const double f3 = 1.0 - fc * (1.0 + m);
    const double cz = cj0 * area;
    DiodeP P{};
    P.isat = isat;
    P.vte = vte;
    P.vcrit = vcrit;
    P.gspr = gs * area;
    P.bv = bv;
    P.has_bv = bv0 != 0.0 ? 1.0 : 0.0;
    P.cz = cz;
    P.cz2 = cz / f2;
    P.dep = fc * vj;
    P.f1 = f1;
    P.f3 = f3;
    P.tt = tt;
    P.vj = vj;
    P.m = m;
    return P;
}

}  // namespace s21
