// Device physics shared by the HIP kernels and the host symbolic pass.
//
// Each model hands its stamps, in the reference's stamp order, to a caller-supplied put(entry, value).
// Entry e is a fixed (row-terminal, col-terminal) pair of the device, named as if source and drain were
// not swapped; a MOS device whose source and drain are swapped at this bias sends reference stamp k to
// entry sigma(k), which is where the reference's `matps[(dr, ..)]` indexing sends it (comps/mos.rs:
// 833-868, 1085-1097).  Entries therefore map statically onto matrix elements.
#pragma once
#ifndef __HIPCC_RTC__  // run-time compilation provides the HIP declarations itself
#include <hip/hip_runtime.h>
#endif

#ifndef __HIPCC_RTC__
#include <cmath>
#endif
#include <cstdint>
#include <type_traits>

namespace s21 {

// ---- Mos1 ------------------------------------------------------------------------------------
// Terminal roles, comps/mos.rs:569-577.
enum { RD = 0, RG = 1, RS = 2, RB = 3, RDP = 4, RSP = 5 };
constexpr int MOS1_NG = 22, MOS1_NB = 4, MOS1_NE = 26;

// Entry e -> (row role, col role); entries 22..25 are RHS entries of one role (col = -1).
__host__ __device__ constexpr int mos1_entry_row(int e) {
    constexpr int r[MOS1_NE] = {RDP, RSP, RDP, RSP, RDP, RSP, RG, RB, RG, RG, RG, RB, RB,
                                RB,  RDP, RSP, RD,  RDP, RD,  RS, RSP, RS, RDP, RSP, RG, RB};
    return r[e];
}
__host__ __device__ constexpr int mos1_entry_col(int e) {
    constexpr int c[MOS1_NE] = {RDP, RSP, RSP, RDP, RG, RG, RG, RB, RB, RDP, RSP, RG, RDP,
                                RSP, RB,  RB,  RDP, RD, RD, RSP, RS, RS, -1,  -1,  -1, -1};
    return c[e];
}
// Source/drain swap of the entry list (an involution).
__host__ __device__ constexpr int mos1_sigma(int e) {
    constexpr int s[MOS1_NE] = {1,  0,  3,  2,  5,  4,  6,  7,  8,  10, 9,  11, 13,
                                12, 15, 14, 19, 20, 21, 16, 17, 18, 23, 22, 24, 25};
    return s[e];
}
// Entries that only carry charge-integration terms (exactly zero for a device without capacitance
// and in the operating point), and entries that only carry the terminal resistances.
__host__ __device__ constexpr bool mos1_entry_tr_only(int e) {
    return e == 6 || e == 8 || e == 9 || e == 10 || e == 11 || e == 24;
}
__host__ __device__ constexpr bool mos1_entry_r_only(int e) { return e >= 16 && e <= 21; }

// Per-device derived parameters used by the DC load (Mos1InternalParams + Mos1Model fields,
// comps/mos.rs:300-317, 107-138).
struct Mos1Core {
    double p, vt0_t, beta, phi_t, gamma, lambda, vtherm, isat_d, isat_s, grd, grs;
};
// Gate capacitances (Meyer + overlap): only these reach the stamps.
struct Mos1Caps {
    double cox, cgs_ov, cgd_ov, cgb_ov;
};
// Committed operating point fields the next transient load reads (comps/mos.rs:757-806).
struct Mos1State {
    double vgs, vgd, vgb, cgs, cgd, cgb;
    double reversed;  // 0 or 1
};

// A sink that declares `raw_stamps` receives Mos1 stamps as stamp(k, reversed, value) with the
// reference's stamp index k, and applies the source/drain swap itself.
template <class P, class = void>
struct RawSink : std::false_type {};
template <class P>
struct RawSink<P, std::void_t<decltype(P::raw_stamps)>> : std::true_type {};

// The operating-point fields Mos1::load_ac reads (comps/mos.rs:914-968): small-signal conductances and
// the intrinsic Meyer capacitances of the last load, with the source/drain orientation.
struct Mos1AcOp {
    double gm, gds, gmbs, gbs, gbd, cgs, cgd, cgb, reversed;
};
constexpr int M1_ACOP_F = 9;

// Backward-Euler companion (analysis.rs:420-428).
struct Integ {
    double g, i, rhs;
};
__host__ __device__ inline Integ be_integ(double dq, double dq_dv, double vguess, double dt) {
    double g = dq_dv / dt;
    double i = dq / dt;
    double rhs = i - g * vguess;
    return Integ{g, i, rhs};
}
// Trapezoidal companion, the formula the reference keeps disabled (analysis.rs:429-435): ip is the charge's
// current at the previous accepted point.
__host__ __device__ inline Integ trap_integ(double dq, double dq_dv, double vguess, double dt, double ip) {
    double g = 2.0 * dq_dv / dt;
    double i = 2.0 * dq / dt - ip;
    double rhs = i - g * vguess;
    return Integ{g, i, rhs};
}
// Integration of one charge: Backward Euler, or trapezoidal when ip (the previous current) is given; the
// new current goes to *i_out when given.
__host__ __device__ inline Integ integ(double dq, double dq_dv, double vguess, double dt, const double* ip,
                                       double* i_out) {
    const Integ r = ip ? trap_integ(dq, dq_dv, vguess, dt, *ip) : be_integ(dq, dq_dv, vguess, dt);
    if (i_out) *i_out = r.i;
    return r;
}

// Mos1::op_stamp (comps/mos.rs:649-893).  Terminal voltages vd, vg, vs, vb.  Calls put(entry, value)
// for the 22 matrix stamps and 4 RHS stamps in the reference's stamp order, stamp k going to entry
// sigma(k) when source and drain are swapped.  TRAN selects the transient branch; with caps == nullptr
// the device has no capacitance and the integration terms are zero.  guess (optional) receives the new
// operating point.
// trap_ip / trap_i (optional, transient only): the previous and new currents of the two integrated gate
// charges (the gate-drain one the stamps carry in tr.gs, then the gate-bulk one) for trapezoidal
// integration; without them the integration is Backward Euler.
template <bool TRAN, class Put>
__host__ __device__ inline void mos1_load(const Mos1Core& P, const Mos1Caps* caps, const Mos1State* op, double vd_,
                                          double vg_, double vs_, double vb_, double gmin, double dt, Put&& put,
                                          Mos1State* guess, Mos1AcOp* acop = nullptr, const double* trap_ip = nullptr,
                                          double* trap_i = nullptr) {
    const double p = P.p;
    const bool reversed = p * (vd_ - vs_) < 0.0;
    const double vd = reversed ? vs_ : vd_;
    const double vs = reversed ? vd_ : vs_;
    const double vgs = p * (vg_ - vs);
    const double vgd = p * (vg_ - vd);
    const double vds = p * (vd - vs);
    const double vgb = p * (vg_ - vb_);
    const double vsb = p * (vs - vb_);
    const double vdb = p * (vd - vb_);
    const double von = vsb > 0.0 ? P.vt0_t + P.gamma * (sqrt(P.phi_t + vsb) - sqrt(P.phi_t)) : P.vt0_t;
    const double vov = vgs - von;
    const double vdsat = fmax(vov, 0.0);
    double ids = 0.0, gm = 0.0, gds = 0.0, gmbs = 0.0;
    if (vov > 0.0) {
        if (vds >= vov) {
            ids = P.beta / 2.0 * (vov * vov) * (1.0 + P.lambda * vds);
            gm = P.beta * vov * (1.0 + P.lambda * vds);
            gds = P.lambda * P.beta / 2.0 * (vov * vov);
        } else {
            ids = P.beta * (vov * vds - (vds * vds) / 2.0) * (1.0 + P.lambda * vds);
            gm = P.beta * vds * (1.0 + P.lambda * vds);
            gds = P.beta * ((vov - vds) * (1.0 + P.lambda * vds) + P.lambda * ((vov * vds) - (vds * vds) / 2.0));
        }
        // With gamma = 0 the quotient is gm * 0 / 2 whatever the root: skip the root and the division.
        if (P.phi_t + vsb > 0.0) gmbs = P.gamma == 0.0 ? gm * P.gamma / 2.0 : gm * P.gamma / 2.0 / sqrt(P.phi_t + vsb);
    }
    const double vt = P.vtherm;
    const double isat_bs = reversed ? P.isat_d : P.isat_s;
    const double isat_bd = reversed ? P.isat_s : P.isat_d;
    // A junction at zero bias (source or drain on the bulk) has exp(-0 / vt) = 1 exactly.
    const double ebs = vsb == 0.0 ? 1.0 : exp(-vsb / vt);
    const double ibs = isat_bs * (ebs - 1.0);
    const double gbs = (isat_bs / vt) * ebs + gmin;
    const double ibs_rhs = ibs + vsb * gbs;
    const double ebd = vdb == 0.0 ? 1.0 : exp(-vdb / vt);
    const double ibd = isat_bd * (ebd - 1.0);
    const double gbd = (isat_bd / vt) * ebd + gmin;
    const double ibd_rhs = ibd + vdb * gbd;

    // Charge integration (only tr.gs -- holding the gate-drain integration, mos.rs:801 -- and tr.gb
    // reach the stamps; tr.gd stays zero).
    double trgs_g = 0.0, trgs_rhs = 0.0, trgb_g = 0.0, trgb_rhs = 0.0;
    const double trgd_g = 0.0, trgd_rhs = 0.0;
    if (caps) {
        const double cox = caps->cox;
        double cgs1, cgd1, cgb1;
        if (vov <= -P.phi_t) {
            cgb1 = cox / 2.0;
            cgs1 = 0.0;
            cgd1 = 0.0;
        } else if (vov <= -P.phi_t / 2.0) {
            cgb1 = -vov * cox / (2.0 * P.phi_t);
            cgs1 = 0.0;
            cgd1 = 0.0;
        } else if (vov <= 0.0) {
            cgb1 = -vov * cox / (2.0 * P.phi_t);
            cgs1 = vov * cox / (1.5 * P.phi_t) + cox / 3.0;
            cgd1 = 0.0;
        } else if (vdsat <= vds) {
            cgs1 = cox / 3.0;
            cgd1 = 0.0;
            cgb1 = 0.0;
        } else {
            const double vddif = 2.0 * vdsat - vds;
            const double vddif1 = vdsat - vds;
            const double vddif2 = vddif * vddif;
            cgd1 = cox * (1.0 - vdsat * vdsat / vddif2) / 3.0;
            cgs1 = cox * (1.0 - vddif1 * vddif1 / vddif2) / 3.0;
            cgb1 = 0.0;
        }
        if (TRAN) {
            const bool same = reversed == (op->reversed != 0.0);
            const double cgd = cgd1 + caps->cgd_ov + (same ? op->cgd : op->cgs);
            const double cgb = cgb1 + caps->cgb_ov + op->cgb;
            const double dqgd = same ? (vgd - op->vgd) * cgd : (vgd - op->vgs) * cgd;
            const Integ a = integ(dqgd, cgd, vgd, dt, trap_ip, trap_i);
            trgs_g = a.g;
            trgs_rhs = a.rhs;
            const Integ b = integ((vgb - op->vgb) * cgb, cgb, vgb, dt, trap_ip ? trap_ip + 1 : nullptr,
                                  trap_i ? trap_i + 1 : nullptr);
            trgb_g = b.g;
            trgb_rhs = b.rhs;
        }
        if (acop) {
            acop->cgs = cgs1;
            acop->cgd = cgd1;
            acop->cgb = cgb1;
        }
        if (guess) {
            guess->vgs = vgs;
            guess->vgd = vgd;
            guess->vgb = vgb;
            guess->cgs = cgs1;
            guess->cgd = cgd1;
            guess->cgb = cgb1;
            guess->reversed = reversed ? 1.0 : 0.0;
        }
    }
    const double irhs = ids - gm * vgs - gds * vds;
    const double grd = P.grd, grs = P.grs;
    if (acop) {
        acop->gm = gm;
        acop->gds = gds;
        acop->gmbs = gmbs;
        acop->gbs = gbs;
        acop->gbd = gbd;
        acop->reversed = reversed ? 1.0 : 0.0;
        if (!caps) acop->cgs = acop->cgd = acop->cgb = 0.0;
    }
    auto E = [&](int k, double v) {
        if constexpr (RawSink<std::remove_reference_t<Put>>::value)
            put.stamp(k, reversed, v);
        else
            put(reversed ? mos1_sigma(k) : k, v);
    };
    E(0, gds + grd + gbd + trgd_g);
    E(1, gm + gds + grs + gbs + gmbs + trgs_g);
    E(2, -gm - gds - gmbs);
    E(3, -gds);
    E(4, gm - trgd_g);
    E(5, -gm - trgs_g);
    E(6, (trgd_g + trgs_g + trgb_g));
    E(7, (gbd + gbs + trgb_g));
    E(8, -trgb_g);
    E(9, -trgd_g);
    E(10, -trgs_g);
    E(11, -trgb_g);
    E(12, -gbd);
    E(13, -gbs);
    E(14, -gbd + gmbs);
    E(15, -gbs - gmbs);
    E(16, -grd);
    E(17, -grd);
    E(18, grd);
    E(19, -grs);
    E(20, -grs);
    E(21, grs);
    E(22, p * (-irhs + ibd_rhs + trgd_rhs));
    E(23, p * (irhs + ibs_rhs + trgs_rhs));
    E(24, -p * (trgs_rhs + trgb_rhs + trgd_rhs));
    E(25, -p * (ibd_rhs + ibs_rhs - trgb_rhs));
}

// Mos1::load_ac (comps/mos.rs:914-968): 23 complex stamps from the committed operating point, sent to
// the DC entry list (stamp k of the reference list goes to entry ac_entry(k), sigma-swapped when the
// operating point is reversed).  The (g, dr) admittance is stamped twice (mos.rs:951 and :954), so
// entry 9 receives two stamps.  put(entry, re, im), or stamp(k-entry, reversed, re, im) for a raw sink.
template <class Put>
__host__ __device__ inline void mos1_load_ac(const Mos1AcOp& op, double grd, double grs, double omega, Put&& put) {
    const double gcgs = omega * op.cgs, gcgd = omega * op.cgd, gcgb = omega * op.cgb;
    const double gm = op.gm, gds = op.gds, gmbs = op.gmbs, gbs = op.gbs, gbd = op.gbd;
    const bool rev = op.reversed != 0.0;
    auto E = [&](int e, double re, double im) {
        if constexpr (RawSink<std::remove_reference_t<Put>>::value)
            put.stamp(e, rev, re, im);
        else
            put(rev ? mos1_sigma(e) : e, re, im);
    };
    E(0, gds + grd + gbd, gcgd);
    E(1, gm + gds + grs + gbs + gmbs, gcgs);
    E(2, -gm - gds - gmbs, 0.0);
    E(3, -gds, 0.0);
    E(4, gm, -gcgd);
    E(5, -gm, -gcgs);
    E(6, 0.0, gcgd + gcgs + gcgb);
    E(7, gbd + gbs, gcgb);
    E(8, 0.0, -gcgb);
    E(9, 0.0, -gcgd);
    E(10, 0.0, -gcgs);
    E(11, 0.0, -gcgb);
    E(9, 0.0, -gcgd);
    E(12, -gbd, 0.0);
    E(13, -gbs, 0.0);
    E(14, -gbd + gmbs, 0.0);
    E(15, -gbs - gmbs, 0.0);
    E(16, -grd, 0.0);
    E(17, -grd, 0.0);
    E(18, grd, 0.0);
    E(19, -grs, 0.0);
    E(20, -grs, 0.0);
    E(21, grs, 0.0);
}

// ---- Mos0 (comps/mos.rs:1051-1099) -------------------------------------------------------------
constexpr int MOS0_NE = 8;  // 6 matrix entries + 2 RHS
__host__ __device__ constexpr int mos0_entry_row(int e) {
    constexpr int r[MOS0_NE] = {RD, RS, RD, RS, RD, RS, RD, RS};
    return r[e];
}
__host__ __device__ constexpr int mos0_entry_col(int e) {
    constexpr int c[MOS0_NE] = {RD, RS, RS, RD, RG, RG, -1, -1};
    return c[e];
}
__host__ __device__ constexpr int mos0_sigma(int e) {
    constexpr int s[MOS0_NE] = {1, 0, 3, 2, 5, 4, 7, 6};
    return s[e];
}
template <class Put>
__host__ __device__ inline void mos0_load(double p, double vd, double vg, double vs, double gmin, Put&& put) {
    const double vth = 0.25, beta = 50e-3, lam = 3e-3;
    const double vds1 = p * (vd - vs);
    const bool reversed = vds1 < 0.0;
    const double vgs = reversed ? p * (vg - vd) : p * (vg - vs);
    const double vds = reversed ? -vds1 : vds1;
    const double vov = vgs - vth;
    double ids = 0.0, gm = 0.0, gds = 0.0;
    if (vov > 0.0) {
        if (vds >= vov) {
            ids = beta / 2.0 * (vov * vov) * (1.0 + lam * vds);
            gm = beta * vov * (1.0 + lam * vds);
            gds = lam * beta / 2.0 * (vov * vov);
        } else {
            ids = beta * (vov * vds - (vds * vds) / 2.0) * (1.0 + lam * vds);
            gm = beta * vds * (1.0 + lam * vds);
            gds = beta * ((vov - vds) * (1.0 + lam * vds) + lam * ((vov * vds) - (vds * vds) / 2.0));
        }
    }
    const double irhs = ids - gm * vgs - gds * vds;
    auto E = [&](int k, double v) { put(reversed ? mos0_sigma(k) : k, v); };
    E(0, gds + gmin);
    E(1, (gm + gds + gmin));
    E(2, -(gm + gds + gmin));
    E(3, -gds - gmin);
    E(4, gm);
    E(5, -gm);
    E(6, -p * irhs);
    E(7, p * irhs);
}

// ---- Diode (comps/diode.rs:228-354) -------------------------------------------------------------
enum { DP_ = 0, DN_ = 1, DR_ = 2 };  // terminals p, n, internal r
constexpr int DIODE_NE = 9;
__host__ __device__ constexpr int diode_entry_row(int e) {
    constexpr int r[DIODE_NE] = {DN_, DR_, DN_, DR_, DP_, DP_, DR_, DR_, DN_};
    return r[e];
}
__host__ __device__ constexpr int diode_entry_col(int e) {
    constexpr int c[DIODE_NE] = {DN_, DN_, DR_, DR_, DP_, DR_, DP_, -1, -1};
    return c[e];
}
struct DiodeP {
    double isat, vte, vcrit, gspr, bv, has_bv, cz, cz2, dep, f1, f3, tt, vj, m;
};
struct DiodeGuess {
    double vd, charge;
};
__host__ __device__ inline double diode_limit(const DiodeP& P, double vnew, double vold) {
    if (vnew <= P.vcrit || fabs(vnew - vold) <= 2.0 * P.vte) return vnew;
    if (vold > 0.0) {
        const double arg = 1.0 + (vnew - vold) / P.vte;
        if (arg > 0.0) return vold + P.vte * log(arg);
        return P.vcrit;
    }
    return P.vte * log(vnew / P.vte);
}
// trap_ip / trap_i (optional): the junction charge's previous and new current for trapezoidal integration.
template <bool TRAN, class Put>
__host__ __device__ inline void diode_load(const DiodeP& P, double vr, double vn, double gmin, double dt,
                                           double prev_guess_vd, double op_charge, Put&& put, DiodeGuess* guess,
                                           const double* trap_ip = nullptr, double* trap_i = nullptr) {
    double vd = vr - vn;
    const bool has_bv = P.has_bv != 0.0;
    if (has_bv && vd < fmin(10.0 * P.vte - P.bv, 0.0)) {
        const double vtemp = diode_limit(P, -P.bv, P.bv - prev_guess_vd);
        vd = vtemp - P.bv;
    } else {
        vd = diode_limit(P, vd, prev_guess_vd);
    }
    double id, gd;
    if (!has_bv || vd >= -P.bv) {
        const double e = exp(vd / P.vte);
        id = P.isat * (e - 1.0) + gmin * vd;
        gd = P.isat * e / P.vte + gmin;
    } else {
        const double e = exp((vd - P.bv) / P.vte);
        id = -P.isat * e + gmin * vd;
        gd = P.isat * e / P.vte + gmin;
    }
    double qd, cd;
    if (vd < P.dep) {
        const double a = 1.0 - vd / P.vj;
        const double s = -P.m * log(a);
        qd = P.tt * P.vj * P.cz * (1.0 - a * s) / (1.0 - P.m);
        cd = P.tt * gd + P.cz * s;
    } else {
        qd = P.tt * id + P.cz * P.f1 + P.cz2 * (P.f3 * (vd - P.dep) + P.m / 2.0 / P.vj * (vd * vd - P.dep * P.dep));
        cd = P.tt + P.cz2 * P.f3 + P.m * vd / P.vj;
    }
    if (TRAN) {
        const Integ c = integ(qd - op_charge, cd, vd, dt, trap_ip, trap_i);
        id += c.i;
        gd += c.g;
    }
    if (guess) {
        guess->vd = vd;
        guess->charge = qd;
    }
    const double irhs = id - vd * gd;
    put(0, gd);
    put(1, -gd);
    put(2, -gd);
    put(3, gd + P.gspr);
    put(4, P.gspr);
    put(5, -P.gspr);
    put(6, -P.gspr);
    put(7, -irhs);
    put(8, irhs);
}

// ---- linear elements (comps/mod.rs:95-346) -------------------------------------------------------
// Resistor entries: (p,p) (n,n) (p,n) (n,p).  Capacitor: the same four plus RHS p, n.
// Vsrc: (p,i) (i,p) (n,i) (i,n) plus RHS i.  Isrc: RHS p, n.
constexpr int R_NE = 4, C_NE = 6, V_NE = 5, I_NE = 2;

// Capacitor transient load: the companion (g, rhs) of the BE integration, or of the trapezoidal one when
// trap_ip (the previous current) is given; the new current goes to *trap_i.
template <class Put>
__host__ __device__ inline void cap_load_tran(double c, double vd, double op_q, double dt, Put&& put,
                                              const double* trap_ip = nullptr, double* trap_i = nullptr) {
    const double q = c * vd;
    const Integ a = integ(q - op_q, c, vd, dt, trap_ip, trap_i);
    put(0, a.g);
    put(1, a.g);
    put(2, -a.g);
    put(3, -a.g);
    put(4, -a.rhs);
    put(5, a.rhs);
}

}  // namespace s21
