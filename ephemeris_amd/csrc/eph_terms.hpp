// Physics terms of one force evaluation, shared by the gfx950 kernels and the
// host (CPU) back end: one source, compiled for both sides.  Each function
// restates a macro or function of the reference expression by expression
// (paths relative to the reference checkout); with FP contraction off every
// operator is one IEEE operation in the written order, so both sides give
// the reference's bits.
#pragma once
#include "../../include/ephemeris_hip.h"
#include "eph_math.hpp"

namespace eph {

// partner data of one body as the main force reads it (also the record
// format of a field system for test particles); 32 doubles.  phi and n are
// phi / c^2 and naccel / c^2 (PREPARE_RELATIVITY, mass.impl:76-80).
struct JMain {
    double rxh, rxl, ryh, ryl, rzh, rzl, GM, sR2;
    double phi, beta2, bx, by, bz, nx, ny, nz;
    double C[9], R2, GLx, GLy, GLz, lum, pad0, pad1;
};

// own-body data of the main force
struct IMain {
    ddv3 r;
    double sR2, phi, beta2, R2, GM, rR2G4c;
    v3 beta, GL;
    m3 C;
};

struct PairGeo { v3 r, b, G; double rr2, rr, rr3, tp_dg; };

// RELATIVITY (mass.impl:82-97)
EPH_HD PairGeo relativity(const IMain &I, const JMain &T, v3 &gaccel) {
    PairGeo p;
    const ddv3 rj{dd{T.rxh, T.rxl}, dd{T.ryh, T.ryl}, dd{T.rzh, T.rzl}};
    p.r = diff(rj, I.r);
    p.rr2 = 1 / smax(dot(p.r, p.r), I.sR2 + T.sR2);
    p.rr = sqrt(p.rr2);
    p.rr3 = p.rr * p.rr2;
    const double tp_dphi = p.rr * T.GM;
    p.tp_dg = p.rr2 * tp_dphi;
    const v3 bj{T.bx, T.by, T.bz}, nj{T.nx, T.ny, T.nz};
    const double rbj = dot(p.r, bj);
    double tp_rbjrr2 = rbj * p.rr;
    tp_rbjrr2 *= tp_rbjrr2;
    const double delta1 = 4 * I.phi + T.phi + I.beta2 + 2 * T.beta2
                        - 4 * dot(I.beta, bj) + (dot(p.r, nj) - 3 * tp_rbjrr2) / 2;
    p.b = bj - I.beta;
    p.G = ((7 * tp_dphi / 2) * nj + (p.tp_dg * delta1) * p.r) + (p.tp_dg * (rbj - dot(p.r, p.b) * 4)) * p.b;
    gaccel += p.G;
    return p;
}

// the bracket of ROTATIONAL_TIDAL_DEFORMATION (mass.impl:111-125)
EPH_HD v3 quad_field(const m3 &C, const PairGeo &p) {
    const v3 Cr = C * p.r;
    return (dot(p.r, Cr) * 5 * p.rr2) * p.r - (Cr + Cr);
}

// LENSE_THIRRING (mass.impl:99-106) and RADIATION_PRESSURE (:108-109)
EPH_HD void lt_rad(const IMain &I, const JMain &T, const PairGeo &p, v3 &da, v3 &dq) {
    const double rcr32 = 2 * T.GM / kC * p.rr3;
    const v3 GLb = cross(I.GL, p.b);
    da -= rcr32 * (GLb - (3 * dot(GLb, p.r) * p.rr2) * p.r);
    dq -= rcr32 * cross(I.GL, cross(p.r, p.b));
    const v3 GLj{T.GLx, T.GLy, T.GLz};
    da += cross(rcr32 * (GLj - (3 * dot(GLj, p.r) * p.rr2) * p.r), p.b);
    da += (T.lum * I.rR2G4c * p.rr3 / I.GM) * p.r;
}

// Every term one pair (i, j) adds to i in the thread-pool order
// (RELATIVITY, ROTATIONAL_TIDAL_DEFORMATION_NANTI_FORCE, LENSE_THIRRING,
// RADIATION_PRESSURE; mass.impl:82-125, mass.cpp:147-155).
struct PairTerms { v3 G, dg, rxdg, gj, lt1, lt2, lt3, rad; double rr, tp_dg; };
EPH_HD PairTerms pair_terms(const IMain &I, const JMain &T) {
    PairTerms o;
    v3 unused{0, 0, 0};
    const PairGeo p = relativity(I, T, unused);
    o.G = p.G;
    o.rr = p.rr;
    o.tp_dg = p.tp_dg;
    o.dg = (T.GM * p.rr3 * p.rr2 * I.R2) * quad_field(I.C, p);
    o.rxdg = cross(p.r, o.dg);
    const m3 Cj{v3{T.C[0], T.C[1], T.C[2]}, v3{T.C[3], T.C[4], T.C[5]}, v3{T.C[6], T.C[7], T.C[8]}};
    o.gj = (T.GM * p.rr3 * p.rr2 * T.R2) * quad_field(Cj, p);
    const double rcr32 = 2 * T.GM / kC * p.rr3;
    const v3 GLb = cross(I.GL, p.b);
    o.lt1 = rcr32 * (GLb - (3 * dot(GLb, p.r) * p.rr2) * p.r);
    o.lt2 = rcr32 * cross(I.GL, cross(p.r, p.b));
    const v3 GLj{T.GLx, T.GLy, T.GLz};
    o.lt3 = cross(rcr32 * (GLj - (3 * dot(GLj, p.r) * p.rr2) * p.r), p.b);
    o.rad = (T.lum * I.rR2G4c * p.rr3 / I.GM) * p.r;
    return o;
}

// Add one pair's terms in the reference order.
struct MainAcc {
    v3 ga{0, 0, 0}, da{0, 0, 0}, dq{0, 0, 0};
    double mind, maxi;
    EPH_HD void add(const PairTerms &t) {
        mind = checked_max_into(mind, t.rr);
        maxi = checked_max_into(maxi, t.tp_dg);
        ga += t.G;
        da += t.dg;
        dq += t.rxdg;
        da += t.gj;
        da -= t.lt1;
        dq -= t.lt2;
        da += t.lt3;
        da += t.rad;
    }
};

// ---------------------------------------------------------------------------
// deformation (mass::deform_by, mass.cpp:78-103)
// ---------------------------------------------------------------------------
// Tidal class of a body: 0 = no tides (k2 = 0, DAMPED_TIDAL_DEFORMATION_MATRIX_NK2),
// 1 = instantaneous tides (tide_delay = 0, _NDELAY), 2 = constant time lag.
EPH_HD int tidal_class(double k2, double tide_delay) { return k2 == 0 ? 0 : (tide_delay == 0 ? 1 : 2); }

struct DeformSum {
    double phi;
    v3 nacc;
    m3 C;
};

// DAMPED_TIDAL_DEFORMATION_MATRIX{,_NDELAY,_NK2} (mass.impl:1-46): what
// partner j adds to body i's phi, naccel and tidal matrix sum.  Own data:
// r, v, sR2, k2, tide_delay, w; partner: r, v, sR2, GM, tide_delay_factor.
EPH_HD void deform_pair(int cls, const ddv3 &ri, const ddv3 &vi, double sR2i, double k2, double td, v3 w,
                        const ddv3 &rj, const ddv3 &vj, double sR2j, double GMj, double tdfj, DeformSum &s) {
    const v3 r = diff(rj, ri);
    const double rr2 = 1 / smax(dot(r, r), sR2i + sR2j);
    const double rr = sqrt(rr2);
    const double tp_dphi = rr * GMj;
    const double tp_dg = rr2 * tp_dphi;
    s.phi -= tp_dphi;
    s.nacc += tp_dg * r;
    if (cls == 0) return;
    const double fmj = -3 * tp_dg;
    if (cls == 1) {
        s.C = s.C - (fmj * k2) * subdiag(outer(rr2 * r, r), 1.0 / 3);
        return;
    }
    const v3 v = diff(vj, vi);
    v3 dw = rr2 * cross(r, cross(w, r) - v);
    double dwt = normalize(dw);
    dwt *= td * tdfj;
    const double ecwt2 = 1 + 4 * dwt * dwt,
                 secwt2 = sqrt(ecwt2),
                 x0 = sqrt((1 + secwt2) / (2 * ecwt2)),
                 y0 = dwt * sqrt(2 / (ecwt2 * (1 + secwt2))),
                 z02 = dwt * dwt * (2 / (ecwt2 + secwt2));
    const v3 dr = x0 * r + y0 * cross(dw, r);
    m3 t = outer(rr2 * dr, dr) - outer(z02 * dw, dw);
    t = adddiag(t, z02 - 1.0 / 3);
    s.C = s.C - (fmj * k2) * t;
}

// Per-body constants of UPDATE_HARMONICS (mass.impl:48-61); the static part
// s C_static(t) s^T does not depend on the spin, so it is formed once.
struct Harmonics {
    double k2r, scale, R2, A;
    m3 Cw;
    EPH_HD void set(double k2r_, double R, double R2_, double A_, double GM, double exJ2, m3 mc, const m3 &s) {
        k2r = k2r_;
        R2 = R2_;
        A = A_;
        scale = R * R2 / (2 * GM);
        mc.x.x += exJ2 / 2;
        mc.y.y += exJ2 / 2;
        mc.z.z -= exJ2;
        Cw = toworld(s, mc);
    }
    // C_potential and GI from the tidal sum and the spin w
    EPH_HD void apply(const m3 &Csum, v3 w, m3 &C, m3 &GI) const {
        const double w2 = dot(w, w);
        C = adddiag(Csum, w2 * k2r / 3);
        C = C - outer(k2r * w, w);
        C = scale * C;
        C = C + Cw;
        GI = (-2.0 / 3 * R2) * subdiag(C, A);
    }
};

// UPDATE_ANGULAR_VELOCITY (mass.impl:63-72); returns should_break.  When GL
// is not finite, gl_from_w is set: the caller stores GL = GI w.
EPH_HD bool spin_step(const m3 &GI, v3 &w, const ddv3 &GL, bool &gl_from_w) {
    if (finite(GL)) {
        gl_from_w = false;
        const v3 oldw = w;
        w = inverse(GI) * todbl(GL);
        const v3 dwv = w - oldw;
        return dot(dwv, dwv) < 1e-18 * dot(oldw, oldw);
    }
    gl_from_w = true;
    return true;
}

// RING_CORRECTION (mass.impl:137-152) of a ring host with final gaccel ga and
// naccel nacc: corrects daccel da and dtorque dq.  GL is the host's GL.
EPH_HD void ring_correction(const eph_ring &mr, double R2, v3 GL, v3 ga, v3 nacc, v3 &da, v3 &dq) {
    da = da - mr.GM_ratio * ((ga + da) + nacc);
    const double rmGL = normalize(GL);
    const v3 ptorque = dot(dq, GL) * GL;
    v3 dtq = dq - ptorque;
    dtq = (rmGL / (rmGL + mr.GL_R2 * R2)) * dtq;
    dq = dtq + ptorque;
}

// The k = 25 spin-gradient correction of the RK12 step (RungeKutta.impl:129-140):
// from GI, GL, the stage's w and torque and the previous Egrad.
EPH_HD v3 egrad_k25(const m3 &GI, v3 GL, v3 w_now, v3 tq_now, v3 egrad, double dt) {
    const m3 ir = inverse(GI);
    const double irdet = det(ir);
    const double dt2 = dt * dt;
    const v3 tq = tq_now * dt2;
    const v3 tqdot = dt2 * (tq_now - egrad);
    v3 j = GL;
    const v3 w = w_now * dt;
    const v3 irtq = ir * tq;
    j = j * dt;
    const v3 t1 = dot(w, j) * (ir * w + (GI * j) * irdet);
    return cross((t1 + cross(w, irtq)) + ir * ((cross(j, irtq) - 2 * cross(w, tq)) + tqdot), w);
}

// ---------------------------------------------------------------------------
// geopotential::sum (src/physics/geopotential.impl:26-106)
// ---------------------------------------------------------------------------
EPH_HD v3 gp_sum(const eph_geopotential &gp, double R, v3 r) {
    double rho2 = r.x * r.x + r.y * r.y;
    double rr2 = 1 / (rho2 + r.z * r.z);
    double rho1 = sqrt(rho2);
    double rr1 = sqrt(rr2);
    double rb1 = R * rr1;
    double rbn = rb1;
    double cphi = rho1 * rr1;
    double sphi = r.z * rr1;
    double cm, sm;
    if (rho2 < 2.2250738585072014e-308) { cm = 1; sm = 0; }
    else { cm = r.x / rho1; sm = r.y / rho1; }
    // Both loops are unrolled to the maximum degree (the loader rejects
    // larger ones), so every index of pnm / pn1m is a constant and the
    // arrays live in registers, not in scratch memory.
    double pnm[EPH_GP_MAX_DEGREE], pn1m[EPH_GP_MAX_DEGREE];
    pnm[0] = 0;
    pn1m[0] = 1;
    double ppn = 1, pn = 1, pn1 = sphi;
    double zx = 0, zz = 0;
    v3 t{0, 0, 0};
    const int Nz = (int)gp.Nz, Nt = (int)gp.Nt;
#pragma unroll
    for (int n = 2; n <= EPH_GP_MAX_DEGREE; ++n) {
        if (n > Nz) break;
        rbn *= rb1;
        const double *JCSn = gp.c + (n * n - 4);
        pn = ((double)(2 * n - 1) * sphi * pn1 - (double)(n - 1) * pn) / (double)n;
        ppn = sphi * ppn + (double)n * pn1;
        double z1 = JCSn[0] * rbn;
        zx += (double)(n + 1) * z1 * pn;
        zz -= z1 * ppn;
        { double sw = pn; pn = pn1; pn1 = sw; }
        if (n > Nt) continue;
        v3 f{0, 0, 0};
        double m1 = (double)(2 * n - 1) * sphi;
        double m2 = (double)(-n) * sphi;
        double smm = 0, cmm = 0, smm1 = 0, cmm1 = 0;
#pragma unroll
        for (int m = 1; m <= EPH_GP_MAX_DEGREE; ++m) {
            if (m > n) break;
            if (m == n) {
                pn1m[m - 1] = 0;
                pnm[m - 1] = (double)(2 * n - 1) * cphi * pnm[m - 2];
            } else
                pnm[m - 1] = (m1 * pn1m[m - 1] - (double)(n + m - 1) * pnm[m - 1]) / (double)(n - m);
            if (m == 1) { smm = sm; cmm = cm; smm1 = 0; cmm1 = 1; }
            else { smm = 2 * cm * smm1 - smm; cmm = 2 * cm * cmm1 - cmm; }
            double t1 = JCSn[2 * m] * smm + JCSn[2 * m - 1] * cmm;
            double t2 = JCSn[2 * m] * cmm - JCSn[2 * m - 1] * smm;
            double ppnm = m2 * pnm[m - 1] + (double)(n + m) * pn1m[m - 1];
            f.x += t1 * pnm[m - 1];
            f.y += (double)m * t2 * pnm[m - 1];
            f.z += t1 * ppnm;
            { double sw = cmm; cmm = cmm1; cmm1 = sw; }
            { double sw = smm; smm = smm1; smm1 = sw; }
            { double sw = pnm[m - 1]; pnm[m - 1] = pn1m[m - 1]; pn1m[m - 1] = sw; }
        }
        f.x *= (double)(n + 1);
        t += rbn * f;
    }
    t.x = zx - cphi * t.x;
    t.z += zz * cphi;
    t = rr2 * t;
    double g1 = t.x * cphi - t.z * sphi;
    return v3{g1 * cm - t.y * sm, g1 * sm + t.y * cm, cphi * t.z + sphi * t.x};
}

// ---------------------------------------------------------------------------
// ring::sum (src/physics/ring.impl:2070-2196).  Tab supplies the FP32
// bicubic correction table and the Pade series of the elliptic integrals:
// static fz(x, y, kh) and xserkk/xserkm/xser0/kserkk/kserkm/kser0(x), each
// reading the reference's constant tables on its side (device or host).
// ---------------------------------------------------------------------------
EPH_HD float fzcorr_tab(const float *tab, float x, float y, int kh) {
    const int N = 32;
    x *= (float)N;
    y *= (float)N;
    // cell indices clamped to the table while still floats, so a non-finite
    // argument (a system gone bad) still reads inside the table; for finite
    // arguments this is the integer clamp of ring.impl:2072-2079
    const float fi = fminf(fmaxf(floorf(x), 0.0f), (float)(N - 1));
    const float fj = fminf(fmaxf(floorf(y) - (float)(kh * N), 0.0f), (float)(N - 1));
    const int i = (int)fi, j = (int)fj + kh * N;
    x -= (float)i;
    y -= (float)j;
    const float *p = tab + 16 * (2 * N * i + j);
    return p[0] + x * (p[4] + x * (p[8] + x * p[12]))
         + y * (p[1] + x * (p[5] + x * (p[9] + x * p[13]))
         + y * (p[2] + x * (p[6] + x * (p[10] + x * p[14]))
         + y * (p[3] + x * (p[7] + x * (p[11] + x * p[15])))));
}
template <int NC>
EPH_HD double padesum(double x, const double (&coef)[NC][2]) {
    double resn = 0, resd = 0;
    int n = NC;
    while (n > 0) {
        --n;
        resn = x * resn + coef[n][0];
        resd = x * resd + coef[n][1];
    }
    return resn / resd;
}

template <class Tab>
EPH_HD v3 ring_sum_t(const eph_ring &rg, double host_R, v3 r) {
    const double pi = kPi;
    bool psign = r.z < 0;
    if (psign) r.z = -r.z;
    const double re2 = r.x * r.x + r.y * r.y;
    const double re = sqrt(re2);
    const double rz2 = r.z * r.z;
    const double r2 = re2 + rz2;
    const double rz = r.z;
    v3 ret{0, 0, 0};
    const int N = (int)rg.N;
    for (int i = 0; i < N; ++i) {
        double R = rg.R_R[i] * host_R;
        double H = rg.H_R[i] * host_R;
        double reR = re * R;
        double R2 = R * R, H2, fzh = 1;
        double k2dre = 4 / (r2 + R2 + 2 * reR);
        if (H > 0) {
            H2 = H * H;
            k2dre -= H2 * (k2dre * k2dre) * (1 - k2dre * rz2) / 48;
            fzh = (re > R ? re - R : 0);
            fzh = 4 * (rz2 + fzh * fzh);
            fzh = sqrt(fzh / (fzh + H2));
        }
        double k2 = reR * k2dre;
        k2dre *= R2 * sqrt(k2dre);
        double phi = atan2(rz, re - R);
        double k = 0, lk = 0, fx;
        bool lkh = k2 >= 0.45;
        if (lkh) {
            k = sqrt(k2);
            lk = log((1 - k2) / (8 * (1 + k)));
            fx = (Tab::xserkk(k) + lk * Tab::xserkm(k)) / (k2 * k2);
        } else {
            fx = Tab::xser0(k2);
        }
        fx *= k2dre;
        double fz;
        double s = sin(phi), c = cos(phi);
        bool kh = k2 >= .5;
        bool exphi;
        if (kh) {
            const double phimax = (pi / 2);
            exphi = c > 0;
            double ck = 1 - k2, lk2 = log(ck / 16), sk = sqrt(ck);
            double cc = exphi ? -c : c;
            double cphi = exphi ? pi - phi : phi;
            fz = cphi - (s / 8) * ((lk2 * (k2 - 5) - 2 * ck) * sk + 2 * cc * (1 + lk2) * ck);
            fz *= Tab::fz((float)((exphi ? phi : pi - phi) * (1 / phimax)), (float)(k2 * 2), 1);
        } else {
            const double phimax = (809. / 512);
            exphi = phi > phimax;
            double cr2 = exphi ? R2 + rz2 : r2;
            double cRdr2 = (exphi ? re2 : R2) / cr2;
            fz = (pi / 8) * (4 - 3 * cRdr2) * (rz / sqrt(cr2) * cRdr2);
            fz *= Tab::fz((float)((exphi ? pi - phi : phi) * (1 / phimax)), (float)(k2 * 2), 0);
        }
        if (exphi) {
            double ek;
            if (lkh) ek = Tab::kserkk(k) + lk * Tab::kserkm(k);
            else ek = Tab::kser0(k2);
            fz = pi - 2 * sqrt(1 - k2) * ek * s - fz;
        }
        double Gs = rg.GsR2[i] / (host_R * host_R);
        fx *= Gs;
        fz *= Gs;
        ret += v3{-4 * fx * r.x, -4 * fx * r.y, -2 * fz * fzh};
    }
    if (psign) ret.z = -ret.z;
    return ret;
}

// world-frame acceleration (per unit GM of the non-point body) at offset r
// from it: geopotential in the body frame s (mass.cpp:186-188) or ring in the
// frame of the body's GL (mass.cpp:200-203)
EPH_HD v3 gp_field_world(const eph_geopotential &gp, double R, const m3 &fmis, v3 r) {
    const v3 lr = tolocal(fmis, r);
    return fmis * gp_sum(gp, R, lr);
}
template <class Tab>
EPH_HD v3 ring_field_world(const eph_ring &rg, double R, v3 migl, v3 r) {
    const v3 x = asc_node(migl), z = unit(migl);
    const m3 fgls{x, cross(z, x), z};
    const v3 lr = tolocal(fgls, r);
    return fgls * ring_sum_t<Tab>(rg, R, lr);
}

}  // namespace eph
