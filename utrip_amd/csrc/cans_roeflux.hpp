// cans_roeflux.hpp -- the linearised Roe Riemann solver of CANS for one
// interface, in the sweep frame (x = normal), for T = double (bit-exact with
// the reference) or T = float (the mixed-precision engines).
//
//   roeflux12: cans1d/hdroe/roeflux_m.f:48-242 (identical body in cans2d)
//   roeflux3 : cans3d/hdroe/roeflux_m.f:33-221 (with vz, bz and the 2006
//              "+1-sp2" term in alphs, :79); without that term it is
//              roeflux_m3 of cans1d/cans2d (cans1d/hdroe/roeflux_m3.f)
// Every expression keeps the reference's association order.
#pragma once
#include "cans_common.hpp"

namespace cans {

template <class T> CANS_DEV T tsqrt(T x);
template <> CANS_DEV double tsqrt<double>(double x) { return __builtin_sqrt(x); }
template <> CANS_DEV float tsqrt<float>(float x) { return __builtin_sqrtf(x); }
template <class T> CANS_DEV T tabs(T x) { return x < T(0) ? -x : (x == T(0) ? T(0) * x : x); }
template <> CANS_DEV double tabs<double>(double x) { return __builtin_fabs(x); }
template <> CANS_DEV float tabs<float>(float x) { return __builtin_fabsf(x); }

// minmod limiter, the tvdminmod.f statement function
// flmt(a,b) = max(0, min(b*sign(1,a), |a|)) * sign(1,a)
CANS_DEV double flmt(double a, double b) {
    return fmaxF(0.0, fminF(b * signF(1.0, a), __builtin_fabs(a))) * signF(1.0, a);
}
CANS_DEV float flmt(float a, float b) {
    return fmaxF(0.0f, fminF(b * signF(1.0f, a), __builtin_fabsf(a))) * signF(1.0f, a);
}

template <class T> struct RoePi {
    static constexpr double pi4d = 4.0 * kPi;          // pi4=4.0d0*pi
    static constexpr double pi4id = 1.0 / pi4d;         // pi4i=1.0d0/pi4
    static constexpr double pi8id = 5.0e-1 * pi4id;     // pi8i=5.0d-1*pi4i
    static constexpr double sqh = 0.70710678118654757;  // sqrt(0.5d0), correctly rounded
};

// out: fro, fee, frx(normal), fry(tangential), fby(tangential B)
template <class T>
CANS_DEV void roeflux12(T gm, T rhol, T prl, T vxl, T vyl, T bxl, T byl, T rhor, T prr, T vxr, T vyr, T bxr,
                        T byr, T *out) {
    const T pi4 = T(RoePi<T>::pi4d), pi4i = T(RoePi<T>::pi4id), pi8i = T(RoePi<T>::pi8id);
    const T one = T(1), half = T(0.5), zero = T(0);
    const T sr0 = tsqrt(rhol);
    const T sr1 = tsqrt(rhor);
    const T sri = one / (sr0 + sr1);
    const T rhobar = sr0 * sr1;
    const T vxbar = (sr0 * vxl + sr1 * vxr) * sri;
    const T vybar = (sr0 * vyl + sr1 * vyr) * sri;
    const T bxbar = (sr0 * bxr + sr1 * bxl) * sri;
    const T bybar = (sr0 * byr + sr1 * byl) * sri;
    const T hl = half * (vxl * vxl + vyl * vyl) + gm * prl / ((gm - one) * rhol) + (bxbar * bxbar + byl * byl) / (pi4 * rhol);
    const T hr = half * (vxr * vxr + vyr * vyr) + gm * prr / ((gm - one) * rhor) + (bxbar * bxbar + byr * byr) / (pi4 * rhor);
    const T hbar = (sr0 * hl + sr1 * hr) * sri;
    const T byave = (byl + byr) / T(2);
    const T delb2 = (gm - T(2)) / (gm - one) * ((byr - byl) * (byr - byl)) * (sri * sri) * pi8i;
    const T vv2 = vxbar * vxbar + vybar * vybar;
    const T bb2 = bxbar * bxbar + bybar * bybar;
    const T cs2 = (gm - one) * (hbar - half * vv2 - delb2 - bb2 * pi4i / rhobar);
    const T astar2 = (gm - one) * (hbar - half * vv2 - delb2) - (gm - T(2)) * bb2 * pi4i / rhobar;
    const T ca2 = bxbar * bxbar / (pi4 * rhobar);
    const T cbr2 = (bybar * bybar) * pi4i / rhobar;
    const T cfast2 = half * (astar2 + tsqrt(cbr2 * (astar2 + cs2 + ca2) + (cs2 - ca2) * (cs2 - ca2)));
    const T cslow2 = cs2 * ca2 / cfast2;
    const T cfast = tsqrt(cfast2);
    const T cslow = tsqrt(cslow2);
    const T ca = tsqrt(ca2);
    const T cs = tsqrt(cs2);
    const T sgr = bybar * bybar - T(1.0e-12);
    const T sp = half + signF(half, sgr);
    const T sqh = T(RoePi<T>::sqh);
    const T betay = sp * bybar * tsqrt(one / (bybar * bybar + one - sp)) + sqh * (one - sp);
    const T betaz = sqh * (one - sp);
    const T sgr2 = (bybar * bybar) / (pi4 * rhobar) + tabs(ca2 - cs2) - T(1.0e-12);
    const T sp2 = half + signF(half, sgr2);
    const T cfca = fmaxF(zero, cfast2 - ca2);
    const T cfcs = fmaxF(zero, cfast2 - cslow2);
    const T cfa = fmaxF(zero, cfast2 - cs2);
    const T alphf = sp2 * tsqrt(cfca / (cfcs + one - sp2)) + one - sp2;
    const T alphs = sp2 * tsqrt(cfa / (cfcs + one - sp2));
    const T sgnbx = signF(one, bxbar);
    const T eeps = (vxr - vxl + tabs(vxr - vxl)) * T(2.5e-1);
    const T elpf = -fmaxF(tabs(vxbar + cfast), eeps);
    const T elmf = -fmaxF(tabs(vxbar - cfast), eeps);
    const T elps = -fmaxF(tabs(vxbar + cslow), eeps);
    const T elms = -fmaxF(tabs(vxbar - cslow), eeps);
    const T elpa = -fmaxF(tabs(vxbar + ca), eeps);
    const T elma = -fmaxF(tabs(vxbar - ca), eeps);
    const T elze = -fmaxF(tabs(vxbar), eeps);
    const T drho = rhor - rhol;
    const T du21 = rhobar * (vxr - vxl);
    const T du31 = rhobar * (vyr - vyl);
    const T du41 = zero;
    const T du6 = byr - byl;
    const T du5 = zero;
    const T t1 = betay * du6 + betaz * du5;
    const T t2 = (prr - prl + (byave * du6) * pi4i + (gm - T(2)) * (bybar * du6) * pi4i) / (gm - one);
    const T t3 = betaz * du6 - betay * du5;
    const T s1 = du21;
    const T s2 = betay * du31 + betaz * du41;
    const T s3 = betaz * du31 - betay * du41;
    const T sq4r = tsqrt(pi4 / rhobar);
    const T p11 = alphs * cfast * sq4r;
    const T p12 = -alphf * cs2 / cfast * sq4r;
    const T p21 = alphf * (cfast2 - cs2 * (gm - T(2)) / (gm - one));
    const T p22 = alphs * (cslow2 - cs2 * (gm - T(2)) / (gm - one));
    const T q11 = alphf * cfast;
    const T q12 = alphs * cslow;
    const T q21 = -alphs * ca * sgnbx;
    const T q22 = alphf * cs * sgnbx;
    const T detp = p11 * p22 - p12 * p21;
    const T detq = q11 * q22 - q12 * q21;
    const T a1 = (p22 * t1 - p12 * t2) / detp, b1 = (q22 * s1 - q12 * s2) / detq;
    const T a2 = (-p21 * t1 + p11 * t2) / detp, b2 = (-q21 * s1 + q11 * s2) / detq;
    const T wpf = half * (a1 + b1);
    const T wmf = half * (a1 - b1);
    const T wps = half * (a2 + b2);
    const T wms = half * (a2 - b2);
    const T sqrp = tsqrt(rhobar * pi4i);
    const T wpa = half * (sqrp * t3 - sgnbx * s3);
    const T wma = half * (sqrp * t3 + sgnbx * s3);
    const T wze = drho - alphf * (wpf + wmf) - alphs * (wps + wms);
    const T fluxlro = rhol * vxl;
    const T fluxlrx = rhol * vxl * vxl + prl + (-(bxbar * bxbar) + byl * byl) * pi8i;
    const T fluxlry = rhol * vxl * vyl - bxbar * byl * pi4i;
    const T fluxlby = vxl * byl - vyl * bxbar;
    const T fluxlee = rhol * vxl * hl - bxbar * (bxbar * vxl + byl * vyl) * pi4i;
    const T fluxrro = rhor * vxr;
    const T fluxrrx = rhor * vxr * vxr + prr + (-(bxbar * bxbar) + byr * byr) * pi8i;
    const T fluxrry = rhor * vxr * vyr - bxbar * byr * pi4i;
    const T fluxrby = vxr * byr - vyr * bxbar;
    const T fluxree = rhor * vxr * hr - bxbar * (bxbar * vxr + byr * vyr) * pi4i;
    const T gq = (gm - T(2)) / (gm - one);
    const T rpfro = alphf;
    const T rpfrx = alphf * (vxbar + cfast);
    const T rpfry = alphf * vybar - alphs * betay * ca * sgnbx;
    const T rpfby = alphs * betay * cfast * sq4r;
    const T rpfee = alphf * (half * vv2 + delb2 + cfast * vxbar + cfast2 / (gm - one) + (cfast2 - cs2) * (gm - T(2)) / (gm - one))
                  - alphs * ca * (betay * vybar) * sgnbx;
    const T rmfro = alphf;
    const T rmfrx = alphf * (vxbar - cfast);
    const T rmfry = alphf * vybar + alphs * betay * ca * sgnbx;
    const T rmfby = rpfby;
    const T rmfee = alphf * (half * vv2 + delb2 - cfast * vxbar + cfast2 / (gm - one) + (cfast2 - cs2) * (gm - T(2)) / (gm - one))
                  + alphs * ca * (betay * vybar) * sgnbx;
    const T rpsro = alphs;
    const T rpsrx = alphs * (vxbar + cslow);
    const T rpsry = alphs * vybar + cs * sgnbx * alphf * betay;
    const T rpsby = -sq4r * cs2 * alphf * betay / cfast;
    const T rpsee = alphs * (half * vv2 + delb2 + cslow * vxbar + cslow2 / (gm - one) + (cslow2 - cs2) * (gm - T(2)) / (gm - one))
                  + alphf * cs * (betay * vybar) * sgnbx;
    const T rmsro = alphs;
    const T rmsrx = alphs * (vxbar - cslow);
    const T rmsry = alphs * vybar - cs * sgnbx * alphf * betay;
    const T rmsby = rpsby;
    const T rmsee = alphs * (half * vv2 + delb2 - cslow * vxbar + cslow2 / (gm - one) + (cslow2 - cs2) * (gm - T(2)) / (gm - one))
                  - alphf * cs * (betay * vybar) * sgnbx;
    (void)gq;
    const T rparo = zero, rparx = zero;
    const T rpary = -sgnbx * betaz;
    const T rpaby = sq4r * betaz;
    const T rpaee = -(betaz * vybar) * sgnbx;
    const T rmaro = zero, rmarx = zero;
    const T rmary = -rpary;
    const T rmaby = rpaby;
    const T rmaee = -rpaee;
    const T rzero = one, rzerx = vxbar, rzery = vybar, rzeby = zero;
    const T rzeee = half * vv2 + delb2;
    const T ef = elpf * wpf, mf = elmf * wmf, es = elps * wps, ms = elms * wms, ea = elpa * wpa, ma = elma * wma,
            ez = elze * wze;
    out[0] = half * (fluxlro + fluxrro + ef * rpfro + mf * rmfro + es * rpsro + ms * rmsro + ea * rparo + ma * rmaro + ez * rzero);
    out[1] = half * (fluxlee + fluxree + ef * rpfee + mf * rmfee + es * rpsee + ms * rmsee + ea * rpaee + ma * rmaee + ez * rzeee);
    out[2] = half * (fluxlrx + fluxrrx + ef * rpfrx + mf * rmfrx + es * rpsrx + ms * rmsrx + ea * rparx + ma * rmarx + ez * rzerx);
    out[3] = half * (fluxlry + fluxrry + ef * rpfry + mf * rmfry + es * rpsry + ms * rmsry + ea * rpary + ma * rmary + ez * rzery);
    out[4] = half * (fluxlby + fluxrby + ef * rpfby + mf * rmfby + es * rpsby + ms * rmsby + ea * rpaby + ma * rmaby + ez * rzeby);
}

// out: fro, fee, frx(normal), fry, frz (tangential), fby, fbz (tangential B).
// FIX2006 = false is roeflux_m3 of cans1d/cans2d (cans1d/hdroe/roeflux_m3.f:
// 52-285, the same body without the "+1-sp2" term in alphs, :124).
template <class T, bool FIX2006 = true>
CANS_DEV void roeflux3(T gm, T rhol, T prl, T vxl, T vyl, T vzl, T bxl, T byl, T bzl, T rhor, T prr, T vxr, T vyr,
                       T vzr, T bxr, T byr, T bzr, T *out) {
    const T pi4 = T(RoePi<T>::pi4d), pi4i = T(RoePi<T>::pi4id), pi8i = T(RoePi<T>::pi8id);
    const T one = T(1), half = T(0.5), zero = T(0), two = T(2);
    const T sr0 = tsqrt(rhol);
    const T sr1 = tsqrt(rhor);
    const T sri = one / (sr0 + sr1);
    const T rhobar = sr0 * sr1;
    const T vxbar = (sr0 * vxl + sr1 * vxr) * sri;
    const T vybar = (sr0 * vyl + sr1 * vyr) * sri;
    const T vzbar = (sr0 * vzl + sr1 * vzr) * sri;
    const T bxbar = (sr0 * bxr + sr1 * bxl) * sri;
    const T bybar = (sr0 * byr + sr1 * byl) * sri;
    const T bzbar = (sr0 * bzr + sr1 * bzl) * sri;
    const T hl = half * (vxl * vxl + vyl * vyl + vzl * vzl) + gm * prl / ((gm - one) * rhol)
               + (bxbar * bxbar + byl * byl + bzl * bzl) / (pi4 * rhol);
    const T hr = half * (vxr * vxr + vyr * vyr + vzr * vzr) + gm * prr / ((gm - one) * rhor)
               + (bxbar * bxbar + byr * byr + bzr * bzr) / (pi4 * rhor);
    const T hbar = (sr0 * hl + sr1 * hr) * sri;
    const T byave = (byl + byr) / two;
    const T bzave = (bzl + bzr) / two;
    const T delb2 = (gm - two) / (gm - one) * ((byr - byl) * (byr - byl) + (bzr - bzl) * (bzr - bzl)) * (sri * sri) * pi8i;
    const T vv2 = vxbar * vxbar + vybar * vybar + vzbar * vzbar;
    const T bb2 = bxbar * bxbar + bybar * bybar + bzbar * bzbar;
    const T cs2 = (gm - one) * (hbar - half * vv2 - delb2 - bb2 * pi4i / rhobar);
    const T astar2 = (gm - one) * (hbar - half * vv2 - delb2) - (gm - two) * bb2 * pi4i / rhobar;
    const T ca2 = bxbar * bxbar / (pi4 * rhobar);
    const T bt2 = bybar * bybar + bzbar * bzbar;
    const T cbr2 = bt2 * pi4i / rhobar;
    const T cfast2 = half * (astar2 + tsqrt(cbr2 * (astar2 + cs2 + ca2) + (cs2 - ca2) * (cs2 - ca2)));
    const T cslow2 = cs2 * ca2 / cfast2;
    const T cfast = tsqrt(cfast2);
    const T cslow = tsqrt(cslow2);
    const T ca = tsqrt(ca2);
    const T cs = tsqrt(cs2);
    const T sgr = bt2 - T(1.0e-12);
    const T sp = half + signF(half, sgr);
    const T sqh = T(RoePi<T>::sqh);
    const T rbt = tsqrt(one / (bt2 + one - sp));
    const T betay = sp * bybar * rbt + sqh * (one - sp);
    const T betaz = sp * bzbar * rbt + sqh * (one - sp);
    const T sgr2 = bt2 / (pi4 * rhobar) + tabs(ca2 - cs2) - T(1.0e-12);
    const T sp2 = half + signF(half, sgr2);
    const T cfca = fmaxF(zero, cfast2 - ca2);
    const T cfcs = fmaxF(zero, cfast2 - cslow2);
    const T cfa = fmaxF(zero, cfast2 - cs2);
    const T alphf = sp2 * tsqrt(cfca / (cfcs + one - sp2)) + one - sp2;
    const T alphs = FIX2006 ? sp2 * tsqrt(cfa / (cfcs + one - sp2)) + one - sp2 : sp2 * tsqrt(cfa / (cfcs + one - sp2));
    const T sgnbx = signF(one, bxbar);
    const T eeps = (vxr - vxl + tabs(vxr - vxl)) * T(2.5e-1);
    const T elpf = -fmaxF(tabs(vxbar + cfast), eeps);
    const T elmf = -fmaxF(tabs(vxbar - cfast), eeps);
    const T elps = -fmaxF(tabs(vxbar + cslow), eeps);
    const T elms = -fmaxF(tabs(vxbar - cslow), eeps);
    const T elpa = -fmaxF(tabs(vxbar + ca), eeps);
    const T elma = -fmaxF(tabs(vxbar - ca), eeps);
    const T elze = -fmaxF(tabs(vxbar), eeps);
    const T drho = rhor - rhol;
    const T du21 = rhobar * (vxr - vxl);
    const T du31 = rhobar * (vyr - vyl);
    const T du41 = rhobar * (vzr - vzl);
    const T du6 = byr - byl;
    const T du5 = bzr - bzl;
    const T t1 = betay * du6 + betaz * du5;
    const T t2 = (prr - prl + (byave * du6 + bzave * du5) * pi4i + (gm - two) * (bybar * du6 + bzbar * du5) * pi4i) / (gm - one);
    const T t3 = betaz * du6 - betay * du5;
    const T s1 = du21;
    const T s2 = betay * du31 + betaz * du41;
    const T s3 = betaz * du31 - betay * du41;
    const T sq4r = tsqrt(pi4 / rhobar);
    const T p11 = alphs * cfast * sq4r;
    const T p12 = -alphf * cs2 / cfast * sq4r;
    const T p21 = alphf * (cfast2 - cs2 * (gm - two) / (gm - one));
    const T p22 = alphs * (cslow2 - cs2 * (gm - two) / (gm - one));
    const T q11 = alphf * cfast;
    const T q12 = alphs * cslow;
    const T q21 = -alphs * ca * sgnbx;
    const T q22 = alphf * cs * sgnbx;
    const T detp = p11 * p22 - p12 * p21;
    const T detq = q11 * q22 - q12 * q21;
    const T a1 = (p22 * t1 - p12 * t2) / detp, b1 = (q22 * s1 - q12 * s2) / detq;
    const T a2 = (-p21 * t1 + p11 * t2) / detp, b2 = (-q21 * s1 + q11 * s2) / detq;
    const T wpf = half * (a1 + b1);
    const T wmf = half * (a1 - b1);
    const T wps = half * (a2 + b2);
    const T wms = half * (a2 - b2);
    const T sqrp = tsqrt(rhobar * pi4i);
    const T wpa = half * (sqrp * t3 - sgnbx * s3);
    const T wma = half * (sqrp * t3 + sgnbx * s3);
    const T wze = drho - alphf * (wpf + wmf) - alphs * (wps + wms);
    const T fluxlro = rhol * vxl;
    const T fluxlrx = rhol * vxl * vxl + prl + (-(bxbar * bxbar) + byl * byl + bzl * bzl) * pi8i;
    const T fluxlry = rhol * vxl * vyl - bxbar * byl * pi4i;
    const T fluxlrz = rhol * vxl * vzl - bxbar * bzl * pi4i;
    const T fluxlby = vxl * byl - vyl * bxbar;
    const T fluxlbz = vxl * bzl - vzl * bxbar;
    const T fluxlee = rhol * vxl * hl - bxbar * (bxbar * vxl + byl * vyl + bzl * vzl) * pi4i;
    const T fluxrro = rhor * vxr;
    const T fluxrrx = rhor * vxr * vxr + prr + (-(bxbar * bxbar) + byr * byr + bzr * bzr) * pi8i;
    const T fluxrry = rhor * vxr * vyr - bxbar * byr * pi4i;
    const T fluxrrz = rhor * vxr * vzr - bxbar * bzr * pi4i;
    const T fluxrby = vxr * byr - vyr * bxbar;
    const T fluxrbz = vxr * bzr - vzr * bxbar;
    const T fluxree = rhor * vxr * hr - bxbar * (bxbar * vxr + byr * vyr + bzr * vzr) * pi4i;
    const T bv = betay * vybar + betaz * vzbar;
    const T rpfro = alphf;
    const T rpfrx = alphf * (vxbar + cfast);
    const T rpfry = alphf * vybar - alphs * betay * ca * sgnbx;
    const T rpfrz = alphf * vzbar - alphs * betaz * ca * sgnbx;
    const T rpfby = alphs * betay * cfast * sq4r;
    const T rpfbz = alphs * betaz * cfast * sq4r;
    const T rpfee = alphf * (half * vv2 + delb2 + cfast * vxbar + cfast2 / (gm - one) + (cfast2 - cs2) * (gm - two) / (gm - one))
                  - alphs * ca * bv * sgnbx;
    const T rmfro = alphf;
    const T rmfrx = alphf * (vxbar - cfast);
    const T rmfry = alphf * vybar + alphs * betay * ca * sgnbx;
    const T rmfrz = alphf * vzbar + alphs * betaz * ca * sgnbx;
    const T rmfby = rpfby;
    const T rmfbz = rpfbz;
    const T rmfee = alphf * (half * vv2 + delb2 - cfast * vxbar + cfast2 / (gm - one) + (cfast2 - cs2) * (gm - two) / (gm - one))
                  + alphs * ca * bv * sgnbx;
    const T rpsro = alphs;
    const T rpsrx = alphs * (vxbar + cslow);
    const T rpsry = alphs * vybar + cs * sgnbx * alphf * betay;
    const T rpsrz = alphs * vzbar + cs * sgnbx * alphf * betaz;
    const T rpsby = -sq4r * cs2 * alphf * betay / cfast;
    const T rpsbz = -sq4r * cs2 * alphf * betaz / cfast;
    const T rpsee = alphs * (half * vv2 + delb2 + cslow * vxbar + cslow2 / (gm - one) + (cslow2 - cs2) * (gm - two) / (gm - one))
                  + alphf * cs * bv * sgnbx;
    const T rmsro = alphs;
    const T rmsrx = alphs * (vxbar - cslow);
    const T rmsry = alphs * vybar - cs * sgnbx * alphf * betay;
    const T rmsrz = alphs * vzbar - cs * sgnbx * alphf * betaz;
    const T rmsby = rpsby;
    const T rmsbz = rpsbz;
    const T rmsee = alphs * (half * vv2 + delb2 - cslow * vxbar + cslow2 / (gm - one) + (cslow2 - cs2) * (gm - two) / (gm - one))
                  - alphf * cs * bv * sgnbx;
    const T rparo = zero, rparx = zero;
    const T rpary = -sgnbx * betaz;
    const T rparz = sgnbx * betay;
    const T rpaby = sq4r * betaz;
    const T rpabz = -sq4r * betay;
    const T rpaee = -(betaz * vybar - betay * vzbar) * sgnbx;
    const T rmaro = zero, rmarx = zero;
    const T rmary = -rpary, rmarz = -rparz;
    const T rmaby = rpaby, rmabz = rpabz;
    const T rmaee = -rpaee;
    const T rzero = one, rzerx = vxbar, rzery = vybar, rzerz = vzbar, rzeby = zero, rzebz = zero;
    const T rzeee = half * vv2 + delb2;
    const T ef = elpf * wpf, mf = elmf * wmf, es = elps * wps, ms = elms * wms, ea = elpa * wpa, ma = elma * wma,
            ez = elze * wze;
    out[0] = half * (fluxlro + fluxrro + ef * rpfro + mf * rmfro + es * rpsro + ms * rmsro + ea * rparo + ma * rmaro + ez * rzero);
    out[1] = half * (fluxlee + fluxree + ef * rpfee + mf * rmfee + es * rpsee + ms * rmsee + ea * rpaee + ma * rmaee + ez * rzeee);
    out[2] = half * (fluxlrx + fluxrrx + ef * rpfrx + mf * rmfrx + es * rpsrx + ms * rmsrx + ea * rparx + ma * rmarx + ez * rzerx);
    out[3] = half * (fluxlry + fluxrry + ef * rpfry + mf * rmfry + es * rpsry + ms * rmsry + ea * rpary + ma * rmary + ez * rzery);
    out[4] = half * (fluxlrz + fluxrrz + ef * rpfrz + mf * rmfrz + es * rpsrz + ms * rmsrz + ea * rparz + ma * rmarz + ez * rzerz);
    out[5] = half * (fluxlby + fluxrby + ef * rpfby + mf * rmfby + es * rpsby + ms * rmsby + ea * rpaby + ma * rmaby + ez * rzeby);
    out[6] = half * (fluxlbz + fluxrbz + ef * rpfbz + mf * rmfbz + es * rpsbz + ms * rmsbz + ea * rpabz + ma * rmabz + ez * rzebz);
}

}  // namespace cans
