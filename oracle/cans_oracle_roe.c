/*
 * cans_oracle_roe.c -- C restatement of the Roe + MUSCL-minmod engines
 * roe_m / roeflux_m / tvdminmod in cans1d, cans2d and cans3d (test
 * infrastructure only; see cans_oracle.h).  Citations are to cans_beta_vary/.
 *
 * Temporaries the reference leaves uninitialised are NaN here, so a result
 * that depends on them shows as NaN: with margin < 3 that is the first
 * interior cell (the MUSCL stencil of face 2 reads half-step values at cell 1,
 * which roe_m never sets in 1D/2D; in 3D vxh,vyh,vzh,prh at cell 1).
 */
#include "cans_oracle.h"
#include "cans_oracle_util.h"

/* One interface of roeflux_m, the 1D/2D body (cans1d/hdroe/roeflux_m.f:48-242,
 * identical in cans2d/hdroe/roeflux_m.f).  Inputs in the sweep frame
 * (x = normal); out[0..4] = fro, fee, frx, fry, fby. */
static void roeflux12(double gm, double rhol, double prl, double vxl, double vyl, double bxl, double byl,
                      double rhor, double prr, double vxr, double vyr, double bxr, double byr, double *out) {
    (void)bxl;
    const double pi = acos(-1.0);
    const double pi4 = 4.0 * pi;
    const double pi4i = 1.0 / pi4;
    const double pi8i = 5.0e-1 * pi4i;
    double sr0 = sqrt(rhol);
    double sr1 = sqrt(rhor);
    double sri = 1.0 / (sr0 + sr1);
    double rhobar = sr0 * sr1;
    double vxbar = (sr0 * vxl + sr1 * vxr) * sri;
    double vybar = (sr0 * vyl + sr1 * vyr) * sri;
    double bxbar = (sr0 * bxr + sr1 * bxl) * sri;
    double bybar = (sr0 * byr + sr1 * byl) * sri;
    double hl = 0.5 * (sq(vxl) + sq(vyl)) + gm * prl / ((gm - 1.0) * rhol) + (sq(bxbar) + sq(byl)) / (pi4 * rhol);
    double hr = 0.5 * (sq(vxr) + sq(vyr)) + gm * prr / ((gm - 1.0) * rhor) + (sq(bxbar) + sq(byr)) / (pi4 * rhor);
    double hbar = (sr0 * hl + sr1 * hr) * sri;
    double byave = (byl + byr) / 2.0;
    double delb2 = (gm - 2.0) / (gm - 1.0) * (sq(byr - byl)) * sq(sri) * pi8i;
    double cs2 = (gm - 1.0) * (hbar - 0.5 * (sq(vxbar) + sq(vybar)) - delb2 - (sq(bxbar) + sq(bybar)) * pi4i / rhobar);
    double astar2 = (gm - 1.0) * (hbar - 0.5 * (sq(vxbar) + sq(vybar)) - delb2)
                  - (gm - 2.0) * (sq(bxbar) + sq(bybar)) * pi4i / rhobar;
    double ca2 = sq(bxbar) / (pi4 * rhobar);
    double cbr2 = (sq(bybar)) * pi4i / rhobar;
    double cfast2 = 0.5 * (astar2 + sqrt(cbr2 * (astar2 + cs2 + ca2) + sq(cs2 - ca2)));
    double cslow2 = cs2 * ca2 / cfast2;
    double cfast = sqrt(cfast2);
    double cslow = sqrt(cslow2);
    double ca = sqrt(ca2);
    double cs = sqrt(cs2);
    double epsi = 1.0e-12;
    double sgr = sq(bybar) - epsi;
    double sp = 0.5 + signF(0.5, sgr);
    double betay = sp * bybar * sqrt(1.0 / (sq(bybar) + 1.0 - sp)) + sqrt(0.5) * (1.0 - sp);
    double betaz = sqrt(0.5) * (1.0 - sp);
    double eps2 = 1.0e-12;
    double sgr2 = (sq(bybar)) / (pi4 * rhobar) + fabs(ca2 - cs2) - eps2;
    double sp2 = 0.5 + signF(0.5, sgr2);
    double cfca = fmaxF(0.0, cfast2 - ca2);
    double cfcs = fmaxF(0.0, cfast2 - cslow2);
    double cfa = fmaxF(0.0, cfast2 - cs2);
    double alphf = sp2 * sqrt(cfca / (cfcs + 1.0 - sp2)) + 1.0 - sp2;
    double alphs = sp2 * sqrt(cfa / (cfcs + 1.0 - sp2));
    double sgnbx = signF(1.0, bxbar);
    double eeps = (vxr - vxl + fabs(vxr - vxl)) * 2.5e-1;
    double elpf = -fmaxF(fabs(vxbar + cfast), eeps);
    double elmf = -fmaxF(fabs(vxbar - cfast), eeps);
    double elps = -fmaxF(fabs(vxbar + cslow), eeps);
    double elms = -fmaxF(fabs(vxbar - cslow), eeps);
    double elpa = -fmaxF(fabs(vxbar + ca), eeps);
    double elma = -fmaxF(fabs(vxbar - ca), eeps);
    double elze = -fmaxF(fabs(vxbar), eeps);
    double drho = rhor - rhol;
    double du21 = rhobar * (vxr - vxl);
    double du31 = rhobar * (vyr - vyl);
    double du41 = 0;
    double du6 = byr - byl;
    double du5 = 0.;
    double t1 = betay * du6 + betaz * du5;
    double t2 = (prr - prl + (byave * du6) * pi4i + (gm - 2.0) * (bybar * du6) * pi4i) / (gm - 1.0);
    double t3 = betaz * du6 - betay * du5;
    double s1 = du21;
    double s2 = betay * du31 + betaz * du41;
    double s3 = betaz * du31 - betay * du41;
    double p11 = alphs * cfast * sqrt(pi4 / rhobar);
    double p12 = -alphf * cs2 / cfast * sqrt(pi4 / rhobar);
    double p21 = alphf * (cfast2 - cs2 * (gm - 2.0) / (gm - 1.0));
    double p22 = alphs * (cslow2 - cs2 * (gm - 2.0) / (gm - 1.0));
    double q11 = alphf * cfast;
    double q12 = alphs * cslow;
    double q21 = -alphs * ca * sgnbx;
    double q22 = alphf * cs * sgnbx;
    double detp = p11 * p22 - p12 * p21;
    double detq = q11 * q22 - q12 * q21;
    double wpf = 0.5 * ((p22 * t1 - p12 * t2) / detp + (q22 * s1 - q12 * s2) / detq);
    double wmf = 0.5 * ((p22 * t1 - p12 * t2) / detp - (q22 * s1 - q12 * s2) / detq);
    double wps = 0.5 * ((-p21 * t1 + p11 * t2) / detp + (-q21 * s1 + q11 * s2) / detq);
    double wms = 0.5 * ((-p21 * t1 + p11 * t2) / detp - (-q21 * s1 + q11 * s2) / detq);
    double wpa = 0.5 * (sqrt(rhobar * pi4i) * t3 - sgnbx * s3);
    double wma = 0.5 * (sqrt(rhobar * pi4i) * t3 + sgnbx * s3);
    double wze = drho - alphf * (wpf + wmf) - alphs * (wps + wms);
    double fluxlro = rhol * vxl;
    double fluxlrx = rhol * vxl * vxl + prl + (-sq(bxbar) + sq(byl)) * pi8i;
    double fluxlry = rhol * vxl * vyl - bxbar * byl * pi4i;
    double fluxlby = vxl * byl - vyl * bxbar;
    double fluxlee = rhol * vxl * hl - bxbar * (bxbar * vxl + byl * vyl) * pi4i;
    double fluxrro = rhor * vxr;
    double fluxrrx = rhor * vxr * vxr + prr + (-sq(bxbar) + sq(byr)) * pi8i;
    double fluxrry = rhor * vxr * vyr - bxbar * byr * pi4i;
    double fluxrby = vxr * byr - vyr * bxbar;
    double fluxree = rhor * vxr * hr - bxbar * (bxbar * vxr + byr * vyr) * pi4i;
    double rpfro = alphf;
    double rpfrx = alphf * (vxbar + cfast);
    double rpfry = alphf * vybar - alphs * betay * ca * sgnbx;
    double rpfby = alphs * betay * cfast * sqrt(pi4 / rhobar);
    double rpfee = alphf * (0.5 * (sq(vxbar) + sq(vybar)) + delb2 + cfast * vxbar + cfast2 / (gm - 1.0)
                            + (cfast2 - cs2) * (gm - 2.0) / (gm - 1.0))
                 - alphs * ca * (betay * vybar) * sgnbx;
    double rmfro = alphf;
    double rmfrx = alphf * (vxbar - cfast);
    double rmfry = alphf * vybar + alphs * betay * ca * sgnbx;
    double rmfby = rpfby;
    double rmfee = alphf * (0.5 * (sq(vxbar) + sq(vybar)) + delb2 - cfast * vxbar + cfast2 / (gm - 1.0)
                            + (cfast2 - cs2) * (gm - 2.0) / (gm - 1.0))
                 + alphs * ca * (betay * vybar) * sgnbx;
    double rpsro = alphs;
    double rpsrx = alphs * (vxbar + cslow);
    double rpsry = alphs * vybar + cs * sgnbx * alphf * betay;
    double rpsby = -sqrt(pi4 / rhobar) * cs2 * alphf * betay / cfast;
    double rpsee = alphs * (0.5 * (sq(vxbar) + sq(vybar)) + delb2 + cslow * vxbar + cslow2 / (gm - 1.0)
                            + (cslow2 - cs2) * (gm - 2.0) / (gm - 1.0))
                 + alphf * cs * (betay * vybar) * sgnbx;
    double rmsro = alphs;
    double rmsrx = alphs * (vxbar - cslow);
    double rmsry = alphs * vybar - cs * sgnbx * alphf * betay;
    double rmsby = rpsby;
    double rmsee = alphs * (0.5 * (sq(vxbar) + sq(vybar)) + delb2 - cslow * vxbar + cslow2 / (gm - 1.0)
                            + (cslow2 - cs2) * (gm - 2.0) / (gm - 1.0))
                 - alphf * cs * (betay * vybar) * sgnbx;
    double rparo = 0.0, rparx = 0.0;
    double rpary = -sgnbx * betaz;
    double rpaby = sqrt(pi4 / rhobar) * betaz;
    double rpaee = -(betaz * vybar) * sgnbx;
    double rmaro = 0.0, rmarx = 0.0;
    double rmary = -rpary;
    double rmaby = rpaby;
    double rmaee = -rpaee;
    double rzero = 1.0, rzerx = vxbar, rzery = vybar, rzeby = 0.0;
    double rzeee = 0.5 * (sq(vxbar) + sq(vybar)) + delb2;
    out[0] = 0.5 * (fluxlro + fluxrro + elpf * wpf * rpfro + elmf * wmf * rmfro + elps * wps * rpsro
                    + elms * wms * rmsro + elpa * wpa * rparo + elma * wma * rmaro + elze * wze * rzero);
    out[1] = 0.5 * (fluxlee + fluxree + elpf * wpf * rpfee + elmf * wmf * rmfee + elps * wps * rpsee
                    + elms * wms * rmsee + elpa * wpa * rpaee + elma * wma * rmaee + elze * wze * rzeee);
    out[2] = 0.5 * (fluxlrx + fluxrrx + elpf * wpf * rpfrx + elmf * wmf * rmfrx + elps * wps * rpsrx
                    + elms * wms * rmsrx + elpa * wpa * rparx + elma * wma * rmarx + elze * wze * rzerx);
    out[3] = 0.5 * (fluxlry + fluxrry + elpf * wpf * rpfry + elmf * wmf * rmfry + elps * wps * rpsry
                    + elms * wms * rmsry + elpa * wpa * rpary + elma * wma * rmary + elze * wze * rzery);
    out[4] = 0.5 * (fluxlby + fluxrby + elpf * wpf * rpfby + elmf * wmf * rmfby + elps * wps * rpsby
                    + elms * wms * rmsby + elpa * wpa * rpaby + elma * wma * rmaby + elze * wze * rzeby);
}

/* One interface of the cans3d roeflux_m (cans3d/hdroe/roeflux_m.f:33-221,
 * with the alphs "+1-sp2" fix of :79 when fix2006) -- without the fix it is
 * roeflux_m3 of cans1d/cans2d (cans1d/hdroe/roeflux_m3.f:52-285, alphs at
 * :124).  out = fro, fee, frx, fry, frz, fby, fbz. */
static void roeflux3(int fix2006, double gm, double rhol, double prl, double vxl, double vyl, double vzl,
                     double bxl, double byl, double bzl, double rhor, double prr, double vxr, double vyr,
                     double vzr, double bxr, double byr, double bzr, double *out) {
    const double pi = acos(-1.0);
    const double pi4 = 4.0 * pi;
    const double pi4i = 1.0 / pi4;
    const double pi8i = 5.0e-1 * pi4i;
    double sr0 = sqrt(rhol);
    double sr1 = sqrt(rhor);
    double sri = 1.0 / (sr0 + sr1);
    double rhobar = sr0 * sr1;
    double vxbar = (sr0 * vxl + sr1 * vxr) * sri;
    double vybar = (sr0 * vyl + sr1 * vyr) * sri;
    double vzbar = (sr0 * vzl + sr1 * vzr) * sri;
    double bxbar = (sr0 * bxr + sr1 * bxl) * sri;
    double bybar = (sr0 * byr + sr1 * byl) * sri;
    double bzbar = (sr0 * bzr + sr1 * bzl) * sri;
    double hl = 0.5 * (sq(vxl) + sq(vyl) + sq(vzl)) + gm * prl / ((gm - 1.0) * rhol)
              + (sq(bxbar) + sq(byl) + sq(bzl)) / (pi4 * rhol);
    double hr = 0.5 * (sq(vxr) + sq(vyr) + sq(vzr)) + gm * prr / ((gm - 1.0) * rhor)
              + (sq(bxbar) + sq(byr) + sq(bzr)) / (pi4 * rhor);
    double hbar = (sr0 * hl + sr1 * hr) * sri;
    double byave = (byl + byr) / 2.0;
    double bzave = (bzl + bzr) / 2.0;
    double delb2 = (gm - 2.0) / (gm - 1.0) * (sq(byr - byl) + sq(bzr - bzl)) * sq(sri) * pi8i;
    double cs2 = (gm - 1.0) * (hbar - 0.5 * (sq(vxbar) + sq(vybar) + sq(vzbar)) - delb2
                               - (sq(bxbar) + sq(bybar) + sq(bzbar)) * pi4i / rhobar);
    double astar2 = (gm - 1.0) * (hbar - 0.5 * (sq(vxbar) + sq(vybar) + sq(vzbar)) - delb2)
                  - (gm - 2.0) * (sq(bxbar) + sq(bybar) + sq(bzbar)) * pi4i / rhobar;
    double ca2 = sq(bxbar) / (pi4 * rhobar);
    double cbr2 = (sq(bybar) + sq(bzbar)) * pi4i / rhobar;
    double cfast2 = 0.5 * (astar2 + sqrt(cbr2 * (astar2 + cs2 + ca2) + sq(cs2 - ca2)));
    double cslow2 = cs2 * ca2 / cfast2;
    double cfast = sqrt(cfast2);
    double cslow = sqrt(cslow2);
    double ca = sqrt(ca2);
    double cs = sqrt(cs2);
    double epsi = 1.0e-12;
    double sgr = sq(bybar) + sq(bzbar) - epsi;
    double sp = 0.5 + signF(0.5, sgr);
    double betay = sp * bybar * sqrt(1.0 / (sq(bybar) + sq(bzbar) + 1.0 - sp)) + sqrt(0.5) * (1.0 - sp);
    double betaz = sp * bzbar * sqrt(1.0 / (sq(bybar) + sq(bzbar) + 1.0 - sp)) + sqrt(0.5) * (1.0 - sp);
    double eps2 = 1.0e-12;
    double sgr2 = (sq(bybar) + sq(bzbar)) / (pi4 * rhobar) + fabs(ca2 - cs2) - eps2;
    double sp2 = 0.5 + signF(0.5, sgr2);
    double cfca = fmaxF(0.0, cfast2 - ca2);
    double cfcs = fmaxF(0.0, cfast2 - cslow2);
    double cfa = fmaxF(0.0, cfast2 - cs2);
    double alphf = sp2 * sqrt(cfca / (cfcs + 1.0 - sp2)) + 1.0 - sp2;
    double alphs = fix2006 ? sp2 * sqrt(cfa / (cfcs + 1.0 - sp2)) + 1.0 - sp2
                           : sp2 * sqrt(cfa / (cfcs + 1.0 - sp2));
    double sgnbx = signF(1.0, bxbar);
    double eeps = (vxr - vxl + fabs(vxr - vxl)) * 2.5e-1;
    double elpf = -fmaxF(fabs(vxbar + cfast), eeps);
    double elmf = -fmaxF(fabs(vxbar - cfast), eeps);
    double elps = -fmaxF(fabs(vxbar + cslow), eeps);
    double elms = -fmaxF(fabs(vxbar - cslow), eeps);
    double elpa = -fmaxF(fabs(vxbar + ca), eeps);
    double elma = -fmaxF(fabs(vxbar - ca), eeps);
    double elze = -fmaxF(fabs(vxbar), eeps);
    double drho = rhor - rhol;
    double du21 = rhobar * (vxr - vxl);
    double du31 = rhobar * (vyr - vyl);
    double du41 = rhobar * (vzr - vzl);
    double du6 = byr - byl;
    double du5 = bzr - bzl;
    double t1 = betay * du6 + betaz * du5;
    double t2 = (prr - prl + (byave * du6 + bzave * du5) * pi4i + (gm - 2.0) * (bybar * du6 + bzbar * du5) * pi4i)
              / (gm - 1.0);
    double t3 = betaz * du6 - betay * du5;
    double s1 = du21;
    double s2 = betay * du31 + betaz * du41;
    double s3 = betaz * du31 - betay * du41;
    double p11 = alphs * cfast * sqrt(pi4 / rhobar);
    double p12 = -alphf * cs2 / cfast * sqrt(pi4 / rhobar);
    double p21 = alphf * (cfast2 - cs2 * (gm - 2.0) / (gm - 1.0));
    double p22 = alphs * (cslow2 - cs2 * (gm - 2.0) / (gm - 1.0));
    double q11 = alphf * cfast;
    double q12 = alphs * cslow;
    double q21 = -alphs * ca * sgnbx;
    double q22 = alphf * cs * sgnbx;
    double detp = p11 * p22 - p12 * p21;
    double detq = q11 * q22 - q12 * q21;
    double wpf = 0.5 * ((p22 * t1 - p12 * t2) / detp + (q22 * s1 - q12 * s2) / detq);
    double wmf = 0.5 * ((p22 * t1 - p12 * t2) / detp - (q22 * s1 - q12 * s2) / detq);
    double wps = 0.5 * ((-p21 * t1 + p11 * t2) / detp + (-q21 * s1 + q11 * s2) / detq);
    double wms = 0.5 * ((-p21 * t1 + p11 * t2) / detp - (-q21 * s1 + q11 * s2) / detq);
    double wpa = 0.5 * (sqrt(rhobar * pi4i) * t3 - sgnbx * s3);
    double wma = 0.5 * (sqrt(rhobar * pi4i) * t3 + sgnbx * s3);
    double wze = drho - alphf * (wpf + wmf) - alphs * (wps + wms);
    double fluxlro = rhol * vxl;
    double fluxlrx = rhol * vxl * vxl + prl + (-sq(bxbar) + sq(byl) + sq(bzl)) * pi8i;
    double fluxlry = rhol * vxl * vyl - bxbar * byl * pi4i;
    double fluxlrz = rhol * vxl * vzl - bxbar * bzl * pi4i;
    double fluxlby = vxl * byl - vyl * bxbar;
    double fluxlbz = vxl * bzl - vzl * bxbar;
    double fluxlee = rhol * vxl * hl - bxbar * (bxbar * vxl + byl * vyl + bzl * vzl) * pi4i;
    double fluxrro = rhor * vxr;
    double fluxrrx = rhor * vxr * vxr + prr + (-sq(bxbar) + sq(byr) + sq(bzr)) * pi8i;
    double fluxrry = rhor * vxr * vyr - bxbar * byr * pi4i;
    double fluxrrz = rhor * vxr * vzr - bxbar * bzr * pi4i;
    double fluxrby = vxr * byr - vyr * bxbar;
    double fluxrbz = vxr * bzr - vzr * bxbar;
    double fluxree = rhor * vxr * hr - bxbar * (bxbar * vxr + byr * vyr + bzr * vzr) * pi4i;
    double rpfro = alphf;
    double rpfrx = alphf * (vxbar + cfast);
    double rpfry = alphf * vybar - alphs * betay * ca * sgnbx;
    double rpfrz = alphf * vzbar - alphs * betaz * ca * sgnbx;
    double rpfby = alphs * betay * cfast * sqrt(pi4 / rhobar);
    double rpfbz = alphs * betaz * cfast * sqrt(pi4 / rhobar);
    double rpfee = alphf * (0.5 * (sq(vxbar) + sq(vybar) + sq(vzbar)) + delb2 + cfast * vxbar + cfast2 / (gm - 1.0)
                            + (cfast2 - cs2) * (gm - 2.0) / (gm - 1.0))
                 - alphs * ca * (betay * vybar + betaz * vzbar) * sgnbx;
    double rmfro = alphf;
    double rmfrx = alphf * (vxbar - cfast);
    double rmfry = alphf * vybar + alphs * betay * ca * sgnbx;
    double rmfrz = alphf * vzbar + alphs * betaz * ca * sgnbx;
    double rmfby = rpfby;
    double rmfbz = rpfbz;
    double rmfee = alphf * (0.5 * (sq(vxbar) + sq(vybar) + sq(vzbar)) + delb2 - cfast * vxbar + cfast2 / (gm - 1.0)
                            + (cfast2 - cs2) * (gm - 2.0) / (gm - 1.0))
                 + alphs * ca * (betay * vybar + betaz * vzbar) * sgnbx;
    double rpsro = alphs;
    double rpsrx = alphs * (vxbar + cslow);
    double rpsry = alphs * vybar + cs * sgnbx * alphf * betay;
    double rpsrz = alphs * vzbar + cs * sgnbx * alphf * betaz;
    double rpsby = -sqrt(pi4 / rhobar) * cs2 * alphf * betay / cfast;
    double rpsbz = -sqrt(pi4 / rhobar) * cs2 * alphf * betaz / cfast;
    double rpsee = alphs * (0.5 * (sq(vxbar) + sq(vybar) + sq(vzbar)) + delb2 + cslow * vxbar + cslow2 / (gm - 1.0)
                            + (cslow2 - cs2) * (gm - 2.0) / (gm - 1.0))
                 + alphf * cs * (betay * vybar + betaz * vzbar) * sgnbx;
    double rmsro = alphs;
    double rmsrx = alphs * (vxbar - cslow);
    double rmsry = alphs * vybar - cs * sgnbx * alphf * betay;
    double rmsrz = alphs * vzbar - cs * sgnbx * alphf * betaz;
    double rmsby = rpsby;
    double rmsbz = rpsbz;
    double rmsee = alphs * (0.5 * (sq(vxbar) + sq(vybar) + sq(vzbar)) + delb2 - cslow * vxbar + cslow2 / (gm - 1.0)
                            + (cslow2 - cs2) * (gm - 2.0) / (gm - 1.0))
                 - alphf * cs * (betay * vybar + betaz * vzbar) * sgnbx;
    double rparo = 0.0, rparx = 0.0;
    double rpary = -sgnbx * betaz;
    double rparz = sgnbx * betay;
    double rpaby = sqrt(pi4 / rhobar) * betaz;
    double rpabz = -sqrt(pi4 / rhobar) * betay;
    double rpaee = -(betaz * vybar - betay * vzbar) * sgnbx;
    double rmaro = 0.0, rmarx = 0.0;
    double rmary = -rpary, rmarz = -rparz;
    double rmaby = rpaby, rmabz = rpabz;
    double rmaee = -rpaee;
    double rzero = 1.0, rzerx = vxbar, rzery = vybar, rzerz = vzbar, rzeby = 0.0, rzebz = 0.0;
    double rzeee = 0.5 * (sq(vxbar) + sq(vybar) + sq(vzbar)) + delb2;
#define RSUM(fl, fr, a, b, c, d, e, f, g) \
    0.5 * (fl + fr + elpf * wpf * a + elmf * wmf * b + elps * wps * c + elms * wms * d + elpa * wpa * e + elma * wma * f + elze * wze * g)
    out[0] = RSUM(fluxlro, fluxrro, rpfro, rmfro, rpsro, rmsro, rparo, rmaro, rzero);
    out[1] = RSUM(fluxlee, fluxree, rpfee, rmfee, rpsee, rmsee, rpaee, rmaee, rzeee);
    out[2] = RSUM(fluxlrx, fluxrrx, rpfrx, rmfrx, rpsrx, rmsrx, rparx, rmarx, rzerx);
    out[3] = RSUM(fluxlry, fluxrry, rpfry, rmfry, rpsry, rmsry, rpary, rmary, rzery);
    out[4] = RSUM(fluxlrz, fluxrrz, rpfrz, rmfrz, rpsrz, rmsrz, rparz, rmarz, rzerz);
    out[5] = RSUM(fluxlby, fluxrby, rpfby, rmfby, rpsby, rmsby, rpaby, rmaby, rzeby);
    out[6] = RSUM(fluxlbz, fluxrbz, rpfbz, rmfbz, rpsbz, rmsbz, rpabz, rmabz, rzebz);
#undef RSUM
}

/* flmt of tvdminmod.f (statement function) */
static inline double flmt(double a, double b) {
    return fmaxF(0.0, fminF(b * signF(1.0, a), fabs(a))) * signF(1.0, a);
}

/* ================================================================== 1D */
/* cans1d/hdroe/tvdminmod.f:33-40 */
static void tvd1(const double *da, double *w1, double *w2, int ix) {
    for (int i = 2; i <= ix - 2; i++) {
        w1[I1(i)] = da[I1(i)] + 0.5 * flmt(da[I1(i + 1)] - da[I1(i)], da[I1(i)] - da[I1(i - 1)]);
        w2[I1(i)] = da[I1(i + 1)] - 0.5 * flmt(da[I1(i + 1)] - da[I1(i)], da[I1(i + 2)] - da[I1(i + 1)]);
    }
}

/* cans1d/hdroe/roe_m.f:2-136 */
void or1d_roe_m(double *ro, double *pr, double *vx, double *vy, double *by,
                const double *bx, const double *bxm, double dt, double gm,
                const double *dx, int ix) {
    const double pi = acos(-1.0);
    const double pi4 = 4.0 * pi;
    const double pi4i = 1.0 / pi4;
    const double pi8i = 5.0e-1 * pi4i;
    double *ee = dalloc(ix), *rx = dalloc(ix), *ry = dalloc(ix);
    double *f[5], *w[6][2];
    for (int q = 0; q < 5; q++) f[q] = dalloc(ix);
    for (int q = 0; q < 6; q++) { w[q][0] = dalloc(ix); w[q][1] = dalloc(ix); }
    double *roh = dalloc(ix), *eeh = dalloc(ix), *rxh = dalloc(ix), *ryh = dalloc(ix), *byh = dalloc(ix);
    double *prh = dalloc(ix), *vxh = dalloc(ix), *vyh = dalloc(ix);
    enum { RO, PR, VX, VY, BX, BY };
    for (int i = 1; i <= ix; i++) { /* :56-63 */
        rx[I1(i)] = ro[I1(i)] * vx[I1(i)];
        ry[I1(i)] = ro[I1(i)] * vy[I1(i)];
        double v2 = sq(vx[I1(i)]) + sq(vy[I1(i)]);
        double b2 = sq(bx[I1(i)]) + sq(by[I1(i)]);
        ee[I1(i)] = pr[I1(i)] / (gm - 1.0) + 0.5 * ro[I1(i)] * v2 + pi8i * b2;
    }
    for (int i = 1; i <= ix - 1; i++) { /* :67-80 */
        w[RO][0][I1(i)] = ro[I1(i)]; w[PR][0][I1(i)] = pr[I1(i)]; w[VX][0][I1(i)] = vx[I1(i)];
        w[VY][0][I1(i)] = vy[I1(i)]; w[BX][0][I1(i)] = bxm[I1(i)]; w[BY][0][I1(i)] = by[I1(i)];
        w[RO][1][I1(i)] = ro[I1(i + 1)]; w[PR][1][I1(i)] = pr[I1(i + 1)]; w[VX][1][I1(i)] = vx[I1(i + 1)];
        w[VY][1][I1(i)] = vy[I1(i + 1)]; w[BX][1][I1(i)] = bxm[I1(i)]; w[BY][1][I1(i)] = by[I1(i + 1)];
    }
    for (int i = 1; i <= ix - 1; i++) {
        double o[5];
        roeflux12(gm, w[RO][0][I1(i)], w[PR][0][I1(i)], w[VX][0][I1(i)], w[VY][0][I1(i)], w[BX][0][I1(i)], w[BY][0][I1(i)],
                  w[RO][1][I1(i)], w[PR][1][I1(i)], w[VX][1][I1(i)], w[VY][1][I1(i)], w[BX][1][I1(i)], w[BY][1][I1(i)], o);
        for (int q = 0; q < 5; q++) f[q][I1(i)] = o[q];
    }
    for (int i = 2; i <= ix - 1; i++) { /* :84-90 */
        roh[I1(i)] = ro[I1(i)] + 0.5 * dt * ((f[0][I1(i - 1)] - f[0][I1(i)]) / dx[I1(i)]);
        eeh[I1(i)] = ee[I1(i)] + 0.5 * dt * ((f[1][I1(i - 1)] - f[1][I1(i)]) / dx[I1(i)]);
        rxh[I1(i)] = rx[I1(i)] + 0.5 * dt * ((f[2][I1(i - 1)] - f[2][I1(i)]) / dx[I1(i)]);
        ryh[I1(i)] = ry[I1(i)] + 0.5 * dt * ((f[3][I1(i - 1)] - f[3][I1(i)]) / dx[I1(i)]);
        byh[I1(i)] = by[I1(i)] + 0.5 * dt * ((f[4][I1(i - 1)] - f[4][I1(i)]) / dx[I1(i)]);
    }
    for (int i = 2; i <= ix - 1; i++) { /* :94-100 */
        vxh[I1(i)] = rxh[I1(i)] / roh[I1(i)];
        vyh[I1(i)] = ryh[I1(i)] / roh[I1(i)];
        double v2 = sq(vxh[I1(i)]) + sq(vyh[I1(i)]);
        double b2 = sq(bx[I1(i)]) + sq(byh[I1(i)]);
        prh[I1(i)] = (gm - 1.0) * (eeh[I1(i)] - 0.5 * roh[I1(i)] * v2 - pi8i * b2);
    }
    tvd1(roh, w[RO][0], w[RO][1], ix);
    tvd1(prh, w[PR][0], w[PR][1], ix);
    tvd1(vxh, w[VX][0], w[VX][1], ix);
    tvd1(vyh, w[VY][0], w[VY][1], ix);
    tvd1(byh, w[BY][0], w[BY][1], ix);
    for (int i = 1; i <= ix - 1; i++) {
        double o[5];
        roeflux12(gm, w[RO][0][I1(i)], w[PR][0][I1(i)], w[VX][0][I1(i)], w[VY][0][I1(i)], w[BX][0][I1(i)], w[BY][0][I1(i)],
                  w[RO][1][I1(i)], w[PR][1][I1(i)], w[VX][1][I1(i)], w[VY][1][I1(i)], w[BX][1][I1(i)], w[BY][1][I1(i)], o);
        for (int q = 0; q < 5; q++) f[q][I1(i)] = o[q];
    }
    for (int i = 3; i <= ix - 2; i++) { /* :117-124 */
        ro[I1(i)] = ro[I1(i)] + dt * ((f[0][I1(i - 1)] - f[0][I1(i)]) / dx[I1(i)]);
        ee[I1(i)] = ee[I1(i)] + dt * ((f[1][I1(i - 1)] - f[1][I1(i)]) / dx[I1(i)]);
        rx[I1(i)] = rx[I1(i)] + dt * ((f[2][I1(i - 1)] - f[2][I1(i)]) / dx[I1(i)]);
        ry[I1(i)] = ry[I1(i)] + dt * ((f[3][I1(i - 1)] - f[3][I1(i)]) / dx[I1(i)]);
        by[I1(i)] = by[I1(i)] + dt * ((f[4][I1(i - 1)] - f[4][I1(i)]) / dx[I1(i)]);
    }
    for (int i = 3; i <= ix - 2; i++) { /* :126-133 */
        vx[I1(i)] = rx[I1(i)] / ro[I1(i)];
        vy[I1(i)] = ry[I1(i)] / ro[I1(i)];
        double v2 = sq(vx[I1(i)]) + sq(vy[I1(i)]);
        double b2 = sq(bx[I1(i)]) + sq(by[I1(i)]);
        pr[I1(i)] = (gm - 1.0) * (ee[I1(i)] - 0.5 * ro[I1(i)] * v2 - pi8i * b2);
    }
    free(ee); free(rx); free(ry);
    for (int q = 0; q < 5; q++) free(f[q]);
    for (int q = 0; q < 6; q++) { free(w[q][0]); free(w[q][1]); }
    free(roh); free(eeh); free(rxh); free(ryh); free(byh); free(prh); free(vxh); free(vyh);
}


/* cans1d/hdroe/roe_m3.f:2-151: roe_m with vz, bz (Bx constant) and the
 * roeflux_m3 flux */
void or1d_roe_m3(double *ro, double *pr, double *vx, double *vy, double *vz,
                 double *by, double *bz, const double *bx, const double *bxm,
                 double dt, double gm, const double *dx, int ix) {
    const double pi = acos(-1.0);
    const double pi4 = 4.0 * pi;
    const double pi4i = 1.0 / pi4;
    const double pi8i = 5.0e-1 * pi4i;
    enum { RO, PR, VX, VY, VZ, BX, BY, BZ, NW };
    enum { FRO, FEE, FRX, FRY, FRZ, FBY, FBZ, NF };
    double *ee = dalloc(ix), *rx = dalloc(ix), *ry = dalloc(ix), *rz = dalloc(ix);
    double *f[NF], *w[NW][2];
    for (int q = 0; q < NF; q++) f[q] = dalloc(ix);
    for (int q = 0; q < NW; q++) { w[q][0] = dalloc(ix); w[q][1] = dalloc(ix); }
    double *roh = dalloc(ix), *eeh = dalloc(ix), *rxh = dalloc(ix), *ryh = dalloc(ix), *rzh = dalloc(ix);
    double *byh = dalloc(ix), *bzh = dalloc(ix), *prh = dalloc(ix), *vxh = dalloc(ix), *vyh = dalloc(ix);
    double *vzh = dalloc(ix);
    for (int i = 1; i <= ix; i++) { /* :59-68 */
        rx[I1(i)] = ro[I1(i)] * vx[I1(i)];
        ry[I1(i)] = ro[I1(i)] * vy[I1(i)];
        rz[I1(i)] = ro[I1(i)] * vz[I1(i)];
        double v2 = sq(vx[I1(i)]) + sq(vy[I1(i)]) + sq(vz[I1(i)]);
        double b2 = sq(bx[I1(i)]) + sq(by[I1(i)]) + sq(bz[I1(i)]);
        ee[I1(i)] = pr[I1(i)] / (gm - 1.0) + 0.5 * ro[I1(i)] * v2 + pi8i * b2;
    }
    const double *prim[NW] = {ro, pr, vx, vy, vz, bxm, by, bz};
    for (int i = 1; i <= ix - 1; i++) /* :70-90 (bx at the face: bxm on both sides) */
        for (int q = 0; q < NW; q++) {
            w[q][0][I1(i)] = prim[q][I1(i)];
            w[q][1][I1(i)] = q == BX ? bxm[I1(i)] : prim[q][I1(i + 1)];
        }
#define FLUX3()                                                                                              \
    for (int i = 1; i <= ix - 1; i++) {                                                                      \
        double o[7];                                                                                         \
        roeflux3(0, gm, w[RO][0][I1(i)], w[PR][0][I1(i)], w[VX][0][I1(i)], w[VY][0][I1(i)], w[VZ][0][I1(i)], \
                 w[BX][0][I1(i)], w[BY][0][I1(i)], w[BZ][0][I1(i)], w[RO][1][I1(i)], w[PR][1][I1(i)],       \
                 w[VX][1][I1(i)], w[VY][1][I1(i)], w[VZ][1][I1(i)], w[BX][1][I1(i)], w[BY][1][I1(i)],       \
                 w[BZ][1][I1(i)], o);                                                                        \
        for (int q = 0; q < NF; q++) f[q][I1(i)] = o[q];                                                     \
    }
    FLUX3()
    for (int i = 2; i <= ix - 1; i++) { /* :94-101 */
        roh[I1(i)] = ro[I1(i)] + 0.5 * dt * ((f[FRO][I1(i - 1)] - f[FRO][I1(i)]) / dx[I1(i)]);
        eeh[I1(i)] = ee[I1(i)] + 0.5 * dt * ((f[FEE][I1(i - 1)] - f[FEE][I1(i)]) / dx[I1(i)]);
        rxh[I1(i)] = rx[I1(i)] + 0.5 * dt * ((f[FRX][I1(i - 1)] - f[FRX][I1(i)]) / dx[I1(i)]);
        ryh[I1(i)] = ry[I1(i)] + 0.5 * dt * ((f[FRY][I1(i - 1)] - f[FRY][I1(i)]) / dx[I1(i)]);
        rzh[I1(i)] = rz[I1(i)] + 0.5 * dt * ((f[FRZ][I1(i - 1)] - f[FRZ][I1(i)]) / dx[I1(i)]);
        byh[I1(i)] = by[I1(i)] + 0.5 * dt * ((f[FBY][I1(i - 1)] - f[FBY][I1(i)]) / dx[I1(i)]);
        bzh[I1(i)] = bz[I1(i)] + 0.5 * dt * ((f[FBZ][I1(i - 1)] - f[FBZ][I1(i)]) / dx[I1(i)]);
    }
    for (int i = 2; i <= ix - 1; i++) { /* :105-113 */
        vxh[I1(i)] = rxh[I1(i)] / roh[I1(i)];
        vyh[I1(i)] = ryh[I1(i)] / roh[I1(i)];
        vzh[I1(i)] = rzh[I1(i)] / roh[I1(i)];
        double v2 = sq(vxh[I1(i)]) + sq(vyh[I1(i)]) + sq(vzh[I1(i)]);
        double b2 = sq(bx[I1(i)]) + sq(byh[I1(i)]) + sq(bzh[I1(i)]);
        prh[I1(i)] = (gm - 1.0) * (eeh[I1(i)] - 0.5 * roh[I1(i)] * v2 - pi8i * b2);
    }
    tvd1(roh, w[RO][0], w[RO][1], ix);
    tvd1(prh, w[PR][0], w[PR][1], ix);
    tvd1(vxh, w[VX][0], w[VX][1], ix);
    tvd1(vyh, w[VY][0], w[VY][1], ix);
    tvd1(byh, w[BY][0], w[BY][1], ix);
    tvd1(vzh, w[VZ][0], w[VZ][1], ix);
    tvd1(bzh, w[BZ][0], w[BZ][1], ix);
    FLUX3()
#undef FLUX3
    for (int i = 3; i <= ix - 2; i++) { /* :128-137 */
        ro[I1(i)] = ro[I1(i)] + dt * ((f[FRO][I1(i - 1)] - f[FRO][I1(i)]) / dx[I1(i)]);
        ee[I1(i)] = ee[I1(i)] + dt * ((f[FEE][I1(i - 1)] - f[FEE][I1(i)]) / dx[I1(i)]);
        rx[I1(i)] = rx[I1(i)] + dt * ((f[FRX][I1(i - 1)] - f[FRX][I1(i)]) / dx[I1(i)]);
        ry[I1(i)] = ry[I1(i)] + dt * ((f[FRY][I1(i - 1)] - f[FRY][I1(i)]) / dx[I1(i)]);
        rz[I1(i)] = rz[I1(i)] + dt * ((f[FRZ][I1(i - 1)] - f[FRZ][I1(i)]) / dx[I1(i)]);
        by[I1(i)] = by[I1(i)] + dt * ((f[FBY][I1(i - 1)] - f[FBY][I1(i)]) / dx[I1(i)]);
        bz[I1(i)] = bz[I1(i)] + dt * ((f[FBZ][I1(i - 1)] - f[FBZ][I1(i)]) / dx[I1(i)]);
    }
    for (int i = 3; i <= ix - 2; i++) { /* :139-148 */
        vx[I1(i)] = rx[I1(i)] / ro[I1(i)];
        vy[I1(i)] = ry[I1(i)] / ro[I1(i)];
        vz[I1(i)] = rz[I1(i)] / ro[I1(i)];
        double v2 = sq(vx[I1(i)]) + sq(vy[I1(i)]) + sq(vz[I1(i)]);
        double b2 = sq(bx[I1(i)]) + sq(by[I1(i)]) + sq(bz[I1(i)]);
        pr[I1(i)] = (gm - 1.0) * (ee[I1(i)] - 0.5 * ro[I1(i)] * v2 - pi8i * b2);
    }
    free(ee); free(rx); free(ry); free(rz);
    for (int q = 0; q < NF; q++) free(f[q]);
    for (int q = 0; q < NW; q++) { free(w[q][0]); free(w[q][1]); }
    free(roh); free(eeh); free(rxh); free(ryh); free(rzh); free(byh); free(bzh);
    free(prh); free(vxh); free(vyh); free(vzh);
}

/* ================================================================== 2D */
/* cans2d/hdroe/tvdminmod.f:11-34 */
static void tvd2(int mdir, const double *da, double *w1, double *w2, int ix, int jx) {
    for (int j = 2; j <= jx - 2; j++)
        for (int i = 2; i <= ix - 2; i++) {
            if (mdir == 1) {
                w1[I2(i, j)] = da[I2(i, j)] + 0.5 * flmt(da[I2(i + 1, j)] - da[I2(i, j)], da[I2(i, j)] - da[I2(i - 1, j)]);
                w2[I2(i, j)] = da[I2(i + 1, j)] - 0.5 * flmt(da[I2(i + 1, j)] - da[I2(i, j)], da[I2(i + 2, j)] - da[I2(i + 1, j)]);
            } else {
                w1[I2(i, j)] = da[I2(i, j)] + 0.5 * flmt(da[I2(i, j + 1)] - da[I2(i, j)], da[I2(i, j)] - da[I2(i, j - 1)]);
                w2[I2(i, j)] = da[I2(i, j + 1)] - 0.5 * flmt(da[I2(i, j + 1)] - da[I2(i, j)], da[I2(i, j + 2)] - da[I2(i, j + 1)]);
            }
        }
}

/* roeflux_m over the 2D w arrays, j=1..jx-1, i=1..ix-1 (cans2d/hdroe/roeflux_m.f:16-17);
 * vn/vt/bn/bt select the sweep frame as the reference's call permutes them */
static void roeflux2_arr(double gm, double *const *w0, double *const *w1, int vn, int vt, int bn, int bt,
                         double *fro, double *fee, double *frn, double *frt, double *fbt, int ix, int jx) {
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix - 1; i++) {
            size_t p = I2(i, j);
            double o[5];
            roeflux12(gm, w0[0][p], w0[1][p], w0[vn][p], w0[vt][p], w0[bn][p], w0[bt][p],
                      w1[0][p], w1[1][p], w1[vn][p], w1[vt][p], w1[bn][p], w1[bt][p], o);
            fro[p] = o[0]; fee[p] = o[1]; frn[p] = o[2]; frt[p] = o[3]; fbt[p] = o[4];
        }
}

/* cans2d/hdroe/roe_m.f:2-208 */
void or2d_roe_m(double *ro, double *pr, double *vx, double *vy, double *bx,
                double *by, double *az, double dt, double gm,
                const double *dx, int ix, const double *dy, int jx) {
    const double pi = acos(-1.0);
    const double pi4 = 4.0 * pi;
    const double pi4i = 1.0 / pi4;
    const double pi8i = 5.0e-1 * pi4i;
    size_t n = (size_t)ix * jx;
    enum { RO, PR, VX, VY, BX, BY, NW };
    double *ee = dalloc(n), *rx = dalloc(n), *ry = dalloc(n);
    double *fro = dalloc(n), *fee = dalloc(n), *frx = dalloc(n), *fry = dalloc(n), *fbx = dalloc(n), *fby = dalloc(n);
    double *roh = dalloc(n), *eeh = dalloc(n), *rxh = dalloc(n), *ryh = dalloc(n);
    double *prh = dalloc(n), *vxh = dalloc(n), *vyh = dalloc(n), *bxh = dalloc(n), *byh = dalloc(n);
    double *w0[NW], *w1[NW];
    for (int q = 0; q < NW; q++) { w0[q] = dalloc(n); w1[q] = dalloc(n); }
    const double *prim[NW] = {ro, pr, vx, vy, bx, by};

    for (size_t p = 0; p < n; p++) { /* :32-40 */
        rx[p] = ro[p] * vx[p];
        ry[p] = ro[p] * vy[p];
        double v2 = sq(vx[p]) + sq(vy[p]);
        double b2 = sq(bx[p]) + sq(by[p]);
        ee[p] = pr[p] / (gm - 1.0) + 0.5 * ro[p] * v2 + pi8i * b2;
    }
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j);
            roh[p] = ro[p]; eeh[p] = ee[p]; rxh[p] = rx[p]; ryh[p] = ry[p]; bxh[p] = bx[p]; byh[p] = by[p];
        }
    /* x sweep, first order */
    for (int j = 1; j <= jx; j++)
        for (int i = 1; i <= ix - 1; i++)
            for (int q = 0; q < NW; q++) { w0[q][I2(i, j)] = prim[q][I2(i, j)]; w1[q][I2(i, j)] = prim[q][I2(i + 1, j)]; }
    roeflux2_arr(gm, w0, w1, VX, VY, BX, BY, fro, fee, frx, fry, fby, ix, jx);
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j), m = I2(i - 1, j);
            roh[p] = roh[p] + 0.5 * dt * ((fro[m] - fro[p]) / dx[I1(i)]);
            eeh[p] = eeh[p] + 0.5 * dt * ((fee[m] - fee[p]) / dx[I1(i)]);
            rxh[p] = rxh[p] + 0.5 * dt * ((frx[m] - frx[p]) / dx[I1(i)]);
            ryh[p] = ryh[p] + 0.5 * dt * ((fry[m] - fry[p]) / dx[I1(i)]);
            byh[p] = byh[p] + 0.5 * dt * ((fby[m] - fby[p]) / dx[I1(i)]);
        }
    /* y sweep, first order */
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix; i++)
            for (int q = 0; q < NW; q++) { w0[q][I2(i, j)] = prim[q][I2(i, j)]; w1[q][I2(i, j)] = prim[q][I2(i, j + 1)]; }
    roeflux2_arr(gm, w0, w1, VY, VX, BY, BX, fro, fee, fry, frx, fbx, ix, jx);
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j), m = I2(i, j - 1);
            roh[p] = roh[p] + 0.5 * dt * ((fro[m] - fro[p]) / dy[I1(j)]);
            eeh[p] = eeh[p] + 0.5 * dt * ((fee[m] - fee[p]) / dy[I1(j)]);
            rxh[p] = rxh[p] + 0.5 * dt * ((frx[m] - frx[p]) / dy[I1(j)]);
            ryh[p] = ryh[p] + 0.5 * dt * ((fry[m] - fry[p]) / dy[I1(j)]);
            bxh[p] = bxh[p] + 0.5 * dt * ((fbx[m] - fbx[p]) / dy[I1(j)]);
        }
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j);
            vxh[p] = rxh[p] / roh[p];
            vyh[p] = ryh[p] / roh[p];
            double v2 = sq(vxh[p]) + sq(vyh[p]);
            double b2 = sq(bxh[p]) + sq(byh[p]);
            prh[p] = (gm - 1.0) * (eeh[p] - 0.5 * roh[p] * v2 - pi8i * b2);
        }
    const double *half[NW] = {roh, prh, vxh, vyh, bxh, byh};
    /* MUSCL x (:155-182) */
    for (int q = 0; q < NW; q++) tvd2(1, half[q], w0[q], w1[q], ix, jx);
    roeflux2_arr(gm, w0, w1, VX, VY, BX, BY, fro, fee, frx, fry, fby, ix, jx);
    for (int j = 3; j <= jx - 2; j++)
        for (int i = 3; i <= ix - 2; i++) {
            size_t p = I2(i, j), m = I2(i - 1, j);
            ro[p] = ro[p] + dt * ((fro[m] - fro[p]) / dx[I1(i)]);
            ee[p] = ee[p] + dt * ((fee[m] - fee[p]) / dx[I1(i)]);
            rx[p] = rx[p] + dt * ((frx[m] - frx[p]) / dx[I1(i)]);
            ry[p] = ry[p] + dt * ((fry[m] - fry[p]) / dx[I1(i)]);
            by[p] = by[p] + dt * ((fby[m] - fby[p]) / dx[I1(i)]);
        }
    /* MUSCL y */
    for (int q = 0; q < NW; q++) tvd2(2, half[q], w0[q], w1[q], ix, jx);
    roeflux2_arr(gm, w0, w1, VY, VX, BY, BX, fro, fee, fry, frx, fbx, ix, jx);
    for (int j = 3; j <= jx - 2; j++)
        for (int i = 3; i <= ix - 2; i++) {
            size_t p = I2(i, j), m = I2(i, j - 1);
            ro[p] = ro[p] + dt * ((fro[m] - fro[p]) / dy[I1(j)]);
            ee[p] = ee[p] + dt * ((fee[m] - fee[p]) / dy[I1(j)]);
            rx[p] = rx[p] + dt * ((frx[m] - frx[p]) / dy[I1(j)]);
            ry[p] = ry[p] + dt * ((fry[m] - fry[p]) / dy[I1(j)]);
            bx[p] = bx[p] + dt * ((fbx[m] - fbx[p]) / dy[I1(j)]);
        }
    for (int j = 3; j <= jx - 2; j++) /* :187-193 */
        for (int i = 3; i <= ix - 2; i++) {
            size_t p = I2(i, j);
            double ez = -vxh[p] * byh[p] + vyh[p] * bxh[p];
            double saz = -ez;
            az[p] = az[p] + dt * saz;
        }
    for (int j = 3; j <= jx - 2; j++)
        for (int i = 3; i <= ix - 2; i++) {
            size_t p = I2(i, j);
            vx[p] = rx[p] / ro[p];
            vy[p] = ry[p] / ro[p];
            double v2 = sq(vx[p]) + sq(vy[p]);
            double b2 = sq(bx[p]) + sq(by[p]);
            pr[p] = (gm - 1.0) * (ee[p] - 0.5 * ro[p] * v2 - pi8i * b2);
        }
    double *tofree[] = {ee, rx, ry, fro, fee, frx, fry, fbx, fby, roh, eeh, rxh, ryh, prh, vxh, vyh, bxh, byh};
    for (size_t t = 0; t < sizeof(tofree) / sizeof(tofree[0]); t++) free(tofree[t]);
    for (int q = 0; q < NW; q++) { free(w0[q]); free(w1[q]); }
}


/* roeflux_m3 over the 2D w arrays, j=1..jx-1, i=1..ix-1, frame (vn, v1, v2,
 * bn, b1, b2) as the reference's calls permute it (cans2d/hdroe/roe_m3.f:85-86,
 * 124-125: x = (x, y, z), y = (y, x, z)) */
static void roeflux23_arr(double gm, double *const *w0, double *const *w1, int vn, int v1, int v2, int bn,
                          int b1, int b2, double *fro, double *fee, double *frn, double *fr1, double *fr2,
                          double *fb1, double *fb2, int ix, int jx) {
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix - 1; i++) {
            size_t p = I2(i, j);
            double o[7];
            roeflux3(0, gm, w0[0][p], w0[1][p], w0[vn][p], w0[v1][p], w0[v2][p], w0[bn][p], w0[b1][p], w0[b2][p],
                     w1[0][p], w1[1][p], w1[vn][p], w1[v1][p], w1[v2][p], w1[bn][p], w1[b1][p], w1[b2][p], o);
            fro[p] = o[0]; fee[p] = o[1]; frn[p] = o[2]; fr1[p] = o[3]; fr2[p] = o[4]; fb1[p] = o[5]; fb2[p] = o[6];
        }
}

/* cans2d/hdroe/roe_m3.f:2-236: roe_m with vz, bz and roeflux_m3 */
void or2d_roe_m3(double *ro, double *pr, double *vx, double *vy, double *vz,
                 double *bx, double *by, double *bz, double *az, double dt,
                 double gm, const double *dx, int ix, const double *dy, int jx) {
    const double pi = acos(-1.0);
    const double pi4 = 4.0 * pi;
    const double pi4i = 1.0 / pi4;
    const double pi8i = 5.0e-1 * pi4i;
    size_t n = (size_t)ix * jx;
    enum { RO, PR, VX, VY, VZ, BX, BY, BZ, NW };
    double *ee = dalloc(n), *rx = dalloc(n), *ry = dalloc(n), *rz = dalloc(n);
    double *fro = dalloc(n), *fee = dalloc(n), *frx = dalloc(n), *fry = dalloc(n), *frz = dalloc(n);
    double *fbx = dalloc(n), *fby = dalloc(n), *fbz = dalloc(n);
    double *roh = dalloc(n), *eeh = dalloc(n), *rxh = dalloc(n), *ryh = dalloc(n), *rzh = dalloc(n);
    double *prh = dalloc(n), *vxh = dalloc(n), *vyh = dalloc(n), *vzh = dalloc(n);
    double *bxh = dalloc(n), *byh = dalloc(n), *bzh = dalloc(n);
    double *w0[NW], *w1[NW];
    for (int q = 0; q < NW; q++) { w0[q] = dalloc(n); w1[q] = dalloc(n); }
    const double *prim[NW] = {ro, pr, vx, vy, vz, bx, by, bz};

    for (size_t p = 0; p < n; p++) { /* :34-44 */
        rx[p] = ro[p] * vx[p];
        ry[p] = ro[p] * vy[p];
        rz[p] = ro[p] * vz[p];
        double v2 = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
        double b2 = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
        ee[p] = pr[p] / (gm - 1.0) + 0.5 * ro[p] * v2 + pi8i * b2;
    }
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j);
            roh[p] = ro[p]; eeh[p] = ee[p]; rxh[p] = rx[p]; ryh[p] = ry[p]; rzh[p] = rz[p];
            bxh[p] = bx[p]; byh[p] = by[p]; bzh[p] = bz[p];
        }
    /* x sweep, first order (:64-99) */
    for (int j = 1; j <= jx; j++)
        for (int i = 1; i <= ix - 1; i++)
            for (int q = 0; q < NW; q++) { w0[q][I2(i, j)] = prim[q][I2(i, j)]; w1[q][I2(i, j)] = prim[q][I2(i + 1, j)]; }
    roeflux23_arr(gm, w0, w1, VX, VY, VZ, BX, BY, BZ, fro, fee, frx, fry, frz, fby, fbz, ix, jx);
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j), m = I2(i - 1, j);
            roh[p] = roh[p] + 0.5 * dt * ((fro[m] - fro[p]) / dx[I1(i)]);
            eeh[p] = eeh[p] + 0.5 * dt * ((fee[m] - fee[p]) / dx[I1(i)]);
            rxh[p] = rxh[p] + 0.5 * dt * ((frx[m] - frx[p]) / dx[I1(i)]);
            ryh[p] = ryh[p] + 0.5 * dt * ((fry[m] - fry[p]) / dx[I1(i)]);
            rzh[p] = rzh[p] + 0.5 * dt * ((frz[m] - frz[p]) / dx[I1(i)]);
            byh[p] = byh[p] + 0.5 * dt * ((fby[m] - fby[p]) / dx[I1(i)]);
            bzh[p] = bzh[p] + 0.5 * dt * ((fbz[m] - fbz[p]) / dx[I1(i)]);
        }
    /* y sweep, first order (:103-139) */
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix; i++)
            for (int q = 0; q < NW; q++) { w0[q][I2(i, j)] = prim[q][I2(i, j)]; w1[q][I2(i, j)] = prim[q][I2(i, j + 1)]; }
    roeflux23_arr(gm, w0, w1, VY, VX, VZ, BY, BX, BZ, fro, fee, fry, frx, frz, fbx, fbz, ix, jx);
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j), m = I2(i, j - 1);
            roh[p] = roh[p] + 0.5 * dt * ((fro[m] - fro[p]) / dy[I1(j)]);
            eeh[p] = eeh[p] + 0.5 * dt * ((fee[m] - fee[p]) / dy[I1(j)]);
            rxh[p] = rxh[p] + 0.5 * dt * ((frx[m] - frx[p]) / dy[I1(j)]);
            ryh[p] = ryh[p] + 0.5 * dt * ((fry[m] - fry[p]) / dy[I1(j)]);
            rzh[p] = rzh[p] + 0.5 * dt * ((frz[m] - frz[p]) / dy[I1(j)]);
            bxh[p] = bxh[p] + 0.5 * dt * ((fbx[m] - fbx[p]) / dy[I1(j)]);
            bzh[p] = bzh[p] + 0.5 * dt * ((fbz[m] - fbz[p]) / dy[I1(j)]);
        }
    for (int j = 2; j <= jx - 1; j++) /* :143-152 */
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j);
            vxh[p] = rxh[p] / roh[p];
            vyh[p] = ryh[p] / roh[p];
            vzh[p] = rzh[p] / roh[p];
            double v2 = sq(vxh[p]) + sq(vyh[p]) + sq(vzh[p]);
            double b2 = sq(bxh[p]) + sq(byh[p]) + sq(bzh[p]);
            prh[p] = (gm - 1.0) * (eeh[p] - 0.5 * roh[p] * v2 - pi8i * b2);
        }
    const double *half[NW] = {roh, prh, vxh, vyh, vzh, bxh, byh, bzh};
    /* MUSCL x (:158-181) */
    for (int q = 0; q < NW; q++) tvd2(1, half[q], w0[q], w1[q], ix, jx);
    roeflux23_arr(gm, w0, w1, VX, VY, VZ, BX, BY, BZ, fro, fee, frx, fry, frz, fby, fbz, ix, jx);
    for (int j = 3; j <= jx - 2; j++)
        for (int i = 3; i <= ix - 2; i++) {
            size_t p = I2(i, j), m = I2(i - 1, j);
            ro[p] = ro[p] + dt * ((fro[m] - fro[p]) / dx[I1(i)]);
            ee[p] = ee[p] + dt * ((fee[m] - fee[p]) / dx[I1(i)]);
            rx[p] = rx[p] + dt * ((frx[m] - frx[p]) / dx[I1(i)]);
            ry[p] = ry[p] + dt * ((fry[m] - fry[p]) / dx[I1(i)]);
            rz[p] = rz[p] + dt * ((frz[m] - frz[p]) / dx[I1(i)]);
            by[p] = by[p] + dt * ((fby[m] - fby[p]) / dx[I1(i)]);
            bz[p] = bz[p] + dt * ((fbz[m] - fbz[p]) / dx[I1(i)]);
        }
    /* MUSCL y (:185-210) */
    for (int q = 0; q < NW; q++) tvd2(2, half[q], w0[q], w1[q], ix, jx);
    roeflux23_arr(gm, w0, w1, VY, VX, VZ, BY, BX, BZ, fro, fee, fry, frx, frz, fbx, fbz, ix, jx);
    for (int j = 3; j <= jx - 2; j++)
        for (int i = 3; i <= ix - 2; i++) {
            size_t p = I2(i, j), m = I2(i, j - 1);
            ro[p] = ro[p] + dt * ((fro[m] - fro[p]) / dy[I1(j)]);
            ee[p] = ee[p] + dt * ((fee[m] - fee[p]) / dy[I1(j)]);
            rx[p] = rx[p] + dt * ((frx[m] - frx[p]) / dy[I1(j)]);
            ry[p] = ry[p] + dt * ((fry[m] - fry[p]) / dy[I1(j)]);
            rz[p] = rz[p] + dt * ((frz[m] - frz[p]) / dy[I1(j)]);
            bx[p] = bx[p] + dt * ((fbx[m] - fbx[p]) / dy[I1(j)]);
            bz[p] = bz[p] + dt * ((fbz[m] - fbz[p]) / dy[I1(j)]);
        }
    for (int j = 3; j <= jx - 2; j++) /* :214-220 */
        for (int i = 3; i <= ix - 2; i++) {
            size_t p = I2(i, j);
            double ez = -vxh[p] * byh[p] + vyh[p] * bxh[p];
            double saz = -ez;
            az[p] = az[p] + dt * saz;
        }
    for (int j = 3; j <= jx - 2; j++) /* :222-232 */
        for (int i = 3; i <= ix - 2; i++) {
            size_t p = I2(i, j);
            vx[p] = rx[p] / ro[p];
            vy[p] = ry[p] / ro[p];
            vz[p] = rz[p] / ro[p];
            double v2 = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
            double b2 = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
            pr[p] = (gm - 1.0) * (ee[p] - 0.5 * ro[p] * v2 - pi8i * b2);
        }
    double *tofree[] = {ee, rx, ry, rz, fro, fee, frx, fry, frz, fbx, fby, fbz, roh, eeh, rxh, ryh, rzh,
                        prh, vxh, vyh, vzh, bxh, byh, bzh};
    for (size_t t = 0; t < sizeof(tofree) / sizeof(tofree[0]); t++) free(tofree[t]);
    for (int q = 0; q < NW; q++) { free(w0[q]); free(w1[q]); }
}

/* ================================================================== 3D */
/* cans3d/hdroe/tvdminmod.f:15-55 (argument order of flmt as in the 3D file) */
static void tvd3(int mdir, const double *da, double *w1, double *w2, int ix, int jx, int kx) {
    for (int k = 2; k <= kx - 2; k++)
        for (int j = 2; j <= jx - 2; j++)
            for (int i = 2; i <= ix - 2; i++) {
                size_t s = mdir == 1 ? 1 : (mdir == 2 ? (size_t)ix : (size_t)ix * jx);
                size_t p = I3(i, j, k);
                w1[p] = da[p] + 0.5 * flmt(da[p] - da[p - s], da[p + s] - da[p]);
                w2[p] = da[p + s] - 0.5 * flmt(da[p + 2 * s] - da[p + s], da[p + s] - da[p]);
            }
}

enum { W_RO, W_PR, W_VX, W_VY, W_VZ, W_BX, W_BY, W_BZ, W_N };

/* roeflux_m over the 3D w arrays for i<=ix-1, j<=jx-1, k<=kx-1
 * (cans3d/hdroe/roeflux_m.f:14-16); (n,t1,t2) is the sweep frame */
static void roeflux3_arr(double gm, double *const *w0, double *const *w1, int vn, int v1, int v2, int bn, int b1, int b2,
                         double *fro, double *fee, double *frn, double *fr1, double *fr2, double *fb1, double *fb2,
                         int ix, int jx, int kx) {
    for (int k = 1; k <= kx - 1; k++)
        for (int j = 1; j <= jx - 1; j++)
            for (int i = 1; i <= ix - 1; i++) {
                size_t p = I3(i, j, k);
                double o[7];
                roeflux3(1, gm, w0[W_RO][p], w0[W_PR][p], w0[vn][p], w0[v1][p], w0[v2][p], w0[bn][p], w0[b1][p], w0[b2][p],
                         w1[W_RO][p], w1[W_PR][p], w1[vn][p], w1[v1][p], w1[v2][p], w1[bn][p], w1[b1][p], w1[b2][p], o);
                fro[p] = o[0]; fee[p] = o[1]; frn[p] = o[2]; fr1[p] = o[3]; fr2[p] = o[4]; fb1[p] = o[5]; fb2[p] = o[6];
            }
}

/* cans3d/hdroe/roe_m.f:2-343 */
void or3d_roe_m(double *ro, double *pr, double *vx, double *vy, double *vz,
                double *bx, double *by, double *bz, double dt, double gm,
                const double *dx, int ix, const double *dy, int jx,
                const double *dz, int kx) {
    const double pi = acos(-1.0);
    const double pi4 = 4.0 * pi;
    const double pi4i = 1.0 / pi4;
    const double pi8i = 5.0e-1 * pi4i;
    size_t n = (size_t)ix * jx * kx;
    double *ee = dalloc(n), *rx = dalloc(n), *ry = dalloc(n), *rz = dalloc(n);
    double *fro = dalloc(n), *fee = dalloc(n), *frx = dalloc(n), *fry = dalloc(n), *frz = dalloc(n);
    double *fbx = dalloc(n), *fby = dalloc(n), *fbz = dalloc(n);
    double *roh = dalloc(n), *eeh = dalloc(n), *rxh = dalloc(n), *ryh = dalloc(n), *rzh = dalloc(n);
    double *bxh = dalloc(n), *byh = dalloc(n), *bzh = dalloc(n);
    double *prh = dalloc(n), *vxh = dalloc(n), *vyh = dalloc(n), *vzh = dalloc(n);
    double *w0[W_N], *w1[W_N];
    for (int q = 0; q < W_N; q++) { w0[q] = dalloc(n); w1[q] = dalloc(n); }
    const double *prim[W_N] = {ro, pr, vx, vy, vz, bx, by, bz};
    const size_t sx = 1, sy = (size_t)ix, sz = (size_t)ix * jx;

    for (size_t p = 0; p < n; p++) { /* :22-33 */
        rx[p] = ro[p] * vx[p];
        ry[p] = ro[p] * vy[p];
        rz[p] = ro[p] * vz[p];
        double v2 = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
        double b2 = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
        ee[p] = pr[p] / (gm - 1.0) + 0.5 * ro[p] * v2 + pi8i * b2;
    }
    for (size_t p = 0; p < n; p++) { /* :34-47 */
        roh[p] = ro[p]; eeh[p] = ee[p]; rxh[p] = rx[p]; ryh[p] = ry[p]; rzh[p] = rz[p];
        bxh[p] = bx[p]; byh[p] = by[p]; bzh[p] = bz[p];
    }
#define FILL1(ilim, jlim, klim, s)                                                   \
    for (int k = 1; k <= (klim); k++)                                                \
        for (int j = 1; j <= (jlim); j++)                                            \
            for (int i = 1; i <= (ilim); i++)                                        \
                for (int q = 0; q < W_N; q++) {                                      \
                    w0[q][I3(i, j, k)] = prim[q][I3(i, j, k)];                      \
                    w1[q][I3(i, j, k)] = prim[q][I3(i, j, k) + (s)];                \
                }
#define HALF(h, fl, s, dd, idx)                                                       \
    h[p] = h[p] + 0.5 * dt * ((fl[p - (s)] - fl[p]) / dd[I1(idx)])
    /* x sweep */
    FILL1(ix - 1, jx, kx, sx)
    roeflux3_arr(gm, w0, w1, W_VX, W_VY, W_VZ, W_BX, W_BY, W_BZ, fro, fee, frx, fry, frz, fby, fbz, ix, jx, kx);
    for (int k = 2; k <= kx - 1; k++)
        for (int j = 2; j <= jx - 1; j++)
            for (int i = 2; i <= ix - 1; i++) {
                size_t p = I3(i, j, k);
                HALF(roh, fro, sx, dx, i); HALF(eeh, fee, sx, dx, i); HALF(rxh, frx, sx, dx, i);
                HALF(ryh, fry, sx, dx, i); HALF(rzh, frz, sx, dx, i); HALF(byh, fby, sx, dx, i);
                HALF(bzh, fbz, sx, dx, i);
            }
    /* y sweep: roeflux_m(fro,fee,fry,frz,frx,fbz,fbx,...,vyw,vzw,vxw,byw,bzw,bxw) */
    FILL1(ix, jx - 1, kx, sy)
    roeflux3_arr(gm, w0, w1, W_VY, W_VZ, W_VX, W_BY, W_BZ, W_BX, fro, fee, fry, frz, frx, fbz, fbx, ix, jx, kx);
    for (int k = 2; k <= kx - 1; k++)
        for (int j = 2; j <= jx - 1; j++)
            for (int i = 2; i <= ix - 1; i++) {
                size_t p = I3(i, j, k);
                HALF(roh, fro, sy, dy, j); HALF(eeh, fee, sy, dy, j); HALF(rxh, frx, sy, dy, j);
                HALF(ryh, fry, sy, dy, j); HALF(rzh, frz, sy, dy, j); HALF(bxh, fbx, sy, dy, j);
                HALF(bzh, fbz, sy, dy, j);
            }
    /* z sweep: roeflux_m(fro,fee,frz,frx,fry,fbx,fby,...,vzw,vxw,vyw,bzw,bxw,byw) */
    FILL1(ix, jx, kx - 1, sz)
    roeflux3_arr(gm, w0, w1, W_VZ, W_VX, W_VY, W_BZ, W_BX, W_BY, fro, fee, frz, frx, fry, fbx, fby, ix, jx, kx);
    for (int k = 2; k <= kx - 1; k++)
        for (int j = 2; j <= jx - 1; j++)
            for (int i = 2; i <= ix - 1; i++) {
                size_t p = I3(i, j, k);
                HALF(roh, fro, sz, dz, k); HALF(eeh, fee, sz, dz, k); HALF(rxh, frx, sz, dz, k);
                HALF(ryh, fry, sz, dz, k); HALF(rzh, frz, sz, dz, k); HALF(bxh, fbx, sz, dz, k);
                HALF(byh, fby, sz, dz, k);
            }
#undef HALF
#undef FILL1
    for (int k = 2; k <= kx - 1; k++)
        for (int j = 2; j <= jx - 1; j++)
            for (int i = 2; i <= ix - 1; i++) {
                size_t p = I3(i, j, k);
                vxh[p] = rxh[p] / roh[p];
                vyh[p] = ryh[p] / roh[p];
                vzh[p] = rzh[p] / roh[p];
                double v2 = sq(vxh[p]) + sq(vyh[p]) + sq(vzh[p]);
                double b2 = sq(bxh[p]) + sq(byh[p]) + sq(bzh[p]);
                prh[p] = (gm - 1.0) * (eeh[p] - 0.5 * roh[p] * v2 - pi8i * b2);
            }
    const double *half[W_N] = {roh, prh, vxh, vyh, vzh, bxh, byh, bzh};
#define FULL(u, fl, s, dd, idx) u[p] = u[p] + dt * ((fl[p - (s)] - fl[p]) / dd[I1(idx)])
    for (int q = 0; q < W_N; q++) tvd3(1, half[q], w0[q], w1[q], ix, jx, kx);
    roeflux3_arr(gm, w0, w1, W_VX, W_VY, W_VZ, W_BX, W_BY, W_BZ, fro, fee, frx, fry, frz, fby, fbz, ix, jx, kx);
    for (int k = 3; k <= kx - 2; k++)
        for (int j = 3; j <= jx - 2; j++)
            for (int i = 3; i <= ix - 2; i++) {
                size_t p = I3(i, j, k);
                FULL(ro, fro, sx, dx, i); FULL(ee, fee, sx, dx, i); FULL(rx, frx, sx, dx, i); FULL(ry, fry, sx, dx, i);
                FULL(rz, frz, sx, dx, i); FULL(by, fby, sx, dx, i); FULL(bz, fbz, sx, dx, i);
            }
    for (int q = 0; q < W_N; q++) tvd3(2, half[q], w0[q], w1[q], ix, jx, kx);
    roeflux3_arr(gm, w0, w1, W_VY, W_VZ, W_VX, W_BY, W_BZ, W_BX, fro, fee, fry, frz, frx, fbz, fbx, ix, jx, kx);
    for (int k = 3; k <= kx - 2; k++)
        for (int j = 3; j <= jx - 2; j++)
            for (int i = 3; i <= ix - 2; i++) {
                size_t p = I3(i, j, k);
                FULL(ro, fro, sy, dy, j); FULL(ee, fee, sy, dy, j); FULL(rx, frx, sy, dy, j); FULL(ry, fry, sy, dy, j);
                FULL(rz, frz, sy, dy, j); FULL(bx, fbx, sy, dy, j); FULL(bz, fbz, sy, dy, j);
            }
    for (int q = 0; q < W_N; q++) tvd3(3, half[q], w0[q], w1[q], ix, jx, kx);
    roeflux3_arr(gm, w0, w1, W_VZ, W_VX, W_VY, W_BZ, W_BX, W_BY, fro, fee, frz, frx, fry, fbx, fby, ix, jx, kx);
    for (int k = 3; k <= kx - 2; k++)
        for (int j = 3; j <= jx - 2; j++)
            for (int i = 3; i <= ix - 2; i++) {
                size_t p = I3(i, j, k);
                FULL(ro, fro, sz, dz, k); FULL(ee, fee, sz, dz, k); FULL(rx, frx, sz, dz, k); FULL(ry, fry, sz, dz, k);
                FULL(rz, frz, sz, dz, k); FULL(bx, fbx, sz, dz, k); FULL(by, fby, sz, dz, k);
            }
#undef FULL
    for (int k = 3; k <= kx - 2; k++)
        for (int j = 3; j <= jx - 2; j++)
            for (int i = 3; i <= ix - 2; i++) {
                size_t p = I3(i, j, k);
                vx[p] = rx[p] / ro[p];
                vy[p] = ry[p] / ro[p];
                vz[p] = rz[p] / ro[p];
                double v2 = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
                double b2 = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
                pr[p] = (gm - 1.0) * (ee[p] - 0.5 * ro[p] * v2 - pi8i * b2);
            }
    double *tofree[] = {ee, rx, ry, rz, fro, fee, frx, fry, frz, fbx, fby, fbz, roh, eeh, rxh, ryh, rzh,
                        bxh, byh, bzh, prh, vxh, vyh, vzh};
    for (size_t t = 0; t < sizeof(tofree) / sizeof(tofree[0]); t++) free(tofree[t]);
    for (int q = 0; q < W_N; q++) { free(w0[q]); free(w1[q]); }
}

/* ================================================================== helpers
 * roeflux_m and tvdminmod with the reference argument lists: w arrays are
 * (n,2) / (ix,jx,2) / (ix,jx,kx,2), the left state in the first half. */
/* cans1d/hdroe/roeflux_m.f:2-3, faces i=1..ix-1 (:48) */
void or1d_roeflux_m(double *fro, double *fee, double *frx, double *fry, double *fby, double gm, const double *row,
                    const double *prw, const double *vxw, const double *vyw, const double *bxw, const double *byw,
                    int ix) {
    for (int i = 1; i <= ix - 1; i++) {
        size_t p = I1(i), q = p + (size_t)ix;
        double o[5];
        roeflux12(gm, row[p], prw[p], vxw[p], vyw[p], bxw[p], byw[p], row[q], prw[q], vxw[q], vyw[q], bxw[q], byw[q],
                  o);
        fro[p] = o[0]; fee[p] = o[1]; frx[p] = o[2]; fry[p] = o[3]; fby[p] = o[4];
    }
}
/* cans2d/hdroe/roeflux_m.f:1-2, faces j=1..jx-1, i=1..ix-1 (:16-17) */
void or2d_roeflux_m(double *fro, double *fee, double *frx, double *fry, double *fby, double gm, const double *row,
                    const double *prw, const double *vxw, const double *vyw, const double *bxw, const double *byw,
                    int ix, int jx) {
    const size_t n = (size_t)ix * jx;
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix - 1; i++) {
            size_t p = I2(i, j), q = p + n;
            double o[5];
            roeflux12(gm, row[p], prw[p], vxw[p], vyw[p], bxw[p], byw[p], row[q], prw[q], vxw[q], vyw[q], bxw[q],
                      byw[q], o);
            fro[p] = o[0]; fee[p] = o[1]; frx[p] = o[2]; fry[p] = o[3]; fby[p] = o[4];
        }
}
/* cans3d/hdroe/roeflux_m.f:6-7, faces k,j,i up to n-1 (:26-28) */
void or3d_roeflux_m(double *fro, double *fee, double *frx, double *fry, double *frz, double *fby, double *fbz,
                    double gm, const double *row, const double *prw, const double *vxw, const double *vyw,
                    const double *vzw, const double *bxw, const double *byw, const double *bzw, int ix, int jx,
                    int kx) {
    const size_t n = (size_t)ix * jx * kx;
    for (int k = 1; k <= kx - 1; k++)
        for (int j = 1; j <= jx - 1; j++)
            for (int i = 1; i <= ix - 1; i++) {
                size_t p = I3(i, j, k), q = p + n;
                double o[7];
                roeflux3(1, gm, row[p], prw[p], vxw[p], vyw[p], vzw[p], bxw[p], byw[p], bzw[p], row[q], prw[q],
                         vxw[q], vyw[q], vzw[q], bxw[q], byw[q], bzw[q], o);
                fro[p] = o[0]; fee[p] = o[1]; frx[p] = o[2]; fry[p] = o[3]; frz[p] = o[4]; fby[p] = o[5];
                fbz[p] = o[6];
            }
}
/* cans1d/hdroe/tvdminmod.f:2 */
void or1d_tvdminmod(const double *da, double *daw, int ix) { tvd1(da, daw, daw + ix, ix); }
/* cans2d/hdroe/tvdminmod.f:1 */
void or2d_tvdminmod(int mdir, const double *da, double *daw, int ix, int jx) {
    tvd2(mdir, da, daw, daw + (size_t)ix * jx, ix, jx);
}
/* cans3d/hdroe/tvdminmod.f:4 */
void or3d_tvdminmod(int mdir, const double *da, double *daw, int ix, int jx, int kx) {
    tvd3(mdir, da, daw, daw + (size_t)ix * jx * kx, ix, jx, kx);
}
