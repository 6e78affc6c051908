/* CPU restatement of the OpenCFD-EC2D residual / implicit-update path.
 * TEST INFRASTRUCTURE ONLY (see ec2d_oracle.h).  Build with
 * -O2 -ffp-contract=off so that every expression is rounded exactly as the
 * reference Fortran evaluates it (left to right, no fused multiply-add).
 * Each function cites the reference lines it restates.
 */
#include "ec2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define LAP 2
#define PRT 0.9
/* Fortran default-real (single precision) literals widened to double */
#define SP(x) ((double)(float)(x))

/* ---- index helpers (Fortran bounds) ------------------------------------ */
#define GI(B, i, j) ((size_t)((i) + 1) + (size_t)((B)->nx + 3) * (size_t)((j) + 1))
#define UI(B, nv, m, i, j) ((size_t)((m) - 1) + (size_t)(nv) * GI(B, i, j))
#define NI(B, i, j) ((size_t)((i) + 1) + (size_t)((B)->nx + 4) * (size_t)((j) + 1))
#define PI_(B, i, j) ((size_t)((i) - 1) + (size_t)(B)->nx * (size_t)((j) - 1))
#define CI(B, i, j) ((size_t)((i) - 1) + (size_t)((B)->nx - 1) * (size_t)((j) - 1))
#define RI(B, nv, m, i, j) ((size_t)((m) - 1) + (size_t)(nv) * CI(B, i, j))
#define DUI(B, nv, m, i, j) ((size_t)((m) - 1) + (size_t)(nv) * PI_(B, i, j))
#define WI(B, i, j) ((size_t)(i) + (size_t)((B)->nx + 1) * (size_t)(j))
#define UNI(B, nv, m, i, j) ((size_t)((m) - 1) + (size_t)(nv) * WI(B, i, j))

/* ------------------------------------------------------------------------
 * reconstruction schemes: sub_Residual.f90:477-487 (minmod), :591-623
 * (scheme_fP), :626-659 (scheme_fm)
 * ---------------------------------------------------------------------- */
static double minmod(double a, double b) {
  if (a * b <= 0.0) return 0.0;
  if (fabs(a) <= fabs(b)) return a;
  return b;
}

static double scheme_fP(double u1, double u2, double u3, double u4, int flag) {
  const double k = 1.0, b = 2.0, k3 = 1.0 / 3.0, ep = 1.0e-6;
  switch (flag) {
    case 0: return u2;
    case 2: return (-u1 + 5.0 * u2 + 2.0 * u3) / 6.0;
    case 1: return u2 + 0.5 * minmod(u2 - u1, u3 - u2);
    case 3: {
      double IS1 = (u2 - u1) * (u2 - u1), IS2 = (u3 - u2) * (u3 - u2);
      double t1 = 1.0e-6 + IS1, t2 = 1.0e-6 + IS2;
      double a1 = 1.0 / (3.0 * (t1 * t1)), a2 = 2.0 / (3.0 * (t2 * t2));
      double w1 = a1 / (a1 + a2), w2 = a2 / (a1 + a2);
      return w1 * (-u1 / 2.0 + 3.0 * u2 / 2.0) + w2 * (u2 / 2.0 + u3 / 2.0);
    }
    case 4:
      return u2 + 0.25 * ((1.0 - k) * minmod(u2 - u1, b * (u3 - u2)) + (1.0 + k) * minmod(u3 - u2, b * (u2 - u1)));
    case 5: {
      double up = u3 - u2, um = u2 - u1;
      double s = (2.0 * up * um + ep) / (up * up + um * um + ep);
      return u2 + 0.25 * s * ((1.0 - k3 * s) * um + (1.0 + k3 * s) * up);
    }
    case 6: {
      double r1 = (u2 - u1 + ep) / (u3 - u2 + ep), r2 = (u3 - u2 + ep) / (u4 - u3 + ep);
      double f1 = 0.8 - 0.175 / r2 + 0.375 * r1;
      double f = fmax(0.0, fmin(fmin(2.0, f1), 2.0 * r1));
      return u2 + 0.5 * f * (u3 - u2);
    }
  }
  return u2;
}

static double scheme_fm(double u1, double u2, double u3, double u4, int flag) {
  const double k = 1.0 / 3.0, b = 1.0, k3 = 1.0 / 3.0, ep = 1.0e-6;
  switch (flag) {
    case 0: return u3;
    case 2: return (2.0 * u2 + 5.0 * u3 - u4) / 6.0;
    case 1: return u3 - 0.5 * minmod(u3 - u2, u4 - u3);
    case 3: {
      double IS1 = (u3 - u4) * (u3 - u4), IS2 = (u2 - u3) * (u2 - u3);
      double t1 = 1.0e-6 + IS1, t2 = 1.0e-6 + IS2;
      double a1 = 1.0 / (3.0 * (t1 * t1)), a2 = 2.0 / (3.0 * (t2 * t2));
      double w1 = a1 / (a1 + a2), w2 = a2 / (a1 + a2);
      return w1 * (-u4 / 2.0 + 3.0 * u3 / 2.0) + w2 * (u3 / 2.0 + u2 / 2.0);
    }
    case 4:
      return u3 - 0.25 * ((1.0 - k) * minmod(u4 - u3, b * (u3 - u2)) + (1.0 + k) * minmod(u3 - u2, b * (u4 - u3)));
    case 5: {
      double up = u4 - u3, um = u3 - u2;
      double s = (2.0 * up * um + ep) / (up * up + um * um + ep);
      return u3 - 0.25 * s * ((1.0 - k3 * s) * up + (1.0 + k3 * s) * um);
    }
    case 6: {
      double r1 = (u4 - u3 + ep) / (u3 - u2 + ep), r2 = (u3 - u2 + ep) / (u2 - u1 + ep);
      double f1 = 0.8 - 0.175 / r2 + 0.375 * r1;
      double f = fmax(0.0, fmin(fmin(2.0, f1), 2.0 * r1));
      return u3 - 0.5 * f * (u3 - u2);
    }
  }
  return u3;
}

/* U0[k][m]: k = point i-2..i+1, m = variable.  sub_Residual.f90:505-514 */
static void recon_original(double U0[4][4], double UL[4], double UR[4], int flag) {
  for (int m = 0; m < 4; m++) {
    UL[m] = scheme_fP(U0[0][m], U0[1][m], U0[2][m], U0[3][m], flag);
    UR[m] = scheme_fm(U0[0][m], U0[1][m], U0[2][m], U0[3][m], flag);
  }
}

/* sub_Residual.f90:517-528 */
static void recon_conservative(double U0[4][4], double UL[4], double UR[4], double gamma, int flag) {
  double QL[4], QR[4];
  for (int m = 0; m < 4; m++) {
    QL[m] = scheme_fP(U0[0][m], U0[1][m], U0[2][m], U0[3][m], flag);
    QR[m] = scheme_fm(U0[0][m], U0[1][m], U0[2][m], U0[3][m], flag);
  }
  UL[0] = QL[0]; UL[1] = QL[1] / UL[0]; UL[2] = QL[2] / UL[0];
  UL[3] = (QL[3] - (UL[1] * QL[1] + UL[2] * QL[2]) * 0.5) * (gamma - 1.0);
  UR[0] = QR[0]; UR[1] = QR[1] / UR[0]; UR[2] = QR[2] / UR[0];
  UR[3] = (QR[3] - (UR[1] * QR[1] + UR[2] * QR[2]) * 0.5) * (gamma - 1.0);
}

/* sub_Residual.f90:531-586 */
static void recon_characteristic(double U0[4][4], double UL[4], double UR[4], double gamma, int flag) {
  double Uh[4], S[4][4], S1[4][4], V0[4][4], VL[4], VR[4], QL[4], QR[4];
  for (int m = 0; m < 4; m++) Uh[m] = 0.5 * (U0[1][m] + U0[2][m]);
  double d1 = Uh[0], u1 = Uh[1] / d1, v1 = Uh[2] / d1;
  double p1 = (Uh[3] - (Uh[1] * u1 + Uh[2] * v1) * 0.5) * (gamma - 1.0);
  double c1 = sqrt(gamma * p1 / d1);
  double V2 = (u1 * u1 + v1 * v1) * 0.5;
  double tmp1 = (gamma - 1.0) / c1;
  double tmp3 = (gamma - 1.0) / (c1 * c1);
  double tmp5 = 1.0 / (2.0 * c1);
  double tmp0 = 1.0 / tmp3;
  S[0][0] = V2 - tmp0; S[0][1] = -u1; S[0][2] = -v1; S[0][3] = 1.0;
  S[1][0] = -v1; S[1][1] = 0.0; S[1][2] = 1.0; S[1][3] = 0.0;
  S[2][0] = -u1 - V2 * tmp1; S[2][1] = 1.0 + tmp1 * u1; S[2][2] = tmp1 * v1; S[2][3] = -tmp1;
  S[3][0] = -u1 + V2 * tmp1; S[3][1] = 1.0 - tmp1 * u1; S[3][2] = -tmp1 * v1; S[3][3] = tmp1;
  S1[0][0] = -tmp3; S1[0][1] = 0.0; S1[0][2] = -tmp5; S1[0][3] = tmp5;
  S1[1][0] = -tmp3 * u1; S1[1][1] = 0.0; S1[1][2] = 0.5 - u1 * tmp5; S1[1][3] = 0.5 + u1 * tmp5;
  S1[2][0] = -tmp3 * v1; S1[2][1] = 1.0; S1[2][2] = -v1 * tmp5; S1[2][3] = v1 * tmp5;
  S1[3][0] = -tmp3 * V2; S1[3][1] = v1; S1[3][2] = (c1 * u1 - V2 - tmp0) * tmp5;
  S1[3][3] = (c1 * u1 + V2 + tmp0) * tmp5;
  for (int k = 0; k < 4; k++)
    for (int m = 0; m < 4; m++) {
      double s = 0.0;
      for (int j = 0; j < 4; j++) s = s + S[m][j] * U0[k][j];
      V0[k][m] = s;
    }
  for (int m = 0; m < 4; m++) {
    VL[m] = scheme_fP(V0[0][m], V0[1][m], V0[2][m], V0[3][m], flag);
    VR[m] = scheme_fm(V0[0][m], V0[1][m], V0[2][m], V0[3][m], flag);
  }
  for (int m = 0; m < 4; m++) {
    double a = 0.0, b = 0.0;
    for (int j = 0; j < 4; j++) {
      a = a + S1[m][j] * VL[j];
      b = b + S1[m][j] * VR[j];
    }
    QL[m] = a;
    QR[m] = b;
  }
  UL[0] = QL[0]; UL[1] = QL[1] / UL[0]; UL[2] = QL[2] / UL[0];
  UL[3] = (QL[3] - (UL[1] * QL[1] + UL[2] * QL[2]) * 0.5) * (gamma - 1.0);
  UR[0] = QR[0]; UR[1] = QR[1] / UR[0]; UR[2] = QR[2] / UR[0];
  UR[3] = (QR[3] - (UR[1] * QR[1] + UR[2] * QR[2]) * 0.5) * (gamma - 1.0);
}

/* ------------------------------------------------------------------------
 * interface fluxes: sub_flux_split.f90.  Q = (rho, u_n, u_t, p).
 * ---------------------------------------------------------------------- */
/* :16-93 (hllc = 0 HLL, 1 HLLC) */
static void flux_hll_hllc(const double QL[4], const double QR[4], double Fx[4], double gamma, int hllc) {
  double dl = QL[0], uul = QL[1], vvl = QL[2], pl = QL[3];
  double dr = QR[0], uur = QR[1], vvr = QR[2], pr = QR[3];
  double UL[4], UR[4], Fl[4], Fr[4];
  UL[0] = dl; UL[1] = dl * uul; UL[2] = dl * vvl; UL[3] = pl / (gamma - 1.0) + dl * (uul * uul + vvl * vvl) * 0.5;
  UR[0] = dr; UR[1] = dr * uur; UR[2] = dr * vvr; UR[3] = pr / (gamma - 1.0) + dr * (uur * uur + vvr * vvr) * 0.5;
  double al = sqrt(gamma * pl / dl), ar = sqrt(gamma * pr / dr);
  double p_pvrs = 0.5 * (pl + pr) - (uur - uul) * (dl + dr) * (al + ar) * 0.125;
  double p_star = fmax(0.0, p_pvrs);
  double qql, qqr;
  if (p_star <= pl) qql = 1.0;
  else qql = sqrt(1.0 + ((gamma + 1.0) / (2.0 * gamma)) * (p_star / pl - 1.0));
  if (p_star <= pr) qqr = 1.0;
  else qqr = sqrt(1.0 + ((gamma + 1.0) / (2.0 * gamma)) * (p_star / pr - 1.0));
  double Sl = uul - al * qql, Sr = uur + ar * qqr;
  Fl[0] = UL[1]; Fl[1] = UL[1] * uul + pl; Fl[2] = UL[2] * uul; Fl[3] = uul * (UL[3] + pl);
  Fr[0] = UR[1]; Fr[1] = UR[1] * uur + pr; Fr[2] = UR[2] * uur; Fr[3] = uur * (UR[3] + pr);
  if (!hllc) {
    if (Sl >= 0) { for (int m = 0; m < 4; m++) Fx[m] = Fl[m]; }
    else if (Sr <= 0) { for (int m = 0; m < 4; m++) Fx[m] = Fr[m]; }
    else for (int m = 0; m < 4; m++) Fx[m] = (Sr * Fl[m] - Sl * Fr[m] + Sl * Sr * (UR[m] - UL[m])) / (Sr - Sl);
    return;
  }
  double S_star = (pr - pl + dl * uul * (Sl - uul) - dr * uur * (Sr - uur)) / (dl * (Sl - uul) - dr * (Sr - uur));
  double tmpl = dl * (Sl - uul) / (Sl - S_star), tmpr = dr * (Sr - uur) / (Sr - S_star);
  double Fsl[4], Fsr[4], tmp;
  Fsl[0] = Fl[0] + Sl * (tmpl - UL[0]);
  Fsl[1] = Fl[1] + Sl * (tmpl * S_star - UL[1]);
  Fsl[2] = Fl[2] + Sl * (tmpl * vvl - UL[2]);
  tmp = UL[3] / dl + (S_star - uul) * (S_star + pl / (dl * (Sl - uul)));
  Fsl[3] = Fl[3] + Sl * (tmpl * tmp - UL[3]);
  Fsr[0] = Fr[0] + Sr * (tmpr - UR[0]);
  Fsr[1] = Fr[1] + Sr * (tmpr * S_star - UR[1]);
  Fsr[2] = Fr[2] + Sr * (tmpr * vvr - UR[2]);
  tmp = UR[3] / dr + (S_star - uur) * (S_star + pr / (dr * (Sr - uur)));
  Fsr[3] = Fr[3] + Sr * (tmpr * tmp - UR[3]);
  const double *src = (Sl >= 0) ? Fl : (Sr <= 0) ? Fr : (S_star >= 0) ? Fsl : Fsr;
  for (int m = 0; m < 4; m++) Fx[m] = src[m];
}

/* :99-136 */
static void flux_sw(const double QL[4], const double QR[4], double Fx[4], double g) {
  double dl = QL[0], uul = QL[1], vvl = QL[2], pl = QL[3];
  double dr = QR[0], uur = QR[1], vvr = QR[2], pr = QR[3];
  double al = sqrt(g * pl / dl), ar = sqrt(g * pr / dr);
  double tmp1 = 2.0 * (g - 1.0);
  double tmp3 = (3.0 - g) / (2.0 * (g - 1.0));
  double E1P = (uul + fabs(uul)) * 0.5, E2P = (uul - al + fabs(uul - al)) * 0.5, E3P = (uul + al + fabs(uul + al)) * 0.5;
  double tmp0 = dl / (2.0 * g);
  double fp[4], fm[4];
  fp[0] = tmp0 * (tmp1 * E1P + E2P + E3P);
  fp[1] = tmp0 * (tmp1 * E1P * uul + E2P * (uul - al) + E3P * (uul + al));
  fp[2] = tmp0 * (tmp1 * E1P * vvl + E2P * vvl + E3P * vvl);
  fp[3] = tmp0 * (E1P * (g - 1.0) * (uul * uul + vvl * vvl) + E2P * ((uul - al) * (uul - al) + vvl * vvl) * 0.5 +
                  E3P * ((uul + al) * (uul + al) + vvl * vvl) * 0.5 + tmp3 * al * al * (E2P + E3P));
  double E1M = (uur - fabs(uur)) * 0.5, E2M = (uur - ar - fabs(uur - ar)) * 0.5, E3M = (uur + ar - fabs(uur + ar)) * 0.5;
  tmp0 = dr / (2.0 * g);
  fm[0] = tmp0 * (tmp1 * E1M + E2M + E3M);
  fm[1] = tmp0 * (tmp1 * E1M * uur + E2M * (uur - ar) + E3M * (uur + ar));
  fm[2] = tmp0 * (tmp1 * E1M * vvr + E2M * vvr + E3M * vvr);
  fm[3] = tmp0 * (E1M * (g - 1.0) * (uur * uur + vvr * vvr) + E2M * ((uur - ar) * (uur - ar) + vvr * vvr) * 0.5 +
                  E3M * ((uur + ar) * (uur + ar) + vvr * vvr) * 0.5 + tmp3 * ar * ar * (E2M + E3M));
  for (int m = 0; m < 4; m++) Fx[m] = fp[m] + fm[m];
}

/* :139-189 */
static void flux_vanleer(const double QL[4], const double QR[4], double Fx[4], double g) {
  double dl = QL[0], uul = QL[1], vvl = QL[2], pl = QL[3];
  double dr = QR[0], uur = QR[1], vvr = QR[2], pr = QR[3];
  double al = sqrt(g * pl / dl), ar = sqrt(g * pr / dr);
  double Ml = uul / al, Mr = uur / ar;
  double fp[4] = {0, 0, 0, 0}, fm[4] = {0, 0, 0, 0};
  if (Ml >= 1.0) {
    fp[0] = dl * uul; fp[1] = dl * uul * uul + pl; fp[2] = dl * uul * vvl;
    fp[3] = uul * (g * pl / (g - 1.0) + 0.5 * dl * (uul * uul + vvl * vvl));
  } else if (fabs(Ml) < 1.0) {
    double Mp = 0.25 * (1.0 + Ml) * (1.0 + Ml);
    double tmp0 = dl * al * Mp;
    double w = (g - 1.0) * uul + 2.0 * al;
    fp[0] = tmp0; fp[1] = tmp0 * w / g; fp[2] = tmp0 * vvl;
    fp[3] = tmp0 * (w * w * 0.5 / (g * g - 1.0) + 0.5 * (vvl * vvl));
  } else if (Ml <= -1.0) {
    fp[0] = fp[1] = fp[2] = fp[3] = 0.0;
  }
  if (Mr >= 1.0) {
    fm[0] = fm[1] = fm[2] = fm[3] = 0.0;
  } else if (fabs(Mr) < 1.0) {
    double Mm = -0.25 * (Mr - 1.0) * (Mr - 1.0);
    double tmp0 = dr * ar * Mm;
    double w = (g - 1.0) * uur - 2.0 * ar;
    fm[0] = tmp0; fm[1] = tmp0 * w / g; fm[2] = tmp0 * vvr;
    fm[3] = tmp0 * (w * w * 0.5 / (g * g - 1.0) + 0.5 * (vvr * vvr));
  } else if (Mr <= -1.0) {
    fm[0] = dr * uur; fm[1] = dr * uur * uur + pr; fm[2] = dr * uur * vvr;
    fm[3] = uur * (g * pr / (g - 1.0) + 0.5 * dr * (uur * uur + vvr * vvr));
  }
  for (int m = 0; m < 4; m++) Fx[m] = fp[m] + fm[m];
}

/* :195-261 (AUSM+) */
static void flux_ausm(const double QL[4], const double QR[4], double Fx[4], double g) {
  double dl = QL[0], uul = QL[1], vvl = QL[2], pl = QL[3];
  double dr = QR[0], uur = QR[1], vvr = QR[2], pr = QR[3];
  double hl = g * pl / ((g - 1.0) * dl) + 0.5 * (uul * uul + vvl * vvl);
  double hr = g * pr / ((g - 1.0) * dr) + 0.5 * (uur * uur + vvr * vvr);
  double al = sqrt(g * pl / dl), ar = sqrt(g * pr / dr);
  double a = 0.5 * (al + ar);
  double Ml = uul / a, Mr = uur / a;
  double Mp4, Mm4, Pp5, Pm5;
  if (fabs(Ml) >= 1.0) Mp4 = 0.5 * (Ml + fabs(Ml));
  else Mp4 = 0.25 * (Ml + 1.0) * (Ml + 1.0) + 0.125 * (Ml * Ml - 1.0) * (Ml * Ml - 1.0);
  if (fabs(Mr) >= 1.0) Mm4 = 0.5 * (Mr - fabs(Mr));
  else Mm4 = -0.25 * (Mr - 1.0) * (Mr - 1.0) - 0.125 * (Mr * Mr - 1.0) * (Mr * Mr - 1.0);
  double M = Mp4 + Mm4;
  double mp = dl * a * fmax(0.0, M);
  double mm = dr * a * fmin(0.0, M);
  double xm = mp + mm;
  double dm = (M >= 0.0) ? a * fabs(M) * dl : a * fabs(M) * dr;
  if (fabs(Ml) >= 1.0) Pp5 = 0.5 * (Ml + fabs(Ml)) / Ml;
  else Pp5 = 0.25 * (Ml + 1.0) * (Ml + 1.0) * (2.0 - Ml) + 3.0 * Ml * (Ml * Ml - 1.0) * (Ml * Ml - 1.0) / 16.0;
  if (fabs(Mr) >= 1.0) Pm5 = 0.5 * (Mr - fabs(Mr)) / Mr;
  else Pm5 = 0.25 * (Mr - 1.0) * (Mr - 1.0) * (2.0 + Mr) - 3.0 * Mr * (Mr * Mr - 1.0) * (Mr * Mr - 1.0) / 16.0;
  double p = pl * Pp5 + pr * Pm5;
  double fp[4] = {1.0, uul, vvl, hl}, fm[4] = {1.0, uur, vvr, hr};
  Fx[0] = 0.5 * xm * (fp[0] + fm[0]) - 0.5 * dm * (fm[0] - fp[0]);
  Fx[1] = 0.5 * xm * (fp[1] + fm[1]) - 0.5 * dm * (fm[1] - fp[1]) + p;
  Fx[2] = 0.5 * xm * (fp[2] + fm[2]) - 0.5 * dm * (fm[2] - fp[2]);
  Fx[3] = 0.5 * xm * (fp[3] + fm[3]) - 0.5 * dm * (fm[3] - fp[3]);
}

/* :265-324 (Roe with Harten entropy fix, delta = 0.1) */
static void flux_roe(const double QL[4], const double QR[4], double Fx[4], double g) {
  const double delt = 0.1;
  double dl = QL[0], uul = QL[1], vvl = QL[2], pl = QL[3];
  double dr = QR[0], uur = QR[1], vvr = QR[2], pr = QR[3];
  double hl = g * pl / ((g - 1.0) * dl) + 0.5 * (uul * uul + vvl * vvl);
  double hr = g * pr / ((g - 1.0) * dr) + 0.5 * (uur * uur + vvr * vvr);
  double Fl[4], Fr[4], lam[4], arr[4];
  Fl[0] = dl * uul; Fl[1] = dl * uul * uul + pl; Fl[2] = dl * uul * vvl; Fl[3] = uul * dl * hl;
  Fr[0] = dr * uur; Fr[1] = dr * uur * uur + pr; Fr[2] = dr * uur * vvr; Fr[3] = uur * dr * hr;
  double tmpp = sqrt(dr / dl);
  double avd = 0.25 * dl * (1.0 + tmpp) * (1.0 + tmpp);
  double avu = (uul + tmpp * uur) / (1.0 + tmpp);
  double avv = (vvl + tmpp * vvr) / (1.0 + tmpp);
  double avh = (hl + tmpp * hr) / (1.0 + tmpp);
  double ava = sqrt((g - 1.0) * (avh - 0.5 * (avu * avu + avv * avv)));
  lam[0] = fabs(avu - ava); lam[1] = fabs(avu); lam[2] = fabs(avu); lam[3] = fabs(avu + ava);
  for (int k = 0; k < 4; k++)
    if (lam[k] < delt) lam[k] = (lam[k] * lam[k] + delt * delt) / (2.0 * delt);
  arr[0] = (avd * ava * (-uur + uul) + pr - pl) / (2.0 * ava * ava);
  arr[1] = dr - dl - (pr - pl) / (ava * ava);
  arr[2] = avd * (vvr - vvl) / ava;
  arr[3] = (avd * ava * (uur - uul) + pr - pl) / (2.0 * ava * ava);
  Fx[0] = 0.5 * (Fl[0] + Fr[0]) - 0.5 * (lam[0] * arr[0] + lam[1] * arr[1] + lam[3] * arr[3]);
  Fx[1] = 0.5 * (Fl[1] + Fr[1]) - 0.5 * (lam[0] * arr[0] * (avu - ava) + lam[1] * arr[1] * avu + lam[3] * arr[3] * (avu + ava));
  Fx[2] = 0.5 * (Fl[2] + Fr[2]) - 0.5 * (lam[0] * arr[0] * avv + lam[1] * arr[1] * avv + lam[2] * arr[2] + lam[3] * arr[3] * avv);
  Fx[3] = 0.5 * (Fl[3] + Fr[3]) - 0.5 * (lam[0] * arr[0] * (avh - avu * ava) + lam[1] * arr[1] * 0.5 * (avu * avu + avv * avv) +
                                        lam[2] * arr[2] * avv + lam[3] * arr[3] * (avh + avu * ava));
}

static void face_flux(int iflux, const double QL[4], const double QR[4], double F0[4], double g) {
  switch (iflux) {
    case 1: flux_sw(QL, QR, F0, g); break;
    case 2: flux_hll_hllc(QL, QR, F0, g, 0); break;
    case 3: flux_hll_hllc(QL, QR, F0, g, 1); break;
    case 5: flux_vanleer(QL, QR, F0, g); break;
    case 6: flux_ausm(QL, QR, F0, g); break;
    case 4: flux_roe(QL, QR, F0, g); break;
    default: F0[0] = F0[1] = F0[2] = F0[3] = 0.0;
  }
}

/* ------------------------------------------------------------------------
 * per-block temporaries of Flow_Var (opencfd_ec2d_v1.5.4.f90:132-135)
 * ---------------------------------------------------------------------- */
typedef struct {
  double *d, *uu, *v, *T, *p, *cc; /* d..p ghosted (-1:nx+1,-1:ny+1), cc (1:nx-1,1:ny-1) */
  double *Fi, *Fj;                  /* (4,1:nx,1:ny) */
} Flow;

static void flow_alloc(const OBlock *B, Flow *f) {
  size_t g = (size_t)(B->nx + 3) * (B->ny + 3);
  f->d = calloc(g, sizeof(double));
  f->uu = calloc(g, sizeof(double));
  f->v = calloc(g, sizeof(double));
  f->T = calloc(g, sizeof(double));
  f->p = calloc(g, sizeof(double));
  f->cc = calloc((size_t)(B->nx - 1) * (B->ny - 1), sizeof(double));
  f->Fi = calloc((size_t)4 * B->nx * B->ny, sizeof(double));
  f->Fj = calloc((size_t)4 * B->nx * B->ny, sizeof(double));
}
static void flow_free(Flow *f) {
  free(f->d); free(f->uu); free(f->v); free(f->T); free(f->p); free(f->cc); free(f->Fi); free(f->Fj);
}
#define FX(B, m, i, j) ((size_t)((m) - 1) + 4 * PI_(B, i, j))

/* sub_Residual.f90:114-152 */
static int comput_duvtpckw(const OMesh *M, const OParams *P, const OBlock *B, Flow *f, int mb, OError *err) {
  int nx = B->nx, ny = B->ny, NV = M->nvar;
  double p00 = 1.0 / (P->gamma * P->Ma * P->Ma);
  for (int j = 1 - LAP; j <= ny + LAP - 1; j++)
    for (int i = 1 - LAP; i <= nx + LAP - 1; i++) {
      size_t g = GI(B, i, j);
      double d = B->U[UI(B, NV, 1, i, j)];
      f->d[g] = d;
      f->uu[g] = B->U[UI(B, NV, 2, i, j)] / d;
      f->v[g] = B->U[UI(B, NV, 3, i, j)] / d;
      f->T[g] = (B->U[UI(B, NV, 4, i, j)] - 0.5 * d * (f->uu[g] * f->uu[g] + f->v[g] * f->v[g])) / (P->Cv * d);
      f->p[g] = p00 * d * f->T[g];
    }
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++) {
      double T = f->T[GI(B, i, j)];
      if (T < 1.0e-5) {
        if (err) { err->code = 1; err->i = i; err->j = j; err->block = mb; err->T = T; }
        return 1;
      }
      f->cc[CI(B, i, j)] = sqrt(T) / P->Ma;
    }
  return 0;
}

/* sub_Residual.f90:670-689 */
static void get_viscous(const OParams *P, OBlock *B, const Flow *f) {
  double Tsb = 110.4 / P->T_inf;
  for (int j = 0; j <= B->ny; j++)
    for (int i = 0; i <= B->nx; i++) {
      double T = f->T[GI(B, i, j)];
      B->Amu[GI(B, i, j)] = 1.0 / P->Re * (1.0 + Tsb) * sqrt(T * T * T) / (Tsb + T);
    }
}

/* sub_turbulence_SA.f90:180-220 */
static void amut_boundary(OBlock *B) {
  int nx = B->nx, ny = B->ny;
  for (int k = 0; k < B->nsub; k++) {
    const int *s = B->bc + 9 * k;
    if (s[6] != -10) continue;
    int face = s[1], ib = s[2], ie = s[3], jb = s[4], je = s[5];
    if (face == 1) for (int j = jb; j <= je - 1; j++) B->Amu_t[GI(B, 0, j)] = -B->Amu_t[GI(B, 1, j)];
    else if (face == 3) for (int j = jb; j <= je - 1; j++) B->Amu_t[GI(B, nx, j)] = -B->Amu_t[GI(B, nx - 1, j)];
    else if (face == 2) for (int i = ib; i <= ie - 1; i++) B->Amu_t[GI(B, i, 0)] = -B->Amu_t[GI(B, i, 1)];
    else if (face == 4) for (int i = ib; i <= ie - 1; i++) B->Amu_t[GI(B, i, ny)] = -B->Amu_t[GI(B, i, ny - 1)];
  }
}

/* x**n for a constant integer n as flang lowers it: a left-to-right chain of
 * multiplications ((x*x)*x)*...  */
static double ipow(double x, int n) {
  double r = x;
  for (int k = 1; k < n; k++) r = r * x;
  return r;
}

/* ------------------------------------------------------------------------
 * Baldwin-Lomax: sub_turbulence_BL.f90:6-275
 * ---------------------------------------------------------------------- */
/* BL_model_1d (:218-275) on one wall-normal line, points 1..ny-1 (1-based
 * arrays of ny+1).  The real*8 parameters AP, Ccp, Ckleb, AKT, AK are
 * initialised from default-real (single precision) literals (:222), so they
 * carry the float values; Cwk is a double literal. */
static void bl_model_1d(int ny, const double *yy, const double *Amu, double *Amu_t, const double *d,
                        const double *u, const double *v, const double *om) {
  const double AP = (double)26.0f, Ccp = (double)1.6f, Ckleb = (double)0.3f, Cwk = 0.25, AKT = (double)0.4f,
               AK = (double)0.0168f;
  double TW = fabs(Amu[1] * om[1]);
  for (int j = 1; j <= ny - 1; j++)
    if (fabs(Amu[j] * om[j]) > TW) TW = fabs(Amu[j] * om[j]);
  const double Ret = sqrt(d[1] * TW) / Amu[1];
  double Fmax = 0.0, etamax = 0.0, Udif = 0.0;
  const int nb = ny - 1; /* Ny_boundary (:236) */
  for (int j = 1; j <= nb; j++) {
    double uu = sqrt(u[j] * u[j] + v[j] * v[j]);
    if (uu > Udif) Udif = uu;
    double etap = yy[j] * Ret;
    double FF = yy[j] * fabs(om[j]) * (1.0 - exp(-etap / AP));
    if (FF > Fmax) {
      Fmax = FF;
      etamax = yy[j];
    }
  }
  const double a = etamax * Fmax, b = Cwk * etamax * Udif * Udif / Fmax;
  const double Fwak = b < a ? b : a; /* dmin1; only Fmax = 0 makes b NaN, and then every Amu_t is 0 */
  int iflag = 0;
  for (int j = 1; j <= ny - 1; j++) {
    double etap = Ret * yy[j];
    double bl = AKT * yy[j] * (1.0 - exp(-etap / AP));
    double visti = d[j] * bl * bl * fabs(om[j]);
    double Fkleb = 1.0 / (1.0 + 5.5 * ipow(Ckleb * yy[j] / etamax, 6));
    double visto = AK * Ccp * d[j] * Fwak * Fkleb;
    if (fabs(visto) < fabs(visti)) iflag = 1;
    Amu_t[j] = iflag == 0 ? visti : visto;
  }
}

/* turbulence_model_BL (:6-205): vorticity at cell centres, then one line per
 * wall cell, walls in subface order; a cell reached by several walls keeps
 * the smallest mu_t; each wall then sets its ghost rows/columns, over the
 * subface's node range (so a wall ending inside a face also copies its last
 * column into the neighbouring interior one, :146-149). */
static void turbulence_bl(OBlock *B, const Flow *f) {
  const int nx = B->nx, ny = B->ny;
  int wall = 0;
  for (int k = 0; k < B->nsub; k++)
    if (B->bc[9 * k + 6] == -10) wall = 1;
  if (!wall) return;
  const size_t gsz = (size_t)(nx + 3) * (ny + 3);
  for (size_t q = 0; q < gsz; q++) B->Amu_t[q] = 0.0;
  double *om = calloc(gsz, sizeof(double));
  int *flag1 = calloc(gsz, sizeof(int));
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++) {
      double xi = B->x1[GI(B, i + 1, j)] - B->x1[GI(B, i - 1, j)];
      double yi = B->y1[GI(B, i + 1, j)] - B->y1[GI(B, i - 1, j)];
      double xj = B->x1[GI(B, i, j + 1)] - B->x1[GI(B, i, j - 1)];
      double yj = B->y1[GI(B, i, j + 1)] - B->y1[GI(B, i, j - 1)];
      double Jac = 1.0 / (xi * yj - xj * yi);
      double ix = Jac * yj, iy = -Jac * xj, jx = -Jac * yi, jy = Jac * xi;
      double ui = f->uu[GI(B, i + 1, j)] - f->uu[GI(B, i - 1, j)];
      double vi = f->v[GI(B, i + 1, j)] - f->v[GI(B, i - 1, j)];
      double uj = f->uu[GI(B, i, j + 1)] - f->uu[GI(B, i, j - 1)];
      double vj = f->v[GI(B, i, j + 1)] - f->v[GI(B, i, j - 1)];
      double vx = vi * ix + vj * jx;
      double uy = ui * iy + uj * jy;
      om[GI(B, i, j)] = fabs(vx - uy);
    }
  const int nmax = (nx > ny ? nx : ny) + 1;
  double *yy = calloc(nmax, sizeof(double)), *a1 = calloc(nmax, sizeof(double)), *t1 = calloc(nmax, sizeof(double));
  double *d1 = calloc(nmax, sizeof(double)), *u1 = calloc(nmax, sizeof(double)), *v1 = calloc(nmax, sizeof(double));
  double *o1 = calloc(nmax, sizeof(double));
  double *At = B->Amu_t;
  for (int k = 0; k < B->nsub; k++) {
    const int *bc = &B->bc[9 * k];
    if (bc[6] != -10) continue;
    const int face = bc[1], ist = bc[2], iend = bc[3], jst = bc[4], jend = bc[5];
    const int along_i = face == 1 || face == 3; /* lines run along i */
    const int n1 = along_i ? nx : ny;           /* Ny of BL_model_1d */
    const int lo = along_i ? jst : ist, hi = along_i ? jend - 1 : iend - 1;
    for (int m = 0; m < nmax; m++) a1[m] = t1[m] = 0.0;
    for (int L = lo; L <= hi; L++) {
      double x0, y0;
      if (face == 1) { x0 = (B->x1[GI(B, 1, L)] + B->x1[GI(B, 0, L)]) * 0.5; y0 = (B->y1[GI(B, 1, L)] + B->y1[GI(B, 0, L)]) * 0.5; }
      else if (face == 3) { x0 = (B->x1[GI(B, nx - 1, L)] + B->x1[GI(B, nx, L)]) * 0.5; y0 = (B->y1[GI(B, nx - 1, L)] + B->y1[GI(B, nx, L)]) * 0.5; }
      else if (face == 2) { x0 = (B->x1[GI(B, L, 1)] + B->x1[GI(B, L, 0)]) * 0.5; y0 = (B->y1[GI(B, L, 1)] + B->y1[GI(B, L, 0)]) * 0.5; }
      else { x0 = (B->x1[GI(B, L, ny - 1)] + B->x1[GI(B, L, ny)]) * 0.5; y0 = (B->y1[GI(B, L, ny - 1)] + B->y1[GI(B, L, ny)]) * 0.5; }
      for (int p = 1; p <= n1 - 1; p++) {
        /* point p of the line, counted from the wall */
        int i, j;
        if (face == 1) { i = p; j = L; }
        else if (face == 3) { i = nx - p; j = L; }
        else if (face == 2) { i = L; j = p; }
        else { i = L; j = ny - p; }
        size_t q = GI(B, i, j);
        double dx = B->x1[q] - x0, dy = B->y1[q] - y0;
        yy[p] = sqrt(dx * dx + dy * dy);
        u1[p] = f->uu[q]; v1[p] = f->v[q]; d1[p] = f->d[q]; o1[p] = om[q];
        a1[p] = B->Amu[q];
      }
      bl_model_1d(n1, yy, a1, t1, d1, u1, v1, o1);
      for (int p = 1; p <= n1 - 1; p++) {
        int i, j;
        if (face == 1) { i = p; j = L; }
        else if (face == 3) { i = nx - p; j = L; }
        else if (face == 2) { i = L; j = p; }
        else { i = L; j = ny - p; }
        size_t q = GI(B, i, j);
        if (flag1[q] == 0) {
          flag1[q] = 1;
          At[q] = t1[p];
        } else {
          At[q] = fmin(At[q], t1[p]);
        }
      }
    }
    /* the wall's ghost rows / columns (:81-84, 111-114, 146-149, 180-183) */
    if (face == 1 || face == 3) {
      for (int i = -1; i <= nx + 1; i++) At[GI(B, i, jst - 1)] = At[GI(B, i, jst)];
      for (int i = -1; i <= nx + 1; i++) At[GI(B, i, jend)] = At[GI(B, i, jend - 1)];
      if (face == 1) {
        for (int j = jst; j <= jend; j++) At[GI(B, nx, j)] = At[GI(B, nx - 1, j)];
        for (int j = jst; j <= jend; j++) At[GI(B, 0, j)] = -At[GI(B, 1, j)];
      } else {
        for (int j = jst; j <= jend; j++) At[GI(B, 0, j)] = At[GI(B, 1, j)];
        for (int j = jst; j <= jend; j++) At[GI(B, nx, j)] = -At[GI(B, nx - 1, j)];
      }
    } else {
      if (face == 2) {
        for (int i = ist; i <= iend; i++) At[GI(B, i, 0)] = -At[GI(B, i, 1)];
        for (int i = ist; i <= iend; i++) At[GI(B, i, ny)] = At[GI(B, i, ny - 1)];
      } else {
        for (int i = ist; i <= iend; i++) At[GI(B, i, ny)] = -At[GI(B, i, ny - 1)];
        for (int i = ist; i <= iend; i++) At[GI(B, i, 0)] = At[GI(B, i, 1)];
      }
      for (int j = -1; j <= ny + 1; j++) At[GI(B, ist - 1, j)] = At[GI(B, ist, j)];
      for (int j = -1; j <= ny + 1; j++) At[GI(B, iend, j)] = At[GI(B, iend - 1, j)];
    }
  }
  free(yy); free(a1); free(t1); free(d1); free(u1); free(v1); free(o1);
  free(om); free(flag1);
}

/* sub_turbulence_SA.f90:4-175 */
static void turbulence_sa(const OMesh *M, const OParams *P, OBlock *B, const Flow *f) {
  const double SA_sigma = 2.0 / 3.0, Cv1 = 7.1, Cb1 = 0.1355, Cb2 = 0.622, SA_k = 0.41;
  const double Cw1 = Cb1 / (SA_k * SA_k) + (1.0 + Cb2) / SA_sigma, Cw2 = 0.3, Cw3 = 2.0, Ct3 = 1.3, Ct4 = 0.5;
  const double Cv1_3 = Cv1 * Cv1 * Cv1;
  const double Cw3_6 = ipow(Cw3, 6);
  const double c_diff = 0.5 * (1.0 + Cb2) / SA_sigma;
  int nx = B->nx, ny = B->ny, NV = M->nvar;
  size_t nvt = (size_t)(nx + 1) * (ny + 1);
  double *vt = malloc(nvt * sizeof(double));
  double *Fv = calloc((size_t)nx * ny, sizeof(double)), *Fv2 = calloc((size_t)nx * ny, sizeof(double));
#define VT(i, j) vt[(size_t)(i) + (size_t)(nx + 1) * (size_t)(j)]
#define FV(A, i, j) A[PI_(B, i, j)]
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      size_t g = GI(B, i, j);
      VT(i, j) = B->U[UI(B, NV, 5, i, j)];
      double X = f->d[g] * VT(i, j) / B->Amu[g];
      double X3 = ipow(X, 3);
      double fv1 = X3 / (X3 + Cv1_3);
      B->Amu_t[g] = fv1 * f->d[g] * VT(i, j);
    }
  amut_boundary(B);
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx; i++) {
      double s0 = B->si[PI_(B, i, j)], n1 = B->ni1[PI_(B, i, j)], n2 = B->ni2[PI_(B, i, j)];
      double vn1 = f->uu[GI(B, i - 1, j)] * n1 + f->v[GI(B, i - 1, j)] * n2;
      double vn2 = f->uu[GI(B, i, j)] * n1 + f->v[GI(B, i, j)] * n2;
      double vfi = -0.5 * ((vn1 + fabs(vn1)) * VT(i - 1, j) + (vn2 - fabs(vn2)) * VT(i, j)) * s0;
      double v0 = c_diff * (VT(i, j) + B->Amu[GI(B, i, j)] / f->d[GI(B, i, j)] + VT(i - 1, j) +
                            B->Amu[GI(B, i - 1, j)] / f->d[GI(B, i - 1, j)]);
      const double *x1 = B->x1, *y1 = B->y1;
      double Dix = x1[GI(B, i, j)] - x1[GI(B, i - 1, j)];
      double Diy = y1[GI(B, i, j)] - y1[GI(B, i - 1, j)];
      double Djx = (x1[GI(B, i - 1, j + 1)] + x1[GI(B, i, j + 1)] - x1[GI(B, i - 1, j - 1)] - x1[GI(B, i, j - 1)]) * 0.25;
      double Djy = (y1[GI(B, i - 1, j + 1)] + y1[GI(B, i, j + 1)] - y1[GI(B, i - 1, j - 1)] - y1[GI(B, i, j - 1)]) * 0.25;
      double Ds = 1.0 / (Dix * Djy - Djx * Diy);
      double Div = VT(i, j) - VT(i - 1, j);
      double Djv = (VT(i - 1, j + 1) + VT(i, j + 1) - VT(i - 1, j - 1) - VT(i, j - 1)) * 0.25;
      double vtx = (Div * Djy - Djv * Diy) * Ds;
      double vty = (-Div * Djx + Djv * Dix) * Ds;
      FV(Fv, i, j) = vfi + v0 * (vtx * n1 + vty * n2) * s0;
      FV(Fv2, i, j) = (vtx * n1 + vty * n2) * s0;
    }
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++) {
      double v0 = (VT(i, j) + B->Amu[GI(B, i, j)] / f->d[GI(B, i, j)]) * Cb2 / SA_sigma;
      B->Res[RI(B, NV, 5, i, j)] = FV(Fv, i + 1, j) - FV(Fv, i, j) - v0 * (FV(Fv2, i + 1, j) - FV(Fv2, i, j));
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx - 1; i++) {
      double s0 = B->sj[PI_(B, i, j)], n1 = B->nj1[PI_(B, i, j)], n2 = B->nj2[PI_(B, i, j)];
      double vn1 = f->uu[GI(B, i, j - 1)] * n1 + f->v[GI(B, i, j - 1)] * n2;
      double vn2 = f->uu[GI(B, i, j)] * n1 + f->v[GI(B, i, j)] * n2;
      double vfi = -0.5 * ((vn1 + fabs(vn1)) * VT(i, j - 1) + (vn2 - fabs(vn2)) * VT(i, j)) * s0;
      double v0 = c_diff * (VT(i, j) + B->Amu[GI(B, i, j)] / f->d[GI(B, i, j)] + VT(i, j - 1) +
                            B->Amu[GI(B, i, j - 1)] / f->d[GI(B, i, j - 1)]);
      const double *x1 = B->x1, *y1 = B->y1;
      double Dix = (x1[GI(B, i + 1, j - 1)] + x1[GI(B, i + 1, j)] - x1[GI(B, i - 1, j - 1)] - x1[GI(B, i - 1, j)]) * 0.25;
      double Diy = (y1[GI(B, i + 1, j - 1)] + y1[GI(B, i + 1, j)] - y1[GI(B, i - 1, j - 1)] - y1[GI(B, i - 1, j)]) * 0.25;
      double Djx = x1[GI(B, i, j)] - x1[GI(B, i, j - 1)];
      double Djy = y1[GI(B, i, j)] - y1[GI(B, i, j - 1)];
      double Ds = 1.0 / (Dix * Djy - Djx * Diy);
      double Div = (VT(i + 1, j - 1) + VT(i + 1, j) - VT(i - 1, j - 1) - VT(i - 1, j)) * 0.25;
      double Djv = VT(i, j) - VT(i, j - 1);
      double vtx = (Div * Djy - Djv * Diy) * Ds;
      double vty = (-Div * Djx + Djv * Dix) * Ds;
      FV(Fv, i, j) = vfi + v0 * (vtx * n1 + vty * n2) * s0;
      FV(Fv2, i, j) = (vtx * n1 + vty * n2) * s0;
    }
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++) {
      double v0 = (VT(i, j) + B->Amu[GI(B, i, j)] / f->d[GI(B, i, j)]) * Cb2 / SA_sigma;
      size_t r = RI(B, NV, 5, i, j);
      B->Res[r] = B->Res[r] + FV(Fv, i, j + 1) - FV(Fv, i, j) - v0 * (FV(Fv2, i, j + 1) - FV(Fv2, i, j));
    }
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++) {
      const double *x1 = B->x1, *y1 = B->y1;
      const double *uu = f->uu, *v = f->v;
      double Dix = (x1[GI(B, i + 1, j)] - x1[GI(B, i - 1, j)]) * 0.5;
      double Diy = (y1[GI(B, i + 1, j)] - y1[GI(B, i - 1, j)]) * 0.5;
      double Djx = (x1[GI(B, i, j + 1)] - x1[GI(B, i, j - 1)]) * 0.5;
      double Djy = (y1[GI(B, i, j + 1)] - y1[GI(B, i, j - 1)]) * 0.5;
      double Ds = 1.0 / (Dix * Djy - Djx * Diy);
      double Diu = (uu[GI(B, i + 1, j)] - uu[GI(B, i - 1, j)]) * 0.5;
      double Div = (v[GI(B, i + 1, j)] - v[GI(B, i - 1, j)]) * 0.5;
      double Dju = (uu[GI(B, i, j + 1)] - uu[GI(B, i, j - 1)]) * 0.5;
      double Djv = (v[GI(B, i, j + 1)] - v[GI(B, i, j - 1)]) * 0.5;
      double vx = (Div * Djy - Djv * Diy) * Ds;
      double uy = (-Diu * Djx + Dju * Dix) * Ds;
      double S = fabs(vx - uy);
      size_t g = GI(B, i, j);
      double vtc = VT(i, j), dwc = B->dw[WI(B, i, j)];
      double X = f->d[g] * vtc / B->Amu[g];
      double X3 = ipow(X, 3);
      double fv1 = X3 / (X3 + Cv1_3);
      double fv2 = 1.0 - X / (1.0 + X * fv1);
      double kd = SA_k * dwc;
      double S1 = fmax(S + fv2 * vtc / (kd * kd), SP(0.3) * S);
      double r = fmin(vtc / (S1 * SA_k * SA_k * dwc * dwc), 10.0);
      double gg = r + Cw2 * (ipow(r, 6) - r);
      double fw = gg * pow((1.0 + Cw3_6) / (ipow(gg, 6) + Cw3_6), 1.0 / 6.0);
      double ft2 = Ct3 * exp(-(Ct4 * X * X));
      double q = vtc / dwc;
      double Q_SA = Cb1 * (1.0 - ft2) * S1 * vtc - (Cw1 * fw - Cb1 * ft2 / (SA_k * SA_k)) * (q * q);
      size_t rr = RI(B, NV, 5, i, j);
      B->Res[rr] = B->Res[rr] + Q_SA * B->vol[PI_(B, i, j)];
    }
#undef VT
#undef FV
  free(vt); free(Fv); free(Fv2);
}

/* sub_turbulence_SST.f90:6-271 */
static void turbulence_sst(const OMesh *M, const OParams *P, OBlock *B, const Flow *f) {
  const double sk1 = 0.85, sw1 = 0.5, beta1 = 0.075, Cw1s = 0.533, sk2 = 1.0, sw2 = 0.856, beta2 = 0.0828, Cw2s = 0.440;
  const double a1s = 0.31, betas = 0.09;
  const double Re = P->Re;
  int nx = B->nx, ny = B->ny, NV = M->nvar;
  size_t g0 = (size_t)(nx + 3) * (ny + 3);
  double *Kt = malloc(g0 * sizeof(double)), *Wt = malloc(g0 * sizeof(double));
  double *f1 = calloc((size_t)nx * ny, sizeof(double)), *Qk = calloc((size_t)nx * ny, sizeof(double));
  double *Qw = calloc((size_t)nx * ny, sizeof(double));
  double *Fk = calloc((size_t)nx * ny, sizeof(double)), *Fw = calloc((size_t)nx * ny, sizeof(double));
  const double *x1 = B->x1, *y1 = B->y1, *uu = f->uu, *v = f->v, *d = f->d;
  double *Amu = B->Amu, *Amu_t = B->Amu_t;
  for (int j = 1 - LAP; j <= ny + LAP - 1; j++)
    for (int i = 1 - LAP; i <= nx + LAP - 1; i++) {
      size_t g = GI(B, i, j);
      Kt[g] = B->U[UI(B, NV, 5, i, j)] / d[g];
      Wt[g] = B->U[UI(B, NV, 6, i, j)] / d[g];
    }
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++) {
      size_t g = GI(B, i, j), pp = PI_(B, i, j);
      double Dix = (x1[GI(B, i + 1, j)] - x1[GI(B, i - 1, j)]) * 0.5;
      double Diy = (y1[GI(B, i + 1, j)] - y1[GI(B, i - 1, j)]) * 0.5;
      double Djx = (x1[GI(B, i, j + 1)] - x1[GI(B, i, j - 1)]) * 0.5;
      double Djy = (y1[GI(B, i, j + 1)] - y1[GI(B, i, j - 1)]) * 0.5;
      double Ds = 1.0 / (Dix * Djy - Djx * Diy);
      double Diu = (uu[GI(B, i + 1, j)] - uu[GI(B, i - 1, j)]) * 0.5;
      double Div = (v[GI(B, i + 1, j)] - v[GI(B, i - 1, j)]) * 0.5;
      double Dik = (Kt[GI(B, i + 1, j)] - Kt[GI(B, i - 1, j)]) * 0.5;
      double Diw = (Wt[GI(B, i + 1, j)] - Wt[GI(B, i - 1, j)]) * 0.5;
      double Dju = (uu[GI(B, i, j + 1)] - uu[GI(B, i, j - 1)]) * 0.5;
      double Djv = (v[GI(B, i, j + 1)] - v[GI(B, i, j - 1)]) * 0.5;
      double Djk = (Kt[GI(B, i, j + 1)] - Kt[GI(B, i, j - 1)]) * 0.5;
      double Djw = (Wt[GI(B, i, j + 1)] - Wt[GI(B, i, j - 1)]) * 0.5;
      double ux = (Diu * Djy - Dju * Diy) * Ds;
      double vx = (Div * Djy - Djv * Diy) * Ds;
      double kx = (Dik * Djy - Djk * Diy) * Ds;
      double wx = (Diw * Djy - Djw * Diy) * Ds;
      double uy = (-Diu * Djx + Dju * Dix) * Ds;
      double vy = (-Div * Djx + Djv * Dix) * Ds;
      double ky = (-Dik * Djx + Djk * Dix) * Ds;
      double wy = (-Diw * Djx + Djw * Dix) * Ds;
      (void)ux; (void)vy;
      Amu[g] = Amu[g] * Re;
      double omega = vx - uy;
      double dw = B->dw[WI(B, i, j)];
      double dw2 = dw * dw, Re2 = Re * Re;
      double arg2 = fmax(2.0 * sqrt(fabs(Kt[g])) / (SP(0.09) * Wt[g] * dw * Re), 500.0 * Amu[g] / (d[g] * Wt[g] * dw2 * Re2));
      double f2 = tanh(arg2 * arg2);
      Amu_t[g] = fmin(fmin(d[g] * Kt[g] / Wt[g], a1s * d[g] * Kt[g] * Re / (f2 * fabs(omega))), SP(100000.0));
      double Kws = 2.0 * (kx * wx + ky * wy) * d[g] * sw2 / (Wt[g] + 1.0e-20);
      double CD_kw = fmax(Kws, 1.0e-20);
      double arg3 = fmax(sqrt(fabs(Kt[g])) / (SP(0.09) * Wt[g] * dw * Re), 500.0 * Amu[g] / (d[g] * Wt[g] * dw2 * Re2));
      double arg1 = fmin(arg3, 4.0 * d[g] * sw2 * Kt[g] / (CD_kw * dw2));
      f1[pp] = tanh(ipow(arg1, 4));
      double Pk = Amu_t[g] * omega * omega;
      double Pk0 = fmin(Pk, 20.0 * betas * Kt[g] * Wt[g] * Re2);
      Qk[pp] = Pk0 / Re - betas * d[g] * Wt[g] * Kt[g] * Re;
      double Cw_ = f1[pp] * Cw1s + (1.0 - f1[pp]) * Cw2s;
      double beta_ = f1[pp] * beta1 + (1.0 - f1[pp]) * beta2;
      Qw[pp] = Cw_ * d[g] * omega * omega / Re - beta_ * d[g] * (Wt[g] * Wt[g]) * Re + (1.0 - f1[pp]) * Kws / Re;
    }
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx; i++) {
      size_t pp = PI_(B, i, j);
      double si = B->si[pp], ni1 = B->ni1[pp], ni2 = B->ni2[pp];
      double un1 = uu[GI(B, i - 1, j)] * ni1 + v[GI(B, i - 1, j)] * ni2;
      double un2 = uu[GI(B, i, j)] * ni1 + v[GI(B, i, j)] * ni2;
      Fk[pp] = -0.5 * ((un1 + fabs(un1)) * Kt[GI(B, i - 1, j)] + (un2 - fabs(un2)) * Kt[GI(B, i, j)]) * si;
      Fw[pp] = -0.5 * ((un1 + fabs(un1)) * Wt[GI(B, i - 1, j)] + (un2 - fabs(un2)) * Wt[GI(B, i, j)]) * si;
      double sK = f1[pp] * sk1 + (1.0 - f1[pp]) * sk2;
      double sW = f1[pp] * sw1 + (1.0 - f1[pp]) * sw2;
      double A1, A2;
      if (i == 1) {
        A1 = Amu[GI(B, i, j)] + sK * Amu_t[GI(B, i, j)];
        A2 = Amu[GI(B, i, j)] + sW * Amu_t[GI(B, i, j)];
      } else if (i == nx) {
        A1 = Amu[GI(B, i - 1, j)] + sK * Amu_t[GI(B, i - 1, j)];
        A2 = Amu[GI(B, i - 1, j)] + sW * Amu_t[GI(B, i - 1, j)];
      } else {
        A1 = (Amu[GI(B, i - 1, j)] + Amu[GI(B, i, j)] + sK * (Amu_t[GI(B, i - 1, j)] + Amu_t[GI(B, i, j)])) * 0.5;
        A2 = (Amu[GI(B, i - 1, j)] + Amu[GI(B, i, j)] + sW * (Amu_t[GI(B, i - 1, j)] + Amu_t[GI(B, i, j)])) * 0.5;
      }
      double Dix = x1[GI(B, i, j)] - x1[GI(B, i - 1, j)];
      double Diy = y1[GI(B, i, j)] - y1[GI(B, i - 1, j)];
      double Djx = (x1[GI(B, i - 1, j + 1)] + x1[GI(B, i, j + 1)] - x1[GI(B, i - 1, j - 1)] - x1[GI(B, i, j - 1)]) * 0.25;
      double Djy = (y1[GI(B, i - 1, j + 1)] + y1[GI(B, i, j + 1)] - y1[GI(B, i - 1, j - 1)] - y1[GI(B, i, j - 1)]) * 0.25;
      double Ds = 1.0 / (Dix * Djy - Djx * Diy);
      double Dik = Kt[GI(B, i, j)] - Kt[GI(B, i - 1, j)];
      double Diw = Wt[GI(B, i, j)] - Wt[GI(B, i - 1, j)];
      double Djk = (Kt[GI(B, i - 1, j + 1)] + Kt[GI(B, i, j + 1)] - Kt[GI(B, i - 1, j - 1)] - Kt[GI(B, i, j - 1)]) * 0.25;
      double Djw = (Wt[GI(B, i - 1, j + 1)] + Wt[GI(B, i, j + 1)] - Wt[GI(B, i - 1, j - 1)] - Wt[GI(B, i, j - 1)]) * 0.25;
      double Kx = (Dik * Djy - Djk * Diy) * Ds, Wx = (Diw * Djy - Djw * Diy) * Ds;
      double Ky = (-Dik * Djx + Djk * Dix) * Ds, Wy = (-Diw * Djx + Djw * Dix) * Ds;
      Fk[pp] = Fk[pp] + A1 * (Kx * ni1 + Ky * ni2) * si / Re;
      Fw[pp] = Fw[pp] + A2 * (Wx * ni1 + Wy * ni2) * si / Re;
    }
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++) {
      B->Res[RI(B, NV, 5, i, j)] = Fk[PI_(B, i + 1, j)] - Fk[PI_(B, i, j)];
      B->Res[RI(B, NV, 6, i, j)] = Fw[PI_(B, i + 1, j)] - Fw[PI_(B, i, j)];
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx - 1; i++) {
      size_t pp = PI_(B, i, j);
      double sj = B->sj[pp], nj1 = B->nj1[pp], nj2 = B->nj2[pp];
      double un1 = uu[GI(B, i, j - 1)] * nj1 + v[GI(B, i, j - 1)] * nj2;
      double un2 = uu[GI(B, i, j)] * nj1 + v[GI(B, i, j)] * nj2;
      Fk[pp] = -0.5 * ((un1 + fabs(un1)) * Kt[GI(B, i, j - 1)] + (un2 - fabs(un2)) * Kt[GI(B, i, j)]) * sj;
      Fw[pp] = -0.5 * ((un1 + fabs(un1)) * Wt[GI(B, i, j - 1)] + (un2 - fabs(un2)) * Wt[GI(B, i, j)]) * sj;
      double sK = f1[pp] * sk1 + (1.0 - f1[pp]) * sk2;
      double sW = f1[pp] * sw1 + (1.0 - f1[pp]) * sw2;
      double A1, A2;
      if (j == 1) {
        A1 = Amu[GI(B, i, j)] + sK * Amu_t[GI(B, i, j)];
        A2 = Amu[GI(B, i, j)] + sW * Amu_t[GI(B, i, j)];
      } else if (j == ny) {
        A1 = Amu[GI(B, i, j - 1)] + sK * Amu_t[GI(B, i, j - 1)];
        A2 = Amu[GI(B, i, j - 1)] + sW * Amu_t[GI(B, i, j - 1)];
      } else {
        A1 = (Amu[GI(B, i, j)] + Amu[GI(B, i, j - 1)] + sK * (Amu_t[GI(B, i, j)] + Amu_t[GI(B, i, j - 1)])) * 0.5;
        A2 = (Amu[GI(B, i, j)] + Amu[GI(B, i, j - 1)] + sW * (Amu_t[GI(B, i, j)] + Amu_t[GI(B, i, j - 1)])) * 0.5;
      }
      double Dix = (x1[GI(B, i + 1, j - 1)] + x1[GI(B, i + 1, j)] - x1[GI(B, i - 1, j - 1)] - x1[GI(B, i - 1, j)]) * 0.25;
      double Diy = (y1[GI(B, i + 1, j - 1)] + y1[GI(B, i + 1, j)] - y1[GI(B, i - 1, j - 1)] - y1[GI(B, i - 1, j)]) * 0.25;
      double Djx = x1[GI(B, i, j)] - x1[GI(B, i, j - 1)];
      double Djy = y1[GI(B, i, j)] - y1[GI(B, i, j - 1)];
      double Ds = 1.0 / (Dix * Djy - Djx * Diy);
      double Dik = (Kt[GI(B, i + 1, j - 1)] + Kt[GI(B, i + 1, j)] - Kt[GI(B, i - 1, j - 1)] - Kt[GI(B, i - 1, j)]) * 0.25;
      double Diw = (Wt[GI(B, i + 1, j - 1)] + Wt[GI(B, i + 1, j)] - Wt[GI(B, i - 1, j - 1)] - Wt[GI(B, i - 1, j)]) * 0.25;
      double Djk = Kt[GI(B, i, j)] - Kt[GI(B, i, j - 1)];
      double Djw = Wt[GI(B, i, j)] - Wt[GI(B, i, j - 1)];
      double Kx = (Dik * Djy - Djk * Diy) * Ds, Wx = (Diw * Djy - Djw * Diy) * Ds;
      double Ky = (-Dik * Djx + Djk * Dix) * Ds, Wy = (-Diw * Djx + Djw * Dix) * Ds;
      Fk[pp] = Fk[pp] + A1 * (Kx * nj1 + Ky * nj2) * sj / Re;
      Fw[pp] = Fw[pp] + A2 * (Wx * nj1 + Wy * nj2) * sj / Re;
    }
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++) {
      size_t r5 = RI(B, NV, 5, i, j), r6 = RI(B, NV, 6, i, j), pp = PI_(B, i, j), g = GI(B, i, j);
      B->Res[r5] = B->Res[r5] + Fk[PI_(B, i, j + 1)] - Fk[pp] + Qk[pp] * B->vol[pp];
      B->Res[r6] = B->Res[r6] + Fw[PI_(B, i, j + 1)] - Fw[pp] + Qw[pp] * B->vol[pp];
      Amu[g] = Amu[g] / Re;
      Amu_t[g] = Amu_t[g] / Re;
    }
  free(Kt); free(Wt); free(f1); free(Qk); free(Qw); free(Fk); free(Fw);
  amut_boundary(B);
}

/* sub_Residual.f90:200-472 */
static void residual_fvm(const OMesh *M, const OParams *P, OBlock *B, Flow *f) {
  int nx = B->nx, ny = B->ny, NV = M->nvar;
  int Scheme = M->scheme, IFlux = M->flux, Recon = M->recon;
  double g = P->gamma;
  int fdm = (B->fvm_fdm == 1);
  const int LF = 4; /* LFDM */
  memset(f->Fi, 0, sizeof(double) * 4 * nx * ny);
  memset(f->Fj, 0, sizeof(double) * 4 * nx * ny);
  double U0[4][4], UL[4], UR[4], QL[4], QR[4], F0[4];
  const double *x1 = B->x1, *y1 = B->y1;
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx; i++) {
      if (fdm && (i >= LF + 2 && i <= nx - LF - 1) && (j >= LF + 2 && j <= ny - LF - 1)) continue;
      size_t pp = PI_(B, i, j);
      double si = B->si[pp], ni1 = B->ni1[pp], ni2 = B->ni2[pp];
      if (Recon == 0) {
        for (int k = 0; k < 4; k++) {
          size_t gk = GI(B, i - 2 + k, j);
          U0[k][0] = f->d[gk]; U0[k][1] = f->uu[gk]; U0[k][2] = f->v[gk]; U0[k][3] = f->p[gk];
        }
        recon_original(U0, UL, UR, Scheme);
      } else {
        for (int k = 0; k < 4; k++)
          for (int m = 0; m < 4; m++) U0[k][m] = B->U[UI(B, NV, m + 1, i - 2 + k, j)];
        if (Recon == 1) recon_conservative(U0, UL, UR, g, Scheme);
        else recon_characteristic(U0, UL, UR, g, Scheme);
      }
      QL[0] = UL[0]; QL[1] = UL[1] * ni1 + UL[2] * ni2; QL[2] = -UL[1] * ni2 + UL[2] * ni1; QL[3] = UL[3];
      QR[0] = UR[0]; QR[1] = UR[1] * ni1 + UR[2] * ni2; QR[2] = -UR[1] * ni2 + UR[2] * ni1; QR[3] = UR[3];
      face_flux(IFlux, QL, QR, F0, g);
      f->Fi[FX(B, 1, i, j)] = -F0[0] * si;
      f->Fi[FX(B, 2, i, j)] = -(F0[1] * ni1 - F0[2] * ni2) * si;
      f->Fi[FX(B, 3, i, j)] = -(F0[1] * ni2 + F0[2] * ni1) * si;
      f->Fi[FX(B, 4, i, j)] = -F0[3] * si;
      if (P->If_viscous == 0) continue;
      double Amu1, Amk1;
      const double *Am = B->Amu, *At = B->Amu_t;
      if (i == 1) {
        Amu1 = Am[GI(B, i, j)] + At[GI(B, i, j)];
        Amk1 = P->Cp * (Am[GI(B, i, j)] / P->Pr + At[GI(B, i, j)] / PRT);
      } else if (i == nx) {
        Amu1 = Am[GI(B, i - 1, j)] + At[GI(B, i - 1, j)];
        Amk1 = P->Cp * (Am[GI(B, i - 1, j)] / P->Pr + At[GI(B, i - 1, j)] / PRT);
      } else {
        Amu1 = (Am[GI(B, i - 1, j)] + Am[GI(B, i, j)] + At[GI(B, i - 1, j)] + At[GI(B, i, j)]) * 0.5;
        Amk1 = P->Cp * ((Am[GI(B, i - 1, j)] + Am[GI(B, i, j)]) / P->Pr + (At[GI(B, i - 1, j)] + At[GI(B, i, j)]) / PRT) * 0.5;
      }
      double Dix = x1[GI(B, i, j)] - x1[GI(B, i - 1, j)];
      double Diy = y1[GI(B, i, j)] - y1[GI(B, i - 1, j)];
      double Djx = (x1[GI(B, i - 1, j + 1)] + x1[GI(B, i, j + 1)] - x1[GI(B, i - 1, j - 1)] - x1[GI(B, i, j - 1)]) * 0.25;
      double Djy = (y1[GI(B, i - 1, j + 1)] + y1[GI(B, i, j + 1)] - y1[GI(B, i - 1, j - 1)] - y1[GI(B, i, j - 1)]) * 0.25;
      double Ds = 1.0 / (Dix * Djy - Djx * Diy);
      const double *uu = f->uu, *v = f->v, *T = f->T;
      double Diu = uu[GI(B, i, j)] - uu[GI(B, i - 1, j)];
      double Div = v[GI(B, i, j)] - v[GI(B, i - 1, j)];
      double DiT = T[GI(B, i, j)] - T[GI(B, i - 1, j)];
      double Dju = (uu[GI(B, i - 1, j + 1)] + uu[GI(B, i, j + 1)] - uu[GI(B, i - 1, j - 1)] - uu[GI(B, i, j - 1)]) * 0.25;
      double Djv = (v[GI(B, i - 1, j + 1)] + v[GI(B, i, j + 1)] - v[GI(B, i - 1, j - 1)] - v[GI(B, i, j - 1)]) * 0.25;
      double DjT = (T[GI(B, i - 1, j + 1)] + T[GI(B, i, j + 1)] - T[GI(B, i - 1, j - 1)] - T[GI(B, i, j - 1)]) * 0.25;
      double ux = (Diu * Djy - Dju * Diy) * Ds, vx = (Div * Djy - Djv * Diy) * Ds, Tx = (DiT * Djy - DjT * Diy) * Ds;
      double uy = (-Diu * Djx + Dju * Dix) * Ds, vy = (-Div * Djx + Djv * Dix) * Ds, Ty = (-DiT * Djx + DjT * Dix) * Ds;
      double u1 = (uu[GI(B, i, j)] + uu[GI(B, i - 1, j)]) * 0.5;
      double v1 = (v[GI(B, i, j)] + v[GI(B, i - 1, j)]) * 0.5;
      double t11 = ((4.0 / 3.0) * ux - (2.0 / 3.0) * vy) * Amu1;
      double t22 = ((4.0 / 3.0) * vy - (2.0 / 3.0) * ux) * Amu1;
      double t12 = (uy + vx) * Amu1;
      double E1 = u1 * t11 + v1 * t12 + Amk1 * Tx;
      double E2 = u1 * t12 + v1 * t22 + Amk1 * Ty;
      f->Fi[FX(B, 2, i, j)] = f->Fi[FX(B, 2, i, j)] + (t11 * ni1 + t12 * ni2) * si;
      f->Fi[FX(B, 3, i, j)] = f->Fi[FX(B, 3, i, j)] + (t12 * ni1 + t22 * ni2) * si;
      f->Fi[FX(B, 4, i, j)] = f->Fi[FX(B, 4, i, j)] + (E1 * ni1 + E2 * ni2) * si;
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx - 1; i++) {
      if (fdm && (i >= LF + 2 && i <= nx - LF - 1) && (j >= LF + 2 && j <= ny - LF - 1)) continue;
      size_t pp = PI_(B, i, j);
      double sj = B->sj[pp], nj1 = B->nj1[pp], nj2 = B->nj2[pp];
      if (Recon == 0) {
        for (int k = 0; k < 4; k++) {
          size_t gk = GI(B, i, j - 2 + k);
          U0[k][0] = f->d[gk]; U0[k][1] = f->uu[gk]; U0[k][2] = f->v[gk]; U0[k][3] = f->p[gk];
        }
        recon_original(U0, UL, UR, P->Iflag_Scheme); /* the global scheme: sub_Residual.f90:367 */
      } else {
        for (int k = 0; k < 4; k++)
          for (int m = 0; m < 4; m++) U0[k][m] = B->U[UI(B, NV, m + 1, i, j - 2 + k)];
        if (Recon == 1) recon_conservative(U0, UL, UR, g, Scheme);
        else recon_characteristic(U0, UL, UR, g, Scheme);
      }
      QL[0] = UL[0]; QL[1] = UL[1] * nj1 + UL[2] * nj2; QL[2] = -UL[1] * nj2 + UL[2] * nj1; QL[3] = UL[3];
      QR[0] = UR[0]; QR[1] = UR[1] * nj1 + UR[2] * nj2; QR[2] = -UR[1] * nj2 + UR[2] * nj1; QR[3] = UR[3];
      face_flux(IFlux, QL, QR, F0, g);
      f->Fj[FX(B, 1, i, j)] = -F0[0] * sj;
      f->Fj[FX(B, 2, i, j)] = -(F0[1] * nj1 - F0[2] * nj2) * sj;
      f->Fj[FX(B, 3, i, j)] = -(F0[1] * nj2 + F0[2] * nj1) * sj;
      f->Fj[FX(B, 4, i, j)] = -F0[3] * sj;
      if (P->If_viscous == 0) continue;
      double Amu1, Amk1;
      const double *Am = B->Amu, *At = B->Amu_t;
      if (j == 1) {
        Amu1 = Am[GI(B, i, j)] + At[GI(B, i, j)];
        Amk1 = P->Cp * (Am[GI(B, i, j)] / P->Pr + At[GI(B, i, j)] / PRT);
      } else if (j == ny) {
        Amu1 = Am[GI(B, i, j - 1)] + At[GI(B, i, j - 1)];
        Amk1 = P->Cp * (Am[GI(B, i, j - 1)] / P->Pr + At[GI(B, i, j - 1)] / PRT);
      } else {
        Amu1 = (Am[GI(B, i, j)] + Am[GI(B, i, j - 1)] + At[GI(B, i, j)] + At[GI(B, i, j - 1)]) * 0.5;
        Amk1 = P->Cp * ((Am[GI(B, i, j - 1)] + Am[GI(B, i, j)]) / P->Pr + (At[GI(B, i, j - 1)] + At[GI(B, i, j)]) / PRT) * 0.5;
      }
      double Dix = (x1[GI(B, i + 1, j - 1)] + x1[GI(B, i + 1, j)] - x1[GI(B, i - 1, j - 1)] - x1[GI(B, i - 1, j)]) * 0.25;
      double Diy = (y1[GI(B, i + 1, j - 1)] + y1[GI(B, i + 1, j)] - y1[GI(B, i - 1, j - 1)] - y1[GI(B, i - 1, j)]) * 0.25;
      double Djx = x1[GI(B, i, j)] - x1[GI(B, i, j - 1)];
      double Djy = y1[GI(B, i, j)] - y1[GI(B, i, j - 1)];
      double Ds = 1.0 / (Dix * Djy - Djx * Diy);
      const double *uu = f->uu, *v = f->v, *T = f->T;
      double Diu = (uu[GI(B, i + 1, j - 1)] + uu[GI(B, i + 1, j)] - uu[GI(B, i - 1, j - 1)] - uu[GI(B, i - 1, j)]) * 0.25;
      double Div = (v[GI(B, i + 1, j - 1)] + v[GI(B, i + 1, j)] - v[GI(B, i - 1, j - 1)] - v[GI(B, i - 1, j)]) * 0.25;
      double DiT = (T[GI(B, i + 1, j - 1)] + T[GI(B, i + 1, j)] - T[GI(B, i - 1, j - 1)] - T[GI(B, i - 1, j)]) * 0.25;
      double Dju = uu[GI(B, i, j)] - uu[GI(B, i, j - 1)];
      double Djv = v[GI(B, i, j)] - v[GI(B, i, j - 1)];
      double DjT = T[GI(B, i, j)] - T[GI(B, i, j - 1)];
      double ux = (Diu * Djy - Dju * Diy) * Ds, vx = (Div * Djy - Djv * Diy) * Ds, Tx = (DiT * Djy - DjT * Diy) * Ds;
      double uy = (-Diu * Djx + Dju * Dix) * Ds, vy = (-Div * Djx + Djv * Dix) * Ds, Ty = (-DiT * Djx + DjT * Dix) * Ds;
      double t11 = ((4.0 / 3.0) * ux - (2.0 / 3.0) * vy) * Amu1;
      double t22 = ((4.0 / 3.0) * vy - (2.0 / 3.0) * ux) * Amu1;
      double t12 = (uy + vx) * Amu1;
      double u1 = (uu[GI(B, i, j)] + uu[GI(B, i, j - 1)]) * 0.5;
      double v1 = (v[GI(B, i, j)] + v[GI(B, i, j - 1)]) * 0.5;
      double E1 = u1 * t11 + v1 * t12 + Amk1 * Tx;
      double E2 = u1 * t12 + v1 * t22 + Amk1 * Ty;
      f->Fj[FX(B, 2, i, j)] = f->Fj[FX(B, 2, i, j)] + (t11 * nj1 + t12 * nj2) * sj;
      f->Fj[FX(B, 3, i, j)] = f->Fj[FX(B, 3, i, j)] + (t12 * nj1 + t22 * nj2) * sj;
      f->Fj[FX(B, 4, i, j)] = f->Fj[FX(B, 4, i, j)] + (E1 * nj1 + E2 * nj2) * sj;
    }
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++) {
      if (fdm && (i >= LF + 1 && i <= nx - LF - 1) && (j >= LF + 1 && j <= ny - LF - 1)) continue;
      for (int m = 1; m <= 4; m++)
        B->Res[RI(B, NV, m, i, j)] = f->Fi[FX(B, m, i + 1, j)] - f->Fi[FX(B, m, i, j)] + f->Fj[FX(B, m, i, j + 1)] - f->Fj[FX(B, m, i, j)];
    }
}

/* ------------------------------------------------------------------------
 * finite-difference WENO5 blocks: sub_Finite_Difference.f90:4-275
 * ---------------------------------------------------------------------- */
static void split_sw(double d, double u, double v, double cc, double fp[4], double fm[4], double A1, double A2, double g) {
  double tmp1 = 2.0 * (g - 1.0);
  double tmp3 = (3.0 - g) / (2.0 * (g - 1.0));
  double vs = A1 * u + A2 * v;
  double ss = sqrt(A1 * A1 + A2 * A2);
  double ak1 = A1 / ss, ak2 = A2 / ss;
  double E1 = vs, E2 = vs - cc * ss, E3 = vs + cc * ss;
  double E1P = (E1 + fabs(E1)) * 0.5, E2P = (E2 + fabs(E2)) * 0.5, E3P = (E3 + fabs(E3)) * 0.5;
  double E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
  double tmp0 = d / (2.0 * g);
  double uc1 = u - cc * ak1, uc2 = u + cc * ak1, vc1 = v - cc * ak2, vc2 = v + cc * ak2;
  double vvc1 = (uc1 * uc1 + vc1 * vc1) / 2.0, vvc2 = (uc2 * uc2 + vc2 * vc2) / 2.0;
  double vv = (g - 1.0) * (u * u + v * v);
  double W2 = tmp3 * cc * cc;
  fp[0] = tmp0 * (tmp1 * E1P + E2P + E3P);
  fp[1] = tmp0 * (tmp1 * E1P * u + E2P * uc1 + E3P * uc2);
  fp[2] = tmp0 * (tmp1 * E1P * v + E2P * vc1 + E3P * vc2);
  fp[3] = tmp0 * (E1P * vv + E2P * vvc1 + E3P * vvc2 + W2 * (E2P + E3P));
  fm[0] = tmp0 * (tmp1 * E1M + E2M + E3M);
  fm[1] = tmp0 * (tmp1 * E1M * u + E2M * uc1 + E3M * uc2);
  fm[2] = tmp0 * (tmp1 * E1M * v + E2M * vc1 + E3M * vc2);
  fm[3] = tmp0 * (E1M * vv + E2M * vvc1 + E3M * vvc2 + W2 * (E2M + E3M));
}

/* v[1..n] (1-based), h[1..n] */
static void fp_weno5(int n, const double *v, double *h) {
  const double ep = 1.0e-6, C03 = 3.0 / 10.0, C13 = 3.0 / 5.0, C23 = 1.0 / 10.0;
  for (int k = 3; k <= n - 2; k++) {
    double a = v[k] - 2.0 * v[k + 1] + v[k + 2], b = 3.0 * v[k] - 4.0 * v[k + 1] + v[k + 2];
    double S0 = 13.0 / 12.0 * (a * a) + 1.0 / 4.0 * (b * b);
    a = v[k - 1] - 2.0 * v[k] + v[k + 1]; b = v[k - 1] - v[k + 1];
    double S1 = 13.0 / 12.0 * (a * a) + 1.0 / 4.0 * (b * b);
    a = v[k - 2] - 2.0 * v[k - 1] + v[k]; b = v[k - 2] - 4.0 * v[k - 1] + 3.0 * v[k];
    double S2 = 13.0 / 12.0 * (a * a) + 1.0 / 4.0 * (b * b);
    double e0 = ep + S0, e1 = ep + S1, e2 = ep + S2;
    double a0 = C03 / (e0 * e0), a1 = C13 / (e1 * e1), a2 = C23 / (e2 * e2);
    double W0 = a0 / (a0 + a1 + a2), W1 = a1 / (a0 + a1 + a2), W2 = a2 / (a0 + a1 + a2);
    double q03 = 1.0 / 3.0 * v[k] + 5.0 / 6.0 * v[k + 1] - 1.0 / 6.0 * v[k + 2];
    double q13 = -(1.0 / 6.0 * v[k - 1]) + 5.0 / 6.0 * v[k] + 1.0 / 3.0 * v[k + 1];
    double q23 = 1.0 / 3.0 * v[k - 2] - 7.0 / 6.0 * v[k - 1] + 11.0 / 6.0 * v[k];
    h[k] = W0 * q03 + W1 * q13 + W2 * q23;
  }
}

static void fm_weno5(int n, const double *v, double *h) {
  const double ep = 1.0e-6, C03 = 3.0 / 10.0, C13 = 3.0 / 5.0, C23 = 1.0 / 10.0;
  for (int k = 3; k <= n - 2; k++) {
    double a = v[k] - 2.0 * v[k - 1] + v[k - 2], b = 3.0 * v[k] - 4.0 * v[k - 1] + v[k - 2];
    double S0 = 13.0 / 12.0 * (a * a) + 1.0 / 4.0 * (b * b);
    a = v[k + 1] - 2.0 * v[k] + v[k - 1]; b = v[k + 1] - v[k - 1];
    double S1 = 13.0 / 12.0 * (a * a) + 1.0 / 4.0 * (b * b);
    a = v[k + 2] - 2.0 * v[k + 1] + v[k]; b = v[k + 2] - 4.0 * v[k + 1] + 3.0 * v[k];
    double S2 = 13.0 / 12.0 * (a * a) + 1.0 / 4.0 * (b * b);
    double e0 = ep + S0, e1 = ep + S1, e2 = ep + S2;
    double a0 = C03 / (e0 * e0), a1 = C13 / (e1 * e1), a2 = C23 / (e2 * e2);
    double W0 = a0 / (a0 + a1 + a2), W1 = a1 / (a0 + a1 + a2), W2 = a2 / (a0 + a1 + a2);
    double q03 = 1.0 / 3.0 * v[k] + 5.0 / 6.0 * v[k - 1] - 1.0 / 6.0 * v[k - 2];
    double q13 = -(1.0 / 6.0 * v[k + 1]) + 5.0 / 6.0 * v[k] + 1.0 / 3.0 * v[k - 1];
    double q23 = 1.0 / 3.0 * v[k + 2] - 7.0 / 6.0 * v[k + 1] + 11.0 / 6.0 * v[k];
    h[k - 1] = W0 * q03 + W1 * q13 + W2 * q23;
  }
}

static void residual_fdm(const OMesh *M, const OParams *P, OBlock *B, Flow *f) {
  int nx = B->nx - 1, ny = B->ny - 1, NV = M->nvar;
  const int LF = 4;
  double g = P->gamma;
  int nmax = (nx > ny ? nx : ny) + 2;
  double *fx1[4], *fx2[4];
  for (int m = 0; m < 4; m++) { fx1[m] = calloc(nmax + 1, sizeof(double)); fx2[m] = calloc(nmax + 1, sizeof(double)); }
  double *h1 = calloc(nmax + 1, sizeof(double)), *h2 = calloc(nmax + 1, sizeof(double));
  double *Ev1 = calloc((size_t)nx * ny * 3, sizeof(double)), *Ev2 = calloc((size_t)nx * ny * 3, sizeof(double));
#define EV(A, i, j, k) A[(size_t)((i) - 1) + (size_t)nx * ((size_t)((j) - 1) + (size_t)ny * (size_t)((k) - 1))]
#define MET(A, i, j) B->A[CI(B, i, j)]
  double fp[4], fm[4];
  for (int j = LF; j <= ny - LF; j++) {
    for (int i = 1; i <= nx; i++) {
      double A1 = MET(Akx, i, j) / MET(Ajac, i, j), A2 = MET(Aky, i, j) / MET(Ajac, i, j);
      size_t gg = GI(B, i, j);
      split_sw(f->d[gg], f->uu[gg], f->v[gg], f->cc[CI(B, i, j)], fp, fm, A1, A2, g);
      for (int m = 0; m < 4; m++) { fx1[m][i] = fp[m]; fx2[m][i] = fm[m]; }
    }
    for (int m = 0; m < 4; m++) {
      fp_weno5(nx, fx1[m], h1);
      fm_weno5(nx, fx2[m], h2);
      for (int i = LF; i <= nx - LF; i++) f->Fi[FX(B, m + 1, i, j)] = -(h1[i] + h2[i]);
    }
  }
  for (int i = LF; i <= nx - LF; i++) {
    for (int j = 1; j <= ny; j++) {
      double A1 = MET(Aix, i, j) / MET(Ajac, i, j), A2 = MET(Aiy, i, j) / MET(Ajac, i, j);
      size_t gg = GI(B, i, j);
      split_sw(f->d[gg], f->uu[gg], f->v[gg], f->cc[CI(B, i, j)], fp, fm, A1, A2, g);
      for (int m = 0; m < 4; m++) { fx1[m][j] = fp[m]; fx2[m][j] = fm[m]; }
    }
    for (int m = 0; m < 4; m++) {
      fp_weno5(ny, fx1[m], h1);
      fm_weno5(ny, fx2[m], h2);
      for (int j = LF; j <= ny - LF; j++) f->Fj[FX(B, m + 1, i, j)] = -(h1[j] + h2[j]);
    }
  }
  if (P->If_viscous == 1) {
    const double *uu = f->uu, *v = f->v, *T = f->T;
    for (int j = 2; j <= ny - 1; j++)
      for (int i = 2; i <= nx - 1; i++) {
        size_t gg = GI(B, i, j);
        double mu0 = B->Amu[gg] + B->Amu_t[gg];
        double Ak0 = P->Cp * (B->Amu[gg] / P->Pr + B->Amu_t[gg] / PRT);
        double ui = 0.5 * (uu[GI(B, i + 1, j)] - uu[GI(B, i - 1, j)]);
        double vi = 0.5 * (v[GI(B, i + 1, j)] - v[GI(B, i - 1, j)]);
        double Ti = 0.5 * (T[GI(B, i + 1, j)] - T[GI(B, i - 1, j)]);
        double uj = 0.5 * (uu[GI(B, i, j + 1)] - uu[GI(B, i, j - 1)]);
        double vj = 0.5 * (v[GI(B, i, j + 1)] - v[GI(B, i, j - 1)]);
        double Tj = 0.5 * (T[GI(B, i, j + 1)] - T[GI(B, i, j - 1)]);
        double kx = MET(Akx, i, j), ky = MET(Aky, i, j), ix = MET(Aix, i, j), iy = MET(Aiy, i, j), jac = MET(Ajac, i, j);
        double ux = ui * kx + uj * ix, vx = vi * kx + vj * ix, Tx = Ti * kx + Tj * ix;
        double uy = ui * ky + uj * iy, vy = vi * ky + vj * iy, Ty = Ti * ky + Tj * iy;
        double t11 = (4.0 / 3.0 * ux - 2.0 / 3.0 * vy) * mu0;
        double t22 = (4.0 / 3.0 * vy - 2.0 / 3.0 * ux) * mu0;
        double t12 = (uy + vx) * mu0;
        double E1 = uu[gg] * t11 + v[gg] * t12 + Ak0 * Tx;
        double E2 = uu[gg] * t12 + v[gg] * t22 + Ak0 * Ty;
        EV(Ev1, i, j, 1) = (kx * t11 + ky * t12) / jac;
        EV(Ev1, i, j, 2) = (kx * t12 + ky * t22) / jac;
        EV(Ev1, i, j, 3) = (kx * E1 + ky * E2) / jac;
        EV(Ev2, i, j, 1) = (ix * t11 + iy * t12) / jac;
        EV(Ev2, i, j, 2) = (ix * t12 + iy * t22) / jac;
        EV(Ev2, i, j, 3) = (ix * E1 + iy * E2) / jac;
      }
    for (int j = 3; j <= ny - 2; j++)
      for (int i = 3; i <= nx - 2; i++)
        for (int k = 1; k <= 3; k++) {
          f->Fi[FX(B, k + 1, i, j)] = f->Fi[FX(B, k + 1, i, j)] + 0.5 * (EV(Ev1, i, j, k) + EV(Ev1, i + 1, j, k));
          f->Fj[FX(B, k + 1, i, j)] = f->Fj[FX(B, k + 1, i, j)] + 0.5 * (EV(Ev2, i, j, k) + EV(Ev2, i, j + 1, k));
        }
  }
  for (int j = LF + 1; j <= ny - LF; j++)
    for (int i = LF + 1; i <= nx - LF; i++)
      for (int m = 1; m <= 4; m++)
        B->Res[RI(B, NV, m, i, j)] = (f->Fi[FX(B, m, i, j)] - f->Fi[FX(B, m, i - 1, j)] + f->Fj[FX(B, m, i, j)] - f->Fj[FX(B, m, i, j - 1)]);
#undef EV
#undef MET
  for (int m = 0; m < 4; m++) { free(fx1[m]); free(fx2[m]); }
  free(h1); free(h2); free(Ev1); free(Ev2);
}

/* sub_NS_singlegrid.f90:352-388 */
static void comput_lvc(const OParams *P, OBlock *B, const Flow *f) {
  for (int j = 1; j <= B->ny - 1; j++)
    for (int i = 1; i <= B->nx - 1; i++) {
      size_t pp = PI_(B, i, j), g = GI(B, i, j);
      double s0 = B->vol[pp];
      double si = 0.5 * (B->si[pp] + B->si[PI_(B, i + 1, j)]);
      double ni1 = 0.5 * (B->ni1[pp] + B->ni1[PI_(B, i + 1, j)]);
      double ni2 = 0.5 * (B->ni2[pp] + B->ni2[PI_(B, i + 1, j)]);
      double sj = 0.5 * (B->sj[pp] + B->sj[PI_(B, i, j + 1)]);
      double nj1 = 0.5 * (B->nj1[pp] + B->nj1[PI_(B, i, j + 1)]);
      double nj2 = 0.5 * (B->nj2[pp] + B->nj2[PI_(B, i, j + 1)]);
      double uni = f->uu[g] * ni1 + f->v[g] * ni2;
      double unj = f->uu[g] * nj1 + f->v[g] * nj2;
      double c = f->cc[CI(B, i, j)];
      B->Lci[pp] = (fabs(uni) + c) * si;
      B->Lcj[pp] = (fabs(unj) + c) * sj;
      double tmp1 = P->gamma / f->d[g] * (B->Amu[g] / P->Pr + B->Amu_t[g] / PRT);
      B->Lvi[pp] = tmp1 * si * si / s0;
      B->Lvj[pp] = tmp1 * sj * sj / s0;
    }
}

/* sub_NS_singlegrid.f90:324-348 */
static void comput_dt(const OParams *P, OBlock *B) {
  for (int j = 1; j <= B->ny - 1; j++)
    for (int i = 1; i <= B->nx - 1; i++) {
      size_t pp = PI_(B, i, j), c = CI(B, i, j);
      if (P->Iflag_local_dt == 0) {
        B->dt[c] = P->dt_global;
      } else {
        double v = P->CFL * B->vol[pp] / (B->Lci[pp] + B->Lcj[pp] + 1.0 * (B->Lvi[pp] + B->Lvj[pp]));
        if (v > P->dtmax) v = P->dtmax;
        if (v < P->dtmin) v = P->dtmin;
        B->dt[c] = v;
      }
    }
}

/* sub_Residual.f90:13-107 (with Residual :158-191) */
int o_residual(OMesh *M, const OParams *P, OError *err) {
  int NV = M->nvar;
  for (int m = 0; m < 6; m++) { M->res_max[m] = 0.0; M->res_rms[m] = 0.0; }
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    Flow f;
    flow_alloc(B, &f);
    if (comput_duvtpckw(M, P, B, &f, mb + 1, err)) { flow_free(&f); return 1; }
    if (P->If_viscous == 1) {
      get_viscous(P, B, &f);
      if (M->turb == 4) {
        /* eddy viscosity supplied by the caller in Amu_t (0:nx,0:ny), as the
         * learned closure supplies it (no reference counterpart); its wall
         * ghosts follow SA's rule (sub_turbulence_SA.f90:200-216) */
        amut_boundary(B);
      } else {
        memset(B->Amu_t, 0, sizeof(double) * (size_t)(B->nx + 3) * (B->ny + 3));
        if (M->turb == 2) turbulence_sa(M, P, B, &f);
        else if (M->turb == 3) turbulence_sst(M, P, B, &f);
        else if (M->turb == 1) turbulence_bl(B, &f);
      }
    }
    residual_fvm(M, P, B, &f);
    if (B->fvm_fdm == 1) residual_fdm(M, P, B, &f);
    if (P->Time_Method == -1) {
      /* dual-time source (Time_dual_LU_SGS = -1), sub_Residual.f90:55-64:
       * Res - (3 U - 4 Un + Un1) vol Sfac, evaluated left to right */
      for (int j = 1; j <= B->ny - 1; j++)
        for (int i = 1; i <= B->nx - 1; i++)
          for (int m = 1; m <= NV; m++) {
            size_t r = RI(B, NV, m, i, j), w = UNI(B, M->nvar_un, m, i, j);
            double t = (3.0 * B->U[UI(B, NV, m, i, j)] - 4.0 * B->Un[w]) + B->Un1[w];
            B->Res[r] = B->Res[r] - (t * B->vol[PI_(B, i, j)]) * P->sfac;
          }
    }
    double rmax[6] = {0}, rrms[6] = {0};
    for (int j = 1; j <= B->ny - 1; j++)
      for (int i = 1; i <= B->nx - 1; i++)
        for (int m = 1; m <= NV; m++) {
          double r = B->Res[RI(B, NV, m, i, j)];
          if (fabs(r) > rmax[m - 1]) rmax[m - 1] = fabs(r);
          rrms[m - 1] = rrms[m - 1] + r * r;
        }
    for (int m = 0; m < NV; m++) {
      M->res_max[m] = fmax(M->res_max[m], rmax[m]);
      M->res_rms[m] = M->res_rms[m] + rrms[m];
    }
    comput_lvc(P, B, &f);
    comput_dt(P, B);
    flow_free(&f);
  }
  for (int m = 0; m < NV; m++) M->res_rms[m] = sqrt(M->res_rms[m] / (1.0 * M->num_cell));
  return 0;
}

/* ------------------------------------------------------------------------
 * LU-SGS: sub_LU_SGS.f90:4-138
 * ---------------------------------------------------------------------- */
static void comput_dfn(int NV, double *DF, const double *U, const double *DU, double n1, double n2, double g) {
  double U2[6];
  for (int m = 0; m < NV; m++) U2[m] = U[m] + DU[m];
  double un1 = (U[1] * n1 + U[2] * n2) / U[0];
  double p1 = (g - 1.0) * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0]);
  double un2 = (U2[1] * n1 + U2[2] * n2) / U2[0];
  double p2 = (g - 1.0) * (U2[3] - 0.5 * (U2[1] * U2[1] + U2[2] * U2[2]) / U2[0]);
  DF[0] = U2[0] * un2 - U[0] * un1;
  DF[1] = (U2[1] * un2 + p2 * n1) - (U[1] * un1 + p1 * n1);
  DF[2] = (U2[2] * un2 + p2 * n2) - (U[2] * un1 + p1 * n2);
  DF[3] = (U2[3] + p2) * un2 - (U[3] + p1) * un1;
  if (NV == 5) DF[4] = U2[4] * un2 - U[4] * un1;
  else if (NV == 6) { DF[4] = U2[4] * un2 - U[4] * un1; DF[5] = U2[5] * un2 - U[5] * un1; }
}

void o_lusgs_block(OMesh *M, const OParams *P, int mblock, double alfa1) {
  OBlock *B = &M->blocks[mblock - 1];
  int NV = M->nvar, nx = B->nx, ny = B->ny;
  const double w = 1.0;
  double alfa[6], dui[6], duj[6], DF[6];
  int visc = P->If_viscous == 1;
  for (int plane = 2; plane <= nx + ny - 2; plane++)
    for (int j = 1; j <= ny - 1; j++) {
      int i = plane - j;
      if (i < 1 || i > nx - 1) continue;
      size_t pp = PI_(B, i, j);
      double a = B->vol[pp] / B->dt[CI(B, i, j)] + w * (B->Lci[pp] + B->Lcj[pp]) + B->vol[pp] * alfa1;
      for (int m = 0; m < 6; m++) alfa[m] = a;
      if (visc) for (int m = 0; m < 6; m++) alfa[m] = alfa[m] + 2.0 * (B->Lvi[pp] + B->Lvj[pp]);
      if (NV == 5) alfa[4] = alfa[4] + B->Lvi[pp] + B->Lvj[pp];
      if (i != 1) {
        comput_dfn(NV, DF, &B->U[UI(B, NV, 1, i - 1, j)], &B->dU[DUI(B, NV, 1, i - 1, j)], B->ni1[pp], B->ni2[pp], P->gamma);
        for (int m = 0; m < NV; m++) {
          double du = B->dU[DUI(B, NV, m + 1, i - 1, j)];
          dui[m] = 0.5 * (DF[m] * B->si[pp] + w * B->Lci[PI_(B, i - 1, j)] * du);
          if (visc) dui[m] = dui[m] + B->Lvi[PI_(B, i - 1, j)] * du;
        }
      } else for (int m = 0; m < 6; m++) dui[m] = 0.0;
      if (j != 1) {
        comput_dfn(NV, DF, &B->U[UI(B, NV, 1, i, j - 1)], &B->dU[DUI(B, NV, 1, i, j - 1)], B->nj1[pp], B->nj2[pp], P->gamma);
        for (int m = 0; m < NV; m++) {
          double du = B->dU[DUI(B, NV, m + 1, i, j - 1)];
          duj[m] = 0.5 * (DF[m] * B->sj[pp] + w * B->Lcj[PI_(B, i, j - 1)] * du);
          if (visc) duj[m] = duj[m] + B->Lvj[PI_(B, i, j - 1)] * du;
        }
      } else for (int m = 0; m < 6; m++) duj[m] = 0.0;
      for (int m = 0; m < NV; m++)
        B->dU[DUI(B, NV, m + 1, i, j)] = (B->Res[RI(B, NV, m + 1, i, j)] + dui[m] + duj[m]) / alfa[m];
    }
  for (int plane = nx + ny - 2; plane >= 2; plane--)
    for (int j = ny - 1; j >= 1; j--) {
      int i = plane - j;
      if (i < 1 || i > nx - 1) continue;
      size_t pp = PI_(B, i, j);
      double a = B->vol[pp] / B->dt[CI(B, i, j)] + w * (B->Lci[pp] + B->Lcj[pp]) + B->vol[pp] * alfa1;
      for (int m = 0; m < 6; m++) alfa[m] = a;
      if (visc) for (int m = 0; m < 6; m++) alfa[m] = alfa[m] + 2.0 * (B->Lvi[pp] + B->Lvj[pp]);
      if (NV == 5) alfa[4] = alfa[4] + B->Lvi[pp] + B->Lvj[pp];
      if (i != nx - 1) {
        comput_dfn(NV, DF, &B->U[UI(B, NV, 1, i + 1, j)], &B->dU[DUI(B, NV, 1, i + 1, j)], B->ni1[pp], B->ni2[pp], P->gamma);
        for (int m = 0; m < NV; m++) {
          double du = B->dU[DUI(B, NV, m + 1, i + 1, j)];
          dui[m] = -0.5 * (DF[m] * B->si[PI_(B, i + 1, j)] - w * B->Lci[PI_(B, i + 1, j)] * du);
          if (visc) dui[m] = dui[m] + B->Lvi[PI_(B, i + 1, j)] * du;
        }
      } else for (int m = 0; m < 6; m++) dui[m] = 0.0;
      if (j != ny - 1) {
        comput_dfn(NV, DF, &B->U[UI(B, NV, 1, i, j + 1)], &B->dU[DUI(B, NV, 1, i, j + 1)], B->nj1[pp], B->nj2[pp], P->gamma);
        for (int m = 0; m < NV; m++) {
          double du = B->dU[DUI(B, NV, m + 1, i, j + 1)];
          duj[m] = -0.5 * (DF[m] * B->sj[PI_(B, i, j + 1)] - w * B->Lcj[PI_(B, i, j + 1)] * du);
          if (visc) duj[m] = duj[m] + B->Lvj[PI_(B, i, j + 1)] * du;
        }
      } else for (int m = 0; m < 6; m++) duj[m] = 0.0;
      for (int m = 0; m < NV; m++) {
        size_t k = DUI(B, NV, m + 1, i, j);
        B->dU[k] = B->dU[k] + (dui[m] + duj[m]) / alfa[m];
      }
    }
}

/* sub_NS_singlegrid.f90:203-211 */
void o_update_lusgs(OMesh *M) {
  int NV = M->nvar;
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    for (int j = 1; j <= B->ny - 1; j++)
      for (int i = 1; i <= B->nx - 1; i++)
        for (int m = 1; m <= NV; m++) B->U[UI(B, NV, m, i, j)] = B->U[UI(B, NV, m, i, j)] + B->dU[DUI(B, NV, m, i, j)];
  }
}

/* sub_NS_singlegrid.f90:30-60 */
void o_force_vt_kw(OMesh *M) {
  int NV = M->nvar;
  if (NV < 5) return;
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    for (int j = 1; j <= B->ny - 1; j++)
      for (int i = 1; i <= B->nx - 1; i++) {
        if (B->U[UI(B, NV, 5, i, j)] < 0) B->U[UI(B, NV, 5, i, j)] = 1.0e-10;
        if (NV == 6 && B->U[UI(B, NV, 6, i, j)] < 0) B->U[UI(B, NV, 6, i, j)] = 1.0e-10;
      }
  }
}

/* sub_NS_singlegrid.f90:84-93 */
void o_euler1_update(OMesh *M) {
  int NV = M->nvar;
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    for (int j = 1; j <= B->ny - 1; j++)
      for (int i = 1; i <= B->nx - 1; i++)
        for (int m = 1; m <= NV; m++) {
          double du = B->Res[RI(B, NV, m, i, j)] / B->vol[PI_(B, i, j)];
          B->U[UI(B, NV, m, i, j)] = B->U[UI(B, NV, m, i, j)] + B->dt[CI(B, i, j)] * du;
        }
  }
}

/* sub_NS_singlegrid.f90:131-139 (Un = U over 0:nx, 0:ny) */
void o_copy_un(OMesh *M) {
  int NV = M->nvar, NU = M->nvar_un;
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    for (int j = 0; j <= B->ny; j++)
      for (int i = 0; i <= B->nx; i++)
        for (int m = 1; m <= NV; m++) B->Un[UNI(B, NU, m, i, j)] = B->U[UI(B, NV, m, i, j)];
  }
}

/* sub_NS_singlegrid.f90:124-126, 155-163 */
void o_rk3_stage(OMesh *M, int krk) {
  const double Ra[3] = {1.0, 3.0 / 4.0, 1.0 / 3.0}, Rb[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0}, Rg[3] = {0.0, 1.0 / 4.0, 2.0 / 3.0};
  int NV = M->nvar, NU = M->nvar_un;
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    for (int j = 1; j <= B->ny - 1; j++)
      for (int i = 1; i <= B->nx - 1; i++)
        for (int m = 1; m <= NV; m++) {
          double du = B->Res[RI(B, NV, m, i, j)] / B->vol[PI_(B, i, j)];
          size_t u = UI(B, NV, m, i, j);
          B->U[u] = Ra[krk - 1] * B->Un[UNI(B, NU, m, i, j)] + Rg[krk - 1] * B->U[u] + B->dt[CI(B, i, j)] * Rb[krk - 1] * du;
        }
  }
}

/* ------------------------------------------------------------------------
 * boundary conditions: sub_boundary.f90:10-34, 306-581, 45-153, 595-612
 * ---------------------------------------------------------------------- */
static void bc_wall(const OMesh *M, const OParams *P, OBlock *B, const int *s) {
  int NV = M->nvar, face = s[1], ib = s[2], ie = s[3], jb = s[4], je = s[5];
  double g = P->gamma, Ma = P->Ma, Tsb = 110.4 / P->T_inf;
  for (int j = jb; j <= je; j++)
    for (int i = ib; i <= ie; i++)
      for (int k = 1; k <= LAP; k++) {
        int i1, j1, i2, j2;
        if (face == 1) { i1 = i + k - 1; j1 = j; i2 = i - k; j2 = j; }
        else if (face == 2) { i1 = i; j1 = j + k - 1; i2 = i; j2 = j - k; }
        else if (face == 3) { i1 = i - k; j1 = j; i2 = i + k - 1; j2 = j; }
        else { i1 = i; j1 = j - k; i2 = i; j2 = j + k - 1; }
        double *U = B->U;
        double d1 = 0;
        if (P->Twall <= 0) {
          U[UI(B, NV, 1, i2, j2)] = U[UI(B, NV, 1, i1, j1)];
          U[UI(B, NV, 2, i2, j2)] = -U[UI(B, NV, 2, i1, j1)];
          U[UI(B, NV, 3, i2, j2)] = -U[UI(B, NV, 3, i1, j1)];
          U[UI(B, NV, 4, i2, j2)] = U[UI(B, NV, 4, i1, j1)];
        } else {
          d1 = U[UI(B, NV, 1, i1, j1)];
          double u1 = U[UI(B, NV, 2, i1, j1)] / d1, v1 = U[UI(B, NV, 3, i1, j1)] / d1;
          double p1 = (U[UI(B, NV, 4, i1, j1)] - 0.5 * d1 * (u1 * u1 + v1 * v1)) * (g - 1.0);
          double T1 = g * Ma * Ma * p1 / d1;
          double p2 = p1, T2 = 2.0 * P->Twall - T1, u2 = -u1, v2 = -v1;
          double d2 = g * Ma * Ma * p2 / T2;
          U[UI(B, NV, 1, i2, j2)] = d2;
          U[UI(B, NV, 2, i2, j2)] = d2 * u2;
          U[UI(B, NV, 3, i2, j2)] = d2 * v2;
          U[UI(B, NV, 4, i2, j2)] = p2 / (g - 1.0) + 0.5 * d2 * (u2 * u2 + v2 * v2);
        }
        if (NV == 5) U[UI(B, NV, 5, i2, j2)] = 0.0;
        if (NV == 6) {
          double d2 = U[UI(B, NV, 1, i2, j2)];
          double T2 = (U[UI(B, NV, 4, i2, j2)] - 0.5 * (U[UI(B, NV, 2, i2, j2)] * U[UI(B, NV, 2, i2, j2)] +
                                                         U[UI(B, NV, 3, i2, j2)] * U[UI(B, NV, 3, i2, j2)]) / d2) / (P->Cv * d2);
          double Amu2 = (1.0 + Tsb) * sqrt(T2 * T2 * T2) / (Tsb + T2);
          double dwv = B->dw[WI(B, i1, j1)];
          double wt = 60.0 * Amu2 / (d2 * 0.075 * (dwv * dwv) * P->Re * P->Re);
          U[UI(B, NV, 5, i2, j2)] = -U[UI(B, NV, 5, i1, j1)];
          U[UI(B, NV, 6, i2, j2)] = (d1 + d2) * wt - U[UI(B, NV, 6, i1, j1)];
        }
      }
}

static void bc_symmetry(const OMesh *M, OBlock *B, const int *s) {
  int NV = M->nvar, face = s[1], ib = s[2], ie = s[3], jb = s[4], je = s[5];
  for (int j = jb; j <= je; j++)
    for (int i = ib; i <= ie; i++) {
      double dx, dy, si, n1, n2;
      if (face == 1 || face == 3) {
        dx = B->x[NI(B, i, j + 1)] - B->x[NI(B, i, j)];
        dy = B->y[NI(B, i, j + 1)] - B->y[NI(B, i, j)];
        si = sqrt(dx * dx + dy * dy);
        n1 = dy / si; n2 = -dx / si;
      } else {
        dx = B->x[NI(B, i + 1, j)] - B->x[NI(B, i, j)];
        dy = B->y[NI(B, i + 1, j)] - B->y[NI(B, i, j)];
        si = sqrt(dx * dx + dy * dy);
        n1 = -dy / si; n2 = dx / si;
      }
      for (int k = 1; k <= LAP; k++) {
        int i1, j1, i2, j2;
        if (face == 1) { i1 = i + k - 1; j1 = j; i2 = i - k; j2 = j; }
        else if (face == 2) { i1 = i; j1 = j + k - 1; i2 = i; j2 = j - k; }
        else if (face == 3) { i1 = i - k; j1 = j; i2 = i + k - 1; j2 = j; }
        else { i1 = i; j1 = j - k; i2 = i; j2 = j + k - 1; }
        double *U = B->U;
        double Vn = U[UI(B, NV, 2, i1, j1)] * n1 + U[UI(B, NV, 3, i1, j1)] * n2;
        U[UI(B, NV, 1, i2, j2)] = U[UI(B, NV, 1, i1, j1)];
        U[UI(B, NV, 2, i2, j2)] = U[UI(B, NV, 2, i1, j1)] - 2.0 * Vn * n1;
        U[UI(B, NV, 3, i2, j2)] = U[UI(B, NV, 3, i1, j1)] - 2.0 * Vn * n2;
        U[UI(B, NV, 4, i2, j2)] = U[UI(B, NV, 4, i1, j1)];
        if (NV >= 5) U[UI(B, NV, 5, i2, j2)] = U[UI(B, NV, 5, i1, j1)];
        if (NV == 6) U[UI(B, NV, 6, i2, j2)] = U[UI(B, NV, 6, i1, j1)];
      }
    }
}

static void bc_farfield(const OMesh *M, const OParams *P, OBlock *B, const int *s) {
  int NV = M->nvar, face = s[1], ib = s[2], ie = s[3], jb = s[4], je = s[5], bct = s[6];
  double g = P->gamma;
  double d_inf = 1.0, u_inf = 1.0 * cos(P->AoA), v_inf = 1.0 * sin(P->AoA), p_inf = 1.0 / (g * P->Ma * P->Ma);
  for (int j = jb; j <= je; j++)
    for (int i = ib; i <= ie; i++) {
      int i1, j1;
      if (face == 1 || face == 2) { i1 = i; j1 = j; }
      else if (face == 3) { i1 = i - 1; j1 = j; }
      else { i1 = i; j1 = j - 1; }
      double *U = B->U;
      double d1 = U[UI(B, NV, 1, i1, j1)], u1 = U[UI(B, NV, 2, i1, j1)] / d1, v1 = U[UI(B, NV, 3, i1, j1)] / d1;
      double p1 = (U[UI(B, NV, 4, i1, j1)] - 0.5 * d1 * (u1 * u1 + v1 * v1)) * (g - 1.0);
      double c1 = sqrt(g * p1 / d1);
      double dx, dy, sl, n1, n2;
      if (face == 1) {
        dx = B->x[NI(B, i, j + 1)] - B->x[NI(B, i, j)]; dy = B->y[NI(B, i, j + 1)] - B->y[NI(B, i, j)];
        sl = sqrt(dx * dx + dy * dy); n1 = -dy / sl; n2 = dx / sl;
      } else if (face == 3) {
        dx = B->x[NI(B, i, j + 1)] - B->x[NI(B, i, j)]; dy = B->y[NI(B, i, j + 1)] - B->y[NI(B, i, j)];
        sl = sqrt(dx * dx + dy * dy); n1 = dy / sl; n2 = -dx / sl;
      } else if (face == 2) {
        dx = B->x[NI(B, i + 1, j)] - B->x[NI(B, i, j)]; dy = B->y[NI(B, i + 1, j)] - B->y[NI(B, i, j)];
        sl = sqrt(dx * dx + dy * dy); n1 = dy / sl; n2 = -dx / sl;
      } else {
        dx = B->x[NI(B, i + 1, j)] - B->x[NI(B, i, j)]; dy = B->y[NI(B, i + 1, j)] - B->y[NI(B, i, j)];
        sl = sqrt(dx * dx + dy * dy); n1 = -dy / sl; n2 = dx / sl;
      }
      double Ma_n = (u1 * n1 + v1 * n2) / c1;
      for (int k = 1; k <= LAP; k++) {
        int i2, j2;
        if (face == 1) { i2 = i - k; j2 = j; }
        else if (face == 2) { i2 = i; j2 = j - k; }
        else if (face == 3) { i2 = i + k - 1; j2 = j; }
        else { i2 = i; j2 = j + k - 1; }
        if (bct == -50 || (bct == -20 && Ma_n >= 0.0)) {
          if (Ma_n > 1.0 || P->P_outlet < 0 || bct == -20) {
            for (int m = 1; m <= NV; m++) U[UI(B, NV, m, i2, j2)] = U[UI(B, NV, m, i1, j1)];
          } else {
            double pb = P->P_outlet;
            double db = d1 + (P->P_outlet - p1) / (c1 * c1);
            double ub = u1 + (p1 - P->P_outlet) / (d1 * c1) * n1;
            double vb = v1 + (p1 - P->P_outlet) / (d1 * c1) * n2;
            double p2 = 2.0 * pb - p1, d2 = 2.0 * db - d1, u2 = 2.0 * ub - u1, v2 = 2.0 * vb - v1;
            U[UI(B, NV, 1, i2, j2)] = d2;
            U[UI(B, NV, 2, i2, j2)] = d2 * u2;
            U[UI(B, NV, 3, i2, j2)] = d2 * v2;
            U[UI(B, NV, 4, i2, j2)] = p2 / (g - 1.0) + 0.5 * d2 * (u2 * u2 + v2 * v2);
            for (int m = 5; m <= NV; m++) U[UI(B, NV, m, i2, j2)] = U[UI(B, NV, m, i1, j1)];
          }
        } else {
          if (P->P_outlet < 0 || Ma_n <= -1.0) {
            U[UI(B, NV, 1, i2, j2)] = d_inf;
            U[UI(B, NV, 2, i2, j2)] = d_inf * u_inf;
            U[UI(B, NV, 3, i2, j2)] = d_inf * v_inf;
            U[UI(B, NV, 4, i2, j2)] = p_inf / (g - 1.0) + 0.5 * d_inf * (u_inf * u_inf + v_inf * v_inf);
          } else {
            double pb = 0.5 * (p1 + p_inf - d1 * c1 * ((u_inf - u1) * n1 + (v_inf - v1) * n2));
            double db = d_inf + (pb - p_inf) / (c1 * c1);
            double ub = u_inf - (p_inf - pb) / (d1 * c1) * n1;
            double vb = v_inf - (p_inf - pb) / (d1 * c1) * n2;
            double p2 = 2.0 * pb - p1, d2 = 2.0 * db - d1, u2 = 2.0 * ub - u1, v2 = 2.0 * vb - v1;
            U[UI(B, NV, 1, i2, j2)] = d2;
            U[UI(B, NV, 2, i2, j2)] = d2 * u2;
            U[UI(B, NV, 3, i2, j2)] = d2 * v2;
            U[UI(B, NV, 4, i2, j2)] = p2 / (g - 1.0) + 0.5 * d2 * (u2 * u2 + v2 * v2);
          }
          if (NV == 5) U[UI(B, NV, 5, i2, j2)] = P->vt_inf / P->Re;
          else if (NV == 6) {
            U[UI(B, NV, 5, i2, j2)] = d_inf * P->Kt_inf;
            U[UI(B, NV, 6, i2, j2)] = d_inf * P->Wt_inf;
          }
        }
      }
    }
}

void o_bc(OMesh *M, const OParams *P) {
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    for (int k = 0; k < B->nsub; k++) {
      const int *s = B->bc + 9 * k;
      int nb = s[6];
      if (nb >= 0) continue;
      if (nb == -10) bc_wall(M, P, B, s);
      else if (nb == -20 || nb == -40 || nb == -50) bc_farfield(M, P, B, s);
      else if (nb == -30) bc_symmetry(M, B, s);
    }
  }
}

static void u_average_corner(int NV, const double *U1, const double *U2, const double *U3, double *U4, double Cv) {
  double d1 = U1[0], uu1 = U1[1] / d1, v1 = U1[2] / d1, T1 = (U1[3] - (uu1 * U1[1] + v1 * U1[2]) * 0.5) / (d1 * Cv);
  double d2 = U2[0], uu2 = U2[1] / d2, v2 = U2[2] / d2, T2 = (U2[3] - (uu2 * U2[1] + v2 * U2[2]) * 0.5) / (d2 * Cv);
  double d3 = U3[0], uu3 = U3[1] / d3, v3 = U3[2] / d3, T3 = (U3[3] - (uu3 * U3[1] + v3 * U3[2]) * 0.5) / (d3 * Cv);
  double d4 = d1 + d3 - d2, uu4 = uu1 + uu3 - uu2, v4 = v1 + v3 - v2, T4 = T1 + T3 - T2;
  U4[0] = d4; U4[1] = d4 * uu4; U4[2] = d4 * v4; U4[3] = d4 * (Cv * T4 + (uu4 * uu4 + v4 * v4) * 0.5);
  if (NV == 5) U4[4] = U1[4] + U3[4] - U2[4];
  if (NV == 6) { U4[4] = U1[4] + U3[4] - U2[4]; U4[5] = U1[5] + U3[5] - U2[5]; }
}

void o_exchange(OMesh *M, const OParams *P) {
  int NV = M->nvar;
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    for (int ks = 0; ks < B->nsub; ks++) {
      const int *s = B->bc + 9 * ks;
      if (s[6] < 0) continue;
      int ib2 = s[2], ie2 = s[3], jb2 = s[4], je2 = s[5];
      int n1 = ie2 - ib2 + je2 - jb2;
      OBlock *B1 = &M->blocks[s[6] - 1];
      const int *s1 = B1->bc + 9 * (s[7] - 1);
      int f1 = s1[1], ib1 = s1[2], ie1 = s1[3], jb1 = s1[4], je1 = s1[5];
      double *Ua = malloc(sizeof(double) * NV * (n1 > 0 ? n1 : 1) * LAP);
#define UA(m, k, l) Ua[(size_t)((m) - 1) + (size_t)NV * ((size_t)((k) - 1) + (size_t)n1 * (size_t)((l) - 1))]
      if (f1 == 1) {
        for (int j = jb1; j <= je1 - 1; j++)
          for (int i = 1; i <= LAP; i++)
            for (int m = 1; m <= NV; m++) UA(m, j - jb1 + 1, i) = B1->U[UI(B1, NV, m, ib1 + i - 1, j)];
      } else if (f1 == 2) {
        for (int j = 1; j <= LAP; j++)
          for (int i = ib1; i <= ie1 - 1; i++)
            for (int m = 1; m <= NV; m++) UA(m, i - ib1 + 1, j) = B1->U[UI(B1, NV, m, i, jb1 + j - 1)];
      } else if (f1 == 3) {
        for (int j = jb1; j <= je1 - 1; j++)
          for (int i = 1; i <= LAP; i++)
            for (int m = 1; m <= NV; m++) UA(m, j - jb1 + 1, i) = B1->U[UI(B1, NV, m, ie1 - LAP - 1 + i, j)];
      } else {
        for (int j = 1; j <= LAP; j++)
          for (int i = ib1; i <= ie1 - 1; i++)
            for (int m = 1; m <= NV; m++) UA(m, i - ib1 + 1, j) = B1->U[UI(B1, NV, m, i, je1 - LAP - 1 + j)];
      }
      int face = s[1], orient = s[8];
      for (int k = 1; k <= n1; k++) {
        int k1 = (orient == 2) ? k : n1 + 1 - k;
        for (int l = 1; l <= LAP; l++)
          for (int m = 1; m <= NV; m++) {
            double val = UA(m, k1, l);
            if (face == 1) B->U[UI(B, NV, m, ib2 - LAP - 1 + l, jb2 + k - 1)] = val;
            else if (face == 2) B->U[UI(B, NV, m, ib2 + k - 1, jb2 - LAP - 1 + l)] = val;
            else if (face == 3) B->U[UI(B, NV, m, ie2 + l - 1, jb2 + k - 1)] = val;
            else B->U[UI(B, NV, m, ib2 + k - 1, je2 + l - 1)] = val;
          }
      }
#undef UA
      free(Ua);
    }
    int nx1 = B->nx, ny1 = B->ny;
    double *U = B->U;
    double Cv = P->Cv;
    u_average_corner(NV, &U[UI(B, NV, 1, 1, 0)], &U[UI(B, NV, 1, 1, 1)], &U[UI(B, NV, 1, 0, 1)], &U[UI(B, NV, 1, 0, 0)], Cv);
    u_average_corner(NV, &U[UI(B, NV, 1, 1, ny1)], &U[UI(B, NV, 1, 1, ny1 - 1)], &U[UI(B, NV, 1, 0, ny1 - 1)], &U[UI(B, NV, 1, 0, ny1)], Cv);
    u_average_corner(NV, &U[UI(B, NV, 1, nx1 - 1, 0)], &U[UI(B, NV, 1, nx1 - 1, 1)], &U[UI(B, NV, 1, nx1, 1)], &U[UI(B, NV, 1, nx1, 0)], Cv);
    u_average_corner(NV, &U[UI(B, NV, 1, nx1 - 1, ny1)], &U[UI(B, NV, 1, nx1 - 1, ny1 - 1)], &U[UI(B, NV, 1, nx1, ny1 - 1)],
                     &U[UI(B, NV, 1, nx1, ny1)], Cv);
  }
}

/* ------------------------------------------------------------------------
 * multigrid transfers: sub_NS_multigrid.f90:156-398 (4 variables)
 * ---------------------------------------------------------------------- */
#define QI(B, m, i, j) ((size_t)((m) - 1) + 4 * CI(B, i, j))
#define DLI(B, m, i, j) ((size_t)((m) - 1) + 4 * GI(B, i, j))

void o_force_function(OMesh *M) {
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    for (int j = 1; j <= B->ny - 1; j++)
      for (int i = 1; i <= B->nx - 1; i++)
        for (int m = 1; m <= 4; m++) B->QF[QI(B, m, i, j)] = B->QF[QI(B, m, i, j)] - B->Res[RI(B, M->nvar, m, i, j)];
  }
}

void o_add_force(OMesh *M) {
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    for (int j = 1; j <= B->ny - 1; j++)
      for (int i = 1; i <= B->nx - 1; i++)
        for (int m = 1; m <= 4; m++) {
          size_t r = RI(B, M->nvar, m, i, j);
          B->Res[r] = B->Res[r] + B->QF[QI(B, m, i, j)];
        }
  }
}

void o_delt_u(OMesh *M) {
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    for (int j = 0; j <= B->ny; j++)
      for (int i = 0; i <= B->nx; i++)
        for (int m = 1; m <= 4; m++)
          B->deltU[DLI(B, m, i, j)] = B->U[UI(B, M->nvar, m, i, j)] - B->Un[UNI(B, M->nvar_un, m, i, j)];
  }
}

void o_set_un(OMesh *M) {
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    for (int j = 0; j <= B->ny; j++)
      for (int i = 0; i <= B->nx; i++)
        for (int m = 1; m <= 4; m++) B->Un[UNI(B, M->nvar_un, m, i, j)] = B->U[UI(B, M->nvar, m, i, j)];
  }
}

/* Init_flow, dual-time stepping: Un = Un1 = U (sub_init.f90:396-409) */
void o_dual_init(OMesh *M) {
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    for (int j = 1; j <= B->ny - 1; j++)
      for (int i = 1; i <= B->nx - 1; i++)
        for (int m = 1; m <= M->nvar; m++) {
          size_t w = UNI(B, M->nvar_un, m, i, j);
          B->Un[w] = B->U[UI(B, M->nvar, m, i, j)];
          B->Un1[w] = B->U[UI(B, M->nvar, m, i, j)];
        }
  }
}

/* end of a dual-time step: Un1 = Un, Un = U (sub_NS_singlegrid.f90:282-299) */
void o_dual_shift(OMesh *M) {
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    for (int j = 1; j <= B->ny - 1; j++)
      for (int i = 1; i <= B->nx - 1; i++)
        for (int m = 1; m <= M->nvar; m++) {
          size_t w = UNI(B, M->nvar_un, m, i, j);
          B->Un1[w] = B->Un[w];
          B->Un[w] = B->U[UI(B, M->nvar, m, i, j)];
        }
  }
}

void o_new_u(OMesh *M) {
  for (int mb = 0; mb < M->nblock; mb++) {
    OBlock *B = &M->blocks[mb];
    for (int j = 1; j <= B->ny - 1; j++)
      for (int i = 1; i <= B->nx - 1; i++)
        for (int m = 1; m <= 4; m++) {
          size_t u = UI(B, M->nvar, m, i, j);
          B->U[u] = B->U[u] + B->deltU[DLI(B, m, i, j)];
        }
  }
}

/* coarse mesh Mc, fine mesh Mf (one level finer); flag 1: U, flag 2: deltU */
void o_prolong(OMesh *Mc, OMesh *Mf, int flag) {
  const double a1 = 9.0 / 16.0, a2 = 3.0 / 16.0, a3 = 1.0 / 16.0;
  for (int mb = 0; mb < Mc->nblock; mb++) {
    OBlock *B1 = &Mc->blocks[mb], *B2 = &Mf->blocks[mb];
    int nx2 = B2->nx, ny2 = B2->ny;
    for (int j = 0; j <= ny2; j++) {
      int ja1 = (j % 2 == 0) ? j / 2 : j / 2 + 1, ja2 = (j % 2 == 0) ? j / 2 + 1 : j / 2;
      for (int i = 0; i <= nx2; i++) {
        int ia1 = (i % 2 == 0) ? i / 2 : i / 2 + 1, ia2 = (i % 2 == 0) ? i / 2 + 1 : i / 2;
        for (int m = 1; m <= 4; m++) {
          if (flag == 1) {
            int N1 = Mc->nvar, N2 = Mf->nvar;
            B2->U[UI(B2, N2, m, i, j)] = a1 * B1->U[UI(B1, N1, m, ia1, ja1)] +
                                         a2 * (B1->U[UI(B1, N1, m, ia2, ja1)] + B1->U[UI(B1, N1, m, ia1, ja2)]) +
                                         a3 * B1->U[UI(B1, N1, m, ia2, ja2)];
          } else {
            B2->deltU[DLI(B2, m, i, j)] = a1 * B1->deltU[DLI(B1, m, ia1, ja1)] +
                                          a2 * (B1->deltU[DLI(B1, m, ia2, ja1)] + B1->deltU[DLI(B1, m, ia1, ja2)]) +
                                          a3 * B1->deltU[DLI(B1, m, ia2, ja2)];
          }
        }
      }
    }
  }
}

/* fine mesh Mf -> coarse mesh Mc; flag 1: U (volume weighted), flag 2: Res summed into QF */
void o_interp2h(OMesh *Mf, OMesh *Mc, int flag) {
  for (int mb = 0; mb < Mf->nblock; mb++) {
    OBlock *B1 = &Mf->blocks[mb], *B2 = &Mc->blocks[mb];
    for (int j = 1; j <= B2->ny - 1; j++)
      for (int i = 1; i <= B2->nx - 1; i++) {
        int i1 = 2 * i - 1, i2 = 2 * i, j1 = 2 * j - 1, j2 = 2 * j;
        for (int m = 1; m <= 4; m++) {
          if (flag == 1) {
            int N1 = Mf->nvar, N2 = Mc->nvar;
            B2->U[UI(B2, N2, m, i, j)] =
                (B1->U[UI(B1, N1, m, i1, j1)] * B1->vol[PI_(B1, i1, j1)] + B1->U[UI(B1, N1, m, i1, j2)] * B1->vol[PI_(B1, i1, j2)] +
                 B1->U[UI(B1, N1, m, i2, j1)] * B1->vol[PI_(B1, i2, j1)] + B1->U[UI(B1, N1, m, i2, j2)] * B1->vol[PI_(B1, i2, j2)]) /
                B2->vol[PI_(B2, i, j)];
          } else {
            int N1 = Mf->nvar;
            B2->QF[QI(B2, m, i, j)] = B1->Res[RI(B1, N1, m, i1, j1)] + B1->Res[RI(B1, N1, m, i1, j2)] +
                                      B1->Res[RI(B1, N1, m, i2, j1)] + B1->Res[RI(B1, N1, m, i2, j2)];
          }
        }
      }
  }
}

/* ------------------------------------------------------------------------
 * pointwise entry points for the closed-form tests of the oracle itself
 * (flux consistency, scheme exactness, WENO5 reproduction, DF identity)
 * ---------------------------------------------------------------------- */
void o_pt_flux(int iflux, const double *QL, const double *QR, double *F, double g) { face_flux(iflux, QL, QR, F, g); }

double o_pt_scheme(int side, const double *u, int flag) {
  return side > 0 ? scheme_fP(u[0], u[1], u[2], u[3], flag) : scheme_fm(u[0], u[1], u[2], u[3], flag);
}

/* kind 0 original, 1 conservative, 2 characteristic; U0[k*4+m] */
void o_pt_recon(int kind, const double *U0, double *UL, double *UR, double g, int flag) {
  double A[4][4];
  for (int k = 0; k < 4; k++)
    for (int m = 0; m < 4; m++) A[k][m] = U0[4 * k + m];
  if (kind == 0) recon_original(A, UL, UR, flag);
  else if (kind == 1) recon_conservative(A, UL, UR, g, flag);
  else recon_characteristic(A, UL, UR, g, flag);
}

/* side > 0: fp_weno5, else fm_weno5; v and h are 0-based arrays of n values */
void o_pt_weno5(int side, int n, const double *v, double *h) {
  if (side > 0) fp_weno5(n, v - 1, h - 1);
  else fm_weno5(n, v - 1, h - 1);
}

void o_pt_split_sw(double d, double u, double v, double cc, double A1, double A2, double g, double *fp, double *fm) {
  split_sw(d, u, v, cc, fp, fm, A1, A2, g);
}

void o_pt_dfn(int NV, double *DF, const double *U, const double *DU, double n1, double n2, double g) {
  comput_dfn(NV, DF, U, DU, n1, n2, g);
}

/* ------------------------------------------------------------------------
 * Wall forces.  The reference has no force integration (only wall Cp in a
 * separate post-processing tool, read_flow-1.2.f90:36-51); this is the
 * definition the device path implements, restated serially:
 *   every wall face (subfaces with neighb = -10, in block / file / index
 *   order) contributes, with n its unit normal pointing into the fluid and s
 *   its length,
 *     pressure  -(p - p_inf) n s, p of the adjacent interior cell,
 *     viscous   (tau . n) s, tau = mu (grad u + grad u^T - 2/3 div u I)
 *               from the face-centred gradient stencil of Residual_FVM
 *               (sub_Residual.f90:301-343, 408-450), mu = Amu + Amu_t of the
 *               adjacent interior cell (the wall-face rule of that routine).
 *   F[0..3] = (Fx, Fy) pressure, (Fx, Fy) viscous, sums in face order.
 * ---------------------------------------------------------------------- */
void o_forces(const OMesh *M, const OParams *P, double *F) {
  double g = P->gamma, p_inf = 1.0 / (g * P->Ma * P->Ma);
  int NV = M->nvar;
  F[0] = F[1] = F[2] = F[3] = 0.0;
  for (int mb = 0; mb < M->nblock; mb++) {
    const OBlock *B = &M->blocks[mb];
    int nx = B->nx, ny = B->ny;
    const double *U = B->U, *x1 = B->x1, *y1 = B->y1;
#define UU(i, j) (U[UI(B, NV, 2, i, j)] / U[UI(B, NV, 1, i, j)])
#define VV(i, j) (U[UI(B, NV, 3, i, j)] / U[UI(B, NV, 1, i, j)])
    for (int k = 0; k < B->nsub; k++) {
      const int *s = B->bc + 9 * k;
      if (s[6] != -10) continue;
      int face = s[1];
      int n = (face == 1 || face == 3) ? s[5] - s[4] : s[3] - s[2];
      for (int t = 0; t < n; t++) {
        int i, j, ic, jc;  /* face index (i, j) and adjacent interior cell (ic, jc) */
        double sgn;
        if (face == 1) { i = 1; j = s[4] + t; ic = 1; jc = j; sgn = 1.0; }
        else if (face == 3) { i = nx; j = s[4] + t; ic = nx - 1; jc = j; sgn = -1.0; }
        else if (face == 2) { i = s[2] + t; j = 1; ic = i; jc = 1; sgn = 1.0; }
        else { i = s[2] + t; j = ny; ic = i; jc = ny - 1; sgn = -1.0; }
        int iface = (face == 1 || face == 3);
        size_t pp = PI_(B, i, j);
        double sl = iface ? B->si[pp] : B->sj[pp];
        double n1 = sgn * (iface ? B->ni1[pp] : B->nj1[pp]);
        double n2 = sgn * (iface ? B->ni2[pp] : B->nj2[pp]);
        size_t uc = GI(B, ic, jc);
        double d = U[UI(B, NV, 1, ic, jc)];
        double u = U[UI(B, NV, 2, ic, jc)] / d, v = U[UI(B, NV, 3, ic, jc)] / d;
        double p = (g - 1.0) * (U[UI(B, NV, 4, ic, jc)] - 0.5 * d * (u * u + v * v));
        double dp = p - p_inf;
        F[0] = F[0] - dp * n1 * sl;
        F[1] = F[1] - dp * n2 * sl;
        if (P->If_viscous == 1) {
          double Dix, Diy, Djx, Djy, Diu, Div, Dju, Djv;
          if (iface) {
            Dix = x1[GI(B, i, j)] - x1[GI(B, i - 1, j)];
            Diy = y1[GI(B, i, j)] - y1[GI(B, i - 1, j)];
            Djx = (x1[GI(B, i - 1, j + 1)] + x1[GI(B, i, j + 1)] - x1[GI(B, i - 1, j - 1)] - x1[GI(B, i, j - 1)]) * 0.25;
            Djy = (y1[GI(B, i - 1, j + 1)] + y1[GI(B, i, j + 1)] - y1[GI(B, i - 1, j - 1)] - y1[GI(B, i, j - 1)]) * 0.25;
            Diu = UU(i, j) - UU(i - 1, j);
            Div = VV(i, j) - VV(i - 1, j);
            Dju = (UU(i - 1, j + 1) + UU(i, j + 1) - UU(i - 1, j - 1) - UU(i, j - 1)) * 0.25;
            Djv = (VV(i - 1, j + 1) + VV(i, j + 1) - VV(i - 1, j - 1) - VV(i, j - 1)) * 0.25;
          } else {
            Dix = (x1[GI(B, i + 1, j - 1)] + x1[GI(B, i + 1, j)] - x1[GI(B, i - 1, j - 1)] - x1[GI(B, i - 1, j)]) * 0.25;
            Diy = (y1[GI(B, i + 1, j - 1)] + y1[GI(B, i + 1, j)] - y1[GI(B, i - 1, j - 1)] - y1[GI(B, i - 1, j)]) * 0.25;
            Djx = x1[GI(B, i, j)] - x1[GI(B, i, j - 1)];
            Djy = y1[GI(B, i, j)] - y1[GI(B, i, j - 1)];
            Diu = (UU(i + 1, j - 1) + UU(i + 1, j) - UU(i - 1, j - 1) - UU(i - 1, j)) * 0.25;
            Div = (VV(i + 1, j - 1) + VV(i + 1, j) - VV(i - 1, j - 1) - VV(i - 1, j)) * 0.25;
            Dju = UU(i, j) - UU(i, j - 1);
            Djv = VV(i, j) - VV(i, j - 1);
          }
          double Ds = 1.0 / (Dix * Djy - Djx * Diy);
          double ux = (Diu * Djy - Dju * Diy) * Ds, vx = (Div * Djy - Djv * Diy) * Ds;
          double uy = (-Diu * Djx + Dju * Dix) * Ds, vy = (-Div * Djx + Djv * Dix) * Ds;
          double mu = B->Amu[uc] + B->Amu_t[uc];
          double div = ux + vy;
          double t11 = mu * (2.0 * ux - 2.0 / 3.0 * div);
          double t12 = mu * (uy + vx);
          double t22 = mu * (2.0 * vy - 2.0 / 3.0 * div);
          F[2] = F[2] + (t11 * n1 + t12 * n2) * sl;
          F[3] = F[3] + (t12 * n1 + t22 * n2) * sl;
        }
      }
    }
#undef UU
#undef VV
  }
}
