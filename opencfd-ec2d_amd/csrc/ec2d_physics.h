// Point-wise physics of the residual path as device functions.
//
// Every expression keeps the operand order of the reference Fortran so that,
// compiled with -ffp-contract=off (no fused multiply-add), the GPU rounds
// exactly as the reference does.  Citations are to the reference sources.
#pragma once
#include <hip/hip_runtime.h>

#define EC2D_PRT 0.9
// Fortran default-real literal widened to double (e.g. 0.3 in
// sub_turbulence_SA.f90:159 is 0.30000001192092896)
#define EC2D_SP(x) ((double)(float)(x))

namespace ec2d {

// a / b, correctly rounded, from y = RN(1/b) (Markstein): for b > 0 and
// normal-range operands; q is faithful after the first refinement, and a
// faithful q refined with RN(1/b) is RN(a/b).  Five dependent operations
// instead of the eleven of a division; used where several quotients share
// a divisor or the divisor is a constant.
__device__ __forceinline__ double div_rcp(double a, double b, double y) {
  double q = a * y;
  double r = __builtin_fma(-b, q, a);
  q = __builtin_fma(r, y, q);
  r = __builtin_fma(-b, q, a);
  return __builtin_fma(r, y, q);
}

// sub_Residual.f90:477-487
__device__ __forceinline__ double minmod(double a, double b) {
  if (a * b <= 0.0) return 0.0;
  if (fabs(a) <= fabs(b)) return a;
  return b;
}

// left state at i+1/2 from u(i-1), u(i), u(i+1), u(i+2): sub_Residual.f90:591-623
__device__ __forceinline__ double scheme_fP(double u1, double u2, double u3, double u4, int flag) {
  const double k3 = 1.0 / 3.0, ep = 1.0e-6;
  switch (flag) {
    case 0: return u2;
    case 1: return u2 + 0.5 * minmod(u2 - u1, u3 - u2);
    case 2: return (-u1 + 5.0 * u2 + 2.0 * u3) / 6.0;
    case 3: {
      double IS1 = (u2 - u1) * (u2 - u1), IS2 = (u3 - u2) * (u3 - u2);
      double t1 = 1.0e-6 + IS1, t2 = 1.0e-6 + IS2;
      double a1 = 1.0 / (3.0 * (t1 * t1)), a2 = 2.0 / (3.0 * (t2 * t2));
      double w1 = a1 / (a1 + a2), w2 = a2 / (a1 + a2);
      return w1 * (-u1 / 2.0 + 3.0 * u2 / 2.0) + w2 * (u2 / 2.0 + u3 / 2.0);
    }
    case 4: {  // MUSCL2 with k = 1, b = 2
      return u2 + 0.25 * ((1.0 - 1.0) * minmod(u2 - u1, 2.0 * (u3 - u2)) + (1.0 + 1.0) * minmod(u3 - u2, 2.0 * (u2 - u1)));
    }
    case 5: {  // MUSCL3 with van Albada limiter
      double up = u3 - u2, um = u2 - u1;
      double s = (2.0 * up * um + ep) / (up * up + um * um + ep);
      return u2 + 0.25 * s * ((1.0 - k3 * s) * um + (1.0 + k3 * s) * up);
    }
    default: {  // 6: OMUSCL2
      double r1 = (u2 - u1 + ep) / (u3 - u2 + ep), r2 = (u3 - u2 + ep) / (u4 - u3 + ep);
      double f1 = 0.8 - 0.175 / r2 + 0.375 * r1;
      double f = fmax(0.0, fmin(fmin(2.0, f1), 2.0 * r1));
      return u2 + 0.5 * f * (u3 - u2);
    }
  }
}

// right state at i+1/2: sub_Residual.f90:626-659
__device__ __forceinline__ double scheme_fm(double u1, double u2, double u3, double u4, int flag) {
  const double k = 1.0 / 3.0, k3 = 1.0 / 3.0, ep = 1.0e-6;
  switch (flag) {
    case 0: return u3;
    case 1: return u3 - 0.5 * minmod(u3 - u2, u4 - u3);
    case 2: return (2.0 * u2 + 5.0 * u3 - u4) / 6.0;
    case 3: {
      double IS1 = (u3 - u4) * (u3 - u4), IS2 = (u2 - u3) * (u2 - u3);
      double t1 = 1.0e-6 + IS1, t2 = 1.0e-6 + IS2;
      double a1 = 1.0 / (3.0 * (t1 * t1)), a2 = 2.0 / (3.0 * (t2 * t2));
      double w1 = a1 / (a1 + a2), w2 = a2 / (a1 + a2);
      return w1 * (-u4 / 2.0 + 3.0 * u3 / 2.0) + w2 * (u3 / 2.0 + u2 / 2.0);
    }
    case 4: {  // MUSCL2 with k = 1/3, b = 1
      return u3 - 0.25 * ((1.0 - k) * minmod(u4 - u3, 1.0 * (u3 - u2)) + (1.0 + k) * minmod(u3 - u2, 1.0 * (u4 - u3)));
    }
    case 5: {
      double up = u4 - u3, um = u3 - u2;
      double s = (2.0 * up * um + ep) / (up * up + um * um + ep);
      return u3 - 0.25 * s * ((1.0 - k3 * s) * up + (1.0 + k3 * s) * um);
    }
    default: {
      double r1 = (u4 - u3 + ep) / (u3 - u2 + ep), r2 = (u3 - u2 + ep) / (u2 - u1 + ep);
      double f1 = 0.8 - 0.175 / r2 + 0.375 * r1;
      double f = fmax(0.0, fmin(fmin(2.0, f1), 2.0 * r1));
      return u3 - 0.5 * f * (u3 - u2);
    }
  }
}

// conservative -> (d, u, v, p) after reconstruction: sub_Residual.f90:526-527
__device__ __forceinline__ void cons_to_prim(const double Q[4], double W[4], double g) {
  W[0] = Q[0];
  W[1] = Q[1] / W[0];
  W[2] = Q[2] / W[0];
  W[3] = (Q[3] - (W[1] * Q[1] + W[2] * Q[2]) * 0.5) * (g - 1.0);
}

// characteristic reconstruction: sub_Residual.f90:531-586.  U0[k][m]: k = i-2..i+1
__device__ __forceinline__ void recon_characteristic(const double U0[4][4], double UL[4], double UR[4], double g,
                                                     int flag) {
  double Uh[4];
  for (int m = 0; m < 4; m++) Uh[m] = 0.5 * (U0[1][m] + U0[2][m]);
  double d1 = Uh[0], u1 = Uh[1] / d1, v1 = Uh[2] / d1;
  double p1 = (Uh[3] - (Uh[1] * u1 + Uh[2] * v1) * 0.5) * (g - 1.0);
  double c1 = sqrt(g * p1 / d1);
  double V2 = (u1 * u1 + v1 * v1) * 0.5;
  double tmp1 = (g - 1.0) / c1;
  double tmp3 = (g - 1.0) / (c1 * c1);
  double tmp5 = 1.0 / (2.0 * c1);
  double tmp0 = 1.0 / tmp3;
  double S[4][4], S1[4][4];
  S[0][0] = V2 - tmp0; S[0][1] = -u1; S[0][2] = -v1; S[0][3] = 1.0;
  S[1][0] = -v1; S[1][1] = 0.0; S[1][2] = 1.0; S[1][3] = 0.0;
  S[2][0] = -u1 - V2 * tmp1; S[2][1] = 1.0 + tmp1 * u1; S[2][2] = tmp1 * v1; S[2][3] = -tmp1;
  S[3][0] = -u1 + V2 * tmp1; S[3][1] = 1.0 - tmp1 * u1; S[3][2] = -tmp1 * v1; S[3][3] = tmp1;
  S1[0][0] = -tmp3; S1[0][1] = 0.0; S1[0][2] = -tmp5; S1[0][3] = tmp5;
  S1[1][0] = -tmp3 * u1; S1[1][1] = 0.0; S1[1][2] = 0.5 - u1 * tmp5; S1[1][3] = 0.5 + u1 * tmp5;
  S1[2][0] = -tmp3 * v1; S1[2][1] = 1.0; S1[2][2] = -v1 * tmp5; S1[2][3] = v1 * tmp5;
  S1[3][0] = -tmp3 * V2; S1[3][1] = v1; S1[3][2] = (c1 * u1 - V2 - tmp0) * tmp5; S1[3][3] = (c1 * u1 + V2 + tmp0) * tmp5;
  double V0[4][4], VL[4], VR[4], QL[4], QR[4];
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
  cons_to_prim(QL, UL, g);
  cons_to_prim(QR, UR, g);
}

// ---- interface fluxes, Q = (rho, u_n, u_t, p): sub_flux_split.f90 -----------
// The quotients keep the reference's operands; those that share a divisor
// (1 + tmpp three times, ava*ava and 2*ava*ava = 2 RN(ava*ava) exactly) or
// divide by a constant (2 delt) are formed with div_rcp from one correctly
// rounded reciprocal, which gives the same RN(a/b).
__device__ __forceinline__ void flux_roe(const double* QL, const double* QR, double* Fx, double g) {
  constexpr double delt = 0.1;  // :271
  constexpr double y2d = 1.0 / (2.0 * delt);
  double dl = QL[0], uul = QL[1], vvl = QL[2], pl = QL[3];
  double dr = QR[0], uur = QR[1], vvr = QR[2], pr = QR[3];
  double hl = g * pl / ((g - 1.0) * dl) + 0.5 * (uul * uul + vvl * vvl);
  double hr = g * pr / ((g - 1.0) * dr) + 0.5 * (uur * uur + vvr * vvr);
  double Fl0 = dl * uul, Fl1 = dl * uul * uul + pl, Fl2 = dl * uul * vvl, Fl3 = uul * dl * hl;
  double Fr0 = dr * uur, Fr1 = dr * uur * uur + pr, Fr2 = dr * uur * vvr, Fr3 = uur * dr * hr;
  double tmpp = sqrt(dr / dl);
  double avd = 0.25 * dl * (1.0 + tmpp) * (1.0 + tmpp);
  const double t1 = 1.0 + tmpp, yt = 1.0 / t1;
  double avu = div_rcp(uul + tmpp * uur, t1, yt);
  double avv = div_rcp(vvl + tmpp * vvr, t1, yt);
  double avh = div_rcp(hl + tmpp * hr, t1, yt);
  double ava = sqrt((g - 1.0) * (avh - 0.5 * (avu * avu + avv * avv)));
  double l0 = fabs(avu - ava), l1 = fabs(avu), l2 = fabs(avu), l3 = fabs(avu + ava);
  if (l0 < delt) l0 = div_rcp(l0 * l0 + delt * delt, 2.0 * delt, y2d);
  if (l1 < delt) l1 = div_rcp(l1 * l1 + delt * delt, 2.0 * delt, y2d);
  if (l2 < delt) l2 = div_rcp(l2 * l2 + delt * delt, 2.0 * delt, y2d);
  if (l3 < delt) l3 = div_rcp(l3 * l3 + delt * delt, 2.0 * delt, y2d);
  const double A = ava * ava, yA = 1.0 / A;
  double a0 = div_rcp(avd * ava * (-uur + uul) + pr - pl, 2.0 * A, 0.5 * yA);
  double a1 = dr - dl - div_rcp(pr - pl, A, yA);
  double a2 = avd * (vvr - vvl) / ava;
  double a3 = div_rcp(avd * ava * (uur - uul) + pr - pl, 2.0 * A, 0.5 * yA);
  Fx[0] = 0.5 * (Fl0 + Fr0) - 0.5 * (l0 * a0 + l1 * a1 + l3 * a3);
  Fx[1] = 0.5 * (Fl1 + Fr1) - 0.5 * (l0 * a0 * (avu - ava) + l1 * a1 * avu + l3 * a3 * (avu + ava));
  Fx[2] = 0.5 * (Fl2 + Fr2) - 0.5 * (l0 * a0 * avv + l1 * a1 * avv + l2 * a2 + l3 * a3 * avv);
  Fx[3] = 0.5 * (Fl3 + Fr3) - 0.5 * (l0 * a0 * (avh - avu * ava) + l1 * a1 * 0.5 * (avu * avu + avv * avv) +
                                     l2 * a2 * avv + l3 * a3 * (avh + avu * ava));
}

__device__ __forceinline__ void flux_ausm(const double* QL, const double* QR, double* Fx, double g) {
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
  Fx[0] = 0.5 * xm * (1.0 + 1.0) - 0.5 * dm * (1.0 - 1.0);
  Fx[1] = 0.5 * xm * (uul + uur) - 0.5 * dm * (uur - uul) + p;
  Fx[2] = 0.5 * xm * (vvl + vvr) - 0.5 * dm * (vvr - vvl);
  Fx[3] = 0.5 * xm * (hl + hr) - 0.5 * dm * (hr - hl);
}

__device__ __forceinline__ void flux_vanleer(const double* QL, const double* QR, double* Fx, double g) {
  double dl = QL[0], uul = QL[1], vvl = QL[2], pl = QL[3];
  double dr = QR[0], uur = QR[1], vvr = QR[2], pr = QR[3];
  double al = sqrt(g * pl / dl), ar = sqrt(g * pr / dr);
  double Ml = uul / al, Mr = uur / ar;
  double fp0 = 0, fp1 = 0, fp2 = 0, fp3 = 0, fm0 = 0, fm1 = 0, fm2 = 0, fm3 = 0;
  if (Ml >= 1.0) {
    fp0 = dl * uul; fp1 = dl * uul * uul + pl; fp2 = dl * uul * vvl;
    fp3 = uul * (g * pl / (g - 1.0) + 0.5 * dl * (uul * uul + vvl * vvl));
  } else if (fabs(Ml) < 1.0) {
    double Mp = 0.25 * (1.0 + Ml) * (1.0 + Ml);
    double tmp0 = dl * al * Mp;
    double w = (g - 1.0) * uul + 2.0 * al;
    fp0 = tmp0; fp1 = tmp0 * w / g; fp2 = tmp0 * vvl;
    fp3 = tmp0 * (w * w * 0.5 / (g * g - 1.0) + 0.5 * (vvl * vvl));
  }
  if (fabs(Mr) < 1.0) {
    double Mm = -0.25 * (Mr - 1.0) * (Mr - 1.0);
    double tmp0 = dr * ar * Mm;
    double w = (g - 1.0) * uur - 2.0 * ar;
    fm0 = tmp0; fm1 = tmp0 * w / g; fm2 = tmp0 * vvr;
    fm3 = tmp0 * (w * w * 0.5 / (g * g - 1.0) + 0.5 * (vvr * vvr));
  } else if (Mr <= -1.0) {
    fm0 = dr * uur; fm1 = dr * uur * uur + pr; fm2 = dr * uur * vvr;
    fm3 = uur * (g * pr / (g - 1.0) + 0.5 * dr * (uur * uur + vvr * vvr));
  }
  Fx[0] = fp0 + fm0; Fx[1] = fp1 + fm1; Fx[2] = fp2 + fm2; Fx[3] = fp3 + fm3;
}

__device__ __forceinline__ void flux_sw(const double* QL, const double* QR, double* Fx, double g) {
  double dl = QL[0], uul = QL[1], vvl = QL[2], pl = QL[3];
  double dr = QR[0], uur = QR[1], vvr = QR[2], pr = QR[3];
  double al = sqrt(g * pl / dl), ar = sqrt(g * pr / dr);
  double tmp1 = 2.0 * (g - 1.0);
  double tmp3 = (3.0 - g) / (2.0 * (g - 1.0));
  double E1P = (uul + fabs(uul)) * 0.5, E2P = (uul - al + fabs(uul - al)) * 0.5, E3P = (uul + al + fabs(uul + al)) * 0.5;
  double tmp0 = dl / (2.0 * g);
  double fp0 = tmp0 * (tmp1 * E1P + E2P + E3P);
  double fp1 = tmp0 * (tmp1 * E1P * uul + E2P * (uul - al) + E3P * (uul + al));
  double fp2 = tmp0 * (tmp1 * E1P * vvl + E2P * vvl + E3P * vvl);
  double fp3 = tmp0 * (E1P * (g - 1.0) * (uul * uul + vvl * vvl) + E2P * ((uul - al) * (uul - al) + vvl * vvl) * 0.5 +
                       E3P * ((uul + al) * (uul + al) + vvl * vvl) * 0.5 + tmp3 * al * al * (E2P + E3P));
  double E1M = (uur - fabs(uur)) * 0.5, E2M = (uur - ar - fabs(uur - ar)) * 0.5, E3M = (uur + ar - fabs(uur + ar)) * 0.5;
  tmp0 = dr / (2.0 * g);
  double fm0 = tmp0 * (tmp1 * E1M + E2M + E3M);
  double fm1 = tmp0 * (tmp1 * E1M * uur + E2M * (uur - ar) + E3M * (uur + ar));
  double fm2 = tmp0 * (tmp1 * E1M * vvr + E2M * vvr + E3M * vvr);
  double fm3 = tmp0 * (E1M * (g - 1.0) * (uur * uur + vvr * vvr) + E2M * ((uur - ar) * (uur - ar) + vvr * vvr) * 0.5 +
                       E3M * ((uur + ar) * (uur + ar) + vvr * vvr) * 0.5 + tmp3 * ar * ar * (E2M + E3M));
  Fx[0] = fp0 + fm0; Fx[1] = fp1 + fm1; Fx[2] = fp2 + fm2; Fx[3] = fp3 + fm3;
}

__device__ __forceinline__ void flux_hll_hllc(const double* QL, const double* QR, double* Fx, double g, int hllc) {
  double dl = QL[0], uul = QL[1], vvl = QL[2], pl = QL[3];
  double dr = QR[0], uur = QR[1], vvr = QR[2], pr = QR[3];
  double UL[4], UR[4], Fl[4], Fr[4];
  UL[0] = dl; UL[1] = dl * uul; UL[2] = dl * vvl; UL[3] = pl / (g - 1.0) + dl * (uul * uul + vvl * vvl) * 0.5;
  UR[0] = dr; UR[1] = dr * uur; UR[2] = dr * vvr; UR[3] = pr / (g - 1.0) + dr * (uur * uur + vvr * vvr) * 0.5;
  double al = sqrt(g * pl / dl), ar = sqrt(g * pr / dr);
  double p_pvrs = 0.5 * (pl + pr) - (uur - uul) * (dl + dr) * (al + ar) * 0.125;
  double p_star = fmax(0.0, p_pvrs);
  double qql = (p_star <= pl) ? 1.0 : sqrt(1.0 + ((g + 1.0) / (2.0 * g)) * (p_star / pl - 1.0));
  double qqr = (p_star <= pr) ? 1.0 : sqrt(1.0 + ((g + 1.0) / (2.0 * g)) * (p_star / pr - 1.0));
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
  if (Sl >= 0) { for (int m = 0; m < 4; m++) Fx[m] = Fl[m]; return; }
  if (Sr <= 0) { for (int m = 0; m < 4; m++) Fx[m] = Fr[m]; return; }
  if (S_star >= 0) {
    double tmpl = dl * (Sl - uul) / (Sl - S_star);
    double tmp = UL[3] / dl + (S_star - uul) * (S_star + pl / (dl * (Sl - uul)));
    Fx[0] = Fl[0] + Sl * (tmpl - UL[0]);
    Fx[1] = Fl[1] + Sl * (tmpl * S_star - UL[1]);
    Fx[2] = Fl[2] + Sl * (tmpl * vvl - UL[2]);
    Fx[3] = Fl[3] + Sl * (tmpl * tmp - UL[3]);
  } else {
    double tmpr = dr * (Sr - uur) / (Sr - S_star);
    double tmp = UR[3] / dr + (S_star - uur) * (S_star + pr / (dr * (Sr - uur)));
    Fx[0] = Fr[0] + Sr * (tmpr - UR[0]);
    Fx[1] = Fr[1] + Sr * (tmpr * S_star - UR[1]);
    Fx[2] = Fr[2] + Sr * (tmpr * vvr - UR[2]);
    Fx[3] = Fr[3] + Sr * (tmpr * tmp - UR[3]);
  }
}

__device__ __forceinline__ void face_flux(int iflux, const double* QL, const double* QR, double* F0, double g) {
  switch (iflux) {
    case 4: flux_roe(QL, QR, F0, g); break;
    case 6: flux_ausm(QL, QR, F0, g); break;
    case 5: flux_vanleer(QL, QR, F0, g); break;
    case 1: flux_sw(QL, QR, F0, g); break;
    case 2: flux_hll_hllc(QL, QR, F0, g, 0); break;
    case 3: flux_hll_hllc(QL, QR, F0, g, 1); break;
    default: F0[0] = F0[1] = F0[2] = F0[3] = 0.0;
  }
}

// comput_DFn (sub_LU_SGS.f90:111-138), F(U+dU).n - F(U).n of the Euler normal
// flux, split in two: flux_n_old gives F(U).n, which depends on the old state
// only, and dfn_from_old the difference for U2 = U + dU.  Both halves evaluate
// the reference's expressions unchanged, so the difference is bit-identical
// to the undivided routine.
template <int NV>
__device__ __forceinline__ void flux_n_old(double* F, const double* U, double n1, double n2, double g) {
  double un1 = (U[1] * n1 + U[2] * n2) / U[0];
  double p1 = (g - 1.0) * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0]);
  F[0] = U[0] * un1;
  F[1] = U[1] * un1 + p1 * n1;
  F[2] = U[2] * un1 + p1 * n2;
  F[3] = (U[3] + p1) * un1;
#pragma unroll
  for (int m = 4; m < NV; m++) F[m] = U[m] * un1;
}

template <int NV>
__device__ __forceinline__ void dfn_from_old(double* DF, const double* U2, const double* F, double n1, double n2,
                                             double g) {
  double un2 = (U2[1] * n1 + U2[2] * n2) / U2[0];
  double p2 = (g - 1.0) * (U2[3] - 0.5 * (U2[1] * U2[1] + U2[2] * U2[2]) / U2[0]);
  DF[0] = U2[0] * un2 - F[0];
  DF[1] = (U2[1] * un2 + p2 * n1) - F[1];
  DF[2] = (U2[2] * un2 + p2 * n2) - F[2];
  DF[3] = (U2[3] + p2) * un2 - F[3];
#pragma unroll
  for (int m = 4; m < NV; m++) DF[m] = U2[m] * un2 - F[m];
}

}  // namespace ec2d
