/*
 * o_coeffs.c -- oracle copy of the per-node-type stencil tables and the
 * scalar reconstruction operators.  TEST INFRASTRUCTURE ONLY.
 *
 * Tables: the values of src/prometeo_metric_coeffs_macros.h:59-552 assembled
 * as in src/prometeo_metric_coeffs.cc:42-209 (node-type order of
 * src/prometeo_metric_coeffs.h:55-67).  Pinned by the exact-equality checks of
 * unitTests/metricTest/operatorsTest_{Periodic,Staggered,Collocated}.rg
 * (transcribed to tests/golden/operators.json).
 *
 * Operators: TENO_Op<-10>, TENOA_Op<-8>, TENOLAD_Op<-16>
 * (src/prometeo_metric.inl:431-1038), TENOsensor::TENOA (:352-403) and
 * KennedyReconstruct (:215-292).  Pinned by closed forms in
 * tests/test_oracle_stencil.py (all b_k = 1 reduces to the 6th-order central
 * interpolation [1,-8,37,37,-8,1]/60; polynomial reproduction; constant data
 * keeps every stencil).  Cut-off decisions on rough data: parity unpinned by
 * any reference vector (none exists); checked by brute-force restatement of
 * the threshold rule on hand-made stencils.
 */
#include <math.h>
#include "htr_oracle.h"

enum { Std = 0, L_S, Lp1_S, Lp2_S, Rm3_S, Rm2_S, Rm1_S, R_S, L_C, Lp1_C, Rm2_C, Rm1_C, R_C };

static const int CP[13][5] = {
   [Std]   = { -2, -1, 1, 2, 3 },
   [L_S]   = {  0,  0, 1, 1, 1 },
   [Lp1_S] = { -1, -1, 1, 2, 3 },
   [Lp2_S] = { -2, -1, 1, 2, 3 },
   [Rm3_S] = { -2, -1, 1, 2, 3 },
   [Rm2_S] = { -2, -1, 1, 2, 2 },
   [Rm1_S] = { -1, -1, 1, 1, 1 },
   [R_S]   = { -1, -1, 0, 0, 0 },
   [L_C]   = {  0,  0, 1, 2, 2 },
   [Lp1_C] = { -1, -1, 1, 2, 2 },
   [Rm2_C] = { -1, -1, 1, 2, 2 },
   [Rm1_C] = { -1, -1, 1, 1, 1 },
   [R_C]   = { -1, -1, 0, 0, 0 },
};

/* rows of 6 weights over [c-2 .. c+3] for stencils 1..4 */
#define ROW_STD1P   0.0, -1.0/6.0, 5.0/6.0, 2.0/6.0, 0.0, 0.0
#define ROW_STD2P   0.0, 0.0, 2.0/6.0, 5.0/6.0, -1.0/6.0, 0.0
#define ROW_STD3P   2.0/6.0, -7.0/6.0, 11.0/6.0, 0.0, 0.0, 0.0
#define ROW_STD4P   0.0, 0.0, 3.0/12.0, 13.0/12.0, -5.0/12.0, 1.0/12.0
#define ROW_STD3M   0.0, 0.0, 0.0, 11.0/6.0, -7.0/6.0, 2.0/6.0
#define ROW_STD4M   1.0/12.0, -5.0/12.0, 13.0/12.0, 3.0/12.0, 0.0, 0.0
#define ROW_ZERO    0.0, 0.0, 0.0, 0.0, 0.0, 0.0
#define ROW_E2      0.0, 0.0, 1.0, 0.0, 0.0, 0.0
#define ROW_E3      0.0, 0.0, 0.0, 1.0, 0.0, 0.0

static const double RECON_P[13][24] = {
   [Std]   = { ROW_STD1P, ROW_STD2P, ROW_STD3P, ROW_STD4P },
   [L_S]   = { ROW_E2, ROW_E2, ROW_E2, ROW_E2 },
   [Lp1_S] = { 0.0, -2.0/4.0, 5.0/4.0, 1.0/4.0, 0.0, 0.0, ROW_STD2P, ROW_ZERO, ROW_STD4P },
   [Lp2_S] = { ROW_STD1P, ROW_STD2P, 1.0, -2.0, 2.0, 0.0, 0.0, 0.0, ROW_STD4P },
   [Rm3_S] = { ROW_STD1P, ROW_STD2P, ROW_STD3P, 0.0, 0.0, 2.0/9.0, 11.0/9.0, -7.0/9.0, 3.0/9.0 },
   [Rm2_S] = { ROW_STD1P, 0.0, 0.0, 1.0/4.0, 5.0/4.0, -2.0/4.0, 0.0, ROW_STD3P, ROW_ZERO },
   [Rm1_S] = { ROW_E3, ROW_E3, ROW_E3, ROW_E3 },
   [R_S]   = { ROW_ZERO, ROW_ZERO, ROW_ZERO, ROW_ZERO },
   [L_C]   = { ROW_E2, ROW_ZERO, ROW_ZERO, ROW_ZERO },
   [Lp1_C] = { ROW_STD1P, ROW_STD2P, ROW_ZERO, ROW_ZERO },
   [Rm2_C] = { ROW_STD1P, ROW_STD2P, ROW_ZERO, ROW_ZERO },
   [Rm1_C] = { ROW_E2, ROW_ZERO, ROW_ZERO, ROW_ZERO },
   [R_C]   = { ROW_ZERO, ROW_ZERO, ROW_ZERO, ROW_ZERO },
};

static const double RECON_M[13][24] = {
   [Std]   = { ROW_STD2P, ROW_STD1P, ROW_STD3M, ROW_STD4M },
   [L_S]   = { ROW_E2, ROW_E2, ROW_E2, ROW_E2 },
   [Lp1_S] = { ROW_STD2P, 0.0, -2.0/4.0, 5.0/4.0, 1.0/4.0, 0.0, 0.0, ROW_STD3M, ROW_ZERO },
   [Lp2_S] = { ROW_STD2P, ROW_STD1P, ROW_STD3M, 3.0/9.0, -7.0/9.0, 11.0/9.0, 2.0/9.0, 0.0, 0.0 },
   [Rm3_S] = { ROW_STD2P, ROW_STD1P, 0.0, 0.0, 0.0, 2.0, -2.0, 1.0, ROW_STD4M },
   [Rm2_S] = { 0.0, 0.0, 1.0/4.0, 5.0/4.0, -2.0/4.0, 0.0, ROW_STD1P, ROW_ZERO, ROW_STD4M },
   [Rm1_S] = { ROW_E3, ROW_E3, ROW_E3, ROW_E3 },
   [R_S]   = { ROW_ZERO, ROW_ZERO, ROW_ZERO, ROW_ZERO },
   [L_C]   = { ROW_E3, ROW_ZERO, ROW_ZERO, ROW_ZERO },
   [Lp1_C] = { ROW_STD2P, ROW_STD1P, ROW_ZERO, ROW_ZERO },
   [Rm2_C] = { ROW_STD2P, ROW_STD1P, ROW_ZERO, ROW_ZERO },
   [Rm1_C] = { ROW_E3, ROW_ZERO, ROW_ZERO, ROW_ZERO },
   [R_C]   = { ROW_ZERO, ROW_ZERO, ROW_ZERO, ROW_ZERO },
};

#define CO_STD { 9.0/20.0, 6.0/20.0, 1.0/20.0, 4.0/20.0 }
static const double COEFF_P[13][4] = {
   [Std] = CO_STD, [L_S] = CO_STD,
   [Lp1_S] = { 2.0/4.0, 1.0/4.0, 0.0, 1.0/4.0 },
   [Lp2_S] = { 47.0/100.0, 27.0/100.0, 10.0/100.0, 16.0/100.0 },
   [Rm3_S] = { 39.0/100.0, 27.0/100.0, 4.0/100.0, 30.0/100.0 },
   [Rm2_S] = { 7.0/16.0, 8.0/16.0, 1.0/16.0, 0.0 },
   [Rm1_S] = CO_STD,
   [R_S]   = { 0.0, 0.0, 0.0, 0.0 },
   [L_C]   = { 1.0, 0.0, 0.0, 0.0 },
   [Lp1_C] = { 1.0/2.0, 1.0/2.0, 0.0, 0.0 },
   [Rm2_C] = { 1.0/2.0, 1.0/2.0, 0.0, 0.0 },
   [Rm1_C] = { 1.0, 0.0, 0.0, 0.0 },
   [R_C]   = { 0.0, 0.0, 0.0, 0.0 },
};
static const double COEFF_M[13][4] = {
   [Std] = CO_STD, [L_S] = CO_STD,
   [Lp1_S] = { 7.0/16.0, 8.0/16.0, 1.0/16.0, 0.0 },
   [Lp2_S] = { 39.0/100.0, 27.0/100.0, 4.0/100.0, 30.0/100.0 },
   [Rm3_S] = { 47.0/100.0, 27.0/100.0, 10.0/100.0, 16.0/100.0 },
   [Rm2_S] = { 2.0/4.0, 1.0/4.0, 0.0, 1.0/4.0 },
   [Rm1_S] = CO_STD,
   [R_S]   = { 0.0, 0.0, 0.0, 0.0 },
   [L_C]   = { 1.0, 0.0, 0.0, 0.0 },
   [Lp1_C] = { 1.0/2.0, 1.0/2.0, 0.0, 0.0 },
   [Rm2_C] = { 1.0/2.0, 1.0/2.0, 0.0, 0.0 },
   [Rm1_C] = { 1.0, 0.0, 0.0, 0.0 },
   [R_C]   = { 0.0, 0.0, 0.0, 0.0 },
};

static const float INTERP[13][2] = {
   {0.5f,0.5f},{1.0f,0.0f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{0.0f,1.0f},
   {0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f} };
static const float GRAD[13][2] = {
   {0.5f,0.5f},{0.0f,2.0f},{1.0f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,1.0f},
   {2.0f,0.0f},{0.0f,1.0f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{1.0f,0.0f} };

static const int KORDER[13] = { 3, 0, 1, 2, 2, 1, 0, 0, 1, 2, 2, 1, 0 };
#define KC6 { 3.0/4.0, -3.0/20.0, 1.0/60.0 }
#define KC4 { 2.0/3.0, -1.0/12.0, 0.0 }
#define KC2 { 0.5, 0.0, 0.0 }
#define KC0 { 0.0, 0.0, 0.0 }
static const double KCOEFF[13][3] = { KC6, KC0, KC2, KC4, KC4, KC2, KC0, KC0, KC2, KC4, KC4, KC2, KC0 };
static const int KNSUM[13] = { 3, 0, 3, 3, 2, 1, 0, 0, 3, 3, 2, 1, 0 };

static const float P10[14] = { 1, 1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9,
                               1e-10, 1e-11, 1e-12, 1e-13 };

int    o_coeff_Cp(int t, int k)           { return CP[t][k]; }
double o_coeff_ReconPlus(int t, int k)    { return RECON_P[t][k]; }
double o_coeff_ReconMinus(int t, int k)   { return RECON_M[t][k]; }
double o_coeff_CoeffsPlus(int t, int k)   { return COEFF_P[t][k]; }
double o_coeff_CoeffsMinus(int t, int k)  { return COEFF_M[t][k]; }
float  o_coeff_Interp(int t, int k)       { return INTERP[t][k]; }
float  o_coeff_Grad(int t, int k)         { return GRAD[t][k]; }
int    o_coeff_KennedyOrder(int t)        { return KORDER[t]; }
double o_coeff_KennedyCoeff(int t, int k) { return KCOEFF[t][k]; }
int    o_coeff_KennedyNSum(int t)         { return KNSUM[t]; }
float  o_coeff_p10(int k)                 { return P10[k]; }

/* ------------------------------------------------------------------------- */
/* constexpr my_10_pow(n) of src/prometeo_metric.inl:60-64: a chain of 0.1*  */
static double ten_pow_chain(int n)
{
   double r = 1.0;
   if (n < 0) for (int k = 0; k < -n; k++) r = 1.0e-1*r;
   else       for (int k = 0; k <  n; k++) r = 1.0e1*r;
   return r;
}

static inline double dmin(double a, double b) { return (b < a) ? b : a; }
static inline double dmax(double a, double b) { return (a < b) ? b : a; }
static inline int imin(int a, int b) { return (b < a) ? b : a; }
static inline int imax(int a, int b) { return (a < b) ? b : a; }

static const double C13 = 13.0/12.0, C23 = 3.0/12.0, C14 = 1.0/36.0, C24 = 13.0/12.0, C34 = 781.0/720.0;

/* smoothness indicators of the plus-side stencils (metric.inl:441-448) */
static void smooth_plus(const double *v, double *s)
{
   const double ym2 = v[0], ym1 = v[1], y = v[2], yp1 = v[3], yp2 = v[4], yp3 = v[5];
   double a1, a2, a3;
   a1 = (ym1 - 2*y   + yp1); a2 = (        ym1 - yp1); s[0] = C13*a1*a1 + C23*a2*a2;
   a1 = (y   - 2*yp1 + yp2); a2 = (3*y - 4*yp1 + yp2); s[1] = C13*a1*a1 + C23*a2*a2;
   a1 = (ym2 - 2*ym1 + y  ); a2 = (3*y - 4*ym1 + ym2); s[2] = C13*a1*a1 + C23*a2*a2;
   a1 = (-11*y + 18*yp1 - 9*yp2 + 2*yp3);
   a2 = (  2*y -  5*yp1 + 4*yp2 -   yp3);
   a3 = (-   y +  3*yp1 - 3*yp2 +   yp3);
   s[3] = C14*a1*a1 + C24*a2*a2 + C34*a3*a3;
}

/* mirrored indicators of the minus side (metric.inl:520-527) */
static void smooth_minus(const double *v, double *s)
{
   const double ym2 = v[0], ym1 = v[1], y = v[2], yp1 = v[3], yp2 = v[4], yp3 = v[5];
   double a1, a2, a3;
   a1 = (yp2 - 2*yp1 + y  ); a2 = (          yp2 - y  ); s[0] = C13*a1*a1 + C23*a2*a2;
   a1 = (yp1 - 2*y   + ym1); a2 = (3*yp1 - 4*y   + ym1); s[1] = C13*a1*a1 + C23*a2*a2;
   a1 = (yp3 - 2*yp2 + yp1); a2 = (3*yp1 - 4*yp2 + yp3); s[2] = C13*a1*a1 + C23*a2*a2;
   a1 = (-11*yp1 + 18*y - 9*ym1 + 2*ym2);
   a2 = (  2*yp1 -  5*y + 4*ym1 -   ym2);
   a3 = (-   yp1 +  3*y - 3*ym1 +   ym2);
   s[3] = C14*a1*a1 + C24*a2*a2 + C34*a3*a3;
}

double o_teno_reconstruct(int op, int side, const double *v, int nType, int *bits)
{
   const double *Coeffs = side ? COEFF_M[nType] : COEFF_P[nType];
   const double *Recon  = side ? RECON_M[nType] : RECON_P[nType];
   double s[4];
   if (side) smooth_minus(v, s); else smooth_plus(v, s);
   const double eps = (op == 0) ? ten_pow_chain(-10) : (op == 1) ? ten_pow_chain(-8) : ten_pow_chain(-16);
   const double tau6 = fabs(s[3] - (s[2] + s[1] + 4*s[0])/6);
   double chi[4], a[4];
   for (int k = 0; k < 4; k++) { chi[k] = tau6/(s[k] + eps); a[k] = pow(1 + chi[k], 6); }

   double cut;
   if (op == 0) {
      for (int k = 0; k < 4; k++) if (Coeffs[k] < 1e-10) a[k] = 0.0;
      cut = 1e-6;
   } else if (op == 1) {
      if (nType == Std) {
         /* Ren sensor adaptation (metric.inl:635-653) */
         const double Ren_r_c = 0.2;
         const double Ren_eps = 0.9*Ren_r_c*1e-16/(1.0 - 0.9*Ren_r_c);
         const double Ren_irc = 1.0/Ren_r_c;
         const double Smooth_pow = 9.5, Shock_pow = 6.0, Diff_pow = Smooth_pow - Shock_pow;
         const double ym2 = v[0], ym1 = v[1], y = v[2], yp1 = v[3], yp2 = v[4], yp3 = v[5];
         const double var2 = fabs(ym1 - ym2);
         const double var3 = fabs(y   - ym1);
         const double var4 = fabs(yp1 - y  );
         const double var5 = fabs(yp2 - yp1);
         const double var6 = fabs(yp3 - yp2);
         double eta = dmin((2*var2*var3 + Ren_eps)/(var2*var2 + var3*var3 + Ren_eps),
                           (2*var3*var4 + Ren_eps)/(var3*var3 + var4*var4 + Ren_eps));
         eta = dmin(eta,   (2*var4*var5 + Ren_eps)/(var4*var4 + var5*var5 + Ren_eps));
         eta = dmin(eta,   (2*var5*var6 + Ren_eps)/(var5*var5 + var6*var6 + Ren_eps));
         const double delta = 1 - dmin(eta*Ren_irc, 1.0);
         const double decay = pow((1 - delta), 4)*(1 + 4*delta);
         const int power = imin(imax((int)floor(Smooth_pow - Diff_pow*(1 - decay)), 0), 13);
         cut = (double)P10[power];
      } else {
         for (int k = 0; k < 4; k++) if (Coeffs[k] < 1e-10) a[k] = 0.0;
         cut = (double)P10[4];
      }
   } else {
      for (int k = 0; k < 4; k++) if (Coeffs[k] < 1e-10) a[k] = 0.0;
      const double Smooth_pow = 12.5, Shock_pow = 3.0, iH = 1e-1, Diff_pow = Smooth_pow - Shock_pow;
      const double chiMax = dmax(chi[0], dmax(chi[1], dmax(chi[2], chi[3])));
      const double theta = 1.0/(1 + chiMax*iH);
      const int power = imin(imax((int)floor(Shock_pow + Diff_pow*theta), 0), 13);
      cut = (double)P10[power];
   }

   double an = 1.0/(a[0] + a[1] + a[2] + a[3]);
   int b[4];
   for (int k = 0; k < 4; k++) b[k] = (a[k]*an < cut) ? 0 : 1;
   if (bits) for (int k = 0; k < 4; k++) bits[k] = b[k];

   const double base = side ? v[3] : v[2];
   double var[4];
   for (int k = 0; k < 4; k++) {
      const double *R = &Recon[6*k];
      var[k] = v[0]*R[0] + v[1]*R[1] + v[2]*R[2] + v[3]*R[3] + v[4]*R[4] + v[5]*R[5] - base;
   }
   double w[4];
   for (int k = 0; k < 4; k++) w[k] = Coeffs[k]*b[k];
   an = 1.0/(w[0] + w[1] + w[2] + w[3]);
   for (int k = 0; k < 4; k++) w[k] = w[k]*an;
   return base + w[0]*var[0] + w[1]*var[1] + w[2]*var[2] + w[3]*var[3];
}

/* TENOsensor::TENOA (metric.inl:352-403): true = smooth */
int o_tenoa_sensor(const double *v, int nType, double Phi)
{
   const double Ren_r_c = 0.2;
   const double Ren_eps = 0.9*Ren_r_c*1e-6/(1.0 - 0.9*Ren_r_c);
   const double Ren_irc = 1.0/Ren_r_c;
   const double Smooth_pow = 12.5, Shock_pow = 1.0, Diff_pow = Smooth_pow - Shock_pow;
   const double eps = 1e-8;
   if (nType == L_S || nType == Rm1_S) return 1;
   if (nType == L_C || nType == Rm1_C) {
      const double var1 = (nType == L_C) ? v[3] - v[2] : v[2] - v[1];
      const double var2 = (nType == L_C) ? v[4] - v[3] : v[3] - v[2];
      const double eta = (fabs(2.0*var1*var2) + Ren_eps)/(var1*var1 + var2*var2 + Ren_eps);
      return (1.0 - dmin(1.0, Ren_irc*eta)) < 0.5;
   }
   double s[4];
   smooth_plus(v, s);
   const double tau6 = fabs(s[3] - (s[2] + s[1] + 4*s[0])/6);
   double a[4];
   for (int k = 0; k < 4; k++) a[k] = pow(1 + tau6/(s[k] + eps), 6);
   const double an = 1.0/(a[0] + a[1] + a[2] + a[3]);
   const double decay = pow((1 - Phi), 12)*(1 + 12*Phi);
   const int power = imin(imax((int)floor(Smooth_pow - Diff_pow*(1 - decay)), 0), 13);
   const double cut = (double)P10[power];
   return (a[0]*an > cut) && (a[1]*an > cut) && (a[2]*an > cut) && (a[3]*an > cut);
}

/* KennedyReconstruct(rho, u, phi) (metric.inl:215-240) */
double o_kennedy_reconstruct3(const double *rho, const double *u, const double *phi, int nType)
{
   if (nType == L_S)   return rho[2]*u[2]*phi[2];
   if (nType == Rm1_S) return rho[3]*u[3]*phi[3];
   double flux = 0.0;
   for (int l = 0; l < KORDER[nType]; l++) {
      const int lp = l + 1;
      double acc = 0.0;
      for (int m = 0; m < lp; m++)
         acc += ((rho[2-m] + rho[2-m+lp])*(u[2-m] + u[2-m+lp])*(phi[2-m] + phi[2-m+lp]));
      flux += KCOEFF[nType][l]*acc;
   }
   return flux*0.25;
}
