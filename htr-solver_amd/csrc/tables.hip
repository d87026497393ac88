// tables.hip -- per-node-type stencil tables (values of
// src/prometeo_metric_coeffs_macros.h:59-552, node order of
// src/prometeo_metric_coeffs.h:55-67) and their upload to __constant__.
#include "common.hpp"
#include "kernels.hpp"

namespace htr {

__constant__ StencilTables c_st;

static StencilTables build_tables()
{
   StencilTables t = {};
   const int cp[13][5] = {
      { -2, -1, 1, 2, 3 }, {  0,  0, 1, 1, 1 }, { -1, -1, 1, 2, 3 }, { -2, -1, 1, 2, 3 },
      { -2, -1, 1, 2, 3 }, { -2, -1, 1, 2, 2 }, { -1, -1, 1, 1, 1 }, { -1, -1, 0, 0, 0 },
      {  0,  0, 1, 2, 2 }, { -1, -1, 1, 2, 2 }, { -1, -1, 1, 2, 2 }, { -1, -1, 1, 1, 1 },
      { -1, -1, 0, 0, 0 } };
   for (int n = 0; n < 13; n++) for (int k = 0; k < 5; k++) t.Cp[n][k] = cp[n][k];

   // the four candidate stencils over [c-2 .. c+3]
   using rows::s1p; using rows::s2p; using rows::s3p; using rows::s4p; using rows::s3m; using rows::s4m;
   const double zr[6]  = { 0, 0, 0, 0, 0, 0 };
   const double e2[6]  = { 0, 0, 1.0, 0, 0, 0 };
   const double e3[6]  = { 0, 0, 0, 1.0, 0, 0 };
   const double lp1a[6] = { 0.0, -2.0/4.0, 5.0/4.0, 1.0/4.0, 0.0, 0.0 };
   const double lp2c[6] = { 1.0, -2.0, 2.0, 0.0, 0.0, 0.0 };
   const double rm3d[6] = { 0.0, 0.0, 2.0/9.0, 11.0/9.0, -7.0/9.0, 3.0/9.0 };
   const double rm2b[6] = { 0.0, 0.0, 1.0/4.0, 5.0/4.0, -2.0/4.0, 0.0 };
   const double lp2d[6] = { 3.0/9.0, -7.0/9.0, 11.0/9.0, 2.0/9.0, 0.0, 0.0 };
   const double rm3c[6] = { 0.0, 0.0, 0.0, 2.0, -2.0, 1.0 };
   const double *P[13][4] = {
      { s1p, s2p, s3p, s4p }, { e2, e2, e2, e2 }, { lp1a, s2p, zr, s4p }, { s1p, s2p, lp2c, s4p },
      { s1p, s2p, s3p, rm3d }, { s1p, rm2b, s3p, zr }, { e3, e3, e3, e3 }, { zr, zr, zr, zr },
      { e2, zr, zr, zr }, { s1p, s2p, zr, zr }, { s1p, s2p, zr, zr }, { e2, zr, zr, zr }, { zr, zr, zr, zr } };
   const double *M[13][4] = {
      { s2p, s1p, s3m, s4m }, { e2, e2, e2, e2 }, { s2p, lp1a, s3m, zr }, { s2p, s1p, s3m, lp2d },
      { s2p, s1p, rm3c, s4m }, { rm2b, s1p, zr, s4m }, { e3, e3, e3, e3 }, { zr, zr, zr, zr },
      { e3, zr, zr, zr }, { s2p, s1p, zr, zr }, { s2p, s1p, zr, zr }, { e3, zr, zr, zr }, { zr, zr, zr, zr } };
   for (int n = 0; n < 13; n++)
      for (int s = 0; s < 4; s++)
         for (int k = 0; k < 6; k++) { t.ReconP[n][6*s+k] = P[n][s][k]; t.ReconM[n][6*s+k] = M[n][s][k]; }

   using rows::std4;
   const double c_lp1p[4] = { 2.0/4.0, 1.0/4.0, 0.0, 1.0/4.0 };
   const double c_lp1m[4] = { 7.0/16.0, 8.0/16.0, 1.0/16.0, 0.0 };
   const double c_lp2p[4] = { 47.0/100.0, 27.0/100.0, 10.0/100.0, 16.0/100.0 };
   const double c_lp2m[4] = { 39.0/100.0, 27.0/100.0, 4.0/100.0, 30.0/100.0 };
   const double c0[4] = { 0, 0, 0, 0 };
   const double c1[4] = { 1.0, 0, 0, 0 };
   const double ch[4] = { 1.0/2.0, 1.0/2.0, 0, 0 };
   const double *CP_[13] = { std4, std4, c_lp1p, c_lp2p, c_lp2m, c_lp1m, std4, c0, c1, ch, ch, c1, c0 };
   const double *CM_[13] = { std4, std4, c_lp1m, c_lp2m, c_lp2p, c_lp1p, std4, c0, c1, ch, ch, c1, c0 };
   for (int n = 0; n < 13; n++) for (int k = 0; k < 4; k++) { t.CoeffP[n][k] = CP_[n][k]; t.CoeffM[n][k] = CM_[n][k]; }

   const float ip[13][2] = { {0.5f,0.5f},{1.0f,0.0f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{0.0f,1.0f},
                             {0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f} };
   const float gr[13][2] = { {0.5f,0.5f},{0.0f,2.0f},{1.0f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,1.0f},
                             {2.0f,0.0f},{0.0f,1.0f},{0.5f,0.5f},{0.5f,0.5f},{0.5f,0.5f},{1.0f,0.0f} };
   for (int n = 0; n < 13; n++) for (int k = 0; k < 2; k++) { t.Interp[n][k] = ip[n][k]; t.Grad[n][k] = gr[n][k]; }

   const int korder[13] = { 3, 0, 1, 2, 2, 1, 0, 0, 1, 2, 2, 1, 0 };
   const int knsum[13]  = { 3, 0, 3, 3, 2, 1, 0, 0, 3, 3, 2, 1, 0 };
   const double k6[3] = { 3.0/4.0, -3.0/20.0, 1.0/60.0 }, k4[3] = { 2.0/3.0, -1.0/12.0, 0.0 };
   const double k2[3] = { 0.5, 0.0, 0.0 }, k0[3] = { 0.0, 0.0, 0.0 };
   const double *kc[13] = { k6, k0, k2, k4, k4, k2, k0, k0, k2, k4, k4, k2, k0 };
   for (int n = 0; n < 13; n++) {
      t.KOrder[n] = korder[n]; t.KNSum[n] = knsum[n];
      for (int k = 0; k < 3; k++) t.KCoeff[n][k] = kc[n][k];
   }
   // float images of the double literals, as the reference's float table
   const double p10d[14] = { 1, 1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9,
                             1e-10, 1e-11, 1e-12, 1e-13 };
   float p10[14];
   for (int k = 0; k < 14; k++) p10[k] = (float)p10d[k];
   for (int k = 0; k < 14; k++) t.p10[k] = p10[k];
   return t;
}

const StencilTables &host_stencil_tables()
{
   static const StencilTables t = build_tables();
   return t;
}

void upload_stencil_tables()
{
   const StencilTables &t = host_stencil_tables();
   (void)hipMemcpyToSymbol(HIP_SYMBOL(c_st), &t, sizeof(t));
}

} // namespace htr
