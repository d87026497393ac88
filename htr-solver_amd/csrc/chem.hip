// chem.hip -- implicit chemistry update (UpdateChemistryTask): per-cell
// 4-stage Kaps-Rentrop Rosenbrock integration of dx/dt = [w(x) + Ct_old].
//
// Reference: src/prometeo_chem.hpp:53-93 (ImplicitSolver::rhs),
// src/prometeo_chem.cc:36-71 (task loop), src/Utils/math_utils.hpp:87-438
// (LUdec, Rosenbrock: FD Jacobian, step, itercheck, stepcheck).
//
// Two linear-solve precisions:
//   HTR_CHEM_FP64     Crout LU with implicit scaling and partial pivoting in
//                     FP64, as the reference;
//   HTR_CHEM_FP32_IR  the matrix equilibrated by powers of two and factorised
//                     in FP32; each of the four stage solves is recovered to
//                     FP64 accuracy by residual-based iterative refinement
//                     against the FP64 matrix (r = b - A x in FP64,
//                     x += LU32 \ r) until the correction is below 1e-15 of
//                     the solution; if a sweep fails to halve the correction
//                     the solve falls back to an FP64 LU of the same matrix
//                     (made once per factorisation), so the result always has
//                     FP64 accuracy.
//
// One thread integrates one cell (cells of a wave march independently; the
// step count varies per cell).  The cell loop is a grid-stride loop over a
// persistent grid sized to the device, so long-running cells do not hold a
// whole launch back more than one cell's worth.
#include "common.hpp"
#include "kernels.hpp"

namespace htr {

template<class MX, int PREC>
struct Rosen {
   static constexpr int NS = MX::NS, N = MX::NEQ;
   const DMix &m;
   const double *Ct;
   double T;

   // ImplicitSolver::rhs
   __device__ void rhs(double *r, const double *x)
   {
      double rhoYi[NS], Yi[NS], w[NS];
#pragma unroll
      for (int i = 0; i < NS; i++) rhoYi[i] = x[i];
      double rho = 0.0;
#pragma unroll
      for (int i = 0; i < NS; i++) rho += rhoYi[i];
      const double irho = 1.0/rho;
#pragma unroll
      for (int i = 0; i < NS; i++) { Yi[i] = irho*rhoYi[i]; Yi[i] = dmax(Yi[i], 1e-60); Yi[i] = dmin(Yi[i], 1.0); }
      const double MixW = MX::MolarWeightFromY(m, Yi);
      const double rhoInv = 1.0/rho;
      double v[3];
#pragma unroll
      for (int i = 0; i < 3; i++) v[i] = x[NS+i]*rhoInv;
      double k2 = 0.0; k2 += v[0]*v[0]; k2 += v[1]*v[1]; k2 += v[2]*v[2];
      const double e = x[NS+3]*rhoInv - 0.5*k2;
      T = MX::TFromInternalEnergy(m, e, T, Yi);
      const double P = MX::PFromRhoAndT(m, rho, MixW, T);
      MX::ProductionRates(m, w, rho, P, T, Yi);
#pragma unroll
      for (int i = 0; i < NS; i++) r[i] = w[i] + Ct[i];
#pragma unroll
      for (int i = NS; i < N; i++) r[i] = Ct[i];
   }

   // ---- LU, column-major A(i,j) = A[i + j*N] ----------------------------------
   template<typename R>
   __device__ static void ludcmp(R *A, int *ind)
   {
      int imax = 0;
      R vv[N];
      for (int i = 0; i < N; i++) {
         R big = 0;
         for (int j = 0; j < N; j++) { const R a = fabs(A[i + j*N]); big = (big < a) ? a : big; }
         vv[i] = (R)1/big;
      }
      for (int j = 0; j < N; j++) {
         for (int i = 0; i < j; i++) {
            R sum = A[i + j*N];
            for (int k = 0; k < i; k++) sum -= A[i + k*N]*A[k + j*N];
            A[i + j*N] = sum;
         }
         R big = 0;
         for (int i = j; i < N; i++) {
            R sum = A[i + j*N];
            for (int k = 0; k < j; k++) sum -= A[i + k*N]*A[k + j*N];
            A[i + j*N] = sum;
            const R dum = vv[i]*fabs(sum);
            if (dum >= big) { big = dum; imax = i; }
         }
         if (j != imax) {
            for (int k = 0; k < N; k++) { const R d = A[imax + k*N]; A[imax + k*N] = A[j + k*N]; A[j + k*N] = d; }
            vv[imax] = vv[j];
         }
         ind[j] = imax;
         if (A[j + j*N] == (R)0) A[j + j*N] = (R)1.0e-20;
         if (j != N-1) {
            const R dum = (R)1/(A[j + j*N]);
            for (int i = j+1; i < N; i++) A[i + j*N] *= dum;
         }
      }
   }

   template<typename R, typename V>
   __device__ static void lubksb(const R *A, const int *ind, V *b)
   {
      int ii = 0;
      for (int i = 0; i < N; i++) {
         const int ip = ind[i];
         V sum = b[ip];
         b[ip] = b[i];
         if (ii != 0)
            for (int j = ii-1; j < i; j++) sum -= (V)A[i + j*N]*b[j];
         else if (sum != (V)0)
            ii = i+1;
         b[i] = sum;
      }
      for (int i = N-1; i > -1; i--) {
         V sum = b[i];
         for (int j = i+1; j < N; j++) sum -= (V)A[i + j*N]*b[j];
         b[i] = sum/(V)A[i + i*N];
      }
   }

   double Jac[N*N];
   double A64[N*N];                                   // FP64: the LU itself; IR: the unfactorised matrix
   float  A32[PREC == HTR_CHEM_FP64 ? 1 : N*N];       // IR: the scaled FP32 LU
   double L64[PREC == HTR_CHEM_FP64 ? 1 : N*N];       // IR: FP64 LU made only when refinement stalls
   double rs[PREC == HTR_CHEM_FP64 ? 1 : N], cs[PREC == HTR_CHEM_FP64 ? 1 : N];
   int    ind[N], ind64[PREC == HTR_CHEM_FP64 ? 1 : N];
   bool   lu64_ready;
   int    fallbacks = 0;
   double err[N], dx[N], xsav[N], dxsav[N];

   __device__ void jacobian(double *x)
   {
      double tmp[N], fp[N], fm[N];
#pragma unroll
      for (int j = 0; j < N; j++) tmp[j] = x[j];
      for (int j = 0; j < N; j++) {
         const double h = x[j]*1.0e-6 + 1.0e-14;
         x[j] = tmp[j] + h;
         const double hp = x[j] - tmp[j];
         rhs(fp, x);
         x[j] = tmp[j] - h;
         const double hm = tmp[j] - x[j];
         rhs(fm, x);
         x[j] = tmp[j];
         for (int i = 0; i < N; i++) Jac[i + j*N] = (fp[i] - fm[i])/(hp + hm);
      }
   }

   // power of two nearest to 1/a (exact scaling)
   __device__ static double pow2_inv(double a)
   {
      if (!(a > 0.0) || !(a < 1.0e308)) return 1.0;
      int e;
      frexp(a, &e);
      return ldexp(1.0, -e);
   }

   __device__ void factor(double dt)
   {
      for (int j = 0; j < N; j++)
         for (int i = 0; i < N; i++) A64[i + j*N] = -Jac[i + j*N];
      for (int i = 0; i < N; i++) A64[i + i*N] += 1.0/((1.0/2.0)*dt);
      if (PREC == HTR_CHEM_FP64) {
         ludcmp<double>(A64, ind);
      } else {
         // equilibrate by powers of two (rows, then columns) before rounding to
         // FP32: the FP32 LU then sees entries of comparable magnitude
         for (int i = 0; i < N; i++) {
            double big = 0.0;
            for (int j = 0; j < N; j++) big = dmax(big, fabs(A64[i + j*N]));
            rs[i] = pow2_inv(big);
         }
         for (int j = 0; j < N; j++) {
            double big = 0.0;
            for (int i = 0; i < N; i++) big = dmax(big, fabs(rs[i]*A64[i + j*N]));
            cs[j] = pow2_inv(big);
         }
         for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) A32[i + j*N] = (float)(rs[i]*A64[i + j*N]*cs[j]);
         ludcmp<float>(A32, ind);
         lu64_ready = false;
      }
   }

   // x = A^-1 r through the scaled FP32 LU
   __device__ void solve32(const double *r, double *x)
   {
      float c[N];
#pragma unroll
      for (int i = 0; i < N; i++) c[i] = (float)(rs[i]*r[i]);
      lubksb<float, float>(A32, ind, c);
#pragma unroll
      for (int i = 0; i < N; i++) x[i] = cs[i]*(double)c[i];
   }

   __device__ void solve(double *b)
   {
      if (PREC == HTR_CHEM_FP64) { lubksb<double, double>(A64, ind, b); return; }
      // FP32 solve + FP64 residual refinement, until the correction is below
      // 1e-15 of the solution; a stalled refinement (the correction shrinks by
      // less than half in a sweep) falls back to an FP64 LU of the same matrix
      double rhs0[N], x[N], r[N], c[N];
#pragma unroll
      for (int i = 0; i < N; i++) rhs0[i] = b[i];
      solve32(rhs0, x);
      bool done = false;
      double prev = 0.0;
      for (int it = 0; it < 8 && !done; it++) {
         double nx = 0.0;
         for (int i = 0; i < N; i++) {
            double s = rhs0[i];
            for (int j = 0; j < N; j++) s -= A64[i + j*N]*x[j];
            r[i] = s;
            nx = dmax(nx, fabs(x[i]));
         }
         solve32(r, c);
         double nc = 0.0;
#pragma unroll
         for (int i = 0; i < N; i++) { x[i] += c[i]; nc = dmax(nc, fabs(c[i])); }
         if (!(nc > 1e-15*nx)) done = true;
         else if (it > 0 && !(nc < 0.5*prev)) break;
         prev = nc;
      }
      if (!done) {
         if (!lu64_ready) {
            for (int k = 0; k < N*N; k++) L64[k] = A64[k];
            ludcmp<double>(L64, ind64);
            lu64_ready = true;
         }
#pragma unroll
         for (int i = 0; i < N; i++) x[i] = rhs0[i];
         lubksb<double, double>(L64, ind64, x);
         fallbacks++;
      }
#pragma unroll
      for (int i = 0; i < N; i++) b[i] = x[i];
   }

   __device__ double step(double *x, double dt, double t0)
   {
      double g1[N], g2[N], g3[N], g4[N];
      factor(dt);
#pragma unroll
      for (int i = 0; i < N; i++) g1[i] = dxsav[i] + dt*(1.0/2.0)*dx[i];
      solve(g1);
#pragma unroll
      for (int i = 0; i < N; i++) x[i] = xsav[i] + 2.0*g1[i];
      rhs(dx, x);
#pragma unroll
      for (int i = 0; i < N; i++) g2[i] = dx[i] + dt*(-3.0/2.0)*dx[i] + (-8.0)*g1[i]/dt;
      solve(g2);
#pragma unroll
      for (int i = 0; i < N; i++) x[i] = xsav[i] + (48.0/25.0)*g1[i] + (6.0/25.0)*g2[i];
      rhs(dx, x);
#pragma unroll
      for (int i = 0; i < N; i++) g3[i] = dx[i] + dt*(121.0/50.0)*dx[i] + ((372.0/25.0)*g1[i] + (12.0/5.0)*g2[i])/dt;
      solve(g3);
#pragma unroll
      for (int i = 0; i < N; i++)
         g4[i] = dx[i] + dt*(29.0/250.0)*dx[i] + ((-112.0/125.0)*g1[i] + (-54.0/125.0)*g2[i] + (-2.0/5.0)*g3[i])/dt;
      solve(g4);
#pragma unroll
      for (int i = 0; i < N; i++)
         x[i] = xsav[i] + (19.0/9.0)*g1[i] + (1.0/2.0)*g2[i] + (25.0/108.0)*g3[i] + (125.0/108.0)*g4[i];
#pragma unroll
      for (int i = 0; i < N; i++)
         err[i] = (17.0/54.0)*g1[i] + (7.0/36.0)*g2[i] + 0.0*g3[i] + (125.0/108.0)*g4[i];
      return t0 + dt;
   }

   __device__ bool itercheck(double &dt, double &dtNext, bool finish)
   {
      double errmax = 0.0;
#pragma unroll
      for (int i = 0; i < N; i++) errmax = dmax(errmax, fabs(err[i]));
      errmax *= (1.0/1e-10);
      if ((errmax < 1.0) || finish) {
         if (errmax > 0.1296) dtNext = 0.9*dt*pow(errmax, -0.25);
         else                 dtNext = 1.5*dt;
         return true;
      }
      dtNext = 0.9*dt*pow(errmax, -1.0/3.0);
      if (dt > 0.0) dt = dmax(dtNext, 0.5*dt);
      else          dt = dmin(dtNext, 0.5*dt);
      return false;
   }

   // returns accepted steps, or -1 if a reference cap was hit
   __device__ int solve_cell(double *x, double dtTry, double DelT)
   {
      bool finish = false;
      double time = 0.0;
      double dt = dtTry;
      for (int st = 0; st < 100000; st++) {
         const double t0 = time;
         double dtNext = 0.0;
         jacobian(x);
         rhs(dx, x);
#pragma unroll
         for (int i = 0; i < N; i++) { xsav[i] = x[i]; dxsav[i] = dx[i]; }
         int jtry;
         for (jtry = 0; jtry < 100; jtry++) {
            time = step(x, dt, t0);
            if (itercheck(dt, dtNext, finish)) break;
         }
         if (jtry == 100) return -1;
         const double Remainder = DelT - time;
         if (finish || (Remainder < 1e-12*DelT)) return st + 1;
         if (dtNext*1.5 > Remainder) { dt = Remainder; finish = true; }
         else                        { dt = dtNext;    finish = false; }
      }
      return -1;
   }
};

template<class MX, int PREC>
__global__ void __launch_bounds__(256)
chem_update_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, double dt,
                   htr_view nsteps, int *__restrict__ fails, int64_t n)
{
   constexpr int N = MX::NEQ;
   for (int64_t i = (int64_t)blockIdx.x*blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x*blockDim.x) {
      const P3 p = exec_point(g, i);
      double x[N], ct[N];
#pragma unroll
      for (int k = 0; k < N; k++) {
         ct[k] = ld<double>(f.Conserved_t, p, k);
         at<double>(f.Conserved_t_old, p, k) = ct[k];
         x[k] = ld<double>(f.Conserved, p, k);
      }
      Rosen<MX, PREC> R{*mix, ct, ld<double>(f.temperature, p)};
      const int st = R.solve_cell(x, dt, dt);
#pragma unroll
      for (int k = 0; k < N; k++) at<double>(f.Conserved, p, k) = x[k];
      at<double>(f.temperature, p) = R.T;
      if (nsteps.ptr) at<int32_t>(nsteps, p) = st;
      if (st < 0) atomicAdd(fails, 1);
   }
}

// Launch shape of the persistent cell loop: threads per block and resident
// blocks per CU (256 CUs); set by htr_chem_launch_config, chosen by the
// autotuner in htr_amd.chem_autotune.
static int g_chem_block = 64;
static int g_chem_blocks_per_cu = 16;

int chem_set_launch(int block, int blocks_per_cu)
{
   if (block != 64 && block != 128 && block != 256) return set_error(HTR_E_ARG, "chem block size %d", block);
   if (blocks_per_cu < 1 || blocks_per_cu > 64) return set_error(HTR_E_ARG, "chem blocks per CU %d", blocks_per_cu);
   g_chem_block = block;
   g_chem_blocks_per_cu = blocks_per_cu;
   return HTR_OK;
}

template<class MX>
int chem_update_launch(const htr_fields &f, const DMix *mix, const Geo &g, double dt, int precision,
                       const htr_view &nsteps, hipStream_t s)
{
   const int64_t n = exec_cells(g);
   if (n == 0) return HTR_OK;
   Workspace *wp = workspace(s);
   if (!wp) return HTR_E_DEVICE;
   Workspace &w = *wp;
   if (int rc0 = zero_counters(w.counter, s)) return rc0;
   const int bs = g_chem_block;
   const unsigned blocks = (unsigned)std::min<int64_t>(grid_for(n, bs), 256*g_chem_blocks_per_cu);
   if (precision == HTR_CHEM_FP64)
      HTR_LAUNCH("chem_update_fp64", (chem_update_kernel<MX, HTR_CHEM_FP64>), dim3(blocks), dim3(bs), 0, s, f, mix, g, dt, nsteps, w.counter, n);
   else
      HTR_LAUNCH("chem_update_fp32ir", (chem_update_kernel<MX, HTR_CHEM_FP32_IR>), dim3(blocks), dim3(bs), 0, s, f, mix, g, dt, nsteps, w.counter, n);
   int rc = check_launch("chem_update_kernel");
   if (rc) return rc;
   int fails = 0;
   if (hipMemcpyAsync(&fails, w.counter, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess)
      return set_error(HTR_E_DEVICE, "chem: copy back failed");
   if (fails) return set_error(HTR_E_NUMERIC, "chem: %d cells hit the Rosenbrock step caps", fails);
   return HTR_OK;
}

template int chem_update_launch<MixAir>(const htr_fields &, const DMix *, const Geo &, double, int, const htr_view &, hipStream_t);
template int chem_update_launch<Mix30>(const htr_fields &, const DMix *, const Geo &, double, int, const htr_view &, hipStream_t);

} // namespace htr
