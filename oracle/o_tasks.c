/*
 * o_tasks.c -- oracle restatement of the point-wise tasks, the sensors, the
 * spectral-radius reduction, the chemistry tasks and the implicit solver.
 * TEST INFRASTRUCTURE ONLY (see htr_oracle.h).
 *
 *   src/prometeo_variables.hpp:62-192   UpdateProperties/Conserved/Primitive
 *   src/prometeo_sensor.inl:52-77       Ducros sensor
 *   src/prometeo_sensor.cc:37-136       Ducros and shock-sensor task loops
 *   src/prometeo_cfl.hpp:73-178         spectral radii
 *   src/prometeo_chem.hpp:53-93         ImplicitSolver::rhs
 *   src/prometeo_chem.cc:36-112         UpdateChemistry, AddChemistrySources
 *   src/Utils/math_utils.hpp:87-438     LUdec, Rosenbrock (Kaps-Rentrop)
 *   src/prometeo_rk.rg:88-112 + src/prometeo_const.rg:54-58  SSP-RK3 update
 *
 * Pinned by: variablesTest (primitive <-> conserved round trip and
 * properties, 1e-8), chemTest (explicit source 1e-5 and the implicit state
 * after 1e-6 s, 1e-5), cflTest (radii 11, 40, 11.42857, 40),
 * CPPmathUtilsTest (LU solution (3,4,-5) at 1e-12; Rosenbrock<3> at t = 25 at
 * 1e-6); Ducros sensor: pure rotation -> 0, pure dilatation -> 1.
 */
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include "htr_oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

static inline double dmax(double a, double b) { return (a < b) ? b : a; }
static inline double dmin(double a, double b) { return (b < a) ? b : a; }

#define RK_OMP _Pragma("omp parallel for collapse(3) schedule(static)")

static inline long L(const o_grid *g, int i, int j, int k)
{ return (long)i + (long)g->nx*((long)j + (long)g->ny*(long)k); }
static inline long NN(const o_grid *g) { return (long)g->nx*g->ny*g->nz; }

/* ------------------------------ variables --------------------------------- */
void o_update_properties(const o_grid *g, const o_fields *f, const o_mix *m)
{
   const long N = NN(g);
   const int ns = m->nSpec;
   RK_OMP
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++) {
            const long p = L(g, i, j, k);
            double X[O_NSP_MAX], Y[O_NSP_MAX], D[O_NSP_MAX];
            for (int s = 0; s < ns; s++) X[s] = f->MolarFracs[s*N + p];
            const double P = f->pressure[p], T = f->temperature[p];
            const double MixW = o_GetMolarWeightFromXi(m, X);
            f->rho[p] = o_GetRho(m, P, T, MixW);
            f->mu[p] = o_GetViscosity(m, T, X);
            f->lam[p] = o_GetHeatConductivity(m, T, X);
            o_GetDiffusivity(m, D, P, T, MixW, X);
            o_GetMassFractions(m, Y, MixW, X);
            for (int s = 0; s < ns; s++) { f->Di[s*N + p] = D[s]; f->MassFracs[s*N + p] = Y[s]; }
            const double gamma = o_GetGamma(m, T, MixW, Y);
            f->SoS[p] = o_GetSpeedOfSound(m, T, gamma, MixW);
         }
}

void o_update_conserved(const o_grid *g, const o_fields *f, const o_mix *m)
{
   const long N = NN(g);
   const int ns = m->nSpec;
   RK_OMP
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++) {
            const long p = L(g, i, j, k);
            double Y[O_NSP_MAX], v[3];
            const double rho = f->rho[p];
            for (int s = 0; s < ns; s++) Y[s] = f->MassFracs[s*N + p];
            for (int s = 0; s < ns; s++) f->Conserved[s*N + p] = rho*Y[s];
            for (int d = 0; d < 3; d++) { v[d] = f->velocity[d*N + p]; f->Conserved[(ns+d)*N + p] = rho*v[d]; }
            double k2 = 0.0; k2 += v[0]*v[0]; k2 += v[1]*v[1]; k2 += v[2]*v[2];
            f->Conserved[(ns+3)*N + p] = rho*(0.5*k2 + o_GetInternalEnergy(m, f->temperature[p], Y));
         }
}

/* prometeo_variables.hpp:170-192; iters (if not NULL) receives the Newton count */
static void primitive_cell(const o_mix *m, const double *x, double *T, double *P,
                           double *X, double *v)
{
   const int ns = m->nSpec;
   double rhoYi[O_NSP_MAX], Yi[O_NSP_MAX];
   for (int i = 0; i < ns; i++) rhoYi[i] = x[i];
   double rho;
   if (m->constprop) rho = rhoYi[0];
   else { rho = 0.0; for (int i = 0; i < ns; i++) rho += rhoYi[i]; }
   if (m->constprop) Yi[0] = rhoYi[0]/rho;
   else {
      const double irho = 1.0/rho;
      for (int i = 0; i < ns; i++) Yi[i] = irho*rhoYi[i];
      for (int i = 0; i < ns; i++) { Yi[i] = dmax(Yi[i], 1e-60); Yi[i] = dmin(Yi[i], 1.0); }
   }
   const double MixW = o_GetMolarWeightFromYi(m, Yi);
   o_GetMolarFractions(m, X, MixW, Yi);
   const double rhoInv = 1.0/rho;
   for (int i = 0; i < 3; i++) v[i] = x[ns+i]*rhoInv;
   double k2 = 0.0; k2 += v[0]*v[0]; k2 += v[1]*v[1]; k2 += v[2]*v[2];
   const double kinetic = 0.5*k2;
   const double e = x[ns+3]*rhoInv - kinetic;
   *T = o_GetTFromInternalEnergy(m, e, *T, Yi);
   *P = o_GetPFromRhoAndT(m, rho, MixW, *T);
}

void o_update_primitive(const o_grid *g, const o_fields *f, const o_mix *m)
{
   const long N = NN(g);
   const int ns = m->nSpec;
   RK_OMP
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++) {
            const long p = L(g, i, j, k);
            double x[O_NEQ_MAX], X[O_NSP_MAX], v[3];
            for (int s = 0; s < ns+4; s++) x[s] = f->Conserved[s*N + p];
            double T = f->temperature[p], P;
            primitive_cell(m, x, &T, &P, X, v);
            f->temperature[p] = T;
            f->pressure[p] = P;
            for (int s = 0; s < ns; s++) f->MolarFracs[s*N + p] = X[s];
            for (int d = 0; d < 3; d++) f->velocity[d*N + p] = v[d];
         }
}

/* ------------------------------- sensors ---------------------------------- */
static inline int wrapm(int c, int off, int n) { return ((c + off) % n + n) % n; }
static inline int wrapp(int c, int off, int n) { return (c + off) % n; }

static void vgrad(const o_grid *g, const o_fields *f, int d, int i, int j, int k, double *out)
{
   const long N = NN(g);
   const long p = L(g, i, j, k);
   const int t = f->nType[d][p];
   const double m = f->m_d[d][p];
   int cm[3] = { i, j, k }, cp[3] = { i, j, k };
   const int n = d == 0 ? g->nx : d == 1 ? g->ny : g->nz;
   cm[d] = wrapm(cm[d], o_coeff_Cp(t, 1), n);
   cp[d] = wrapp(cp[d], o_coeff_Cp(t, 2), n);
   const long pm = L(g, cm[0], cm[1], cm[2]), pq = L(g, cp[0], cp[1], cp[2]);
   for (int c = 0; c < 3; c++)
      out[c] = ((double)o_coeff_Grad(t, 0)*(f->velocity[c*N + p] - f->velocity[c*N + pm]) +
                (double)o_coeff_Grad(t, 1)*(f->velocity[c*N + pq] - f->velocity[c*N + p]))*m;
}

void o_ducros_sensor(const o_grid *g, const o_fields *f, double vorticityScale)
{
   const double eps = dmax(vorticityScale*vorticityScale, 1e-6);
   RK_OMP
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++) {
            double gx[3], gy[3], gz[3];
            vgrad(g, f, 0, i, j, k, gx);
            vgrad(g, f, 1, i, j, k, gy);
            vgrad(g, f, 2, i, j, k, gz);
            const double div = gx[0] + gy[1] + gz[2];
            const double div2 = div*div;
            const double omz = gx[1] - gy[0];
            const double omy = gy[2] - gz[1];
            const double omx = gz[0] - gx[2];
            const double om2 = omx*omx + omy*omy + omz*omz;
            f->DucrosSensor[L(g, i, j, k)] = div2/(div2 + om2 + eps);
         }
}

void o_shock_sensor(const o_grid *g, const o_fields *f, int dir, const o_mix *m)
{
   const long N = NN(g);
   const int n = dir == 0 ? g->nx : dir == 1 ? g->ny : g->nz;
   RK_OMP
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++) {
            const long p = L(g, i, j, k);
            const int t = f->nType[dir][p];
            long q[6];
            const int offs[6] = { o_coeff_Cp(t, 0), o_coeff_Cp(t, 1), 0, o_coeff_Cp(t, 2), o_coeff_Cp(t, 3), o_coeff_Cp(t, 4) };
            for (int s = 0; s < 6; s++) {
               int c[3] = { i, j, k };
               c[dir] = (s < 2) ? wrapm(c[dir], offs[s], n) : wrapp(c[dir], offs[s], n);
               q[s] = L(g, c[0], c[1], c[2]);
            }
            double Phi = f->DucrosSensor[q[0]];
            for (int s = 1; s < 6; s++) Phi = dmax(Phi, f->DucrosSensor[q[s]]);
            int sensor = 1;
            for (int h = 0; h < m->nSpec; h++) {
               if (!sensor) break;
               double y[6];
               for (int s = 0; s < 6; s++) y[s] = f->Conserved[h*N + q[s]];
               sensor = sensor && o_tenoa_sensor(y, t, Phi);
            }
            f->shockSensor[dir][p] = (uint8_t)sensor;
         }
}

/* --------------------------- spectral radius ------------------------------ */
double o_max_spectral_radius(const o_grid *g, const o_fields *f, const o_mix *m)
{
   const long N = NN(g);
   const int ns = m->nSpec;
   double r = 0.0;
   #pragma omp parallel for collapse(3) reduction(max:r) schedule(static)
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++) {
            const long p = L(g, i, j, k);
            const double dcsi = f->m_d[0][p], deta = f->m_d[1][p], dzet = f->m_d[2][p];
            const double SoS = f->SoS[p];
            double rr = (fabs(f->velocity[p]) + SoS)*dcsi;
            rr = dmax(rr, (fabs(f->velocity[N + p]) + SoS)*deta);
            rr = dmax(rr, (fabs(f->velocity[2*N + p]) + SoS)*dzet);
            /* viscous */
            const double nu = f->mu[p]/f->rho[p];
            double rv = nu*dcsi*dcsi;
            rv = dmax(rv, nu*deta*deta);
            rv = dmax(rv, nu*dzet*dzet);
            rv *= 4;
            rr = dmax(rr, rv);
            /* heat conduction */
            double Y[O_NSP_MAX];
            for (int s = 0; s < ns; s++) Y[s] = f->MassFracs[s*N + p];
            const double cp = o_GetHeatCapacity(m, f->temperature[p], Y);
            const double DifT = f->lam[p]/(cp*f->rho[p]);
            double rt = DifT*dcsi*dcsi;
            rt = dmax(rt, DifT*deta*deta);
            rt = dmax(rt, DifT*dzet*dzet);
            rt *= 4;
            rr = dmax(rr, rt);
            /* species diffusion */
            double rs = 0.0;
            for (int s = 0; s < ns; s++) {
               const double D = f->Di[s*N + p];
               rs = dmax(rs, D*dcsi*dcsi);
               rs = dmax(rs, D*deta*deta);
               rs = dmax(rs, D*dzet*dzet);
            }
            rs *= 4;
            rr = dmax(rr, rs);
            if (rr > r) r = rr;
         }
   return r;
}

/* ------------------------------ chemistry --------------------------------- */
void o_chem_source(const o_grid *g, const o_fields *f, const o_mix *m)
{
   const long N = NN(g);
   const int ns = m->nSpec;
   RK_OMP
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++) {
            const long p = L(g, i, j, k);
            double Y[O_NSP_MAX], w[O_NSP_MAX];
            for (int s = 0; s < ns; s++) Y[s] = f->MassFracs[s*N + p];
            o_GetProductionRates(m, w, f->rho[p], f->pressure[p], f->temperature[p], Y);
            for (int s = 0; s < ns; s++) f->Conserved_t[s*N + p] += w[s];
         }
}

/* ---------------------------- LU (Crout, NR) ------------------------------ */
/* A column-major n x n: A(i,j) = A[i + j*n]  (my_array.hpp:209) */
#define AA(i, j) A[(i) + (j)*n]
static void ludcmp(int n, double *A, int *ind)
{
   int imax = 0;
   double vv[O_NEQ_MAX];
   for (int i = 0; i < n; i++) {
      double big = 0.0;
      for (int j = 0; j < n; j++) big = dmax(fabs(AA(i, j)), big);
      vv[i] = 1.0/big;
   }
   for (int j = 0; j < n; j++) {
      for (int i = 0; i < j; i++) {
         double sum = AA(i, j);
         for (int k = 0; k < i; k++) sum -= AA(i, k)*AA(k, j);
         AA(i, j) = sum;
      }
      double big = 0.0;
      for (int i = j; i < n; i++) {
         double sum = AA(i, j);
         for (int k = 0; k < j; k++) sum -= AA(i, k)*AA(k, j);
         AA(i, j) = sum;
         const double dum = vv[i]*fabs(sum);
         if (dum >= big) { big = dum; imax = i; }
      }
      if (j != imax) {
         for (int k = 0; k < n; k++) {
            const double dum = AA(imax, k);
            AA(imax, k) = AA(j, k);
            AA(j, k) = dum;
         }
         vv[imax] = vv[j];
      }
      ind[j] = imax;
      if (AA(j, j) == 0.0) AA(j, j) = 1.0e-20;
      if (j != n-1) {
         const double dum = 1.0/(AA(j, j));
         for (int i = j+1; i < n; i++) AA(i, j) *= dum;
      }
   }
}

static void lubksb(int n, const double *A, const int *ind, double *b)
{
   int ii = 0;
   for (int i = 0; i < n; i++) {
      const int ip = ind[i];
      double sum = b[ip];
      b[ip] = b[i];
      if (ii != 0)
         for (int j = ii-1; j < i; j++) sum -= AA(i, j)*b[j];
      else if (sum != 0.0)
         ii = i+1;
      b[i] = sum;
   }
   for (int i = n-1; i > -1; i--) {
      double sum = b[i];
      for (int j = i+1; j < n; j++) sum -= AA(i, j)*b[j];
      b[i] = sum/AA(i, i);
   }
}
#undef AA

void o_lu_solve(int n, const double *Ain, double *b)
{
   double A[O_NEQ_MAX*O_NEQ_MAX];
   int ind[O_NEQ_MAX];
   memcpy(A, Ain, sizeof(double)*n*n);
   ludcmp(n, A, ind);
   lubksb(n, A, ind, b);
}

/* --------------------------- Rosenbrock solver ---------------------------- */
typedef void (*rhs_fn)(void *ctx, double *r, const double *x);
/* Jacobian at x after rhs(x) (analytic mode) */
typedef void (*jac_fn)(void *ctx, double *J, const double *x);

typedef struct {
   int n;
   rhs_fn rhs;
   jac_fn jac;                 /* 0: the reference's central differences */
   double tol;                 /* step tolerance (the reference's 1e-10) */
   void *ctx;
   double err[O_NEQ_MAX], dx[O_NEQ_MAX], xsav[O_NEQ_MAX], dxsav[O_NEQ_MAX];
   double Jac[O_NEQ_MAX*O_NEQ_MAX], LU[O_NEQ_MAX*O_NEQ_MAX];
   int ind[O_NEQ_MAX];
} rb_t;

static const int    RB_MAXSTP = 100000;
static const int    RB_MAXITS = 100;
#define RB_TOL    1e-10
#define RB_SAFETY 0.9
#define RB_GROW   1.5
#define RB_PGROW  (-0.25)
#define RB_SHRNK  0.5
#define RB_PSHRNK (-1.0/3.0)
#define RB_ERRCON 0.1296
#define RB_GAM    (1.0/2.0)
#define RB_A21    2.0
#define RB_A31    (48.0/25.0)
#define RB_A32    (6.0/25.0)
#define RB_C21    (-8.0)
#define RB_C31    (372.0/25.0)
#define RB_C32    (12.0/5.0)
#define RB_C41    (-112.0/125.0)
#define RB_C42    (-54.0/125.0)
#define RB_C43    (-2.0/5.0)
#define RB_B1     (19.0/9.0)
#define RB_B2     (1.0/2.0)
#define RB_B3     (25.0/108.0)
#define RB_B4     (125.0/108.0)
#define RB_E1     (17.0/54.0)
#define RB_E2     (7.0/36.0)
#define RB_E3     0.0
#define RB_E4     (125.0/108.0)
#define RB_C1X    (1.0/2.0)
#define RB_C2X    (-3.0/2.0)
#define RB_C3X    (121.0/50.0)
#define RB_C4X    (29.0/250.0)
#define RB_A2X    1.0
#define RB_A3X    (3.0/5.0)

/* math_utils.hpp:247-271 */
static void rb_jacobian(rb_t *s, double *x)
{
   const int n = s->n;
   const double EPS = 1.0e-6, DEL = 1.0e-14;
   double tmp[O_NEQ_MAX], fp[O_NEQ_MAX], fm[O_NEQ_MAX];
   memcpy(tmp, x, sizeof(double)*n);
   for (int j = 0; j < n; j++) {
      const double h = x[j]*EPS + DEL;
      x[j] = tmp[j] + h;
      const double hp = x[j] - tmp[j];
      s->rhs(s->ctx, fp, x);
      x[j] = tmp[j] - h;
      const double hm = tmp[j] - x[j];
      s->rhs(s->ctx, fm, x);
      x[j] = tmp[j];
      for (int i = 0; i < n; i++) s->Jac[i + j*n] = (fp[i] - fm[i])/(hp + hm);
   }
}

/* math_utils.hpp:273-333 */
static double rb_step(rb_t *s, double *x, double dt, double t0)
{
   const int n = s->n;
   double g1[O_NEQ_MAX], g2[O_NEQ_MAX], g3[O_NEQ_MAX], g4[O_NEQ_MAX];
   for (int i = 0; i < n; i++) {
      for (int j = 0; j < n; j++) s->LU[i + j*n] = -s->Jac[i + j*n];
      s->LU[i + i*n] += 1.0/(RB_GAM*dt);
   }
   ludcmp(n, s->LU, s->ind);
   for (int i = 0; i < n; i++) g1[i] = s->dxsav[i] + dt*RB_C1X*s->dx[i];
   lubksb(n, s->LU, s->ind, g1);
   for (int i = 0; i < n; i++) x[i] = s->xsav[i] + RB_A21*g1[i];
   double time = t0 + RB_A2X*dt;
   s->rhs(s->ctx, s->dx, x);
   for (int i = 0; i < n; i++) g2[i] = s->dx[i] + dt*RB_C2X*s->dx[i] + RB_C21*g1[i]/dt;
   lubksb(n, s->LU, s->ind, g2);
   for (int i = 0; i < n; i++) x[i] = s->xsav[i] + RB_A31*g1[i] + RB_A32*g2[i];
   time = t0 + RB_A3X*dt;
   s->rhs(s->ctx, s->dx, x);
   for (int i = 0; i < n; i++) g3[i] = s->dx[i] + dt*RB_C3X*s->dx[i] + (RB_C31*g1[i] + RB_C32*g2[i])/dt;
   lubksb(n, s->LU, s->ind, g3);
   for (int i = 0; i < n; i++) g4[i] = s->dx[i] + dt*RB_C4X*s->dx[i] + (RB_C41*g1[i] + RB_C42*g2[i] + RB_C43*g3[i])/dt;
   lubksb(n, s->LU, s->ind, g4);
   for (int i = 0; i < n; i++) x[i] = s->xsav[i] + RB_B1*g1[i] + RB_B2*g2[i] + RB_B3*g3[i] + RB_B4*g4[i];
   for (int i = 0; i < n; i++) s->err[i] = RB_E1*g1[i] + RB_E2*g2[i] + RB_E3*g3[i] + RB_E4*g4[i];
   time = t0 + dt;
   (void)time;
   return t0 + dt;
}

/* math_utils.hpp:337-363 */
static int rb_itercheck(rb_t *s, double *dt, double *dtNext, int finish)
{
   double errmax = 0.0;
   for (int i = 0; i < s->n; i++) errmax = dmax(errmax, fabs(s->err[i]));
   errmax *= (1.0/s->tol);
   if ((errmax < 1.0) || finish) {
      if (errmax > RB_ERRCON) *dtNext = RB_SAFETY*(*dt)*pow(errmax, RB_PGROW);
      else                    *dtNext = RB_GROW*(*dt);
      return 1;
   }
   *dtNext = RB_SAFETY*(*dt)*pow(errmax, RB_PSHRNK);
   if (*dt > 0.0) *dt = dmax(*dtNext, RB_SHRNK*(*dt));
   else           *dt = dmin(*dtNext, RB_SHRNK*(*dt));
   return 0;
}

/* math_utils.hpp:367-386 */
static int rb_stepcheck(double *dt, double dtNext, double DelT, double time, int *finish)
{
   const double Remainder = DelT - time;
   if (*finish || (Remainder < 1e-12*DelT)) return 1;
   if (dtNext*1.5 > Remainder) { *dt = Remainder; *finish = 1; }
   else                        { *dt = dtNext;    *finish = 0; }
   return 0;
}

/* math_utils.hpp:208-242; returns the number of accepted steps, or -1 when a
 * step or iteration cap of the reference is reached */
static int rb_solve(rb_t *s, double *x, double dtTry, double DelT)
{
   int finish = 0;
   double time = 0.0;
   double dt = dtTry;
   for (int st = 0; st < RB_MAXSTP; st++) {
      const double t0 = time;
      double dtNext = 0.0;
      if (s->jac) {
         /* analytic Jacobian (not the reference's order: the right-hand side
          * first, so that the Jacobian sees the converged temperature) */
         s->rhs(s->ctx, s->dx, x);
         s->jac(s->ctx, s->Jac, x);
      } else {
         rb_jacobian(s, x);
         s->rhs(s->ctx, s->dx, x);
      }
      memcpy(s->xsav, x, sizeof(double)*s->n);
      memcpy(s->dxsav, s->dx, sizeof(double)*s->n);
      int jtry;
      for (jtry = 0; jtry < RB_MAXITS; jtry++) {
         time = rb_step(s, x, dt, t0);
         if (rb_itercheck(s, &dt, &dtNext, finish)) break;
      }
      if (jtry == RB_MAXITS) return -1;
      if (rb_stepcheck(&dt, dtNext, DelT, time, &finish)) return st + 1;
   }
   return -1;
}

/* CPPmathUtilsTest.cc:32-39 problem */
static void testprob_rhs(void *ctx, double *r, const double *x)
{
   (void)ctx;
   r[0] = -0.013*x[0] - 1000.0*x[0]*x[2];
   r[1] = -2500.0*x[1]*x[2];
   r[2] = -0.013*x[0] - 1000.0*x[0]*x[2] - 2500.0*x[1]*x[2];
}

void o_rosenbrock_testprob(double *x, double dtTry, double DelT, int32_t *nsteps)
{
   rb_t *s = (rb_t*)calloc(1, sizeof(rb_t));
   s->n = 3; s->rhs = testprob_rhs; s->ctx = 0; s->tol = RB_TOL;
   const int st = rb_solve(s, x, dtTry, DelT);
   if (nsteps) *nsteps = st;
   free(s);
}

/* ImplicitSolver::rhs (prometeo_chem.hpp:66-88) */
typedef struct { const o_mix *m; const double *Ct_old; double *T; } chem_ctx;

static void chem_rhs(void *vctx, double *r, const double *x)
{
   chem_ctx *c = (chem_ctx*)vctx;
   const o_mix *m = c->m;
   const int ns = m->nSpec;
   double rhoYi[O_NSP_MAX], Yi[O_NSP_MAX], w[O_NSP_MAX];
   for (int i = 0; i < ns; i++) rhoYi[i] = x[i];
   double rho;
   if (m->constprop) rho = rhoYi[0];
   else { rho = 0.0; for (int i = 0; i < ns; i++) rho += rhoYi[i]; }
   if (m->constprop) Yi[0] = rhoYi[0]/rho;
   else {
      const double irho = 1.0/rho;
      for (int i = 0; i < ns; i++) Yi[i] = irho*rhoYi[i];
      for (int i = 0; i < ns; i++) { Yi[i] = dmax(Yi[i], 1e-60); Yi[i] = dmin(Yi[i], 1.0); }
   }
   const double MixW = o_GetMolarWeightFromYi(m, Yi);
   const double rhoInv = 1.0/rho;
   double v[3];
   for (int i = 0; i < 3; i++) v[i] = x[ns+i]*rhoInv;
   double k2 = 0.0; k2 += v[0]*v[0]; k2 += v[1]*v[1]; k2 += v[2]*v[2];
   const double e = x[ns+3]*rhoInv - 0.5*k2;
   *c->T = o_GetTFromInternalEnergy(m, e, *c->T, Yi);
   const double P = o_GetPFromRhoAndT(m, rho, MixW, *c->T);
   o_GetProductionRates(m, w, rho, P, *c->T, Yi);
   for (int i = 0; i < ns; i++) r[i] = w[i] + c->Ct_old[i];
   for (int i = ns; i < ns+4; i++) r[i] = c->Ct_old[i];
}

static void chem_jac(void *vctx, double *J, const double *x)
{
   chem_ctx *c = (chem_ctx*)vctx;
   o_chem_jacobian_analytic(c->m, x, *c->T, J);
}

int o_chem_cell_ex(const o_mix *m, double *x, const double *Ct_old, double *temperature, double dt,
                   int jacobian, double tol)
{
   rb_t *s = (rb_t*)calloc(1, sizeof(rb_t));
   chem_ctx c = { m, Ct_old, temperature };
   s->n = m->nSpec + 4; s->rhs = chem_rhs; s->ctx = &c;
   s->jac = (jacobian == O_JAC_ANALYTIC) ? chem_jac : 0;
   s->tol = tol;
   const int st = rb_solve(s, x, dt, dt);
   free(s);
   return st;
}

int o_chem_cell(const o_mix *m, double *x, const double *Ct_old, double *temperature, double dt)
{ return o_chem_cell_ex(m, x, Ct_old, temperature, dt, O_JAC_FD, RB_TOL); }

void o_chem_jacobian(const o_mix *m, const double *x0, const double *Ct_old, double *temperature,
                     int jacobian, double *J)
{
   rb_t *s = (rb_t*)calloc(1, sizeof(rb_t));
   chem_ctx c = { m, Ct_old, temperature };
   double x[O_NEQ_MAX], r[O_NEQ_MAX];
   s->n = m->nSpec + 4; s->rhs = chem_rhs; s->ctx = &c;
   memcpy(x, x0, sizeof(double)*s->n);
   if (jacobian == O_JAC_ANALYTIC) { chem_rhs(&c, r, x); o_chem_jacobian_analytic(m, x, *temperature, J); }
   else { rb_jacobian(s, x); memcpy(J, s->Jac, sizeof(double)*s->n*s->n); }
   free(s);
}

void o_chem_update_ex(const o_grid *g, const o_fields *f, double dt, const o_mix *m, int32_t *nsteps,
                      int jacobian, double tol)
{
   const long N = NN(g);
   const int neq = m->nSpec + 4;
   #pragma omp parallel for collapse(3) schedule(dynamic, 64)
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++) {
            const long p = L(g, i, j, k);
            double x[O_NEQ_MAX], old[O_NEQ_MAX];
            for (int s = 0; s < neq; s++) {
               old[s] = f->Conserved_t[s*N + p];
               f->Conserved_t_old[s*N + p] = old[s];
               x[s] = f->Conserved[s*N + p];
            }
            double T = f->temperature[p];
            const int st = o_chem_cell_ex(m, x, old, &T, dt, jacobian, tol);
            for (int s = 0; s < neq; s++) f->Conserved[s*N + p] = x[s];
            f->temperature[p] = T;
            if (nsteps) nsteps[p] = st;
         }
}

void o_chem_update(const o_grid *g, const o_fields *f, double dt, const o_mix *m, int32_t *nsteps)
{ o_chem_update_ex(g, f, dt, m, nsteps, O_JAC_FD, RB_TOL); }

/* ------------------------------- RK update -------------------------------- */
void o_rk_update(const o_grid *g, const o_fields *f, int nEq, int stage, double dt)
{
   static const double RKC[3][3] = { { 1.0, 0.0, 1.0 }, { 3.0/4.0, 1.0/4.0, 1.0/4.0 }, { 1.0/3.0, 2.0/3.0, 2.0/3.0 } };
   const double *c = RKC[stage];
   const long N = NN(g);
   RK_OMP
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++) {
            const long p = L(g, i, j, k);
            for (int s = 0; s < nEq; s++)
               f->Conserved[s*N + p] = c[0]*f->Conserved_old[s*N + p] + c[1]*f->Conserved[s*N + p]
                                     + c[2]*f->Conserved_t[s*N + p]*dt;
         }
}

void o_full_rhs(const o_grid *g, const o_fields *f, int scheme, double RK0, double RK1,
                double deltaTime, int chem_source, const o_mix *m)
{
   const long N = NN(g);
   const int neq = m->nSpec + 4;
   RK_OMP
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++)
            for (int s = 0; s < neq; s++) f->Conserved_t[s*N + L(g, i, j, k)] = 0.0;
   if (chem_source) o_chem_source(g, f, m);
   for (int d = 2; d >= 0; d--) o_euler_flux(g, f, d, scheme, RK0, RK1, deltaTime, m, 0);
   for (int d = 2; d >= 0; d--) o_diffusion_flux(g, f, d, m);
}

/* ------------------------------------------------------------------------
 * Boundary ghost cells (src/prometeo_bc.rg:1159-1265): SetDirichletBC,
 * SetAdiabaticWallBC, SetIsothermalWallBC; SetNSCBC_InflowBC and
 * SetNSCBC_OutflowBC (src/prometeo_bc.cc:82-189).  The interior neighbour of a
 * ghost is c_int = c + 1 along the normal for a Neg side, c - 1 for a Pos
 * side (mk_cint of the reference).
 * Pinned by tests/test_oracle_pins.py (wall and Dirichlet ghost values,
 * zero mass flux through a wall).
 */
void o_set_bc(const o_grid *g, const o_fields *f, int kind, int side, double Pbc, const o_mix *m)
{
   if (kind == 4) {                    /* SetNSCBC_OutflowBC: UpdatePrimitive over the rect */
      o_update_primitive(g, f, m);
      return;
   }
   const int ns = m->nSpec;
   const int dir = side/2, inward = (side % 2 == 0) ? 1 : -1;
   const long N = (long)g->nx*g->ny*g->nz;
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++) {
            int c[3] = { i, j, k };
            const long p = (long)i + (long)g->nx*((long)j + (long)g->ny*(long)k);
            c[dir] += inward;
            const long q = (long)c[0] + (long)g->nx*((long)c[1] + (long)g->ny*(long)c[2]);
            if (kind == 3) {
               for (int s = 0; s < ns; s++) f->MolarFracs[s*N + p] = f->MolarFracs_profile[s*N + p];
               f->temperature[p] = f->temperature_profile[p];
               for (int s = 0; s < 3; s++) f->velocity[s*N + p] = f->velocity_profile[s*N + p];
               if (fabs(f->velocity_profile[dir*N + p]) >= f->SoS[p]) {
                  f->pressure[p] = Pbc;        /* supersonic: everything imposed */
               } else {
                  /* pressure from the transported density and the imposed T and X */
                  double rho = 0.0, X[O_NSP_MAX];
                  if (m->constprop) rho = f->Conserved[p];
                  else for (int s = 0; s < ns; s++) rho += f->Conserved[s*N + p];
                  for (int s = 0; s < ns; s++) X[s] = f->MolarFracs_profile[s*N + p];
                  f->pressure[p] = o_GetPFromRhoAndT(m, rho, o_GetMolarWeightFromXi(m, X), f->temperature_profile[p]);
               }
               continue;
            }
            if (kind == 0) {
               for (int s = 0; s < ns; s++) f->MolarFracs[s*N + p] = f->MolarFracs_profile[s*N + p];
               for (int s = 0; s < 3; s++) f->velocity[s*N + p] = f->velocity_profile[s*N + p];
               f->temperature[p] = f->temperature_profile[p];
               f->pressure[p] = Pbc;
            } else {
               for (int s = 0; s < ns; s++) f->MolarFracs[s*N + p] = f->MolarFracs[s*N + q];
               f->pressure[p] = f->pressure[q];
               f->temperature[p] = (kind == 1) ? f->temperature[q] : f->temperature_profile[p];
               for (int s = 0; s < 3; s++) f->velocity[s*N + p] = 0.0;
            }
         }
}

/* ------------------------------------------------------------------------
 * NSCBC (characteristic) flux corrections of the boundary cells.
 *   src/prometeo_rhs.inl:1288-1352  InflowMinusSide::addLODIfluxes
 *   src/prometeo_rhs.inl:1354-1420  InflowPlusSide::addLODIfluxes
 *   src/prometeo_rhs.inl:1424-1525  OutflowMinusSide::addLODIfluxes
 *   src/prometeo_rhs.inl:1527-1631  OutflowPlusSide::addLODIfluxes
 *   src/prometeo_rhs.cc:529-575, 680-730  task loops over the BC rect
 *   src/prometeo_metric.inl:89-152  getDeriv, getDerivLeftBC, getDerivRightBC
 *   src/prometeo_stat.rg:223-234    CalculateMaxMachNumber
 *   src/Utils/PointDomain_helper.hpp:76-134  tangential directions, p_int
 * The interior point of boundary cell p is p + 1 along the normal on a Minus
 * side, p - 1 on a Plus side.  Derivatives at a point wrap periodically in
 * the box, as warpPeriodic does over Fluid_bounds.
 * Pinned by tests/test_oracle_pins.py (closed forms of reduced states: a
 * uniform state at the far-field pressure leaves the RHS unchanged, a
 * pressure offset gives the relaxation wave, a supersonic inlet clears the
 * RHS, prescribed dT/dt and du/dt give the inlet waves, a normal pressure
 * gradient gives the outgoing acoustic wave; every side and direction).
 */
static inline double deriv_left(int t, double x, double xp1, double m)
{ return ((double)o_coeff_Grad(t, 1)*(xp1 - x))*m; }
static inline double deriv_right(int t, double xm1, double x, double m)
{ return ((double)o_coeff_Grad(t, 0)*(x - xm1))*m; }

/* getDeriv<d>(q, p, comp, nType_d[p], m_d_d[p], bounds) on a component field */
static double deriv_at(const o_grid *g, const o_fields *f, const double *q, int d, const int *c)
{
   const long p = L(g, c[0], c[1], c[2]);
   const int t = f->nType[d][p];
   const double m = f->m_d[d][p];
   int cm[3] = { c[0], c[1], c[2] }, cp[3] = { c[0], c[1], c[2] };
   const int n = d == 0 ? g->nx : d == 1 ? g->ny : g->nz;
   cm[d] = wrapm(cm[d], o_coeff_Cp(t, 1), n);
   cp[d] = wrapp(cp[d], o_coeff_Cp(t, 2), n);
   const double xm1 = q[L(g, cm[0], cm[1], cm[2])], x = q[p], xp1 = q[L(g, cp[0], cp[1], cp[2])];
   return ((double)o_coeff_Grad(t, 0)*(x - xm1) + (double)o_coeff_Grad(t, 1)*(xp1 - x))*m;
}

static inline double species_W(const o_mix *m, int s) { return m->constprop ? 8.3144598/m->R : m->sp[s].W; }

void o_nscbc_inflow(const o_grid *g, const o_fields *f, int side, const o_mix *m)
{
   const int ns = m->nSpec, neq = ns + 4;
   const int dir = side/2, minus = (side % 2 == 0);
   const long N = NN(g);
   const double *vN = f->velocity + dir*N;
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++) {
            int c[3] = { i, j, k };
            const long p = L(g, i, j, k);
            c[dir] += minus ? 1 : -1;
            const long q = L(g, c[0], c[1], c[2]);
            const double SoS = f->SoS[p], rho = f->rho[p], T = f->temperature[p];
            const double un = vN[p];
            if (minus ? (un >= SoS) : (un <= -SoS)) {
               /* supersonic inlet */
               for (int l = 0; l < neq; l++) f->Conserved_t[l*N + p] = 0.0;
               continue;
            }
            double Y[O_NSP_MAX], LS[O_NSP_MAX];
            for (int s = 0; s < ns; s++) Y[s] = f->MassFracs[s*N + p];
            const int t = f->nType[dir][p];
            const double md = f->m_d[dir][p];
            const double MixW = o_GetMolarWeightFromYi(m, Y);
            const double Cp = o_GetHeatCapacity(m, T, Y);
            const double lambda = un;
            double L1, LN;
            if (minus) {
               const double lambda_1 = un - SoS;
               const double dp_dn = deriv_left(t, f->pressure[p], f->pressure[q], md);
               const double du_dn = deriv_left(t, vN[p], vN[q], md);
               L1 = lambda_1*(dp_dn - rho*SoS*du_dn);
               for (int s = 0; s < ns; s++)
                  LS[s] = lambda*deriv_left(t, f->MassFracs[s*N + p], f->MassFracs[s*N + q], md);
               LN = L1 - 2*rho*SoS*f->dudtBoundary[dir*N + p];
            } else {
               const double lambda_N = un + SoS;
               const double dp_dn = deriv_right(t, f->pressure[q], f->pressure[p], md);
               const double du_dn = deriv_right(t, vN[q], vN[p], md);
               LN = lambda_N*(dp_dn + rho*SoS*du_dn);
               for (int s = 0; s < ns; s++)
                  LS[s] = lambda*deriv_right(t, f->MassFracs[s*N + q], f->MassFracs[s*N + p], md);
               L1 = LN + 2*rho*SoS*f->dudtBoundary[dir*N + p];
            }
            double L2 = f->dTdtBoundary[p]/T + (LN + L1)/(2*rho*Cp*T);
            for (int s = 0; s < ns; s++) L2 -= MixW/species_W(m, s)*LS[s];
            L2 *= -rho*SoS*SoS;
            const double d1 = (0.5*(L1 + LN) - L2)/(SoS*SoS);
            for (int s = 0; s < ns; s++) f->Conserved_t[s*N + p] -= (d1*Y[s] + rho*LS[s]);
         }
}

void o_nscbc_outflow(const o_grid *g, const o_fields *f, int side, double MaxMach, double LengthScale,
                     double PInf, const o_mix *m)
{
   const int ns = m->nSpec;
   const int dir = side/2, minus = (side % 2 == 0);
   const int iN = dir, iT1 = (dir == 0) ? 1 : 0, iT2 = (dir == 2) ? 1 : 2;
   const int irU = ns, irE = ns + 3;
   const long N = NN(g);
   const double *v[3] = { f->velocity, f->velocity + N, f->velocity + 2*N };
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++) {
            const int cb[3] = { i, j, k };
            int ci[3] = { i, j, k };
            const long p = L(g, i, j, k);
            ci[dir] += minus ? 1 : -1;
            const long q = L(g, ci[0], ci[1], ci[2]);
            const int t = f->nType[iN][p];
            const double md = f->m_d[iN][p];
            const double SoS = f->SoS[p], T = f->temperature[p], rho = f->rho[p], P = f->pressure[p];
            const double un = v[iN][p];
            double Y[O_NSP_MAX], LS[O_NSP_MAX];
            for (int s = 0; s < ns; s++) Y[s] = f->MassFracs[s*N + p];
            double dp_dn, dun_dn, duT1dn, duT2dn, drho_dn;
            if (minus) {
               dp_dn  = deriv_left(t, P, f->pressure[q], md);
               dun_dn = deriv_left(t, un, v[iN][q], md);
               duT1dn = deriv_left(t, v[iT1][p], v[iT1][q], md);
               duT2dn = deriv_left(t, v[iT2][p], v[iT2][q], md);
               drho_dn = deriv_left(t, rho, f->rho[q], md);
            } else {
               dp_dn  = deriv_right(t, f->pressure[q], P, md);
               dun_dn = deriv_right(t, v[iN][q], un, md);
               duT1dn = deriv_right(t, v[iT1][q], v[iT1][p], md);
               duT2dn = deriv_right(t, v[iT2][q], v[iT2][p], md);
               drho_dn = deriv_right(t, f->rho[q], rho, md);
            }
            const double lambda_1 = un - SoS;
            const double lambda   = un;
            const double lambda_s = minus ? dmin(un, 0.0) : dmax(un, 0.0);
            const double lambda_N = un + SoS;
            const double sigma = 0.25;
            const double K = (MaxMach < 0.99) ? sigma*(1.0 - MaxMach*MaxMach)*SoS/LengthScale
                                              : sigma*(SoS - (un*un)/SoS)/LengthScale;
            double L1, LN;
            if (minus) {
               L1 = lambda_1*(dp_dn - rho*SoS*dun_dn);
               LN = (lambda_N < 0) ? lambda_N*(dp_dn + rho*SoS*dun_dn) : K*(P - PInf);
            } else {
               L1 = (lambda_1 > 0) ? lambda_1*(dp_dn - rho*SoS*dun_dn) : K*(P - PInf);
               LN = lambda_N*(dp_dn + rho*SoS*dun_dn);
            }
            const double LM = lambda*(dp_dn - SoS*SoS*drho_dn);
            for (int s = 0; s < ns; s++)
               LS[s] = lambda_s*(minus ? deriv_left(t, Y[s], f->MassFracs[s*N + q], md)
                                       : deriv_right(t, f->MassFracs[s*N + q], Y[s], md));
            const double d1 = (0.5*(L1 + LN) - LM)/(SoS*SoS);
            double dM[3];
            dM[iN]  = (LN - L1)/(2*rho*SoS);
            dM[iT1] = lambda*duT1dn;
            dM[iT2] = lambda*duT2dn;
            const double dN = LM/(SoS*SoS);
            /* viscous terms */
            double vGradT1[3], vGradT2[3];
            for (int c = 0; c < 3; c++) {
               vGradT1[c] = deriv_at(g, f, v[c], iT1, cb);
               vGradT2[c] = deriv_at(g, f, v[c], iT2, cb);
            }
            const double tauNN_bnd = f->mu[p]*(4*dun_dn - 2*vGradT1[iT1] - 2*vGradT2[iT2])/3;
            const double tauNN_int = f->mu[q]*(4*deriv_at(g, f, v[iN], iN, ci)
                                               - 2*deriv_at(g, f, v[iT1], iT1, ci)
                                               - 2*deriv_at(g, f, v[iT2], iT2, ci))/3;
            double dtau_dn, vh;
            if (minus) {
               dtau_dn = deriv_left(t, tauNN_bnd, tauNN_int, md);
               vh = deriv_left(t, un*tauNN_bnd, v[iN][q]*tauNN_int, md);
            } else {
               dtau_dn = deriv_right(t, tauNN_int, tauNN_bnd, md);
               vh = deriv_right(t, v[iN][q]*tauNN_int, un*tauNN_bnd, md);
            }
            const double viscous_heating = vh + duT1dn*f->mu[p]*(duT1dn + vGradT1[iN]) +
                                           duT2dn*f->mu[p]*(duT2dn + vGradT2[iN]);
            /* RHS update */
            double *Ct = f->Conserved_t;
            const double *Cs = f->Conserved;
            const double cpT = o_GetHeatCapacity(m, T, Y)*T;
            for (int s = 0; s < ns; s++) Ct[s*N + p] -= (d1*Y[s] + rho*LS[s]);
            Ct[(irU+iN)*N + p]  -= (v[iN][p]*d1 + rho*dM[iN] - dtau_dn);
            Ct[(irU+iT1)*N + p] -= (v[iT1][p]*d1 + rho*dM[iT1]);
            Ct[(irU+iT2)*N + p] -= (v[iT2][p]*d1 + rho*dM[iT2]);
            Ct[irE*N + p] -= ((Cs[irE*N + p] + P)*d1/rho
                              + Cs[(irU+iN)*N + p]*dM[iN]
                              + Cs[(irU+iT1)*N + p]*dM[iT1]
                              + Cs[(irU+iT2)*N + p]*dM[iT2]
                              + cpT*dN
                              - viscous_heating);
            const double cpT_Wmix = cpT*o_GetMolarWeightFromYi(m, Y);
            for (int s = 0; s < ns; s++) {
               const double dEdYi = o_GetSpeciesEnthalpy(m, s, T) - cpT_Wmix/species_W(m, s);
               Ct[irE*N + p] -= (rho*LS[s]*dEdYi);
            }
         }
}

/* RHS update of the far-field LODI fluxes (no viscous terms),
 * src/prometeo_rhs.inl:1713-1735, 1759-1779 */
static void lodi_update(const o_fields *f, long N, long p, const o_mix *m, int iN, int iT1, int iT2,
                        const double *Y, const double *LS, const double *u, double rho, double P,
                        double T, double cpT, double MixW, double d1, const double *dM, double dN)
{
   const int ns = m->nSpec, irU = ns, irE = ns + 3;
   double *Ct = f->Conserved_t;
   const double *Cs = f->Conserved;
   for (int s = 0; s < ns; s++) Ct[s*N + p] -= (d1*Y[s] + rho*LS[s]);
   Ct[(irU+iN)*N + p]  -= (u[iN]*d1 + rho*dM[iN]);
   Ct[(irU+iT1)*N + p] -= (u[iT1]*d1 + rho*dM[iT1]);
   Ct[(irU+iT2)*N + p] -= (u[iT2]*d1 + rho*dM[iT2]);
   Ct[irE*N + p] -= ((Cs[irE*N + p] + P)*d1/rho
                     + Cs[(irU+iN)*N + p]*dM[iN]
                     + Cs[(irU+iT1)*N + p]*dM[iT1]
                     + Cs[(irU+iT2)*N + p]*dM[iT2]
                     + cpT*dN);
   const double cpT_MixW = cpT*MixW;
   for (int s = 0; s < ns; s++) {
      const double dEdYi = o_GetSpeciesEnthalpy(m, s, T) - cpT_MixW/species_W(m, s);
      Ct[irE*N + p] -= (rho*LS[s]*dEdYi);
   }
}

void o_nscbc_farfield(const o_grid *g, const o_fields *f, int side, double PInf, double MaxMach,
                      double LengthScale, const o_mix *m)
{
   const int ns = m->nSpec;
   const int dir = side/2, minus = (side % 2 == 0);
   const int iN = dir, iT1 = (dir == 0) ? 1 : 0, iT2 = (dir == 2) ? 1 : 2;
   const int irU = ns, irE = ns + 3;
   const long N = NN(g);
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++) {
            int ci[3] = { i, j, k };
            const long p = L(g, i, j, k);
            ci[dir] += minus ? 1 : -1;
            const long q = L(g, ci[0], ci[1], ci[2]);
            const int t = f->nType[iN][p];
            const double md = f->m_d[iN][p];
            const double SoS = f->SoS[p], T = f->temperature[p], rho = f->rho[p], P = f->pressure[p];
            double u[3], vInf[3], Y[O_NSP_MAX], LS[O_NSP_MAX];
            for (int c = 0; c < 3; c++) { u[c] = f->velocity[c*N + p]; vInf[c] = f->velocity_profile[c*N + p]; }
            for (int s = 0; s < ns; s++) Y[s] = f->MassFracs[s*N + p];
            const double TInf = f->temperature_profile[p];
            const double un = u[iN];
            /* characteristic velocities (entropy and species are not extrapolated) */
            const double lambda_1 = un - SoS;
            const double lambda   = un;
            const double lambda_N = un + SoS;
            double dp_dn, dun_dn;
            if (minus) {
               dp_dn  = deriv_left(t, P, f->pressure[q], md);
               dun_dn = deriv_left(t, un, f->velocity[iN*N + q], md);
            } else {
               dp_dn  = deriv_right(t, f->pressure[q], P, md);
               dun_dn = deriv_right(t, f->velocity[iN*N + q], un, md);
            }
            const double MixW = o_GetMolarWeightFromYi(m, Y);
            const double cpT = o_GetHeatCapacity(m, T, Y)*T;
            const double MaInflow = minus ? dmax(lambda/SoS, 0.0) : dmax(-lambda/SoS, 0.0);
            const double sigma = 0.25;
            double *Ct = f->Conserved_t;
            const double *Cs = f->Conserved;
            if (MaInflow > 0) {
               const double coeff = 39;
               const double Kp = sigma*(1 + coeff*MaInflow*MaInflow)*SoS/LengthScale;
               if (MaInflow > 1) {
                  /* supersonic inflow: relax every conserved variable to the far field */
                  const double rhoTarget = o_GetRho(m, PInf, TInf, MixW);
                  for (int s = 0; s < ns; s++)
                     Ct[s*N + p] += Kp*(rhoTarget*f->MolarFracs_profile[s*N + p]*species_W(m, s)/MixW - Cs[s*N + p]);
                  for (int c = 0; c < 3; c++)
                     Ct[(irU+c)*N + p] += Kp*(rhoTarget*vInf[c] - Cs[(irU+c)*N + p]);
                  double v2 = 0.0;
                  v2 += vInf[0]*vInf[0]; v2 += vInf[1]*vInf[1]; v2 += vInf[2]*vInf[2];
                  Ct[irE*N + p] += Kp*(rhoTarget*(0.5*v2 + o_GetInternalEnergy(m, TInf, Y)) - Cs[irE*N + p]);
               } else {
                  /* subsonic inflow */
                  const double K = sigma*(coeff + 1)*MaInflow*MaInflow*SoS/LengthScale;
                  double L1, LN;
                  if (minus) {
                     L1 = lambda_1*(dp_dn - rho*SoS*dun_dn);
                     LN = Kp*(P - PInf);
                  } else {
                     L1 = Kp*(P - PInf);
                     LN = lambda_N*(dp_dn + rho*SoS*dun_dn);
                  }
                  for (int s = 0; s < ns; s++)
                     LS[s] = K*(Y[s] - f->MolarFracs_profile[s*N + p]*species_W(m, s)/MixW);
                  double L2 = K*(TInf/T - 1) + (LN + L1)/(2*rho*cpT);
                  for (int s = 0; s < ns; s++) L2 -= MixW/species_W(m, s)*LS[s];
                  L2 *= -rho*SoS*SoS;
                  const double d1 = (0.5*(L1 + LN) - L2)/(SoS*SoS);
                  double dM[3];
                  dM[iN]  = (LN - L1)/(2*rho*SoS);
                  dM[iT1] = K*(u[iT1] - vInf[iT1]);
                  dM[iT2] = K*(u[iT2] - vInf[iT2]);
                  const double dN = L2/(SoS*SoS);
                  lodi_update(f, N, p, m, iN, iT1, iT2, Y, LS, u, rho, P, T, cpT, MixW, d1, dM, dN);
               }
            } else {
               /* outflow */
               const double Kout = (MaxMach < 0.99) ? sigma*(1.0 - MaxMach*MaxMach)*SoS/LengthScale
                                                    : sigma*(SoS - (un*un)/SoS)/LengthScale;
               double L1, L2, LN, duT1, duT2;
               if (minus) {
                  L1 = lambda_1*(dp_dn - rho*SoS*dun_dn);
                  L2 = lambda*(dp_dn - SoS*SoS*deriv_left(t, rho, f->rho[q], md));
                  for (int s = 0; s < ns; s++)
                     LS[s] = lambda*deriv_left(t, Y[s], f->MassFracs[s*N + q], md);
                  LN = (lambda_N < 0) ? lambda_N*(dp_dn + rho*SoS*dun_dn) : Kout*(P - PInf);
                  duT1 = deriv_left(t, u[iT1], f->velocity[iT1*N + q], md);
                  duT2 = deriv_left(t, u[iT2], f->velocity[iT2*N + q], md);
               } else {
                  L1 = (lambda_1 > 0) ? lambda_1*(dp_dn - rho*SoS*dun_dn) : Kout*(P - PInf);
                  L2 = lambda*(dp_dn - SoS*SoS*deriv_right(t, f->rho[q], rho, md));
                  for (int s = 0; s < ns; s++)
                     LS[s] = lambda*deriv_right(t, f->MassFracs[s*N + q], Y[s], md);
                  LN = lambda_N*(dp_dn + rho*SoS*dun_dn);
                  duT1 = deriv_right(t, f->velocity[iT1*N + q], u[iT1], md);
                  duT2 = deriv_right(t, f->velocity[iT2*N + q], u[iT2], md);
               }
               const double d1 = (0.5*(L1 + LN) - L2)/(SoS*SoS);
               double dM[3];
               dM[iN]  = (LN - L1)/(2*rho*SoS);
               dM[iT1] = lambda*duT1;
               dM[iT2] = lambda*duT2;
               const double dN = L2/(SoS*SoS);
               lodi_update(f, N, p, m, iN, iT1, iT2, Y, LS, u, rho, P, T, cpT, MixW, d1, dM, dN);
            }
         }
}

double o_max_mach(const o_grid *g, const o_fields *f, int dir)
{
   const long N = NN(g);
   double acc = -HUGE_VAL;
   for (int k = g->lo[2]; k <= g->hi[2]; k++)
      for (int j = g->lo[1]; j <= g->hi[1]; j++)
         for (int i = g->lo[0]; i <= g->hi[0]; i++) {
            const long p = L(g, i, j, k);
            acc = dmax(acc, fabs(f->velocity[dir*N + p])/f->SoS[p]);
         }
   return acc;
}

/* threads of the OpenMP loops (the CPU baseline's one-core reading) */
int o_set_threads(int n)
{
#ifdef _OPENMP
   if (n > 0) omp_set_num_threads(n);
   return omp_get_max_threads();
#else
   (void)n;
   return 1;
#endif
}
