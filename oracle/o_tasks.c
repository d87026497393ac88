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

typedef struct {
   int n;
   rhs_fn rhs;
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
   errmax *= (1.0/RB_TOL);
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
      rb_jacobian(s, x);
      s->rhs(s->ctx, s->dx, x);
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
   s->n = 3; s->rhs = testprob_rhs; s->ctx = 0;
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

int o_chem_cell(const o_mix *m, double *x, const double *Ct_old, double *temperature, double dt)
{
   rb_t *s = (rb_t*)calloc(1, sizeof(rb_t));
   chem_ctx c = { m, Ct_old, temperature };
   s->n = m->nSpec + 4; s->rhs = chem_rhs; s->ctx = &c;
   const int st = rb_solve(s, x, dt, dt);
   free(s);
   return st;
}

void o_chem_update(const o_grid *g, const o_fields *f, double dt, const o_mix *m, int32_t *nsteps)
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
            const int st = o_chem_cell(m, x, old, &T, dt);
            for (int s = 0; s < neq; s++) f->Conserved[s*N + p] = x[s];
            f->temperature[p] = T;
            if (nsteps) nsteps[p] = st;
         }
}

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
 * SetAdiabaticWallBC, SetIsothermalWallBC.  The interior neighbour of a
 * ghost is c_int = c + 1 along the normal for a Neg side, c - 1 for a Pos
 * side (mk_cint of the reference).
 * Pinned by tests/test_oracle_pins.py (wall and Dirichlet ghost values,
 * zero mass flux through a wall).
 */
void o_set_bc(const o_grid *g, const o_fields *f, int kind, int side, double Pbc, const o_mix *m)
{
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
