// pointwise.hip -- the cell-local tasks of the hot path and the spectral
// radius reduction.
//
//   UpdatePropertiesFromPrimitive / UpdateConservedFromPrimitive /
//   UpdatePrimitiveFromConserved   src/prometeo_variables.hpp:62-192
//   UpdateDucrosSensor             src/prometeo_sensor.inl:52-77, sensor.cc:37-80
//   UpdateShockSensor<dir>         src/prometeo_sensor.cc:82-136
//   AddChemistrySources            src/prometeo_chem.cc:78-112
//   CalculateMaxSpectralRadius     src/prometeo_cfl.hpp:73-178 (wave64 butterfly
//                                  max, one partial per block, one-block tail)
//   RK stage update                src/prometeo_rk.rg:48-112
#include "common.hpp"
#include "kernels.hpp"

namespace htr {

#define HTR_POINT_LOOP(g, n) \
   const int64_t tid = (int64_t)blockIdx.x*blockDim.x + threadIdx.x; \
   if (tid >= (n)) return; \
   const P3 p = exec_point(g, tid);

// ------------------------------------------------------------------ variables
// Per-cell failure flags (the optional htr_diag.flags view)
constexpr uint32_t FLAG_NAN = 1u;       // NaN in p, T or Conserved (DetectNaN, src/prometeo.rg:361-376)
constexpr uint32_t FLAG_NEWTON = 2u;    // temperature Newton stopped at its iteration cap
constexpr uint32_t FLAG_NEG_RHOY = 4u;  // a partial density was negative before the clipping of Y

// UpdatePropertiesFromPrimitive of one cell (variables.hpp:62-111) from P, T, X
template<class MX>
__device__ __forceinline__ void properties_to(const htr_fields &f, const DMix &m, const P3 &p, double P, double T,
                                              const double *X)
{
   constexpr int NS = MX::NS;
   double Y[NS], D[NS];
   const double MixW = MX::MolarWeightFromX(m, X);
   at<double>(f.rho, p) = MX::Rho(m, P, T, MixW);
   at<double>(f.mu, p) = MX::Viscosity(m, T, X);
   at<double>(f.lam, p) = MX::HeatConductivity(m, T, X);
   MX::Diffusivity(m, D, P, T, MixW, X);
   MX::MassFractions(m, Y, MixW, X);
#pragma unroll
   for (int i = 0; i < NS; i++) { at<double>(f.Di, p, i) = D[i]; at<double>(f.MassFracs, p, i) = Y[i]; }
   const double gamma = MX::Gamma(m, T, MixW, Y);
   at<double>(f.SoS, p) = MX::SpeedOfSound(m, T, gamma, MixW);
}

template<class MX>
struct CellPrim { double T, P, X[MX::NS], v[3]; };

// UpdatePrimitiveFromConserved of one cell (variables.hpp:160-192) from the
// conserved vector q and the temperature guess; returns failure flags
template<class MX>
__device__ __forceinline__ uint32_t primitive_of(const DMix &m, const double *q, double Tguess, CellPrim<MX> &o)
{
   constexpr int NS = MX::NS;
   uint32_t flags = 0;
   double Yi[NS];
   double rho;
   if (MX::CP) {
      rho = q[0];
      Yi[0] = q[0]/rho;
   } else {
      rho = 0.0;
#pragma unroll
      for (int i = 0; i < NS; i++) { rho += q[i]; if (q[i] < 0.0) flags |= FLAG_NEG_RHOY; }
      const double irho = 1.0/rho;
#pragma unroll
      for (int i = 0; i < NS; i++) { Yi[i] = irho*q[i]; Yi[i] = dmax(Yi[i], 1e-60); Yi[i] = dmin(Yi[i], 1.0); }
   }
   const double MixW = MX::MolarWeightFromY(m, Yi);
   MX::MolarFractions(m, o.X, MixW, Yi);
   const double rhoInv = 1.0/rho;
#pragma unroll
   for (int i = 0; i < 3; i++) o.v[i] = q[NS+i]*rhoInv;
   double k2 = 0.0; k2 += o.v[0]*o.v[0]; k2 += o.v[1]*o.v[1]; k2 += o.v[2]*o.v[2];
   const double e = q[NS+3]*rhoInv - 0.5*k2;
   int iters = 0;
   o.T = MX::TFromInternalEnergy(m, e, Tguess, Yi, &iters);
   if (!MX::CP && iters >= 1000) flags |= FLAG_NEWTON;
   o.P = MX::PFromRhoAndT(m, rho, MixW, o.T);
   return flags;
}

template<class MX>
__device__ __forceinline__ void store_primitive(const htr_fields &f, const P3 &p, const CellPrim<MX> &o)
{
   at<double>(f.temperature, p) = o.T;
   at<double>(f.pressure, p) = o.P;
#pragma unroll
   for (int i = 0; i < MX::NS; i++) at<double>(f.MolarFracs, p, i) = o.X[i];
#pragma unroll
   for (int i = 0; i < 3; i++) at<double>(f.velocity, p, i) = o.v[i];
}

template<class MX>
__global__ void __launch_bounds__(256)
properties_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, int64_t n)
{
   HTR_POINT_LOOP(g, n)
   constexpr int NS = MX::NS;
   double X[NS];
#pragma unroll
   for (int i = 0; i < NS; i++) X[i] = ld<double>(f.MolarFracs, p, i);
   properties_to<MX>(f, *mix, p, ld<double>(f.pressure, p), ld<double>(f.temperature, p), X);
}

template<class MX>
__global__ void __launch_bounds__(256)
conserved_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, int64_t n)
{
   HTR_POINT_LOOP(g, n)
   constexpr int NS = MX::NS;
   const DMix &m = *mix;
   double Y[NS], v[3];
   const double rho = ld<double>(f.rho, p);
#pragma unroll
   for (int i = 0; i < NS; i++) { Y[i] = ld<double>(f.MassFracs, p, i); at<double>(f.Conserved, p, i) = rho*Y[i]; }
#pragma unroll
   for (int i = 0; i < 3; i++) { v[i] = ld<double>(f.velocity, p, i); at<double>(f.Conserved, p, NS+i) = rho*v[i]; }
   double k2 = 0.0; k2 += v[0]*v[0]; k2 += v[1]*v[1]; k2 += v[2]*v[2];
   at<double>(f.Conserved, p, NS+3) = rho*(0.5*k2 + MX::InternalEnergy(m, ld<double>(f.temperature, p), Y));
}

template<class MX>
__global__ void __launch_bounds__(256)
primitive_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, int64_t n)
{
   HTR_POINT_LOOP(g, n)
   double q[MX::NEQ];
#pragma unroll
   for (int i = 0; i < MX::NEQ; i++) q[i] = ld<double>(f.Conserved, p, i);
   CellPrim<MX> o;
   primitive_of<MX>(*mix, q, ld<double>(f.temperature, p), o);
   store_primitive<MX>(f, p, o);
}

// One SSP-RK3 stage of a cell in one pass (src/prometeo_rk.rg:88-112, then
// UpdatePrimitiveFromConserved and UpdatePropertiesFromPrimitive of
// UpdatePrimitivesAndBCsFromConserved, src/prometeo.rg:428-478, without the
// boundary tasks): Conserved = c0 Conserved_old + c1 Conserved + c2
// Conserved_t dt, then primitives and properties from the updated vector held
// in registers -- the same operations as the three separate tasks.
template<class MX>
__global__ void __launch_bounds__(256)
rk_stage_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, double c0, double c1, double c2,
                DtArg dta, htr_view flags, int64_t n)
{
   HTR_POINT_LOOP(g, n)
   constexpr int NEQ = MX::NEQ;
   const double dt = dta.get();
   double q[NEQ];
   bool bad = false;
#pragma unroll
   for (int i = 0; i < NEQ; i++) {
      q[i] = c0*ld<double>(f.Conserved_old, p, i) + c1*ld<double>(f.Conserved, p, i) + c2*ld<double>(f.Conserved_t, p, i)*dt;
      at<double>(f.Conserved, p, i) = q[i];
      bad |= isnan(q[i]);
   }
   CellPrim<MX> o;
   uint32_t fl = primitive_of<MX>(*mix, q, ld<double>(f.temperature, p), o);
   store_primitive<MX>(f, p, o);
   properties_to<MX>(f, *mix, p, o.P, o.T, o.X);
   if (bad || isnan(o.P) || isnan(o.T)) fl |= FLAG_NAN;
   if (flags.ptr) at<uint32_t>(flags, p) = fl;
}

// -------------------------------------------------------------------- sensors
template<int DD>
__device__ inline void vgrad(const htr_fields &f, const Geo &g, const P3 &p, double *out)
{
   const int t = ld<int>(f.nType[DD], p);
   const double mm = ld<double>(f.m_d[DD], p);
   const P3 a = wminus(g, DD, p, c_st.Cp[t][1]);
   const P3 b = wplus(g, DD, p, c_st.Cp[t][2]);
   const double g0 = (double)c_st.Grad[t][0], g1 = (double)c_st.Grad[t][1];
#pragma unroll
   for (int i = 0; i < 3; i++) {
      const double x = ld<double>(f.velocity, p, i);
      out[i] = (g0*(x - ld<double>(f.velocity, a, i)) + g1*(ld<double>(f.velocity, b, i) - x))*mm;
   }
}

__global__ void __launch_bounds__(256)
ducros_kernel(const htr_fields f, const Geo g, double eps, int64_t n)
{
   HTR_POINT_LOOP(g, n)
   double gx[3], gy[3], gz[3];
   vgrad<0>(f, g, p, gx);
   vgrad<1>(f, g, p, gy);
   vgrad<2>(f, g, p, gz);
   const double div = gx[0] + gy[1] + gz[2];
   const double div2 = div*div;
   const double omz = gx[1] - gy[0];
   const double omy = gy[2] - gz[1];
   const double omx = gz[0] - gx[2];
   const double om2 = omx*omx + omy*omy + omz*omz;
   at<double>(f.DucrosSensor, p) = div2/(div2 + om2 + eps);
}

template<class MX, int DIR>
__global__ void __launch_bounds__(256)
shock_sensor_kernel(const htr_fields f, const Geo g, int64_t n)
{
   HTR_POINT_LOOP(g, n)
   const int t = ld<int>(f.nType[DIR], p);
   P3 q[6];
   q[0] = wminus(g, DIR, p, c_st.Cp[t][0]);
   q[1] = wminus(g, DIR, p, c_st.Cp[t][1]);
   q[2] = p;
   q[3] = wplus(g, DIR, p, c_st.Cp[t][2]);
   q[4] = wplus(g, DIR, p, c_st.Cp[t][3]);
   q[5] = wplus(g, DIR, p, c_st.Cp[t][4]);
   double Phi = ld<double>(f.DucrosSensor, q[0]);
#pragma unroll
   for (int k = 1; k < 6; k++) Phi = dmax(Phi, ld<double>(f.DucrosSensor, q[k]));
   bool sensor = true;
   for (int h = 0; h < MX::NS && sensor; h++) {
      double y[6];
#pragma unroll
      for (int k = 0; k < 6; k++) y[k] = ld<double>(f.Conserved, q[k], h);
      sensor = tenoa_sensor(y, t, Phi);
   }
   at<uint8_t>(f.shockSensor[DIR], p) = sensor ? 1 : 0;
}

// ------------------------------------------------------------------ chemistry
template<class MX>
__global__ void __launch_bounds__(256)
chem_source_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, int zero, int64_t n)
{
   HTR_POINT_LOOP(g, n)
   constexpr int NS = MX::NS;
   double Y[NS], w[NS];
#pragma unroll
   for (int i = 0; i < NS; i++) Y[i] = ld<double>(f.MassFracs, p, i);
   const double rho = ld<double>(f.rho, p), P = ld<double>(f.pressure, p), T = ld<double>(f.temperature, p);
   if constexpr (!MX::CP && NS <= 8) {
      // rates, concentrations and Gibbs energies in this thread's LDS columns
      // (3 NS 2 KiB per block)
      __shared__ double sp[3*NS*256];
      const SpeciesCol W{sp + threadIdx.x, 256}, C{sp + NS*256 + threadIdx.x, 256}, G{sp + 2*NS*256 + threadIdx.x, 256};
      MX::ProductionRatesIn(*mix, W, C, G, rho, P, T, Y);
#pragma unroll
      for (int i = 0; i < NS; i++) w[i] = W[i];
   } else {
      MX::ProductionRates(*mix, w, rho, P, T, Y);
   }
   if (zero) {
      // first term of the stage's RHS: Conserved_t = 0 + source, the other
      // components cleared (the zeroing of InitializeTimeDerivatives fused)
#pragma unroll
      for (int i = 0; i < NS; i++) at<double>(f.Conserved_t, p, i) = 0.0 + w[i];
#pragma unroll
      for (int i = NS; i < MX::NEQ; i++) at<double>(f.Conserved_t, p, i) = 0.0;
   } else {
#pragma unroll
      for (int i = 0; i < NS; i++) at<double>(f.Conserved_t, p, i) += w[i];
   }
}

// ----------------------------------------------------------- spectral radius
__device__ inline double wave_max(double v)
{
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o, 64));
   return v;
}

template<class MX>
__global__ void __launch_bounds__(256)
spectral_radius_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, int64_t n,
                       double *__restrict__ partial)
{
   const int64_t tid = (int64_t)blockIdx.x*blockDim.x + threadIdx.x;
   double r = 0.0;
   const int64_t stride = (int64_t)gridDim.x*blockDim.x;
   for (int64_t i = tid; i < n; i += stride) {
      const P3 p = exec_point(g, i);
      const double dcsi = ld<double>(f.m_d[0], p), deta = ld<double>(f.m_d[1], p), dzet = ld<double>(f.m_d[2], p);
      const double SoS = ld<double>(f.SoS, p);
      double rr = (fabs(ld<double>(f.velocity, p, 0)) + SoS)*dcsi;
      rr = dmax(rr, (fabs(ld<double>(f.velocity, p, 1)) + SoS)*deta);
      rr = dmax(rr, (fabs(ld<double>(f.velocity, p, 2)) + SoS)*dzet);
      const double rho = ld<double>(f.rho, p);
      const double nu = ld<double>(f.mu, p)/rho;
      double rv = nu*dcsi*dcsi;
      rv = dmax(rv, nu*deta*deta);
      rv = dmax(rv, nu*dzet*dzet);
      rv *= 4;
      rr = dmax(rr, rv);
      double Y[MX::NS];
#pragma unroll
      for (int s = 0; s < MX::NS; s++) Y[s] = ld<double>(f.MassFracs, p, s);
      const double cp = MX::HeatCapacity(*mix, ld<double>(f.temperature, p), Y);
      const double DifT = ld<double>(f.lam, p)/(cp*rho);
      double rt = DifT*dcsi*dcsi;
      rt = dmax(rt, DifT*deta*deta);
      rt = dmax(rt, DifT*dzet*dzet);
      rt *= 4;
      rr = dmax(rr, rt);
      double rs = 0.0;
#pragma unroll
      for (int s = 0; s < MX::NS; s++) {
         const double D = ld<double>(f.Di, p, s);
         rs = dmax(rs, D*dcsi*dcsi);
         rs = dmax(rs, D*deta*deta);
         rs = dmax(rs, D*dzet*dzet);
      }
      rs *= 4;
      rr = dmax(rr, rs);
      if (rr > r) r = rr;
   }
   __shared__ double sm[4];
   r = wave_max(r);
   if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = r;
   __syncthreads();
   if (threadIdx.x == 0) partial[blockIdx.x] = dmax(dmax(sm[0], sm[1]), dmax(sm[2], sm[3]));
}

__global__ void __launch_bounds__(256)
reduce_max_kernel(const double *__restrict__ partial, int n, double *__restrict__ out)
{
   double r = 0.0;
   for (int i = threadIdx.x; i < n; i += blockDim.x) r = dmax(r, partial[i]);
   __shared__ double sm[4];
   r = wave_max(r);
   if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = r;
   __syncthreads();
   if (threadIdx.x == 0) *out = dmax(dmax(sm[0], sm[1]), dmax(sm[2], sm[3]));
}

// CalculateMaxMachNumber (src/prometeo_stat.rg:223-234): max |u_dir|/SoS,
// grid-stride, one partial per block (the Mach number is >= 0, so the block
// maxima start from 0)
__global__ void __launch_bounds__(256)
max_mach_kernel(const htr_fields f, const Geo g, int dir, int64_t n, double *__restrict__ partial)
{
   const int64_t tid = (int64_t)blockIdx.x*blockDim.x + threadIdx.x;
   const int64_t stride = (int64_t)gridDim.x*blockDim.x;
   double r = 0.0;
   for (int64_t i = tid; i < n; i += stride) {
      const P3 p = exec_point(g, i);
      r = dmax(r, fabs(ld<double>(f.velocity, p, dir))/ld<double>(f.SoS, p));
   }
   __shared__ double sm[4];
   r = wave_max(r);
   if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = r;
   __syncthreads();
   if (threadIdx.x == 0) partial[blockIdx.x] = dmax(dmax(sm[0], sm[1]), dmax(sm[2], sm[3]));
}

// ------------------------------------------------------------------------ RK
__global__ void __launch_bounds__(256)
rk_update_kernel(const htr_fields f, const Geo g, int nEq, double c0, double c1, double c2,
                 double dt, int64_t n)
{
   HTR_POINT_LOOP(g, n)
   for (int i = 0; i < nEq; i++)
      at<double>(f.Conserved, p, i) = c0*ld<double>(f.Conserved_old, p, i) + c1*ld<double>(f.Conserved, p, i)
                                    + c2*ld<double>(f.Conserved_t, p, i)*dt;
}

__global__ void __launch_bounds__(256)
rk_init_kernel(const htr_fields f, const Geo g, int nEq, int64_t n)
{
   HTR_POINT_LOOP(g, n)
   for (int i = 0; i < nEq; i++) at<double>(f.Conserved_old, p, i) = ld<double>(f.Conserved, p, i);
}

__global__ void __launch_bounds__(256)
zero_rhs_kernel(const htr_fields f, const Geo g, int nEq, int64_t n)
{
   HTR_POINT_LOOP(g, n)
   for (int i = 0; i < nEq; i++) at<double>(f.Conserved_t, p, i) = 0.0;
}

// clears the device counters (a kernel rather than a memset, so the call
// sequence stays valid inside a captured graph)
__global__ void zero_counters_kernel(int *counter)
{
   if (threadIdx.x < 4) counter[threadIdx.x] = 0;
}

int zero_counters(int *counter, hipStream_t s)
{
   HTR_LAUNCH("zero_counters", zero_counters_kernel, dim3(1), dim3(64), 0, s, counter);
   return check_launch("zero_counters_kernel");
}

// -------------------------------------------------------- boundary conditions
// Ghost-cell primitives of one boundary (src/prometeo_bc.rg:1159-1265):
// kind 0 SetDirichletBC, 1 SetAdiabaticWallBC, 2 SetIsothermalWallBC,
// 3 SetNSCBC_InflowBC (src/prometeo_bc.cc:82-136; kind 4, SetNSCBC_OutflowBC,
// is the primitive recompute over the rect).  The interior neighbour q of
// ghost p is one cell along dir, inward (+1 for a Neg side, -1 for a Pos
// side); the host checks that it lies inside Fluid_bounds.
template<class MX>
__global__ void __launch_bounds__(256)
bc_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, int kind, int dir, int inward, double Pbc,
          int64_t n)
{
   HTR_POINT_LOOP(g, n)
   constexpr int NS = MX::NS;
   if (kind == 3) {
      double X[NS];
#pragma unroll
      for (int i = 0; i < NS; i++) { X[i] = ld<double>(f.MolarFracs_profile, p, i); at<double>(f.MolarFracs, p, i) = X[i]; }
      const double Tp = ld<double>(f.temperature_profile, p);
      at<double>(f.temperature, p) = Tp;
#pragma unroll
      for (int i = 0; i < 3; i++) at<double>(f.velocity, p, i) = ld<double>(f.velocity_profile, p, i);
      if (fabs(ld<double>(f.velocity_profile, p, dir)) >= ld<double>(f.SoS, p)) {
         at<double>(f.pressure, p) = Pbc;           // supersonic: everything imposed
      } else {
         // pressure of the transported density at the imposed T and X
         double rho = 0.0;
#pragma unroll
         for (int i = 0; i < NS; i++) rho += ld<double>(f.Conserved, p, i);
         at<double>(f.pressure, p) = MX::PFromRhoAndT(*mix, rho, MX::MolarWeightFromX(*mix, X), Tp);
      }
      return;
   }
   P3 q = p;
   setc(g, q, dir, p.c[dir] + inward);
   if (kind == 0) {
#pragma unroll
      for (int i = 0; i < NS; i++) at<double>(f.MolarFracs, p, i) = ld<double>(f.MolarFracs_profile, p, i);
#pragma unroll
      for (int i = 0; i < 3; i++) at<double>(f.velocity, p, i) = ld<double>(f.velocity_profile, p, i);
      at<double>(f.temperature, p) = ld<double>(f.temperature_profile, p);
      at<double>(f.pressure, p) = Pbc;
      return;
   }
#pragma unroll
   for (int i = 0; i < NS; i++) at<double>(f.MolarFracs, p, i) = ld<double>(f.MolarFracs, q, i);
   at<double>(f.pressure, p) = ld<double>(f.pressure, q);
   at<double>(f.temperature, p) = (kind == 1) ? ld<double>(f.temperature, q) : ld<double>(f.temperature_profile, p);
#pragma unroll
   for (int i = 0; i < 3; i++) at<double>(f.velocity, p, i) = 0.0;
}

// ------------------------------------------------------------------ launchers
#define HTR_LAUNCH_POINT(kern, ...)                                                       \
   const int64_t n = exec_cells(g);                                                       \
   if (n == 0) return HTR_OK;                                                             \
   HTR_LAUNCH(#kern, kern, dim3(grid_for(n, 256)), dim3(256), 0, s, __VA_ARGS__, n);     \
   return check_launch(#kern);

template<class MX> int properties_launch(const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s)
{ HTR_LAUNCH_POINT(properties_kernel<MX>, f, mix, g) }
template<class MX> int conserved_launch(const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s)
{ HTR_LAUNCH_POINT(conserved_kernel<MX>, f, mix, g) }
template<class MX> int primitive_launch(const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s)
{ HTR_LAUNCH_POINT(primitive_kernel<MX>, f, mix, g) }
template<class MX> int chem_source_launch(const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s, bool zero)
{ HTR_LAUNCH_POINT(chem_source_kernel<MX>, f, mix, g, zero ? 1 : 0) }

int ducros_launch(const htr_fields &f, const Geo &g, double vorticityScale, hipStream_t s)
{
   const double eps = dmax(vorticityScale*vorticityScale, 1e-6);
   HTR_LAUNCH_POINT(ducros_kernel, f, g, eps)
}

template<class MX> int shock_sensor_launch(int dir, const htr_fields &f, const Geo &g, hipStream_t s)
{
   if (dir == 0) { HTR_LAUNCH_POINT((shock_sensor_kernel<MX, 0>), f, g) }
   if (dir == 1) { HTR_LAUNCH_POINT((shock_sensor_kernel<MX, 1>), f, g) }
   if (dir == 2) { HTR_LAUNCH_POINT((shock_sensor_kernel<MX, 2>), f, g) }
   return set_error(HTR_E_ARG, "shock sensor: bad direction %d", dir);
}

int rk_update_launch(const htr_fields &f, const Geo &g, int nEq, int stage, double dt, hipStream_t s)
{
   static const double C[3][3] = { { 1.0, 0.0, 1.0 }, { 3.0/4.0, 1.0/4.0, 1.0/4.0 }, { 1.0/3.0, 2.0/3.0, 2.0/3.0 } };
   if (stage < 0 || stage > 2) return set_error(HTR_E_ARG, "rk: bad stage %d", stage);
   HTR_LAUNCH_POINT(rk_update_kernel, f, g, nEq, C[stage][0], C[stage][1], C[stage][2], dt)
}

template<class MX>
int rk_stage_launch(const htr_fields &f, const DMix *mix, const Geo &g, int stage, const DtArg &dt,
                    const htr_view &flags, hipStream_t s)
{
   static const double C[3][3] = { { 1.0, 0.0, 1.0 }, { 3.0/4.0, 1.0/4.0, 1.0/4.0 }, { 1.0/3.0, 2.0/3.0, 2.0/3.0 } };
   if (stage < 0 || stage > 2) return set_error(HTR_E_ARG, "rk: bad stage %d", stage);
   HTR_LAUNCH_POINT(rk_stage_kernel<MX>, f, mix, g, C[stage][0], C[stage][1], C[stage][2], dt, flags)
}

int rk_init_launch(const htr_fields &f, const Geo &g, int nEq, hipStream_t s)
{ HTR_LAUNCH_POINT(rk_init_kernel, f, g, nEq) }

template<class MX>
int bc_launch(const htr_fields &f, const DMix *mix, const Geo &g, int kind, int dir, int inward, double Pbc,
              hipStream_t s)
{
   if (kind == 4) return primitive_launch<MX>(f, mix, g, s);
   HTR_LAUNCH_POINT(bc_kernel<MX>, f, mix, g, kind, dir, inward, Pbc)
}

int zero_rhs_launch(const htr_fields &f, const Geo &g, int nEq, hipStream_t s)
{ HTR_LAUNCH_POINT(zero_rhs_kernel, f, g, nEq) }

template<class MX>
int spectral_radius_launch(const htr_fields &f, const DMix *mix, const Geo &g, double *result, hipStream_t s)
{
   const int64_t n = exec_cells(g);
   Workspace *wp = workspace(s);
   if (!wp) return HTR_E_DEVICE;
   Workspace &w = *wp;
   const int64_t blocks = std::min<int64_t>(std::max<int64_t>(grid_for(n, 256), 1), 4096);
   HTR_LAUNCH("spectral_radius", spectral_radius_kernel<MX>, dim3((unsigned)blocks), dim3(256), 0, s, f, mix, g, n, w.partial);
   int rc = check_launch("spectral_radius_kernel");
   if (rc) return rc;
   HTR_LAUNCH("reduce_max", reduce_max_kernel, dim3(1), dim3(256), 0, s, w.partial, (int)blocks, w.partial + Workspace::kPartial - 1);
   rc = check_launch("reduce_max_kernel");
   if (rc) return rc;
   if (hipMemcpyAsync(w.result_host, w.partial + Workspace::kPartial - 1, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess)
      return set_error(HTR_E_DEVICE, "spectral radius: copy back failed");
   *result = *w.result_host;
   return HTR_OK;
}

int max_mach_launch(int dir, const htr_fields &f, const Geo &g, double *result, hipStream_t s)
{
   const int64_t n = exec_cells(g);
   Workspace *wp = workspace(s);
   if (!wp) return HTR_E_DEVICE;
   Workspace &w = *wp;
   const int64_t blocks = std::min<int64_t>(std::max<int64_t>(grid_for(n, 256), 1), 4096);
   HTR_LAUNCH("max_mach", max_mach_kernel, dim3((unsigned)blocks), dim3(256), 0, s, f, g, dir, n, w.partial);
   int rc = check_launch("max_mach_kernel");
   if (rc) return rc;
   HTR_LAUNCH("reduce_max", reduce_max_kernel, dim3(1), dim3(256), 0, s, w.partial, (int)blocks, w.partial + Workspace::kPartial - 1);
   rc = check_launch("reduce_max_kernel");
   if (rc) return rc;
   if (hipMemcpyAsync(w.result_host, w.partial + Workspace::kPartial - 1, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess)
      return set_error(HTR_E_DEVICE, "max Mach: copy back failed");
   *result = *w.result_host;
   return HTR_OK;
}

// ---- spectral radius without allocation --------------------------------------
// The partial maxima of htr_max_spectral_radius_noalloc live in a file-scope
// device array: one double per block of the first pass, which takes at most
// kRadiusBlocks blocks (grid-stride), so every rect the library accepts fits.
// Each GPU holds its own copy until the process ends.  In host code the
// array's name is not a device address: the address on a device comes from
// hipGetSymbolAddress, looked up once per device with that device current,
// and only that address reaches the kernels.  Calls that overlap on one GPU
// share the array and race; every launch rewrites the partials it reads, so
// nothing needs clearing between calls.
constexpr int kRadiusBlocks = 4096;
static __device__ __align__(256) double scratch[kRadiusBlocks];

static std::mutex g_scratch_mu;
static double *g_scratch_dev[64];   // scratch address per device ordinal

static int radius_scratch(double **out)
{
   int dev = -1;
   if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(HTR_E_DEVICE, "no HIP device");
   std::lock_guard<std::mutex> lk(g_scratch_mu);
   if (!g_scratch_dev[dev]) {
      void *p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
         return set_error(HTR_E_DEVICE, "radius scratch: symbol address lookup failed on device %d", dev);
      g_scratch_dev[dev] = static_cast<double*>(p);
   }
   *out = g_scratch_dev[dev];
   return HTR_OK;
}

// n = 0 (empty rect): the tail alone writes the maximum over no cells, 0.
// Launched directly (no timing events: nothing is created on this path).
template<class MX>
int spectral_radius_noalloc_launch(const htr_fields &f, const DMix *mix, const Geo &g, int64_t n, double *d_result,
                                   hipStream_t s)
{
   double *partial = nullptr;
   int rc = radius_scratch(&partial);
   if (rc) return rc;
   const int64_t blocks = std::min<int64_t>(grid_for(n, 256), kRadiusBlocks);
   if (blocks > kRadiusBlocks) return set_error(HTR_E_UNSUPPORTED, "radius scratch holds %d partials", kRadiusBlocks);
   if (n > 0) {
      hipLaunchKernelGGL(spectral_radius_kernel<MX>, dim3((unsigned)blocks), dim3(256), 0, s, f, mix, g, n, partial);
      if ((rc = check_launch("spectral_radius_kernel"))) return rc;
   }
   hipLaunchKernelGGL(reduce_max_kernel, dim3(1), dim3(256), 0, s, partial, n > 0 ? (int)blocks : 0, d_result);
   return check_launch("reduce_max_kernel");
}

#define HTR_INST_POINT(MX)                                                                          \
   template int properties_launch<MX>(const htr_fields &, const DMix *, const Geo &, hipStream_t);  \
   template int conserved_launch<MX>(const htr_fields &, const DMix *, const Geo &, hipStream_t);   \
   template int primitive_launch<MX>(const htr_fields &, const DMix *, const Geo &, hipStream_t);   \
   template int chem_source_launch<MX>(const htr_fields &, const DMix *, const Geo &, hipStream_t, bool); \
   template int shock_sensor_launch<MX>(int, const htr_fields &, const Geo &, hipStream_t);         \
   template int spectral_radius_launch<MX>(const htr_fields &, const DMix *, const Geo &, double *, hipStream_t); \
   template int spectral_radius_noalloc_launch<MX>(const htr_fields &, const DMix *, const Geo &, int64_t, double *, \
                                                   hipStream_t);                                                   \
   template int rk_stage_launch<MX>(const htr_fields &, const DMix *, const Geo &, int, const DtArg &,           \
                                    const htr_view &, hipStream_t);                                               \
   template int bc_launch<MX>(const htr_fields &, const DMix *, const Geo &, int, int, int, double, hipStream_t);
HTR_INST_POINT(MixCP)
HTR_INST_POINT(MixAir)
HTR_INST_POINT(Mix30)

} // namespace htr
