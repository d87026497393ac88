// diffusion.hip -- viscous, heat and multicomponent (Fick + correction
// velocity) flux divergence along one direction.
//
// Reference: src/prometeo_rhs.inl:1058-1282 (GetSigma, GetDiffusionFlux,
// updateRHSSpan), src/prometeo_rhs.cu:461-622.  One pass per direction:
// line segments walked face by face, the divergence fused (below).
#include "common.hpp"
#include "kernels.hpp"
#include <type_traits>

namespace htr {

template<class MX, int DIR>
struct Diffusion {
   static constexpr int NS = MX::NS, NEQ = MX::NEQ;
   static constexpr int D1 = (DIR == 0) ? 1 : 0;   // getT1
   static constexpr int D2 = (DIR == 2) ? 1 : 2;   // getT2
   const htr_fields &f;
   const DMix &m;
   const Geo &g;

   __device__ double vel(const P3 &p, int i) const { return ld<double>(f.velocity, p, i); }

   template<int DD>
   __device__ void deriv3(const P3 &p, double *out) const
   {
      const int t = ld<int>(f.nType[DD], p);
      const double mm = ld<double>(f.m_d[DD], p);
      const P3 pM1 = wminus(g, DD, p, c_st.Cp[t][1]);
      const P3 pP1 = wplus(g, DD, p, c_st.Cp[t][2]);
      const double g0 = (double)c_st.Grad[t][0], g1 = (double)c_st.Grad[t][1];
#pragma unroll
      for (int i = 0; i < 3; i++) {
         const double x = vel(p, i);
         out[i] = (g0*(x - vel(pM1, i)) + g1*(vel(pP1, i) - x))*mm;
      }
   }

   // face flux of cell p with the carried transverse gradients g[0] (left)
   // and those of q = p + 1 formed here into g[1]; g[1] moves to g[0]
   __device__ void flux(double *Flux, const P3 &p, double g1[2][3], double g2[2][3]) const
   {
      const P3 q = wplus(g, DIR, p, 1);
      deriv3<D1>(q, g1[1]);
      deriv3<D2>(q, g2[1]);
      flux_at(Flux, p, q, g1, g2);
#pragma unroll
      for (int i = 0; i < 3; i++) { g1[0][i] = g1[1][i]; g2[0][i] = g2[1][i]; }
   }

   // The cell values the species part of a face reads, loaded once per cell:
   // a line walk hands the right cell's values on as the next face's left.
   struct SpCell {
      double X[NS], Di[NS], rY[NS];
      double rho, T, lam, W;      // W: molar weight of the mixture from X
   };
   __device__ void sp_load(SpCell &c, const P3 &p) const
   {
#pragma unroll
      for (int i = 0; i < NS; i++) {
         c.X[i] = ld<double>(f.MolarFracs, p, i);
         c.Di[i] = ld<double>(f.Di, p, i);
         c.rY[i] = ld<double>(f.Conserved, p, i);
      }
      c.rho = ld<double>(f.rho, p);
      c.T = ld<double>(f.temperature, p);
      c.lam = ld<double>(f.lam, p);
      c.W = MX::MolarWeightFromX(m, c.X);
   }

   // GetDiffusionFlux (rhs.inl:1140-1218) at the face between cells P and Q
   // (node type t and staggered metric m_s of the left cell), species part:
   // the species fluxes into Flux[0..NS-1] and the heat flux (conduction +
   // species enthalpy transport) returned
   __device__ double species_cells(double *Flux, int t, double m_s, const SpCell &P, const SpCell &Q) const
   {
      const double rho_s = interp2(t, P.rho, Q.rho);
      const double iMixW_s = 1.0/interp2(t, P.W, Q.W);
      const double T_s = interp2(t, P.T, Q.T);
      double heatFlux = interp2(t, P.lam, Q.lam)*m_s*(P.T - Q.T);
      double ViCorr = 0.0;
      double hs[NS];
#pragma unroll
      for (int i = 0; i < NS; i++) {
         hs[i] = MX::SpeciesEnthalpy(m, i, T_s);
         const double YiVi = interp2(t, P.Di[i], Q.Di[i])*m_s*(P.X[i] - Q.X[i])*MX::SpeciesW(m, i)*iMixW_s;
         Flux[i] = -rho_s*YiVi;
         heatFlux += rho_s*YiVi*hs[i];
         ViCorr += YiVi;
      }
#pragma unroll
      for (int i = 0; i < NS; i++) {
         const double corr = interp2(t, P.rY[i], Q.rY[i])*ViCorr;
         Flux[i] += corr;
         heatFlux -= corr*hs[i];
      }
      return heatFlux;
   }

   __device__ double species_at(double *Flux, const P3 &p, const P3 &q) const
   {
      SpCell P, Q;
      sp_load(P, p);
      sp_load(Q, q);
      return species_cells(Flux, ld<int>(f.nType[DIR], p), ld<double>(f.m_s[DIR], p), P, Q);
   }

   // the cell values the momentum part of a face reads
   struct StCell { double v[3], mu; };
   __device__ void st_load(StCell &c, const P3 &p) const
   {
#pragma unroll
      for (int i = 0; i < 3; i++) c.v[i] = vel(p, i);
      c.mu = ld<double>(f.mu, p);
   }

   __device__ double stress_at(double *Flux, const P3 &p, const P3 &q, const double g1[2][3],
                               const double g2[2][3]) const
   {
      StCell P, Q;
      st_load(P, p);
      st_load(Q, q);
      return stress_cells(Flux, ld<int>(f.nType[DIR], p), ld<double>(f.m_s[DIR], p), P, Q, g1, g2);
   }

   // GetDiffusionFlux, momentum part: the viscous stress into Flux[0..2]
   // and u.sigma returned (GetSigma, rhs.inl:1058-1137)
   __device__ double stress_cells(double *Flux, int t, double m_s, const StCell &P, const StCell &Q,
                                  const double g1[2][3], const double g2[2][3]) const
   {
      const double mu_s = interp2(t, P.mu, Q.mu);
      const double *vp = P.v, *vq = Q.v;
      double dn[3], s[3];
#pragma unroll
      for (int i = 0; i < 3; i++) dn[i] = m_s*(vq[i] - vp[i]);
      if (DIR == 0) {
         const double dUdY = interp2(t, g1[0][0], g1[1][0]);
         const double dVdY = interp2(t, g1[0][1], g1[1][1]);
         const double dUdZ = interp2(t, g2[0][0], g2[1][0]);
         const double dWdZ = interp2(t, g2[0][2], g2[1][2]);
         s[0] = mu_s*(4*dn[0] - 2*dVdY - 2*dWdZ)/3;
         s[1] = mu_s*(dn[1] + dUdY);
         s[2] = mu_s*(dn[2] + dUdZ);
      } else if (DIR == 1) {
         const double dUdX = interp2(t, g1[0][0], g1[1][0]);
         const double dVdX = interp2(t, g1[0][1], g1[1][1]);
         const double dVdZ = interp2(t, g2[0][1], g2[1][1]);
         const double dWdZ = interp2(t, g2[0][2], g2[1][2]);
         s[0] = mu_s*(dn[0] + dVdX);
         s[1] = mu_s*(4*dn[1] - 2*dUdX - 2*dWdZ)/3;
         s[2] = mu_s*(dn[2] + dVdZ);
      } else {
         const double dUdX = interp2(t, g1[0][0], g1[1][0]);
         const double dWdX = interp2(t, g1[0][2], g1[1][2]);
         const double dVdY = interp2(t, g2[0][1], g2[1][1]);
         const double dWdY = interp2(t, g2[0][2], g2[1][2]);
         s[0] = mu_s*(dn[0] + dWdX);
         s[1] = mu_s*(dn[1] + dWdY);
         s[2] = mu_s*(4*dn[2] - 2*dUdX - 2*dVdY)/3;
      }

#pragma unroll
      for (int i = 0; i < 3; i++) Flux[i] = s[i];
      double uSigma = 0.0;
#pragma unroll
      for (int i = 0; i < 3; i++) uSigma += interp2(t, vp[i], vq[i])*s[i];
      return uSigma;
   }

   // GetDiffusionFlux (rhs.inl:1140-1218) at the face between p and q
   __device__ void flux_at(double *Flux, const P3 &p, const P3 &q, const double g1[2][3], const double g2[2][3]) const
   {
      const double heatFlux = species_at(Flux, p, q);
      const double uSigma = stress_at(Flux + NS, p, q, g1, g2);
      Flux[NS+3] = (uSigma - heatFlux);
   }
};

// Faces walked along the line, divergence fused.  A thread takes one line
// and one segment of L faces [j0, j0 + L) (a segment other than the first
// also forms face j0 - 1), in the face order of the Euler passes: face j = 0
// is the minus face of the line's first cell p0, at pm1 = p0 + offM1(nType);
// face j >= 1 is the plus face of cell elo + j - 1.  The transverse velocity
// gradients entering a face's left value are carried from the previous
// face's right cell, as the reference does (rhs.inl:1212-1214): the face
// cell itself, except for j = 1, where they belong to pm1 + 1 (different
// from p0 only when offM1(nType[p0]) = 0, a staggered wall node).  Face j's
// flux is differenced with face j-1's into Conserved_t of the cell between
// them with m_d (rhs.inl:1264-1279).
// The left cell's values of a face are the previous face's right cell's,
// handed on in registers (reloaded at face 1, whose left cell p0 follows pm1
// only when offM1(nType[p0]) = -1).
// Two component groups, each walked by its own launch: G = 0 the species
// fluxes, whose heat flux (conduction + species enthalpy transport) goes to
// a face buffer H; G = 1 the viscous stress and the energy flux
// u.sigma - H, read back per face.  Every value is the one the combined flux
// forms; the split only halves the registers a thread holds.
template<class MX, int G>
struct DiffGroup {
   static constexpr int LO = (G == 0) ? 0 : MX::NS;   // first conserved component
   static constexpr int NC = (G == 0) ? MX::NS : 4;    // components of the group
};

template<class MX, int DIR, int G>
__device__ __forceinline__ void diff_face(const Diffusion<MX, DIR> &D, double *Fc, const P3 &p, const P3 &q,
                                          const double g1[2][3], const double g2[2][3], double *__restrict__ H,
                                          uint32_t fid, bool owner)
{
   if constexpr (G == 0) {
      const double heat = D.species_at(Fc, p, q);
      if (owner) H[fid] = heat;
   } else {
      const double uSigma = D.stress_at(Fc, p, q, g1, g2);
      Fc[3] = (uSigma - H[fid]);
   }
}

template<class MX, int DIR, int G>
__global__ void __launch_bounds__(256)
diffusion_walk_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, double *__restrict__ H, int L,
                      uint32_t nseg, uint32_t nunits)
{
   const uint32_t u = blockIdx.x*blockDim.x + threadIdx.x;
   if (u >= nunits) return;
   constexpr int A = Faces<DIR>::A, B = Faces<DIR>::B;
   using K = DiffGroup<MX, G>;
   uint32_t line, seg;
   if (DIR == 0) { line = u/nseg; seg = u - line*nseg; }
   else {
      const uint32_t nl = (uint32_t)g.esz[A]*(uint32_t)g.esz[B];
      seg = u/nl; line = u - seg*nl;
   }
   const int j0 = (int)seg*L;
   const int j1 = min(j0 + L, g.esz[DIR] + 1);
   int c[3];
   c[A] = g.elo[A] + (int)(line % (uint32_t)g.esz[A]);
   c[B] = g.elo[B] + (int)(line / (uint32_t)g.esz[A]);
   c[DIR] = g.elo[DIR];
   const P3 p0 = mkp(g, c[0], c[1], c[2]);
   const P3 pm1 = wminus(g, DIR, p0, c_st.Cp[ld<int>(f.nType[DIR], p0)][1]);
   int j = (j0 == 0) ? 0 : j0 - 1;
   P3 cell = pm1, gl = pm1;
   if (j > 0) {
      cell = p0;
      setc(g, cell, DIR, g.elo[DIR] + j - 1);
      gl = (j == 1) ? wplus(g, DIR, pm1, 1) : cell;
   }
   using D_t = Diffusion<MX, DIR>;
   D_t D{f, *mix, g};
   double g1[2][3], g2[2][3], Fm[K::NC], Fc[K::NC];
   // species group: the left cell's values, handed on from the previous
   // face's right cell (reloaded at face 1, whose left cell p0 follows pm1
   // only when offM1(nType[p0]) = -1)
   using Cell = std::conditional_t<G == 0, typename D_t::SpCell, typename D_t::StCell>;
   auto load = [&](Cell &c, const P3 &p) {
      if constexpr (G == 0) D.sp_load(c, p); else D.st_load(c, p);
   };
   Cell P;
   auto face = [&](const P3 &p, int jj, double *F) {
      const P3 q = wplus(g, DIR, p, 1);
      Cell Q;
      load(Q, q);
      const int t = ld<int>(f.nType[DIR], p);
      const double m_s = ld<double>(f.m_s[DIR], p);
      const uint32_t fid = Faces<DIR>::id(g, line, jj);
      if constexpr (G == 0) {
         const double heat = D.species_cells(F, t, m_s, P, Q);
         if (jj >= j0) H[fid] = heat;
      } else {
         D.template deriv3<D_t::D1>(q, g1[1]);
         D.template deriv3<D_t::D2>(q, g2[1]);
         F[3] = (D.stress_cells(F, t, m_s, P, Q, g1, g2) - H[fid]);
#pragma unroll
         for (int i = 0; i < 3; i++) { g1[0][i] = g1[1][i]; g2[0][i] = g2[1][i]; }
      }
      P = Q;
   };
   load(P, cell);
   if constexpr (G == 1) {
      D.template deriv3<D_t::D1>(gl, g1[0]);
      D.template deriv3<D_t::D2>(gl, g2[0]);
   }
   face(cell, j, Fm);
   for (j++; j < j1; j++) {
      cell = (j == 1) ? p0 : wplus(g, DIR, cell, 1);
      if (j == 1) load(P, cell);
      face(cell, j, Fc);
      const double m = ld<double>(f.m_d[DIR], cell);
#pragma unroll
      for (int i = 0; i < K::NC; i++) {
         at<double>(f.Conserved_t, cell, K::LO + i) += (Fc[i] - Fm[i])*m;
         Fm[i] = Fc[i];
      }
   }
}

// x sweep: lanes on consecutive faces.  The faces of all rows are numbered
// row-major (face j of row r is r (esz + 1) + j, its face id), and a wave
// takes XD_OUT = 62 consecutive ones (lanes 2..63), so a wave may span the
// end of one row and the start of the next: every lane forms the transverse
// gradients of its right cell q, and takes those of its left cell and its
// FluxM from lane l-1 (wave shuffles), so lanes 0 and 1 serve as the halo;
// lane l-1 holds face j-1 of the same row whenever j >= 1, and face 0 takes
// its left gradients at pm1 itself (rhs.inl:1253-1254) and adds no
// divergence.
constexpr int XD_OUT = 62;

template<class MX, int G>
__global__ void __launch_bounds__(256)
diffusion_x_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, double *__restrict__ H,
                   uint32_t nfx, uint32_t nwaves)
{
   using D_t = Diffusion<MX, 0>;
   using K = DiffGroup<MX, G>;
   const uint32_t wave = (blockIdx.x*blockDim.x + threadIdx.x) >> 6;
   if (wave >= nwaves) return;                                      // whole waves
   const int lane = threadIdx.x & 63;
   const int esz = g.esz[0];
   const int gf = (int)(wave*XD_OUT) + lane - 2;                    // face id of this lane
   const bool live = gf >= 0 && (uint32_t)gf < nfx;
   uint32_t row = 0;
   int j = esz + 1;                                                  // (no face)
   if (live) {
      row = (uint32_t)gf/(uint32_t)(esz + 1);
      j = gf - (int)(row*(uint32_t)(esz + 1));
   }
   const P3 prow = mkp(g, g.elo[0], g.elo[1] + (int)(row % (uint32_t)g.esz[1]),
                       g.elo[2] + (int)(row/(uint32_t)g.esz[1]));
   P3 p = wplus(g, 0, prow, j - 1);                                  // left cell of face j
   if (j == 0) p = wminus(g, 0, prow, c_st.Cp[ld<int>(f.nType[0], prow)][1]);
   const P3 q = wplus(g, 0, p, 1);
   D_t D{f, *mix, g};
   double g1[2][3], g2[2][3];
   if constexpr (G == 1) {
#pragma unroll
      for (int i = 0; i < 3; i++) { g1[1][i] = 0.0; g2[1][i] = 0.0; }
      if (live) {
         D.template deriv3<D_t::D1>(q, g1[1]);
         D.template deriv3<D_t::D2>(q, g2[1]);
      }
#pragma unroll
      for (int i = 0; i < 3; i++) { g1[0][i] = __shfl_up(g1[1][i], 1, 64); g2[0][i] = __shfl_up(g2[1][i], 1, 64); }
   }
   double Fc[K::NC];
#pragma unroll
   for (int i = 0; i < K::NC; i++) Fc[i] = 0.0;
   if (lane >= 1 && live && j >= 0) {
      if constexpr (G == 1) {
         if (j == 0) {
            D.template deriv3<D_t::D1>(p, g1[0]);
            D.template deriv3<D_t::D2>(p, g2[0]);
         }
      }
      diff_face<MX, 0, G>(D, Fc, p, q, g1, g2, H, (uint32_t)gf, lane >= 2);
   }
   double Fm[K::NC];
#pragma unroll
   for (int i = 0; i < K::NC; i++) Fm[i] = __shfl_up(Fc[i], 1, 64);
   if (lane >= 2 && live && j >= 1) {
      const double m = ld<double>(f.m_d[0], p);
#pragma unroll
      for (int i = 0; i < K::NC; i++) at<double>(f.Conserved_t, p, K::LO + i) += (Fc[i] - Fm[i])*m;
   }
}

template<class MX, int DIR>
static int launch_diff(const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s)
{
   const int64_t ncells = exec_cells(g);
   if (ncells == 0) return HTR_OK;
   const int64_t nfaces = Faces<DIR>::count(g);
   if (nfaces >= (1ll << 31)) return set_error(HTR_E_UNSUPPORTED, "diffusion: too many faces");
   Workspace *wp = workspace(s);
   if (!wp) return HTR_E_DEVICE;
   int rc = workspace_faces(*wp, nfaces, 0, 0);
   if (rc) return rc;
   double *H = wp->faces;                          // heat flux per face
   if constexpr (DIR == 0) {
      // a lane's face id, below nwaves*XD_OUT + 62 <= nfaces + 123, stays an int
      if (nfaces + 128 >= (1ll << 31)) return set_error(HTR_E_UNSUPPORTED, "diffusion: too many faces");
      const int64_t nwaves = (nfaces + XD_OUT - 1)/XD_OUT;
      HTR_LAUNCH("diffusion_species_x", (diffusion_x_kernel<MX, 0>), dim3(grid_for(nwaves*64, 256)), dim3(256), 0, s,
                 f, mix, g, H, (uint32_t)nfaces, (uint32_t)nwaves);
      if ((rc = check_launch("diffusion_x_kernel"))) return rc;
      HTR_LAUNCH("diffusion_stress_x", (diffusion_x_kernel<MX, 1>), dim3(grid_for(nwaves*64, 256)), dim3(256), 0, s,
                 f, mix, g, H, (uint32_t)nfaces, (uint32_t)nwaves);
      return check_launch("diffusion_x_kernel");
   } else {
      const int L = walk_segment(DIR, (int64_t)g.esz[Faces<DIR>::A]*g.esz[Faces<DIR>::B], (int64_t)g.esz[DIR] + 1);
      const int64_t nseg = ((int64_t)g.esz[DIR] + 1 + L - 1)/L;
      const int64_t nunits = nseg*g.esz[Faces<DIR>::A]*g.esz[Faces<DIR>::B];
      if (nunits >= (1ll << 31)) return set_error(HTR_E_UNSUPPORTED, "diffusion: too many line segments");
      HTR_LAUNCH(HTR_DIR_NAME("diffusion_species", DIR), (diffusion_walk_kernel<MX, DIR, 0>),
                 dim3(grid_for(nunits, 256)), dim3(256), 0, s, f, mix, g, H, L, (uint32_t)nseg, (uint32_t)nunits);
      if ((rc = check_launch("diffusion_walk_kernel"))) return rc;
      HTR_LAUNCH(HTR_DIR_NAME("diffusion_stress", DIR), (diffusion_walk_kernel<MX, DIR, 1>),
                 dim3(grid_for(nunits, 256)), dim3(256), 0, s, f, mix, g, H, L, (uint32_t)nseg, (uint32_t)nunits);
      return check_launch("diffusion_walk_kernel");
   }
}

template<class MX>
int diffusion_dispatch(int dir, const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s)
{
   if (dir == 0) return launch_diff<MX, 0>(f, mix, g, s);
   if (dir == 1) return launch_diff<MX, 1>(f, mix, g, s);
   if (dir == 2) return launch_diff<MX, 2>(f, mix, g, s);
   return set_error(HTR_E_ARG, "diffusion: bad direction %d", dir);
}

template int diffusion_dispatch<MixCP>(int, const htr_fields &, const DMix *, const Geo &, hipStream_t);
template int diffusion_dispatch<MixAir>(int, const htr_fields &, const DMix *, const Geo &, hipStream_t);

} // namespace htr
