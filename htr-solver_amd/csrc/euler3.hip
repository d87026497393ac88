// euler3.hip -- the hybrid (or pure Kennedy-Gruber) Euler flux divergence of
// all three directions for the full RHS, the x and y sweeps fused.
//
// Reference: src/prometeo_rhs.inl:42-135 (ComputeKGSums, KGFluxReconstruction),
// :767-889 (updateRHSSpan of the hybrid task), src/prometeo_rhs.rg:281-297
// (Euler Z, then Y, then X into the same Conserved_t).
//
// The per-direction sweeps (euler.hip) read the inputs and read-modify-write
// Conserved_t once per direction and component group.  Here, after the TENO
// passes of the three directions (each into its own face buffer), the z sweep
// is the line walk of euler.hip, and the x and y sweeps are one pass per
// group: a block owns a tile of 64 x-cells by 4 y-rows of a plane,
//   plane      the cell values the x and y Kennedy-Gruber faces read (rho, u,
//              v and the group's pair values), for the tile plus a 3-cell
//              halo in x and y (70 x 10 cells), staged in LDS once;
//   faces      each thread forms the + face and the - face of its cell along
//              y and along x from the plane (the - face is the + face of the
//              previous cell, or the reference's primed face at the first cell
//              of the rect: the same function of the same cells, so the same
//              value the sweep carries as FluxM); faces the hybrid sensor
//              sends to TENO are read from the TENO pass's face buffer;
//   update     Conserved_t of the group's components is read once, receives
//              the y, then the x difference (the reference's order after z)
//              and is written once.
// Pair sums enter in the reference's order (l outer, m inner) and every face
// value is formed by the same operations as the line walks, so the result is
// bit-identical to the three sweeps.
#include "common.hpp"
#include "kernels.hpp"

namespace htr {

constexpr int T3X = 64, T3Y = 4, T3H = 3;                    // tile and halo
constexpr int T3LX = T3X + 2*T3H, T3LY = T3Y + 2*T3H;        // 70 x 10 plane cells
constexpr int T3N = T3LX*T3LY;

// Values of a plane cell: 0 rho, 1 u, 2 v, then the group's pair values
// (G = 0: Y_0 .. Y_NS-1; G = 1: w, H, p).
template<class MX, int G>
struct Plane3 {
   static constexpr int NS = MX::NS;
   static constexpr int NV = (G == 0) ? 3 + NS : 6;
   __device__ static void load(double *o, const htr_fields &f, const htr_view &hv, const P3 &p)
   {
      o[0] = ld<double>(f.rho, p);
      o[1] = ld<double>(f.velocity, p, 0);
      o[2] = ld<double>(f.velocity, p, 1);
      if constexpr (G == 0) {
#pragma unroll
         for (int i = 0; i < NS; i++) o[3 + i] = ld<double>(f.MassFracs, p, i);
      } else {
         o[3] = ld<double>(f.velocity, p, 2);
         o[4] = ld<double>(hv, p);
         o[5] = ld<double>(f.pressure, p);
      }
   }
   // index of pair value i (G = 1: u, v, w, H, then the pressure sum's p)
   __host__ __device__ static constexpr int pv(int i) { return (G == 0) ? 3 + i : 1 + i; }
   static_assert(pv(KGGroup<MX, G>::NP - 1) < NV, "pair values lie in the plane record");
};

__device__ __forceinline__ int wrap1(int c, int n)
{
   while (c >= n) c -= n;
   while (c < 0) c += n;
   return c;
}

// Kennedy-Gruber flux of the group at the face of plane cell q (LDS index)
// along D (0 x, 1 y), node type t: KGFluxReconstruction (rhs.inl:79-135)
// from the pair sums S(a, b) (rhs.inl:42-75) of the plane's cells.
template<class MX, int G, int D>
__device__ __forceinline__ void kg_face_plane(const double *L, int q, int t, double *Flux)
{
   using K = KGGroup<MX, G>;
   using PL = Plane3<MX, G>;
   constexpr int st = (D == 0) ? 1 : T3LX;
   auto pair = [&](double *S, int a, int b) {
      // each pair's plane reads next to their use: hoisted, the reads of the
      // six stencil cells of both faces would hold some 100 VGPRs
      __builtin_amdgcn_sched_barrier(0);
      const double rhom = L[a] + L[b];
      const double vm = -(L[(1 + D)*T3N + a] + L[(1 + D)*T3N + b]);
#pragma unroll
      for (int i = 0; i < K::NC; i++) S[i] = rhom*vm*(L[PL::pv(i)*T3N + a] + L[PL::pv(i)*T3N + b]);
      if constexpr (G == 1) S[4] = L[PL::pv(4)*T3N + a] + L[PL::pv(4)*T3N + b];
   };
   const int order = c_st.KOrder[t];
   double fl[K::NP], N[K::NP], M[K::NP];
   pair(N, q, q + st);                                               // l = 0: S(p,p+1)
   {
      const double c = c_st.KCoeff[t][0];
#pragma unroll
      for (int i = 0; i < K::NP; i++) fl[i] = (order > 0) ? c*N[i] : 0.0;
   }
   if (order > 1) {                                                  // l = 1: S(p,p+2) + S(p-1,p+1)
      pair(N, q, q + 2*st);
      pair(M, q - st, q + st);
      const double c = c_st.KCoeff[t][1];
#pragma unroll
      for (int i = 0; i < K::NP; i++) fl[i] += c*(N[i] + M[i]);
   }
   if (order > 2) {                                                  // l = 2: S(p,p+3) + S(p-1,p+2) + S(p-2,p+1)
      double R[K::NP];
      pair(N, q, q + 3*st);
      pair(M, q - st, q + 2*st);
      pair(R, q - 2*st, q + st);
      const double c = c_st.KCoeff[t][2];
#pragma unroll
      for (int i = 0; i < K::NP; i++) fl[i] += c*((N[i] + M[i]) + R[i]);
   }
#pragma unroll
   for (int i = 0; i < K::NC; i++) Flux[i] = 0.25*fl[i];
   if constexpr (G == 1) Flux[D] -= fl[4];
}

// staggered-node flux of the group at cell p (rhs.inl:92-110)
template<class MX, int G, int DIR>
__device__ __forceinline__ void staggered_flux(const htr_fields &f, const Geo &g, const P3 &p, int t, double *Flux)
{
   using K = KGGroup<MX, G>;
   const P3 q = (t == L_S_node) ? p : wplus(g, DIR, p, c_st.Cp[t][2]);
   const double u = ld<double>(f.velocity, q, DIR);
#pragma unroll
   for (int i = 0; i < K::NC; i++) Flux[i] = -(ld<double>(f.Conserved, q, K::LO + i)*u);
   if constexpr (G == 1) {
      const double pr = ld<double>(f.pressure, q);
      Flux[DIR] -= pr;
      Flux[3] -= pr*u;
   }
}

// line index of cell p for the face ids of the sweep along DIR
template<int DIR>
__device__ __forceinline__ uint32_t line_of(const Geo &g, const P3 &p)
{
   constexpr int A = Faces<DIR>::A, B = Faces<DIR>::B;
   return (uint32_t)(p.c[A] - g.elo[A]) + (uint32_t)(p.c[B] - g.elo[B])*(uint32_t)g.esz[A];
}

// face of cell q along D in {0, 1} (plane index qi): Kennedy-Gruber, the
// staggered-node flux, or the TENO flux the TENO pass of D wrote for face j
template<class MX, bool ALL_KG, int G, int D>
__device__ __forceinline__ void plane_face(const htr_fields &f, const Geo &g, const double *L,
                                           const double *__restrict__ F, const P3 &q, int qi, uint32_t line, int j,
                                           double *Fl)
{
   using K = KGGroup<MX, G>;
   const int t = ld<int>(f.nType[D], q);
   if (!ALL_KG && !hybrid_kg<D>(f, g, q, t)) {
      const uint32_t fid = Faces<D>::id(g, line, j);
#pragma unroll
      for (int i = 0; i < K::NC; i++) Fl[i] = face_ld<MX::NEQ>(F, fid, K::LO + i);
   } else if (t == L_S_node || t == Rm1_S_node) {
      staggered_flux<MX, G, D>(f, g, q, t, Fl);
   } else {
      kg_face_plane<MX, G, D>(L, qi, t, Fl);
   }
}

// adds (F(p) - F(p-)) m_e along D in {0, 1} to acc
template<class MX, bool ALL_KG, int G, int D>
__device__ __forceinline__ void plane_dir(const htr_fields &f, const Geo &g, const double *L,
                                          const double *__restrict__ F, const P3 &p, int pi, double *acc)
{
   using K = KGGroup<MX, G>;
   constexpr int st = (D == 0) ? 1 : T3LX;
   const uint32_t line = line_of<D>(g, p);
   const int j = p.c[D] - g.elo[D] + 1;                   // the + face of p is face j
   double Fp[K::NC], Fm[K::NC];
   plane_face<MX, ALL_KG, G, D>(f, g, L, F, p, pi, line, j, Fp);
   // the - face: the + face of p - 1, or at the first cell of the rect the
   // primed face at p + offM1(nType[p]) (rhs.inl:800-817)
   const int off = (j == 1) ? c_st.Cp[ld<int>(f.nType[D], p)][1] : -1;
   const P3 m = wplus(g, D, p, off);
   plane_face<MX, ALL_KG, G, D>(f, g, L, F, m, pi + off*st, line, j - 1, Fm);
   const double me = ld<double>(f.m_e[D], p);
#pragma unroll
   for (int i = 0; i < K::NC; i++) acc[i] += (Fp[i] - Fm[i])*me;
}

// Blocks: (x tile, y tile, z segment), x fastest; 256 threads = 64 x-cells
// (lanes) by 4 y-rows (waves).  Every thread stages plane cells and reaches
// every barrier; only threads whose cell lies in the rect form faces.
template<class MX, bool ALL_KG, int G>
__global__ void __launch_bounds__(256)
euler_kg_xy_kernel(const htr_fields f, const Geo g, const htr_view hv, const double *__restrict__ Fx,
                   const double *__restrict__ Fy, int Lz, uint32_t ntx, uint32_t nty)
{
   using K = KGGroup<MX, G>;
   using PL = Plane3<MX, G>;
   constexpr int NV = PL::NV;
   __shared__ double L[NV*T3N];
   const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
   uint32_t b = blockIdx.x;
   const uint32_t tx = b % ntx;
   b /= ntx;
   const uint32_t ty = b % nty, tz = b/nty;
   const int x0 = g.elo[0] + (int)tx*T3X, y0 = g.elo[1] + (int)ty*T3Y;
   const int zs = g.elo[2] + (int)tz*Lz;
   const int ze = min(zs + Lz, g.elo[2] + g.esz[2]);
   const int x = x0 + lane, y = y0 + wv;
   const bool active = x < g.elo[0] + g.esz[0] && y < g.elo[1] + g.esz[1];
   const int pi = (T3H + wv)*T3LX + T3H + lane;                       // plane index of the own cell
   for (int z = zs; z < ze; z++) {
      __syncthreads();                         // the previous plane's readers are done
      for (int s = threadIdx.x; s < T3N; s += 256) {
         const int lx = s % T3LX, ly = s/T3LX;
         const P3 q = mkp(g, wrap1(x0 - T3H + lx, g.bsz[0]), wrap1(y0 - T3H + ly, g.bsz[1]), z);
         double v[NV];
         PL::load(v, f, hv, q);
#pragma unroll
         for (int k = 0; k < NV; k++) L[k*T3N + s] = v[k];
      }
      __syncthreads();
      if (active) {
         const P3 p = mkp(g, x, y, z);
         double acc[K::NC];
#pragma unroll
         for (int i = 0; i < K::NC; i++) acc[i] = ld<double>(f.Conserved_t, p, K::LO + i);
         plane_dir<MX, ALL_KG, G, 1>(f, g, L, Fy, p, pi, acc);       // Y
         plane_dir<MX, ALL_KG, G, 0>(f, g, L, Fx, p, pi, acc);       // X
#pragma unroll
         for (int i = 0; i < K::NC; i++) at<double>(f.Conserved_t, p, K::LO + i) = acc[i];
      }
   }
}

// Segment length along z: the longest of 64 .. 8 planes that still gives at
// least 2048 blocks (eight per CU), as a longer segment repeats less warm-up.
static int kg3_segment(int64_t tiles, int64_t nz)
{
   int Lz = 64;
   while (Lz > 8 && tiles*((nz + Lz - 1)/Lz) < 2048) Lz /= 2;
   return Lz;
}

template<class MX>
int euler3_dispatch(int scheme, const htr_fields &f, const DMix *mix, const Geo &g, double RK0, double RK1,
                    const DtArg &dt, hipStream_t s)
{
   if (exec_cells(g) == 0) return HTR_OK;
   const int64_t nf[3] = { Faces<0>::count(g), Faces<1>::count(g), Faces<2>::count(g) };
   const int64_t nmax = std::max(nf[0], std::max(nf[1], nf[2]));
   if (nmax*MX::NEQ >= (1ll << 31)) return set_error(HTR_E_UNSUPPORTED, "euler: too many faces");
   Workspace *wp = workspace(s);
   if (!wp) return HTR_E_DEVICE;
   Workspace &w = *wp;
   const bool hyb = scheme == HTR_EULER_HYBRID;
   int rc = workspace_faces(w, hyb ? (nf[0] + nf[1] + nf[2])*MX::NEQ : 0, hyb ? 2*nmax : 0,
                            hyb ? teno_doubles<MX>(nmax) : 0);
   if (rc) return rc;
   htr_view hv = {};
   hv.ptr = w.hbuf;                             // total enthalpy, formed by the caller
   hv.sx = 8;
   double *F[3] = { w.faces, w.faces + nf[0]*MX::NEQ, w.faces + (nf[0] + nf[1])*MX::NEQ };
   const htr_view dec = {};
   if (hyb) {
      if ((rc = teno_faces<MX, 2>(f, mix, g, hv, RK0, RK1, dt, w, nf[2], F[2], dec, s))) return rc;
      if ((rc = teno_faces<MX, 1>(f, mix, g, hv, RK0, RK1, dt, w, nf[1], F[1], dec, s))) return rc;
      if ((rc = teno_faces<MX, 0>(f, mix, g, hv, RK0, RK1, dt, w, nf[0], F[0], dec, s))) return rc;
   }
   // Z: the line walks of the z sweep (group by group) into Conserved_t
   {
      const int Lw = walk_segment(2, (int64_t)g.esz[0]*g.esz[1], (int64_t)g.esz[2] + 1);
      const int64_t nseg = ((int64_t)g.esz[2] + 1 + Lw - 1)/Lw;
      const int64_t nunits = nseg*g.esz[0]*g.esz[1];
      if (nunits >= (1ll << 31)) return set_error(HTR_E_UNSUPPORTED, "euler: too many line segments");
      if (hyb) {
         HTR_LAUNCH("euler_kg_species_z", (euler_kg_walk_kernel<MX, 2, false, 0>), dim3(grid_for(nunits, 256)),
                    dim3(256), 0, s, f, g, hv, F[2], nf[2], Lw, (uint32_t)nseg, (uint32_t)nunits, dec);
         if ((rc = check_launch("euler_kg_walk_kernel"))) return rc;
         HTR_LAUNCH("euler_kg_flow_z", (euler_kg_walk_kernel<MX, 2, false, 1>), dim3(grid_for(nunits, 256)),
                    dim3(256), 0, s, f, g, hv, F[2], nf[2], Lw, (uint32_t)nseg, (uint32_t)nunits, dec);
      } else {
         HTR_LAUNCH("euler_kg_species_z", (euler_kg_walk_kernel<MX, 2, true, 0>), dim3(grid_for(nunits, 256)),
                    dim3(256), 0, s, f, g, hv, F[2], nf[2], Lw, (uint32_t)nseg, (uint32_t)nunits, dec);
         if ((rc = check_launch("euler_kg_walk_kernel"))) return rc;
         HTR_LAUNCH("euler_kg_flow_z", (euler_kg_walk_kernel<MX, 2, true, 1>), dim3(grid_for(nunits, 256)),
                    dim3(256), 0, s, f, g, hv, F[2], nf[2], Lw, (uint32_t)nseg, (uint32_t)nunits, dec);
      }
      if ((rc = check_launch("euler_kg_walk_kernel"))) return rc;
   }
   // Y and X: one plane pass per group
   const uint32_t ntx = (uint32_t)((g.esz[0] + T3X - 1)/T3X), nty = (uint32_t)((g.esz[1] + T3Y - 1)/T3Y);
   const int Lz = kg3_segment((int64_t)ntx*nty, g.esz[2]);
   const int64_t nb = (int64_t)ntx*nty*((g.esz[2] + Lz - 1)/Lz);
   if (nb >= (1ll << 31)) return set_error(HTR_E_UNSUPPORTED, "euler: too many tiles");
   if (hyb) {
      HTR_LAUNCH("euler_kg_xy_species", (euler_kg_xy_kernel<MX, false, 0>), dim3((unsigned)nb), dim3(256), 0, s,
                 f, g, hv, F[0], F[1], Lz, ntx, nty);
      if ((rc = check_launch("euler_kg_xy_kernel"))) return rc;
      HTR_LAUNCH("euler_kg_xy_flow", (euler_kg_xy_kernel<MX, false, 1>), dim3((unsigned)nb), dim3(256), 0, s,
                 f, g, hv, F[0], F[1], Lz, ntx, nty);
   } else {
      HTR_LAUNCH("euler_kg_xy_species", (euler_kg_xy_kernel<MX, true, 0>), dim3((unsigned)nb), dim3(256), 0, s,
                 f, g, hv, F[0], F[1], Lz, ntx, nty);
      if ((rc = check_launch("euler_kg_xy_kernel"))) return rc;
      HTR_LAUNCH("euler_kg_xy_flow", (euler_kg_xy_kernel<MX, true, 1>), dim3((unsigned)nb), dim3(256), 0, s,
                 f, g, hv, F[0], F[1], Lz, ntx, nty);
   }
   return check_launch("euler_kg_xy_kernel");
}

template int euler3_dispatch<MixCP>(int, const htr_fields &, const DMix *, const Geo &, double, double, const DtArg &,
                                    hipStream_t);
template int euler3_dispatch<MixAir>(int, const htr_fields &, const DMix *, const Geo &, double, double,
                                     const DtArg &, hipStream_t);

} // namespace htr
