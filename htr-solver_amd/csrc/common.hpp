// common.hpp -- field views, periodic wrapping and launch helpers shared by
// the HTR hot-path kernels.
//
// Device-side addressing.  The host rebases every view to the low corner of
// Fluid_bounds (kernel coordinates run over [0, bsz) in each direction) and
// checks that the spatial strides of every view are an integer multiple k of
// one common triple (cx, cy, cz): that covers structure-of-arrays fields
// (k = 1, components sc = N apart) as well as the reference's per-point
// arrays (VecNEq, VecNSp, Vec3: k = number of components, sc = 1).  A point
// carries one 32-bit offset o = x cx + y cy + z cz, and an access is
//   field base + comp*sc*sizeof(T)  (uniform)  +  o * k*sizeof(T)  (32-bit),
// so the backend uses a scalar base and one VGPR of offset per access instead
// of a 64-bit address per (point, field).  In the kernels' copy of a view,
// sx holds k*sizeof(T) in bytes; sy and sz are unused.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/htr_amd.h"
#include "mix.hpp"
#include "stencil.hpp"

namespace htr {

// Box geometry relative to the low corner of Fluid_bounds.
struct Geo {
   int32_t  bsz[3];   // Fluid_bounds extent
   int32_t  elo[3];   // execution rect, low corner
   int32_t  esz[3];   // execution rect, extent
   uint32_t cb[3];    // common stride triple (cx, cy, cz)
};

// Time step of a launch: the host value dt, or, when radius is set, the
// ConstantCFL step (cfl / max spectral radius) read on the device from the
// result of a device-side reduction and multiplied by scale, in the
// reference's rounding order: dt = cfl/radius, then dt*scale
// (src/prometeo.rg:1165, :1419).
struct DtArg {
   double dt;
   const double *radius;
   double cfl, scale;
   __device__ double get() const { return radius ? (cfl/(*radius))*scale : dt; }
};

inline DtArg host_dt(double dt) { DtArg a; a.dt = dt; a.radius = nullptr; a.cfl = 0.0; a.scale = 0.0; return a; }

struct P3 {
   int32_t  c[3];
   uint32_t o;        // x cx + y cy + z cz
};

__device__ inline P3 mkp(const Geo &g, int x, int y, int z)
{
   P3 p;
   p.c[0] = x; p.c[1] = y; p.c[2] = z;
   p.o = (uint32_t)x*g.cb[0] + (uint32_t)y*g.cb[1] + (uint32_t)z*g.cb[2];
   return p;
}

__device__ inline void setc(const Geo &g, P3 &p, int d, int v)
{
   p.o += (uint32_t)(v - p.c[d])*g.cb[d];
   p.c[d] = v;
}

template<typename T>
__device__ inline const T *lp(const htr_view &v, const P3 &p, int comp)
{
   return reinterpret_cast<const T*>(reinterpret_cast<const char*>(v.ptr) + comp*v.sc*(int64_t)sizeof(T) +
                                     p.o*(uint32_t)v.sx);
}

template<typename T>
__device__ inline T ld(const htr_view &v, const P3 &p, int comp = 0) { return *lp<T>(v, p, comp); }

template<typename T>
__device__ inline T &at(const htr_view &v, const P3 &p, int comp = 0) { return *const_cast<T*>(lp<T>(v, p, comp)); }

// warpPeriodic (src/Utils/PointDomain_helper.hpp:157-188): the point moved by
// off along d, wrapped into Fluid_bounds
__device__ inline P3 wplus(const Geo &g, int d, P3 p, int off)
{
   int c = p.c[d] + off;
   while (c >= g.bsz[d]) c -= g.bsz[d];
   while (c < 0) c += g.bsz[d];
   setc(g, p, d, c);
   return p;
}

__device__ inline P3 wminus(const Geo &g, int d, P3 p, int off) { return wplus(g, d, p, off); }

inline int64_t exec_cells(const Geo &g) { return (int64_t)g.esz[0]*g.esz[1]*g.esz[2]; }

// XCD-aware block order: the dispatcher hands workgroups to the eight XCDs
// round robin (block b runs on XCD b % 8), and each XCD has its own L2.  The
// logical index makes blocks with consecutive logical indices share an XCD
// (XCD x holds logical blocks [x n/8, (x+1) n/8)), so work items that touch
// the same cache lines -- neighbouring chunks of a face list -- meet in one
// L2.  Identity when n is not a multiple of 8.
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t n)
{
   if (n & 7u) return b;
   return (b & 7u)*(n >> 3) + (b >> 3);
}

// point of the execution rect for a flat index, x fastest (the host keeps
// the cell count below 2^31)
__device__ inline P3 exec_point(const Geo &g, int64_t i64)
{
   uint32_t i = (uint32_t)i64;
   const int x = g.elo[0] + (int)(i % (uint32_t)g.esz[0]);
   i /= (uint32_t)g.esz[0];
   const int y = g.elo[1] + (int)(i % (uint32_t)g.esz[1]);
   const int z = g.elo[2] + (int)(i / (uint32_t)g.esz[1]);
   return mkp(g, x, y, z);
}

} // namespace htr
