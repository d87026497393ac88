// common.hpp -- field views, periodic wrapping and launch helpers shared by
// the HTR hot-path kernels.
//
// Device-side addressing.  The host rebases every view to the low corner of
// Fluid_bounds, so kernel coordinates run over [0, bsz) in each direction,
// and checks that all FP64 cell fields share one set of spatial strides.  A
// point then carries a single 32-bit byte offset that is valid for every
// FP64 field: a load is (uniform field base + component offset) + (per-lane
// 32-bit offset), which the backend emits as a global load with a scalar
// base and one VGPR of offset instead of a 64-bit address per field.
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
   uint32_t sb[3];    // byte strides shared by the FP64 cell fields
};

struct P3 {
   int32_t  c[3];
   uint32_t o;        // byte offset of the point in the FP64 fields
};

__device__ inline P3 mkp(const Geo &g, int x, int y, int z)
{
   P3 p;
   p.c[0] = x; p.c[1] = y; p.c[2] = z;
   p.o = (uint32_t)x*g.sb[0] + (uint32_t)y*g.sb[1] + (uint32_t)z*g.sb[2];
   return p;
}

__device__ inline void setc(const Geo &g, P3 &p, int d, int v)
{
   p.o += (uint32_t)(v - p.c[d])*g.sb[d];
   p.c[d] = v;
}

template<typename T>
__device__ inline const T *lp(const htr_view &v, const P3 &p, int comp)
{
   if constexpr (sizeof(T) == 8)
      return reinterpret_cast<const T*>(reinterpret_cast<const char*>(v.ptr) + comp*v.sc*8 + p.o);
   else
      return reinterpret_cast<const T*>(v.ptr) + ((int64_t)p.c[0]*v.sx + (int64_t)p.c[1]*v.sy +
                                                   (int64_t)p.c[2]*v.sz + comp*v.sc);
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
