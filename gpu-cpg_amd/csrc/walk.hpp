// One lane's walk over a parent path's suffix chain: the enumeration of
// the parent's children shared by the single-query and multi-query PFXT.
//
// A child of a parent with cost c takes one deviation edge u->w at a vertex
// u of the parent's suffix chain; its cost is ((c + distf[w]) + w_e) -
// distf[u] in binary32, left to right (gpucpg/gpucpg.cu:2670-2676,
// :3271-3276).  The walk reads the deviation index of the SFXT: a chain
// vertex whose smallest delta exceeds X = (U - c) + slack is skipped
// through its next-smaller pointer, and its delta-sorted deviation list is
// cut at the first delta > X.  The slack bounds the rounding gap between
// a delta and the binary32 recurrence, so no child with cost within the
// bound is skipped; the exact comparison decides emission.
#pragma once

#include "common.hpp"

namespace cpg {

struct ChainWalk {
  int mode;   // 0: at a chain vertex, 1: scanning its deviations, 2: done
  int v, j, je, sv, steps;
  float c, X, dv;

  // mag = 2 max|distf| + max|w| + 1 (a bound on the terms' magnitudes).
  __device__ inline void start(bool active, float cost, int to, float U, bool bounded, float mag) {
    mode = active && to != -1 ? 0 : 2;
    v = to;
    j = je = 0;
    sv = -1;
    steps = 0;
    c = cost;
    X = bounded ? (U - cost) + ldexpf(fabsf(U) + fabsf(cost) + mag, -17) : __builtin_huge_valf();
    dv = 0.0f;
  }

  // One step.  Returns true when it found a child with cost < U (cost <= U
  // when !strict); *cc, *cw and *from receive its cost, its deviation head
  // w and the deviation vertex u.  A walk visits each chain vertex and each
  // deviation at most once, so more than step_limit steps means a corrupt
  // index: the walk stops and *err is raised.  *nonmono counts the children
  // whose binary32 cost comes out below the parent's (a deviation whose
  // int-scaled delta is 0, or a negative weight): where one sits at the
  // bound, a path just below U can descend from a parent at or above U,
  // which is never expanded.
  __device__ inline bool step(const int4* vrec, const int2* devinfo, const int4* devs, float U,
                              bool strict, int step_limit, int* err, unsigned long long* hops,
                              unsigned long long* devv, unsigned long long* nonmono, float* cc,
                              int* cw, int* from) {
    if (mode == 2) return false;
    if (++steps > step_limit) {
      atomicExch(err, 1);
      mode = 2;
      return false;
    }
    *from = v;
    if (mode == 0) {
      // the vertex record and its deviation range are loaded together (one
      // round trip where the walk stops here, not two); j and je are set
      // either way, and only read once the walk scans this vertex
      const int4 r = vrec[v];
      const int2 di = devinfo[v];
      ++*hops;
      j = di.x;
      je = di.x + di.y;
      if (__int_as_float(r.x) > X) {
        v = r.y;
        if (v == -1) mode = 2;
      } else {
        sv = r.z;
        dv = __int_as_float(r.w);
        mode = 1;
      }
      return false;
    }
    bool leave = j >= je;
    bool emit = false;
    if (!leave) {
      const int4 d = devs[j];
      ++*devv;
      if (__int_as_float(d.w) > X) {
        leave = true;
      } else {
        const float x = ((c + __int_as_float(d.z)) + __int_as_float(d.y)) - dv;
        ++j;
        if (x < c) ++*nonmono;
        if (strict ? x < U : x <= U) {
          emit = true;
          *cc = x;
          *cw = d.x;
        }
      }
    }
    if (leave) {
      v = sv;
      mode = v == -1 ? 2 : 0;
    }
    return emit;
  }
};

}  // namespace cpg
