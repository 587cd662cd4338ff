// Staggered-grid geometry shared by host and device code.
//
// Restates the index arithmetic of fd::discretization
// (fd/discretization.h:143-156), util::math::floor / modulo
// (util/math.h:158-168) and the per-marker stencil set-up of ib::sweep
// (ib/sweep.h:38-77) and ib::indexing (ib/indexing.h:10-122).
//
// Floating-point contraction is disabled in these helpers: the cell a marker
// falls in is decided by floor() of a rounded expression, and fusing a
// multiply-add would move markers that sit within an ulp of a cell face.
#pragma once
#include <cmath>
#include "ibm.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define IBM_HD __host__ __device__ inline
#else
#define IBM_HD inline
#endif

namespace ibm {

// floor() as the reference writes it (util/math.h:158-162): truncation toward
// zero, corrected by one for negative non-integers.  The reference truncates
// through a cast to long; trunc() is the same value wherever that cast is
// defined (|x| < 2^63) and a single instruction on the GPU.
IBM_HD double ref_floor(double x)
{
    return trunc(x - (x < 0 && (-x != trunc(-x))));
}

// modulo(x, s) = x - s * floor(x / s) (util/math.h:164-168).
IBM_HD double ref_modulo(double x, double s)
{
#pragma clang fp contract(off)
    return x - s * ref_floor(x / s);
}

// units(x): position measured in mesh widths from the grid origin, wrapped
// into the periodic box (fd/discretization.h:143-149, fd/dimension.h:52-57).
template <class A>
IBM_HD double axis_units(const A& a, double x)
{
#pragma clang fp contract(off)
    double s = a.shift;
    double r = a.resolution;
    double y = x - s / r;
    if (!a.solid) y = ref_modulo(y, a.length);
    return y * r + s;
}

// Periodic wrap of a grid index (fd/discretization.h:151-156).
template <class A>
IBM_HD int axis_index(const A& a, int i)
{
    return a.solid ? i : (i + a.cells) % a.cells;
}

// Base cell of the 4-point stencil and the offset d = floor(diff) - diff in
// (-1, 0] (ib/sweep.h:49-77 with clamp(r) = floor(r) for an even stencil).
// `key` is the reference's clamped sort index (ib/indexing.h:10-25 applied at
// ib/sweep.h:70-74) and `base` the decomposed cell it stands for
// (ib/indexing.h:103-115): base = key - solid.
struct stencil_axis {
    int base;
    double d;
};

template <class A>
IBM_HD stencil_axis axis_stencil(const A& a, double x)
{
#pragma clang fp contract(off)
    double diff = axis_units(a, x) - a.shift;
    double fl = ref_floor(diff);
    int i = (int) fl;
    int j = axis_index(a, i);
    int lo = -a.solid;
    int w = a.points + a.solid;           // weight of the clamped index
    int k = j - lo;
    k = k < 0 ? 0 : (k >= w ? w - 1 : k);
    stencil_axis s;
    s.base = k - a.solid;
    s.d = fl - diff;
    return s;
}

// Grid index of stencil point base + o (o in {-1, 0, 1, 2}) for spreading:
// periodic wrap, then -1 outside [0, points) (ib/indexing.h:35-62).
template <class A>
IBM_HD int axis_spread_index(const A& a, int p)
{
    int j = axis_index(a, p);
    return (0 <= j && j < a.points) ? j : -1;
}

// Same for interpolation: periodic wrap, then clamp into [0, points)
// (ib/interpolate.h:16-22, ib/indexing.h:10-25).
template <class A>
IBM_HD int axis_interp_index(const A& a, int p)
{
    int j = axis_index(a, p);
    return j < 0 ? 0 : (j >= a.points ? a.points - 1 : j);
}

} // namespace ibm
