// Exact band spectrum of masks of rectangles and simple polygons on the device
// (FP64 evaluation, FP32 band boxes in the layout the field kernels read).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace huvk {

// cells: ncells x (x0, x1, y0, y1, w_re, w_im), nm, disjoint, w = value - background.
// X, Y: ncells x (2B + 1) axis coefficients (FP64 complex).
hipError_t launch_rect_tables(const double* cells, int64_t ncells, int n, double L, int B,
                              double2* X, double2* Y, hipStream_t s);
// coff: nclips + 1 cell offsets; bg: nclips x (re, im).  Writes the (2BB+1)^2
// box Mq of every clip (and its column-paired copy Mp when non-null).
hipError_t launch_rect_band(const double2* X, const double2* Y, const int64_t* coff,
                            const double* bg, int nclips, int n, int B, int BB, float2* Mq,
                            float2* Mp, hipStream_t s);

// Largest band half-width of the polygon pass: one edge's two order tables
// (2 x (2B + 1) entries of 32 bytes) must fit 64 KiB of LDS.
constexpr int kPolyMaxBand = 511;

// Simple polygons added onto the box written by launch_rect_band.  edges:
// (dx, dy, cx, cy, re, im) per edge, (re, im) = orientation sign x (value -
// background); eoff: nclips + 1 edge offsets; pdc: per clip sum of
// (value - background) x area (nm^2).
hipError_t launch_poly_band(const double* edges, const int64_t* eoff, const double* pdc,
                            int nclips, int n, double L, int B, int BB, float2* Mq, float2* Mp,
                            hipStream_t s);

}  // namespace huvk
