// Fused six-kernel evaluation schedule for power-of-two grids.
//
// Spectral data between the z-line kernels and the plane kernels lives in a
// chunked, parity-split layout.  The y axis is split by parity (the radix-2
// decimation is done inside the z-line kernels): for a z-transformed field
// x~(ix, iy, kz) the two parts are
//   a_r(kz, ix, m) = x~(ix, m, kz) + (-1)^r x~(ix, m + Q, kz),  Q = Ny/2
// (the r = 1 part also twiddled), and x the same way by kx parity p (half
// rows i and i + Nx/2 combine to p = 0 and, twiddled, to p = 1).  Element
// (kz, r, i, p, m) is stored at
//   ((i * (Q/8) + m/8) * (Nz/2+1) + kz) * 32 + r * 16 + p * 8 + m % 8,
// so both parts of 8 consecutive m share a 256-byte row per kz (128 bytes per
// r in FP32) and the data of the 32 lines at x = i, i + Nx/2 and y = m .. m+7,
// m+Q .. m+Q+7 form one contiguous chunk.  A plane kernel owns one
// (kz, r, p) sub-plane and reads or writes half segments; the z-line kernels
// own chunks and read and write whole segments.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace cdft {

struct FastGeom {
  int nx, ny, nz, nzh, q, mb;  // q = ny/2, mb = q/16
  int ncu;                     // compute units of the device (blocks of a balanced launch)
  long N, M;
  double dkx, dky, dkz, dv, inv_n, R;
};

// True when the fused schedule has kernels for this grid.
bool fast_supported(int nx, int ny, int nz, int precision);
FastGeom make_fast_geom(const Geom& g);

// FP64 and FP32 twiddle tables of length kTwLen.
void fast_twiddles(void** tw64, void** tw32);

// Mean-field weight: real parts of the FP64 spectrum in plane-major order
// [kz][kx][ky] scattered into the plane kernels' sub-plane order
// [kz][r][p][s][k] (kx = 2 k + p, ky = r + 2 s), on the device.
void launch_fast_mf_layout(const cx<double>* spec, double* out, const Geom& g, hipStream_t stream);

// Blocks of the point kernel (rows of its partial sums) for FP64 or FP32 storage.
int fast_point_blocks(const FastGeom& g, bool fp64);
int fast_bwd_blocks(const FastGeom& g);

// z-R2C of rho and its y- and x-parity split (the x twiddle left to
// kf_rho_xy) into the channel layout; D0 (nx * ny doubles) receives the FP64
// kz = 0 plane (the line sums) in the same row order.
template <typename S>
void launch_fast_zfwd(const double* rho, cx<S>* out, double* D0, const FastGeom& g, const void* tw64, AliasMap amap,
                      hipStream_t st);
// FP64 spectrum rho_hat of the y-parity parts (M complex doubles, order
// [kz][r][p][s][k], kx = 2 k + p, ky = r + 2 s), read by the forward plane kernel, without
// its (0, 0, 0) coefficient; sum_out[0] receives that coefficient (sum of rho).
template <typename S>
void launch_fast_rho_xy(const cx<S>* A0, const double* D0, void* XR, double* sum_out, const FastGeom& g,
                        const void* tw64, hipStream_t st);
// FP64 radial FMT weights (w3, w2) of every element in the plane kernels'
// order (M complex doubles), computed once per context.
void launch_fast_wtable(void* WT, const FastGeom& g, hipStream_t st);
template <typename S, typename CV>
void launch_fast_plane_fwd(const void* XR, const void* WT, cx<S>* SF, const double* mfB, const FastGeom& g,
                           const void* tw64, const void* twv, ChannelSet cs, hipStream_t st);
template <typename S, typename CV>
void launch_fast_zpoint(cx<S>* SF, const double* rho, const double* vext, double* lf, double* partials,
                        const FastGeom& g, const void* tw64, const void* twv, ChannelSet cs, FmtParams fp, double mu,
                        AliasMap amap, const double* shift_sum, DcTable dc, hipStream_t st);
template <typename S, typename CV>
void launch_fast_plane_bwd(const cx<S>* SB, const void* WT, cx<S>* P, const FastGeom& g, const void* twv,
                           hipStream_t st);
// With fin_rows > 0 one extra block also reduces fin_rows rows of partial
// sums into res[0..7] (as launch_finalize does).
template <typename S, typename CV>
void launch_fast_zinv(const cx<S>* P, double* out, const double* add, double scale, const FastGeom& g,
                      const void* twv, const double* partials, int fin_rows, double* res, hipStream_t st);

// Generic FP64 3-D transforms of `count` real fields (N apart) on the fused
// kernels: forward into whole spectra (M apart) in the sub-plane order
// [kz][r][p][s][k] (kx = 2 k + p, ky = r + 2 s), unnormalised; and the inverse,
// out = scale * c2r(spec) + add (add nullable, N apart).  tmp holds count * M
// complex values of the channel layout.  The real fields are FP64 or FP32
// (stored in that type; the arithmetic is FP64 throughout).
template <typename I>
void launch_fields_forward(const I* in, cx<double>* spec, cx<double>* tmp, int count, const FastGeom& g,
                           const void* tw64, hipStream_t st);
template <typename O>
void launch_fields_inverse(const cx<double>* spec, O* out, const double* add, double scale, cx<double>* tmp,
                           int count, const FastGeom& g, const void* tw64, hipStream_t st);

}  // namespace cdft
