// Kernel-side shared types and host launchers of the evaluation pipeline.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_lds.hpp"
#include "fmt_point.hpp"

namespace cdft {

// Grid geometry as seen by the kernels (grid.hpp:26-63 of the reference).
struct Geom {
  int nx, ny, nz, nzh;
  long N, M, nxy;
  double dx, dv;
  double dkx, dky, dkz;
  double inv_n;
  double R;  // hard-sphere radius of the species
};

// Forward channels: 11 FMT channels when `hs`, then the mean-field channel
// when `mf`; nf = 11*hs + mf.  Backward channels nb = 11*hs.
struct ChannelSet {
  int hs, mf, nf, nb;
};

// Density parametrisation applied on load (minimization.hpp:45-59):
// kind 0: rho = x; 1: rho = rho_min + x^2; 2: rho_min + range x^2/(1+x^2).
struct AliasMap {
  int kind;
  double rho_min, range;
  __host__ __device__ double apply(double x) const {
    if (kind == 0) return x;
    const double x2 = x * x;
    return kind == 1 ? rho_min + x2 : rho_min + range * x2 / (1.0 + x2);
  }
  // d rho / d x (minimization.hpp:82-98)
  __host__ __device__ double deriv(double x) const {
    if (kind == 0) return 1.0;
    if (kind == 1) return 2.0 * x;
    const double x2 = x * x;
    return 2.0 * range * x / ((1.0 + x2) * (1.0 + x2));
  }
};

// Mean-density split: the transforms carry rho - c with c = shift_sum[0] / N
// (an FP64 sum of rho); each forward channel then gets back its exact k = 0
// contribution dc.v[ch] * c in FP64.  A null shift_sum means c = 0.
struct DcTable {
  double v[12];
};

// The mean split subtracts shift_sum[0] / N, or with sum_rows the sum of
// column 0 of launch_sum's sum_n rows, which every block reduces itself (block
// 0 stores it in sum_out[0], zeros in sum_out[1..7], as launch_finalize would).
template <typename S, typename C>
void launch_zr2c(const double* in, cx<S>* out, const Geom& g, const FftPlan& pz, const cx<C>* twz, AliasMap amap,
                 const double* shift_sum, hipStream_t st, const double* sum_rows = nullptr, int sum_n = 0,
                 double* sum_out = nullptr);
// z-R2C of `nfields` real fields (N apart) into spectra (M apart), one launch,
// no density map or mean split.
template <typename S, typename C>
void launch_zr2c_fields(const double* in, int nfields, cx<S>* out, const Geom& g, const FftPlan& pz,
                        const cx<C>* twz, hipStream_t st);
// With fin_rows set, block 0 also reduces fin_n rows of 8 partial sums written
// by an earlier kernel into fin_res (launch_finalize's result), first.
template <typename S, typename C>
void launch_zc2r(const cx<S>* Sp, int nch, double* out, const double* add, double scale, const Geom& g,
                 const FftPlan& pz, const cx<C>* twz, const double* shift_sum, DcTable dc, hipStream_t st,
                 const double* fin_rows = nullptr, int fin_n = 0, double* fin_res = nullptr);
// FP64 block sums of amap(in) into partials[b*8]; finalize with launch_finalize.
int sum_blocks(const Geom& g);
void launch_sum(const double* in, AliasMap amap, double* partials, const Geom& g, hipStream_t st);
template <typename S, typename C>
void launch_lines(cx<S>* Sp, long nlines, const FftPlan& p, const cx<C>* tw, int sg, hipStream_t st);
// FP32-stored lines in one launch: blocks entirely inside lines [lo0, lo1) in
// FP32 arithmetic, the others in FP64.
void launch_lines_split(cx<float>* Sp, long nlines, const FftPlan& p, const cx<double>* twh, const cx<float>* twl,
                        long lo0, long lo1, int sg, hipStream_t st);
template <typename S, typename C>
void launch_xlines(cx<S>* Sp, int nch, const Geom& g, const FftPlan& px, const cx<C>* twx, int sg, hipStream_t st);
// wt: FP64 radial weights (w3, w2) per element, from launch_wtable_general.
// The scalar channels' x transforms run in C, the vector and tensor ones in CV.
// With `red` (hard spheres) SF receives the reduced channel set of the kg_zpoint
// point kernel (zpoint_reduced): no v_z, T_zz, the mean field after T_yz.
template <typename S, typename C, typename CV>
void launch_xfwd(const cx<S>* S0, cx<S>* SF, const S* mf, const cx<double>* wt, const Geom& g, const FftPlan& px,
                 const cx<C>* twx, const cx<CV>* twv, ChannelSet cs, bool red, hipStream_t st);
// True when the point kernel of a z extent nz reads and writes the reduced
// channel sets (10 or 9 forward channels, 9 backward).
bool zpoint_reduced(int nz, const ChannelSet& cs);
template <typename S, typename C>
void launch_xbwd(const cx<S>* SB, cx<S>* SO, const cx<double>* wt, const Geom& g, const FftPlan& px,
                 const cx<C>* twx, int nb, hipStream_t st);
void launch_wtable_general(cx<double>* wt, const Geom& g, hipStream_t st);
template <typename S>
int zpoint_blocks(const Geom& g, const FftPlan& pz);
template <typename S, typename CF, typename CB>
void launch_zpoint(cx<S>* SF, const double* rho, const double* vext, double* local_force, double* partials,
                   const Geom& g, const FftPlan& pz, const cx<CF>* twf, const cx<CB>* twb, ChannelSet cs,
                   FmtParams fp, double mu, AliasMap amap, const double* shift_sum, DcTable dc, hipStream_t st);
// res[0..7] = deterministic compensated column sums of nrows rows of 8 doubles.
void launch_finalize(const double* partials, int nrows, double* res, hipStream_t st);

}  // namespace cdft
