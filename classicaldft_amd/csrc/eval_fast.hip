// Fused six-kernel evaluation schedule (gfx950) for power-of-two grids.
//
//   kf_zfwd       per line quad (the z lines at x = i, i + NX/2 and y = m, m + NY/2):
//                 z-R2C of rho (FP64), y- and x-parity split -> A0
//   kf_rho_xy     per (kz, r, p): y- and x-FFT of rho's sub-plane (FP64) -> rho_hat
//   kf_plane_fwd  per (kz, r, p): for each of the 10 plane channels: weight
//                 times rho_hat on the kx = 2k + p sub-plane, its x- and
//                 y-inverse (NX/2 x NY/2 points), written in the channel layout
//   kf_zpoint     per line quad: x-parity and y-parity combine + z-C2R of
//                 every channel (n_v,z from eta's spectrum, T_zz by the trace
//                 identity), pointwise Phi/dPhi + local terms + partial sums,
//                 z-R2C of the derivative channels + y- and x-parity split
//                 (in place, 9 out)
//   kf_plane_bwd  per (kz, r, p): y- and x-FFT of the 9 derivative channels'
//                 sub-planes, conj-weighted accumulation, x- and y-inverse of
//                 the sum -> partial
//   kf_zinv       per line quad: sum of the partials, x-parity and y-parity
//                 combine, z-C2R, * dV + local force           -> force
//
// Parity splits.  A transform of length 2L along y or x is split into two of
// length L by its first radix-2 stage, so that the plane kernels work on
// independent (kz, r, p) sub-planes: along y by ky parity r (rows y and
// y + NY/2 combine to r = 0 and (difference) * exp(-2 pi i y / NY) to r = 1,
// in the z-line kernels), along x by kx parity p.  For x the inverse
// transforms end with
//   x(i) = Z0(i) + exp(2 pi i i / NX) Z1(i),  x(i + NX/2) = Z0(i) - exp(...) Z1(i)
// (Zp the NX/2-point inverse of the kx = 2k + p values) and the forward one
// starts with
//   d0(i) = d(i) + d(i + NX/2),  d1(i) = exp(-2 pi i i / NX) (d(i) - d(i + NX/2)),
// both done by the z-line kernels, whose blocks own the lines at x = i and
// i + NX/2 (one twiddle per block).  So no plane block reads another's
// sub-plane or transforms more than its own quarter of a kz plane.
//
// Line transforms run in slots of one wavefront (slot_fft.hpp); a block
// synchronises only where data crosses wavefronts (the plane transposes and
// the chunk staging of the z-line kernels).
//
// Precision: with FP32 storage (mixed) the forward transform of rho and the
// inverse transforms of the scalar channels (eta, n2, mean field) are computed
// in FP64 -- a smooth density concentrates its spectrum in a few large low-k
// values whose FP32 rounding would add up coherently in Omega -- while the
// vector and tensor channels and the whole backward path use FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "cdft.h"
#include "device_common.hpp"
#include "eval_fast.hpp"
#include "fmt_point.hpp"
#include "slot_fft.hpp"

namespace cdft {

namespace {

constexpr int kBT = 1024;            // threads of a plane kernel block
// Waves of a kf_zpoint block (one line pair each): 4 with FP32 storage (two
// blocks share a CU and overlap each other's barriers), 8 with FP64 storage,
// whose 4-wave block alone would fill most of a CU's LDS.
template <typename S>
__host__ __device__ constexpr int zp_waves() {
  return sizeof(S) == 8 ? 8 : 4;
}

// Blocks are dispatched round-robin over the 8 XCDs (block b on XCD b % 8);
// this maps consecutive logical blocks onto one XCD so that they share its
// L2.  XCD x runs n/8 blocks, plus one for x < n % 8; its blocks take the
// logical numbers after those of XCDs 0 .. x-1 (a bijection of [0, n)).
__host__ __device__ constexpr int xcd_remap(int b, int n) {
  return (b & 7) * (n >> 3) + ((b & 7) < (n & 7) ? (b & 7) : (n & 7)) + (b >> 3);
}

__host__ __device__ constexpr int align16(int b) { return (b + 15) & ~15; }

// Channel-layout offset of element (kz, r, i, p, m) of the arrays between
// the z-line kernels and the plane kernels (rho's parts, the channels, the
// backward partials): half row i of kx-parity part p.  A 32-element row
// [r][p][m % 8] holds both parts of 8 columns, so a z-line block, which
// combines the parts of its columns, reads or writes one 128-byte segment per
// (kz, r); a plane block reads or writes half segments.
__device__ __forceinline__ long chan_idx(int kz, int r, int i, int p, int m, int mb8, int nzh) {
  return ((static_cast<long>(i) * mb8 + (m >> 3)) * nzh + kz) * 32 + r * 16 + p * 8 + (m & 7);
}

// Plane kernels' order of a half spectrum: element (kx = 2 k + p, ky = r + 2 s, kz)
// of sub-plane u = (kz * 2 + r) * 2 + p at u * Q * NX/2 + s * NX/2 + k.
__host__ __device__ inline long hp_idx(int kz, int r, int p, int s, int k, int q, int nxh) {
  return ((((static_cast<long>(kz) * 2 + r) * 2 + p) * q + s) * nxh) + k;
}

// Channel bookkeeping of the fused schedule.  Two identities of the FMT
// weights (weights.hpp:110-137,164-178) spare four plane transforms per evaluation:
//   n_v,z = i kz w3 rho_hat = i kz eta_hat, which the z-line kernel forms
//     from eta's z-spectrum (kz is fixed along a z-line's spectrum), and
//   T_zz = n2 - T_xx - T_yy, since sum_i (aik2 k_i^2 + w3/R) = w2;
// and, for the back-convolution (hard_sphere.hpp:115-134), their adjoints:
//   dPhi/dT_zz folds into the n2, T_xx and T_yy derivatives before the
//     transforms, and the z-vector derivative's z-spectrum, times -i kz,
//     into eta's.
// So the forward plane kernel produces 10 channels (eta, n2, mf, v_x, v_y,
// T_xx, T_xy, T_xz, T_yy, T_yz) and the backward one consumes 9.
//
// Forward slot positions of the z-line kernel (round pos / NS, slot pos % NS):
// the channel staged there, kVzFromEta for the slot that builds n_v,z from
// eta's staged spectrum (round 0, slot 0), -1 for an idle slot.  The FP64
// (scalar) channels eta, n2 and the mean field come first.
//   nf 12: eta n2 mf vz* vx vy Txx Txy Txz Tyy Tyz      nf 11: eta n2 vz* vx vy Txx Txy Txz Tyy Tyz
constexpr int kVzFromEta = -2;
__host__ __device__ constexpr int zf_chan(int nf, int pos) {
  return nf == 1    ? (pos == 0 ? 11 : -1)
         : nf == 12 ? (pos < 2 ? pos : pos == 2 ? 11 : pos == 3 ? kVzFromEta : pos < 6 ? pos - 2 : pos < 11 ? pos - 1 : -1)
         : nf == 11 ? (pos < 2 ? pos : pos == 2 ? kVzFromEta : pos < 5 ? pos - 1 : pos < 10 ? pos : -1)
                    : -1;
}
// positions holding a channel value
__host__ __device__ constexpr int zf_count(int nf) { return nf == 12 ? 11 : (nf == 11 ? 10 : nf); }
// position of channel ch (-1: T_zz, formed pointwise)
__host__ __device__ constexpr int zpos(int nf, int ch) {
  return nf == 1    ? (ch == 11 ? 0 : -1)
         : nf == 12 ? (ch < 2 ? ch : ch == 11 ? 2 : ch == 4 ? 3 : ch < 4 ? ch + 2 : ch < 10 ? ch + 1 : -1)
                    : (ch < 2 ? ch : ch == 4 ? 2 : ch < 4 ? ch + 1 : ch < 10 ? ch : -1);
}
// Backward positions: the derivative channel R2C-transformed at position b;
// b = 1 (the z vector) is folded into b = 0 (eta) in the same round.
constexpr int kZB = 10;
__host__ __device__ constexpr int zb_chan(int b) { return b == 0 ? 0 : b == 1 ? 4 : b < 5 ? b - 1 : b < 10 ? b : -1; }
__host__ __device__ constexpr bool zscalar(int ch) { return ch == 0 || ch == 1 || ch == 11; }
// Storage slot of channel ch (in the channel arrays and the k = 0 table): the
// mean field is slot 11 after the FMT channels, or slot 0 without them.
__host__ __device__ constexpr int chan_store(int nf, int ch) { return nf == 1 ? 0 : ch; }
// FP64 rounds: those with a scalar channel (or n_v,z, formed next to eta)
__host__ __device__ constexpr bool zround64(int nf, int ns, int jr) {
  for (int g = 0; g < ns; ++g) {
    const int c = zf_chan(nf, jr * ns + g);
    if (c == kVzFromEta || (c >= 0 && zscalar(c))) return true;
  }
  return false;
}
// Plane channels: forward (eta, n2, mf first), backward (v_z and T_zz folded away).
__host__ __device__ constexpr int fwd_nplane(int nf) { return nf == 12 ? 10 : (nf == 11 ? 9 : nf); }
__host__ __device__ constexpr int fwd_channel(int nf, int pos) {
  return nf == 1 ? 11 : (nf == 12 ? (pos < 2 ? pos : pos == 2 ? 11 : pos < 5 ? pos - 1 : pos) : (pos < 4 ? pos : pos + 1));
}
constexpr int kBwdPlaneCh = 9;
__host__ __device__ constexpr int bwd_channel(int i) { return i < 4 ? i : i + 1; }

template <typename C>
__device__ __forceinline__ cx<C> cneg(cx<C> a) {
  return {-a.re, -a.im};
}

// Bytes of a twiddle table in LDS (kTwLen complex values).
template <typename C>
__host__ __device__ constexpr int tw_bytes() {
  return kTwLen * static_cast<int>(sizeof(cx<C>));
}
// Block-wide copy of N elements by NT threads: element i = tid + NT k is
// load(i), then store(i, value).  Every load of a thread is issued before its
// first store, so the loads' memory latencies overlap; a rolled loop would
// wait for each load in turn.
template <int NT, int N, typename T, typename LD, typename ST>
__device__ __forceinline__ void block_copy(LD&& load, ST&& store) {
  constexpr int K = (N + NT - 1) / NT;
  T r[K];
  const int tid = static_cast<int>(threadIdx.x);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = tid + NT * k;
    if (N % NT == 0 || i < N) r[k] = load(i);
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = tid + NT * k;
    if (N % NT == 0 || i < N) store(i, r[k]);
  }
}

// Copies a twiddle table into LDS with NT threads (the caller synchronises
// before first use): twiddle reads then cost an LDS access instead of an
// L1/L2 round trip.
template <typename C, int NT>
__device__ __forceinline__ cx<C>* stage_twiddles(unsigned char* dst, const cx<C>* __restrict__ src) {
  cx<C>* d = reinterpret_cast<cx<C>*>(dst);
  constexpr int n = kTwLen;
  block_copy<NT, n, cx<C>>([&](int i) { return src[i]; }, [&](int i, cx<C> v) { d[i] = v; });
  return d;
}
// Copies only the length-L table (L values) and returns it.
template <typename C, int NT, int L>
__device__ __forceinline__ cx<C>* stage_twiddle_table(unsigned char* dst, const cx<C>* __restrict__ src) {
  cx<C>* d = reinterpret_cast<cx<C>*>(dst);
  const cx<C>* s = tw_for<L>(src);
  block_copy<NT, L, cx<C>>([&](int i) { return s[i]; }, [&](int i, cx<C> v) { d[i] = v; });
  return d;
}

// ============================================================ kf_zfwd
template <int NZ, typename S>
__host__ __device__ constexpr int zfwd_tw_off() {
  return align16((NZ / 2 + 1) * 32 * static_cast<int>(sizeof(cx<double>))) +
         16 * slot_buf_len<NZ>() * static_cast<int>(sizeof(cx<double>));
}
template <int NZ, typename S>
__host__ __device__ constexpr int zfwd_smem() {
  return zfwd_tw_off<NZ, S>() + tw_bytes<double>();
}

// Block (2 NZ threads, 16 slots): half row i and 8 columns m = 8 mc + c, the
// line pairs at x = i and x = i + NX/2.  z-R2C (FP64) of every line pair, the
// y-parity split, then the x-parity split of the half row: part p = 0 is the
// sum, p = 1 the difference of the two lines (its twiddle exp(-2 pi i i / NX)
// is applied by kf_rho_xy, in FP64).  The block's output is one channel-layout
// chunk.  The kz = 0 coefficients (the line sums, which carry the mean density)
// are real; they are also written in FP64 to D0 in the same row order
// (D0[chunk * 32 + r * 16 + p * 8 + c]): kf_rho_xy takes the kz = 0 plane from
// there, so its FP32 rounding never meets the mean.  Field f = blockIdx.y
// reads rho + f N and writes out + f M (D0 nullable: FP64 storage needs none).
// The input fields are FP64, or FP32 (I = float).
template <typename S, int NZ, typename I = double>
__global__ void __launch_bounds__(2 * NZ) kf_zfwd(const I* __restrict__ rho, cx<S>* __restrict__ out, FastGeom g,
                                                  const cx<double>* __restrict__ tw, AliasMap amap,
                                                  double* __restrict__ D0) {
  constexpr int T = NZ / 8, NZH = NZ / 2 + 1, XS = slot_buf_len<NZ>();
  extern __shared__ __align__(16) unsigned char smem[];
  cx<double>* cb = reinterpret_cast<cx<double>*>(smem);  // [xh][r][c][kz]
  cx<double>* xbuf = reinterpret_cast<cx<double>*>(smem + align16(NZH * 32 * static_cast<int>(sizeof(cx<double>))));
  const int slot = threadIdx.x / T, lane = threadIdx.x % T;
  const int xh = slot >> 3, c = slot & 7;
  const int mb8 = g.q >> 3;
  const long blk = blockIdx.x;
  const int i = static_cast<int>(blk / mb8), mc = static_cast<int>(blk % mb8);
  const long la = static_cast<long>(i + xh * (g.nx / 2)) * g.ny + mc * 8 + c, lb = la + g.q;
  rho += blockIdx.y * g.N;
  out += blockIdx.y * g.M;
  const cx<double>* twl = stage_twiddles<double, 2 * NZ>(smem + zfwd_tw_off<NZ, S>(), tw);
  // every load is issued before the first use, so their latencies overlap
  I ra[8], rb[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    ra[q] = rho[la * NZ + lane + T * q];
    rb[q] = rho[lb * NZ + lane + T * q];
  }
  __syncthreads();
  cx<double> v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = {amap.apply(static_cast<double>(ra[q])), amap.apply(static_cast<double>(rb[q]))};
  cx<double>* xb = xbuf + slot * XS;
  slot_fft<double, NZ, T, -1>(v, xb, lane, tw_for<NZ>(twl));
  wave_sync();
#pragma unroll
  for (int e = 0; e < 8; ++e) xb[xpad(lane + T * e)] = v[e];
  wave_sync();
#pragma unroll
  for (int e = 0; e <= 4; ++e) {
    const int k = lane + T * e;
    if (k <= NZ / 2) {
      // spectra of the packed lines (Z = a + i b), then the y-parity split
      const cx<double> zk = v[e], zm = xb[xpad((NZ - k) & (NZ - 1))];
      const cx<double> A = {0.5 * (zk.re + zm.re), 0.5 * (zk.im - zm.im)};
      const cx<double> B = {0.5 * (zk.im + zm.im), -0.5 * (zk.re - zm.re)};
      cb[((xh * 2 + 0) * 8 + c) * NZH + k] = cadd(A, B);
      cb[((xh * 2 + 1) * 8 + c) * NZH + k] = csub(A, B);
    }
  }
  __syncthreads();
  // row element j = kz * 32 + r * 16 + p * 8 + c of the chunk
  cx<S>* o = out + blk * static_cast<long>(NZH * 32);
  constexpr int K = (NZH * 32 + 2 * NZ - 1) / (2 * NZ);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int j = static_cast<int>(threadIdx.x) + 2 * NZ * k;
    if (j < NZH * 32) {
      const int kz = j >> 5, r = (j >> 4) & 1, p = (j >> 3) & 1, cc = j & 7;
      const cx<double> a = cb[(r * 8 + cc) * NZH + kz], b = cb[((2 + r) * 8 + cc) * NZH + kz];
      const cx<double> w = p ? csub(a, b) : cadd(a, b);
      o[j] = cvt<S>(w);
      if (kz == 0 && D0) D0[blk * 32 + j] = w.re;
    }
  }
}

// ============================================================ kf_zinv
// A block (2 NZ threads, 16 slots) owns half row i and 8 consecutive m: the
// line pairs (x, m) / (x, m + NY/2) at x = i and x = i + NX/2, whose inputs
// are one channel-layout row per kz.
template <int NZ>
__host__ __device__ constexpr int zinv_threads() {
  return 2 * NZ;
}
template <int NZ, typename C>
__host__ __device__ constexpr int zinv_tw_off() {
  return align16((NZ / 2 + 1) * 32 * static_cast<int>(sizeof(cx<C>))) +
         16 * slot_buf_len<NZ>() * static_cast<int>(sizeof(cx<C>));
}
template <int NZ, typename C>
__host__ __device__ constexpr int zinv_smem() {
  return zinv_tw_off<NZ, C>() + tw_bytes<C>();
}

// Builds the packed C2R input Z[lane + T q] of line pair (m, m + Q) from the
// two y-parity parts b0, b1 held at part(r, kz): lines m = b0 + b1,
// m + Q = b0 - b1; kz = 0 and Nz/2 are projected onto their real parts
// (the Hermitian extension of the half spectrum).
template <typename C, int NZ, typename F>
__device__ __forceinline__ void build_c2r(cx<C> (&v)[8], int lane, F&& part) {
  constexpr int T = NZ / 8;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = lane + T * q;
    const bool mirror = k > NZ / 2;
    const int kk = mirror ? NZ - k : k;
    const cx<C> b0 = part(0, kk), b1 = part(1, kk);
    cx<C> A = cadd(b0, b1), B = csub(b0, b1);
    if (kk == 0 || kk == NZ / 2) {
      A.im = C(0);
      B.im = C(0);
    }
    if (mirror) {
      A.im = -A.im;
      B.im = -B.im;
    }
    v[q] = {A.re - B.im, A.im + B.re};
  }
}

__host__ __device__ inline bool bwd_unit_split(int unit, int nunits, int nblk);

// With fin_rows > 0 the grid has one block more, block 0, which does the
// final reduction of the point kernel's partial sums (reduce_rows) while the
// other blocks transform: the evaluation then needs no separate finalisation
// launch, and the reduction, dispatched first, overlaps the whole kernel.
template <typename S, typename C, int NZ, typename O = double>
__global__ void __launch_bounds__(zinv_threads<NZ>()) kf_zinv(const cx<S>* __restrict__ P, O* __restrict__ out,
                                                               const double* __restrict__ add, double scale,
                                                               FastGeom g, const cx<C>* __restrict__ tw,
                                                               int bwd_blocks, const double* __restrict__ partials,
                                                               int fin_rows, double* __restrict__ res) {
  constexpr int NT = zinv_threads<NZ>(), T = NZ / 8, NZH = NZ / 2 + 1, XS = slot_buf_len<NZ>();
  extern __shared__ __align__(16) unsigned char smem[];
  if (fin_rows > 0 && blockIdx.x == 0) {
    double* sh = reinterpret_cast<double*>(smem);
    reduce_rows<NT>(partials, fin_rows, res, sh, sh + NT);
    return;
  }
  cx<C>* cb = reinterpret_cast<cx<C>*>(smem);  // [xh][r][m8][kz]
  cx<C>* xbuf = reinterpret_cast<cx<C>*>(smem + align16(NZH * 32 * static_cast<int>(sizeof(cx<C>))));
  const int slot = threadIdx.x / T, lane = threadIdx.x % T;
  const int xh = slot >> 3, ms = slot & 7;
  const int mb8 = g.q >> 3;
  const long blk = static_cast<long>(blockIdx.x) - (fin_rows > 0 ? 1 : 0);
  const int i = static_cast<int>(blk / mb8), mc = static_cast<int>(blk % mb8);
  const int m = mc * 8 + ms;
  const long la = static_cast<long>(i + xh * (g.nx / 2)) * g.ny + m, lb = la + g.q;
  // partial sets of the backward plane kernel at set * M, in the channel
  // layout: the block's rows hold parts p = 0, 1 of half row i, columns 8 mc ..
  const cx<S>* p0 = P + blockIdx.y * g.M + blk * static_cast<long>(NZH * 32);
  out += blockIdx.y * g.N;
  if (add) add += blockIdx.y * g.N;
  const cx<C>* twl = stage_twiddles<C, NT>(smem + zinv_tw_off<NZ, C>(), tw);
  const int nunits = 4 * NZH;
  {
    // exp(2 pi i i / NX), read from global memory (the LDS copy is not yet published)
    cx<C> w = tw[g.nx - 4 + i];
    w.im = -w.im;
    constexpr int NPR = NZH * 16, K = (NPR + NT - 1) / NT;  // (kz, r, column) pairs of parts
    cx<C> e[K], o[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int j = static_cast<int>(threadIdx.x) + NT * k;
      if (j < NPR) {
        const int kz = j >> 4, r = (j >> 3) & 1, c = j & 7;
        const int u = (kz * 2 + r) * 2;  // sub-plane (kz, r, p = 0)
        const int off = kz * 32 + r * 16 + c;
        e[k] = cvt<C>(p0[off]);
        o[k] = cvt<C>(p0[off + 8]);
        if (bwd_unit_split(u, nunits, bwd_blocks)) e[k] = cadd(e[k], cvt<C>(p0[g.M + off]));
        if (bwd_unit_split(u + 1, nunits, bwd_blocks)) o[k] = cadd(o[k], cvt<C>(p0[g.M + off + 8]));
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int j = static_cast<int>(threadIdx.x) + NT * k;
      if (j < NPR) {
        const int kz = j >> 4, rc = j & 15;  // rc = r * 8 + column
        const cx<C> wo = cmul(o[k], w);
        cb[rc * NZH + kz] = cadd(e[k], wo);
        cb[(16 + rc) * NZH + kz] = csub(e[k], wo);
      }
    }
  }
  // the local force of the lane's points, loaded while the transform runs
  double ad[2][8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int z = lane + T * e;
    ad[0][e] = add ? add[la * NZ + z] : 0.0;
    ad[1][e] = add ? add[lb * NZ + z] : 0.0;
  }
  __syncthreads();
  cx<C> v[8];
  build_c2r<C, NZ>(v, lane, [&](int r, int kz) { return cb[((xh * 2 + r) * 8 + ms) * NZH + kz]; });
  slot_fft<C, NZ, T, +1>(v, xbuf + slot * XS, lane, tw_for<NZ>(twl));
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int z = lane + T * e;
    out[la * NZ + z] = static_cast<O>(static_cast<double>(v[e].re) * scale + ad[0][e]);
    out[lb * NZ + z] = static_cast<O>(static_cast<double>(v[e].im) * scale + ad[1][e]);
  }
}

// ============================================================ kf_zpoint
template <int NZ, typename S>
__host__ __device__ constexpr int zpoint_stage_bytes() {
  return align16((64 / (NZ / 8)) * 2 * zp_waves<S>() * (NZ / 2 + 1) * static_cast<int>(sizeof(cx<S>)));
}
template <int NZ, typename S>
__host__ __device__ constexpr int zpoint_tw_off() {
  return zpoint_stage_bytes<NZ, S>() +
         zp_waves<S>() * (64 / (NZ / 8)) * slot_buf_len<NZ>() * static_cast<int>(sizeof(cx<double>));
}
// the length-NZ twiddle tables of the FP64 and of the CV rounds
template <int NZ, typename S, typename CV>
__host__ __device__ constexpr int zpoint_smem() {
  return zpoint_tw_off<NZ, S>() + NZ * static_cast<int>(sizeof(cx<double>) + sizeof(cx<CV>));
}

// One forward round: slot `slot` of wave `wave` transforms channel position
// jr*NS + slot of line pair `wave` (C2R), then the wave redistributes the
// results so that lane lid holds points z = lid + 64 j of both lines.  The
// n_v,z slot of round 0 transforms i kz times eta's staged spectrum.
template <typename C, typename S, int NZ, int NF, int JR, int PPL, int NWV>
__device__ __forceinline__ void zpoint_fwd_round(S (&wv)[NWV][PPL], const cx<S>* stg, unsigned char* xwave, int wave,
                                                 int lid, const cx<C>* __restrict__ tw, double dkz) {
  constexpr int T = NZ / 8, NS = 64 / T, NZH = NZ / 2 + 1, XS = slot_buf_len<NZ>();
  const int slot = lid / T, lane = lid % T;
  cx<C>* xw = reinterpret_cast<cx<C>*>(xwave);
  const bool vz = NF >= 11 && JR == 0 && slot == zpos(NF, 4);
  const int src = vz ? zpos(NF, 0) : slot;
  const C dk = static_cast<C>(dkz);
  cx<C> v[8];
  build_c2r<C, NZ>(v, lane, [&](int r, int kz) {
    const cx<C> b = cvt<C>(stg[((src * 2 + r) * zp_waves<S>() + wave) * NZH + kz]);
    const C k = static_cast<C>(kz) * dk;
    return vz ? cx<C>{-b.im * k, b.re * k} : b;
  });
  slot_fft<C, NZ, T, +1>(v, xw + slot * XS, lane, tw);
  wave_sync();
#pragma unroll
  for (int e = 0; e < 8; ++e) xw[slot * (8 * T) + lane + T * e] = v[e];
  wave_sync();
#pragma unroll
  for (int gs = 0; gs < NS; ++gs) {
    const int pos = JR * NS + gs;
    if (pos < zf_count(NF)) {
#pragma unroll
      for (int j = 0; j < PPL / 2; ++j) {
        const cx<C> t = xw[gs * (8 * T) + lid + 64 * j];
        wv[pos][2 * j] = static_cast<S>(t.re);
        wv[pos][2 * j + 1] = static_cast<S>(t.im);
      }
    }
  }
}

// One backward round: the derivative channels at positions JR*NS + slot
// (wv[b] holds channel zb_chan(b)) of line pair `wave` go through the z-R2C
// and the y-parity split into the staging buffer.
template <typename C, typename S, int NZ, int JR, int PPL, int NWV>
__device__ __forceinline__ void zpoint_bwd_round(const S (&wv)[NWV][PPL], cx<S>* stg, unsigned char* xwave, int wave,
                                                 int lid, const cx<C>* __restrict__ tw) {
  constexpr int T = NZ / 8, NS = 64 / T, NZH = NZ / 2 + 1, XS = slot_buf_len<NZ>();
  const int slot = lid / T, lane = lid % T;
  cx<C>* xw = reinterpret_cast<cx<C>*>(xwave);
  wave_sync();
#pragma unroll
  for (int gs = 0; gs < NS; ++gs) {
    const int c = JR * NS + gs;
#pragma unroll
    for (int j = 0; j < PPL / 2; ++j) {
      cx<C> t = {C(0), C(0)};
      if (c < kZB) t = {static_cast<C>(wv[c][2 * j]), static_cast<C>(wv[c][2 * j + 1])};
      xw[gs * (8 * T) + lid + 64 * j] = t;
    }
  }
  wave_sync();
  cx<C> v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = xw[slot * (8 * T) + lane + T * q];
  wave_sync();
  slot_fft<C, NZ, T, -1>(v, xw + slot * XS, lane, tw);
  wave_sync();
#pragma unroll
  for (int e = 0; e < 8; ++e) xw[slot * XS + xpad(lane + T * e)] = v[e];
  wave_sync();
#pragma unroll
  for (int e = 0; e <= 4; ++e) {
    const int k = lane + T * e;
    if (k <= NZ / 2) {
      const cx<C> zk = v[e], zm = xw[slot * XS + xpad((NZ - k) & (NZ - 1))];
      const cx<C> A = {C(0.5) * (zk.re + zm.re), C(0.5) * (zk.im - zm.im)};
      const cx<C> B = {C(0.5) * (zk.im + zm.im), C(-0.5) * (zk.re - zm.re)};
      stg[((slot * 2 + 0) * zp_waves<S>() + wave) * NZH + k] = cvt<S>(cadd(A, B));
      stg[((slot * 2 + 1) * zp_waves<S>() + wave) * NZH + k] = cvt<S>(csub(A, B));
    }
  }
}

// FM: the FMT model fixed at compile time (its branches fold away), or
// kFmtRuntime to read it from the parameters.
constexpr int kFmtRuntime = -100;
template <typename S, typename CV, int NZ, int NF, int NB, int FM>
__global__ void __launch_bounds__(64 * zp_waves<S>(), 2) kf_zpoint(cx<S>* SF, const double* __restrict__ rho_in,
                                                 const double* __restrict__ vext, double* __restrict__ lf,
                                                 double* __restrict__ partials, FastGeom g,
                                                 const cx<double>* __restrict__ tw64, const cx<CV>* __restrict__ twv,
                                                 FmtParams fp, double mu, AliasMap amap,
                                                 const double* __restrict__ shift_sum, DcTable dc) {
  constexpr int T = NZ / 8, NS = 64 / T, NZH = NZ / 2 + 1, XS = slot_buf_len<NZ>();
  constexpr int kZW = zp_waves<S>(), kZP = 64 * kZW, MW = kZW / 2;  // waves, threads, columns
  constexpr int PPL = NZ / 32;  // points per lane: z = lid + 64 j of both lines
  constexpr int NWV = zf_count(NF) > 0 ? zf_count(NF) : 1;
  constexpr int NRF = (zf_count(NF) + NS - 1) / NS, NRB = NB > 0 ? (kZB + NS - 1) / NS : 0;
  static_assert(NB == 0 || (NF >= 11 && NWV >= kZB), "derivative channels need the FMT positions");
  constexpr bool HS = NF >= 11;
  constexpr bool MF = (NF == 12 || NF == 1);
  extern __shared__ __align__(16) unsigned char smem[];
  cx<S>* stg = reinterpret_cast<cx<S>*>(smem);  // [gs][r][line pair][kz]
  // the wave index is uniform over the wave: kept in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lid = threadIdx.x & 63;
  unsigned char* xwave = smem + zpoint_stage_bytes<NZ, S>() + wave * NS * XS * static_cast<int>(sizeof(cx<double>));
  // Block: half row i and the MW columns m0 .. m0 + MW - 1 (m0 = 8 mc + MW mq);
  // wave (xh, mm) = (wave / MW, wave % MW) owns the line pair at
  // x = i + xh NX/2, m = m0 + mm.  The blocks of one 8-column group are
  // consecutive, so they share an XCD's L2 and each channel row is fetched once.
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int mb8 = g.q >> 3;
  const int mq = logical % (8 / MW);
  const long grp = logical / (8 / MW);
  const int irow = static_cast<int>(grp / mb8), mc = static_cast<int>(grp % mb8);
  const int xh = wave / MW, mm = wave % MW;
  const int m = mc * 8 + MW * mq + mm;
  const long la = static_cast<long>(irow + xh * (g.nx / 2)) * g.ny + m, lb = la + g.q;
  // channel rows of half row i at columns m0 (part p = 0; part 1 is 8 further)
  const long cbase = (static_cast<long>(irow) * mb8 + mc) * (NZH * 32) + MW * mq;
  const double shift = shift_sum ? shift_sum[0] * g.inv_n : 0.0;
  S wv[NWV][PPL];
  // the length-NZ twiddle tables in LDS; the first barrier below publishes them
  const cx<double>* tw64l = stage_twiddle_table<double, kZP, NZ>(smem + zpoint_tw_off<NZ, S>(), tw64);
  const cx<CV>* twvl = stage_twiddle_table<CV, kZP, NZ>(
      smem + zpoint_tw_off<NZ, S>() + NZ * static_cast<int>(sizeof(cx<double>)), twv);
  if constexpr (NRF == 0) __syncthreads();

  // ---- forward rounds: stage NS channels of the block's line pairs, C2R.
  // The loads of round JR + 1 are issued into registers before round JR's
  // transforms, so their memory latency overlaps that work.
  // Pair j = tid + kZP k of a round: (slot gs, part r, column mm, kz).  It
  // loads the kx-parity parts E (p = 0) and O (p = 1) of half row i and stages
  // E + w O for the line pair at x = i and E - w O for x = i + NX/2,
  // w = exp(2 pi i i / NX); its offsets in a channel's chunk and in the
  // staging buffer are the same every round and computed once.
  constexpr int NSP = NS * NZH * 2 * MW, KPF = (NSP + kZP - 1) / kZP;
  constexpr int XH1 = MW * NZH;  // staging distance of the x = i + NX/2 line pair
  cx<S> pfe[KPF], pfo[KPF];
  int goff[KPF], soff[KPF], gsk[KPF];
#pragma unroll
  for (int k = 0; k < KPF; ++k) {
    const int j = min(static_cast<int>(threadIdx.x) + kZP * k, NSP - 1);
    const int c = j % MW, r = (j / MW) & 1, t = j / (2 * MW);
    const int kz = t % NZH, gs = t / NZH;
    goff[k] = (kz * 2 + r) * 16 + c;
    soff[k] = ((gs * 2 + r) * kZW + c) * NZH + kz;  // line pair (xh = 0, mm = c)
    gsk[k] = gs;
  }
  const cx<S>* sfc = SF + cbase;
  // (every lane loads, from channel 0 where its slot stages nothing: a load
  // under a condition would make the transforms wait for it)
  auto load_round = [&](int jr) {
#pragma unroll
    for (int k = 0; k < KPF; ++k) {
      const int ch = zf_chan(NF, jr * NS + gsk[k]);
      const cx<S>* src = sfc + static_cast<long>(ch >= 0 ? chan_store(NF, ch) : 0) * g.M + goff[k];
      pfe[k] = src[0];
      pfo[k] = src[8];
    }
  };
  // the combine runs in the round's arithmetic (FP64 for the scalar channels)
  const cx<double> wx = {tw64[g.nx - 4 + irow].re, -tw64[g.nx - 4 + irow].im};
  auto store_round = [&](auto jrc) {
    constexpr int jr = decltype(jrc)::value;
    using C = typename std::conditional<zround64(NF, NS, jr), double, CV>::type;
    const cx<C> w = cvt<C>(wx);
#pragma unroll
    for (int k = 0; k < KPF; ++k) {
      if (NSP % kZP == 0 || static_cast<int>(threadIdx.x) + kZP * k < NSP) {
        cx<S> a = {S(0), S(0)}, b = {S(0), S(0)};
        if (zf_chan(NF, jr * NS + gsk[k]) >= 0) {
          const cx<C> e = cvt<C>(pfe[k]), wo = cmul(cvt<C>(pfo[k]), w);
          a = cvt<S>(cadd(e, wo));
          b = cvt<S>(csub(e, wo));
        }
        stg[soff[k]] = a;
        stg[soff[k] + XH1] = b;
      }
    }
  };
  if constexpr (NRF > 0) {
    load_round(0);
    store_round(std::integral_constant<int, 0>{});
    __syncthreads();
  }
#define CDFT_ZF(JR)                                                                                   \
  if constexpr ((JR) < NRF) {                                                                         \
    if constexpr ((JR) + 1 < NRF) load_round((JR) + 1);                                               \
    if constexpr (zround64(NF, NS, JR))                                                               \
      zpoint_fwd_round<double, S, NZ, NF, JR, PPL>(wv, stg, xwave, wave, lid, tw64l, g.dkz);          \
    else                                                                                              \
      zpoint_fwd_round<CV, S, NZ, NF, JR, PPL>(wv, stg, xwave, wave, lid, twvl, g.dkz);               \
    __syncthreads();                                                                                  \
    if constexpr ((JR) + 1 < NRF) {                                                                   \
      store_round(std::integral_constant<int, (JR) + 1>{});                                           \
      __syncthreads();                                                                                \
    }                                                                                                 \
  }
  CDFT_ZF(0)
  CDFT_ZF(1)
  CDFT_ZF(2)
  CDFT_ZF(3)
  CDFT_ZF(4)
  CDFT_ZF(5)
#undef CDFT_ZF
  static_assert(NRF <= 6 && NRB <= 6, "too many z rounds");

  // ---- pointwise: one point at a time, channel values in and out of the
  // lane's register arrays
  // per-lane sums of its PPL points (plain FP64: a handful of terms); the
  // sums over lanes, waves and blocks follow, the last one compensated
  double acc[5] = {0, 0, 0, 0, 0};  // ideal, external, hs, mf, mass
  FmtParams fpk = fp;
  if constexpr (FM != kFmtRuntime) fpk.model = FM;
  if constexpr (FM == FMT_WHITEBEAR_II) {  // esFMT(1, -1) (models.hpp:326-357)
    fpk.A = 1.0;
    fpk.B = -1.0;
  }
  const double dv = g.dv;
#pragma unroll
  for (int pt = 0; pt < PPL; ++pt) {
    const int z = lid + 64 * (pt >> 1);
    const long p = ((pt & 1) ? lb : la) * NZ + z;
    S cur[NWV];
#pragma unroll
    for (int i = 0; i < NWV; ++i) {
      S t = wv[i][0];
#pragma unroll
      for (int j = 1; j < PPL; ++j) t = (pt == j) ? wv[i][j] : t;
      cur[i] = t;
    }
    S outv[kZB];
#pragma unroll
    for (int i = 0; i < kZB; ++i) outv[i] = S(0);
    const double rho = amap.apply(rho_in[p]);
    const double lr = flog(fmax(rho, 1e-300));
    double lfv = (lr - mu) * dv;  // ideal_gas.hpp:27-33
    acc[0] += rho * (lr - 1.0);
    acc[4] += rho;
    if (vext) {  // external_field.hpp:16-31
      const double V = vext[p];
      lfv += V * dv;
      acc[1] += rho * V;
    }
    if constexpr (MF) {  // mean_field.hpp:259-266 (self pair)
      const double conv = static_cast<double>(cur[zpos(NF, 11)]) + dc.v[chan_store(NF, 11)] * shift;
      lfv += conv * dv;
      acc[3] += 0.5 * rho * conv;
    }
    if constexpr (HS) {
      double w[11], gg[11];
#pragma unroll
      for (int i = 0; i < 10; ++i) w[i] = static_cast<double>(cur[zpos(NF, i)]) + dc.v[i] * shift;
      w[10] = w[1] - w[5] - w[8];  // T_zz = n2 - T_xx - T_yy
      // FP32 storage: the reduced White-Bear tensor algebra (fewer operations)
      const double phi = fmt_phi_grad_t<double, sizeof(S) == 4>(fpk, w, gg);
      acc[2] += phi;
      // adjoint of the T_zz identity
      gg[1] += gg[10];
      gg[5] -= gg[10];
      gg[8] -= gg[10];
#pragma unroll
      for (int b = 0; b < kZB; ++b) outv[b] = static_cast<S>(gg[zb_chan(b)]);
    }
    lf[p] = lfv;
    if constexpr (NB > 0) {
#pragma unroll
      for (int b = 0; b < kZB; ++b)
#pragma unroll
        for (int j = 0; j < PPL; ++j) wv[b][j] = (pt == j) ? outv[b] : wv[b][j];
    }
  }
  // Partial sums: each wave reduces its lanes and parks the totals in LDS;
  // the block's row is written after the next block barrier (the first
  // backward round's), so no barrier is spent on it.
  __shared__ double wsum[kZW][8];
  {
    // through the wave's own LDS exchange region (free between the rounds):
    // lanes 0..39 add 8 lanes' values of one sum each, lanes 0..4 the 8 of those
    double* red = reinterpret_cast<double*>(xwave);
    wave_sync();
#pragma unroll
    for (int i = 0; i < 5; ++i) red[i * 64 + lid] = acc[i];
    wave_sync();
    double t = 0.0;
    if (lid < 40) {
      const double* src = red + (lid >> 3) * 64 + (lid & 7) * 8;
#pragma unroll
      for (int j = 0; j < 8; ++j) t += src[j];
    }
    wave_sync();
    if (lid < 40) red[lid] = t;
    wave_sync();
    if (lid < 8) {
      double u = 0.0;
      if (lid < 5)
#pragma unroll
        for (int j = 0; j < 8; ++j) u += red[lid * 8 + j];
      wsum[wave][lid] = u * dv;
    }
  }
  auto write_row = [&]() {
    if (threadIdx.x < 8) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < kZW; ++w) t += wsum[w][threadIdx.x];
      partials[static_cast<long>(blockIdx.x) * 8 + threadIdx.x] = t;
    }
  };
  if constexpr (NRB == 0) {
    __syncthreads();
    write_row();
  }

  // ---- backward rounds: R2C of NS derivative channels, split, coalesced store
  // eta's derivative spectrum takes -i kz times the z vector's (positions 0 and 1)
  auto stage_out = [&](int jr) {
    // w' = exp(-2 pi i i / NX): part 0 = d(i) + d(i + NX/2), part 1 = w' (d(i) - d(i + NX/2))
    const cx<CV> wb = {static_cast<CV>(wx.re), static_cast<CV>(-wx.im)};
    const CV dkc = static_cast<CV>(g.dkz);
    cx<S> va[KPF], vb[KPF];
#pragma unroll
    for (int k = 0; k < KPF; ++k) {
      cx<CV> d0 = cvt<CV>(stg[soff[k]]), d1 = cvt<CV>(stg[soff[k] + XH1]);
      if (jr == 0 && gsk[k] == 0) {
        // position 1 of the same (r, line pair, kz)
        constexpr int G1 = 2 * kZW * NZH;
        const cx<CV> z0 = cvt<CV>(stg[soff[k] + G1]), z1 = cvt<CV>(stg[soff[k] + G1 + XH1]);
        const CV kk = static_cast<CV>(goff[k] >> 5) * dkc;  // kz = goff >> 5
        d0 = {d0.re + kk * z0.im, d0.im - kk * z0.re};
        d1 = {d1.re + kk * z1.im, d1.im - kk * z1.re};
      }
      va[k] = cvt<S>(cadd(d0, d1));
      vb[k] = cvt<S>(cmul(csub(d0, d1), wb));
    }
#pragma unroll
    for (int k = 0; k < KPF; ++k) {
      const int b = jr * NS + gsk[k];
      if ((NSP % kZP == 0 || static_cast<int>(threadIdx.x) + kZP * k < NSP) && b < kZB && b != 1) {
        cx<S>* o = SF + static_cast<long>(zb_chan(b)) * g.M + cbase + goff[k];
        o[0] = va[k];
        o[8] = vb[k];
      }
    }
  };
#define CDFT_ZB(JR)                                                                                           \
  if constexpr ((JR) < NRB) {                                                                                 \
    zpoint_bwd_round<CV, S, NZ, JR, PPL>(wv, stg, xwave, wave, lid, twvl);                                    \
    __syncthreads();                                                                                          \
    if constexpr ((JR) == 0) write_row();                                                                     \
    stage_out(JR);                                                                                            \
    __syncthreads();                                                                                          \
  }
  CDFT_ZB(0)
  CDFT_ZB(1)
  CDFT_ZB(2)
  CDFT_ZB(3)
  CDFT_ZB(4)
  CDFT_ZB(5)
#undef CDFT_ZB
}

// ============================================================ plane kernels
// A plane kernel block (1024 threads) works on one (kz, r, p) sub-plane of
// NX/2 half rows i and Q = NY/2 columns.  Slot shapes per phase (all 1024
// lanes busy at 128 x 128 grids):
//   y-FFT of the rows (NX/2 rows, Q points):      Q/4 lanes x 4 elements
//   x-FFT of the columns (Q columns, NX/2 points): NX/8 lanes x 4 elements
// FP64 slot exchanges go through the buffers one real part at a time, so a
// slot-buffer element is 8 bytes in every phase.
template <int NX, int Q>
struct PlaneCfg {
  static constexpr int BT = kBT;
  static constexpr int NY = 2 * Q, NXH = NX / 2;
  static constexpr int TYH = Q / 4, SBYH = BT / TYH, RPH = (NXH + SBYH - 1) / SBYH;
  static constexpr int TX = NX / 8, SBX = BT / TX, CPS = (Q + SBX - 1) / SBX;
  static constexpr int XBYTES = 8;
  static_assert(TYH <= 64 && TX <= 64 && TYH >= 1, "unsupported plane shape");
};

// Weight factors of one element in arithmetic C: w3, w2 and
// aik2 = (w2 - 3 w3/R) / k^2 (0 at k = 0), from which every FMT channel
// weight follows without a division (weights.hpp:91-184): ch 0 w3, 1 w2,
// 2..4 i k_a w3, 5..10 aik2 k_i k_j + delta_ij w3/R.
template <typename C>
struct WFac {
  C w3, w2, aik2;
};
// From the FP64 radial weights w = (w3, w2) of the weight table.
template <typename C>
__device__ __forceinline__ WFac<C> wfactors(cx<double> w, double kx, double ky, double kz, double invR) {
  const C k2 = static_cast<C>(kx * kx + ky * ky + kz * kz);
  const C num = static_cast<C>(w.im - 3.0 * w.re * invR);
  C a;
  if constexpr (sizeof(C) == 4)
    a = num * __builtin_amdgcn_rcpf(k2);  // 1-ulp reciprocal: FP32 weights carry 6e-8 anyway
  else
    a = num / k2;
  return {static_cast<C>(w.re), static_cast<C>(w.im), k2 > C(0) ? a : C(0)};
}

// FP64 radial FMT weights (w3, w2) (weights.hpp:91-108,159-162, times 1/N) of every
// element of the half spectrum, in the plane kernels' sub-plane order (hp_idx).
// They depend only on the grid and the radius: computed once.
__global__ void __launch_bounds__(256) kf_wtable(cx<double>* __restrict__ WT, FastGeom g) {
  const long n = g.M;
  const int nxh = g.nx / 2;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    const int k = static_cast<int>(i % nxh);
    const long t = i / nxh;
    const int s = static_cast<int>(t % g.q);
    const long u = t / g.q;
    const int kx = 2 * k + static_cast<int>(u & 1), r = static_cast<int>((u >> 1) & 1), kz = static_cast<int>(u >> 2);
    const double kxv = kvec(kx, g.nx, g.dkx), kyv = kvec(r + 2 * s, g.ny, g.dky), kzv = kz * g.dkz;
    double w3, w2;
    radial(sqrt(kxv * kxv + kyv * kyv + kzv * kzv), g.R, g.inv_n, &w3, &w2);
    WT[i] = {w3, w2};
  }
}
// Forward weight of channel ch for one element (weights.hpp:110-137,164-178):
// ch 0 w3, 1 w2, 2..4 i k_a w3, 5..10 aik2 k_i k_j + delta_ij w3/R, by cases
// on the channel (uniform over a block).
template <typename C>
__device__ __forceinline__ cx<C> chan_weight(int ch, C kx, C ky, C kz, const WFac<C>& f, C invR) {
  if (ch == 0) return {f.w3, C(0)};
  if (ch == 1) return {f.w2, C(0)};
  if (ch <= 4) return {C(0), (ch == 2 ? kx : (ch == 3 ? ky : kz)) * f.w3};
  const int q = ch - 5;
  const C ki = q < 3 ? kx : (q < 5 ? ky : kz);
  const C kj = (q == 0) ? kx : (q == 1 || q == 3) ? ky : kz;
  const bool diag = (q == 0 || q == 3 || q == 5);
  return {f.aik2 * ki * kj + (diag ? f.w3 * invR : C(0)), C(0)};
}

// The FP64 backward plane kernel's channel weight as a linear form in the
// element's weight factors: W = (a3 w3 + a2 w2 + at aik2 k_i k_j) + i av k_i w3,
// k_i = ix kx + iy ky + iz kz, k_j = jx kx + jy ky + jz kz, coefficients set up
// once per channel (`adj` conjugates it: the odd vector channels change sign,
// hard_sphere.hpp:118-134).
template <typename C>
struct ChanForm {
  C a3, a2, at, av, ix, iy, iz, jx, jy, jz;
};
template <typename C>
__device__ __forceinline__ ChanForm<C> chan_form(int ch, bool adj, C invR) {
  ChanForm<C> f{C(0), C(0), C(0), C(0), C(0), C(0), C(0), C(0), C(0), C(0)};
  int i = 0, j = 0;
  if (ch == 0) {
    f.a3 = C(1);
  } else if (ch == 1) {
    f.a2 = C(1);
  } else if (ch <= 4) {
    f.av = adj ? C(-1) : C(1);
    i = ch - 2;
  } else {
    const int q = ch - 5;
    i = q < 3 ? 0 : (q < 5 ? 1 : 2);
    j = (q == 0) ? 0 : ((q == 1 || q == 3) ? 1 : 2);
    f.at = C(1);
    if (q == 0 || q == 3 || q == 5) f.a3 = invR;
  }
  f.ix = i == 0 ? C(1) : C(0);
  f.iy = i == 1 ? C(1) : C(0);
  f.iz = i == 2 ? C(1) : C(0);
  f.jx = j == 0 ? C(1) : C(0);
  f.jy = j == 1 ? C(1) : C(0);
  f.jz = j == 2 ? C(1) : C(0);
  return f;
}
template <typename C>
__device__ __forceinline__ cx<C> apply_form(const ChanForm<C>& f, C kx, C ky, C kz, const WFac<C>& w) {
  const C ki = f.ix * kx + (f.iy * ky + f.iz * kz);
  const C kj = f.jx * kx + (f.jy * ky + f.jz * kz);
  return {f.a3 * w.w3 + f.a2 * w.w2 + f.at * w.aik2 * ki * kj, f.av * ki * w.w3};
}

// ---------------------------------------------------------------- kf_rho_xy
// FP64 y- and x-FFT of rho's parts: block (kz, r, p) owns one sub-plane of
// NX/2 half rows i and Q columns m (the channel-layout chunks of kf_zfwd),
// applies the parity twiddles exp(-2 pi i (r m / NY + p i / NX)), transforms
// the rows (Q points) and then the columns (NX/2 points) through a transpose
// taken one real part at a time, and writes rho_hat(kx = 2 k + p, ky = r + 2 s, kz)
// at XR[hp_idx(kz, r, p, s, k)] (complex FP64).  512-thread blocks with 75 KB of
// LDS run two to a CU, so the 4 (Nz/2 + 1) blocks take one round.
template <int NX, int Q>
struct RhoCfg {
  static constexpr int NT = 512, NXH = NX / 2, NY = 2 * Q;
  static constexpr int TY = Q / 8, SY = NT / TY, RY = (NXH + SY - 1) / SY;    // row transforms
  static constexpr int TX = NXH / 8, SX = NT / TX, CX = (Q + SX - 1) / SX;    // column transforms
  static constexpr int NXHP = NXH + 1;
  static constexpr int tb_bytes = align16(Q * NXHP * 8);  // one real part of the transpose
  static constexpr int xb_bytes() {
    const int a = align16(SY * slot_buf_len<Q>() * 8), b = align16(SX * slot_buf_len<NXH>() * 8);
    return a > b ? a : b;
  }
  static constexpr int tw_off = tb_bytes + xb_bytes();
  static constexpr int smem = tw_off + tw_bytes<double>();
  static_assert(TY >= 1 && TX >= 1 && TY <= 64 && TX <= 64, "unsupported plane shape");
};

// The (0, 0, 0) coefficient, the sum of rho, goes to sum_out[0] and is
// replaced by 0: the channels are transformed without their mean, which the
// point kernel adds back in FP64 (the k = 0 weights times the mean density).
// With sum_out and D0 null (generic field transforms in FP64) the kz = 0 plane
// comes from A0 and the spectrum is written whole; field f = blockIdx.y.
template <typename S, int NX, int Q>
__global__ void __launch_bounds__(512) kf_rho_xy(const cx<S>* __restrict__ A0, const double* __restrict__ D0,
                                                 cx<double>* __restrict__ XR, double* __restrict__ sum_out,
                                                 FastGeom g, const cx<double>* __restrict__ tw) {
  using RC = RhoCfg<NX, Q>;
  constexpr int NXH = RC::NXH, NXHP = RC::NXHP, TY = RC::TY, SY = RC::SY, RY = RC::RY;
  constexpr int TX = RC::TX, SX = RC::SX, CX = RC::CX, NZT = 32;
  extern __shared__ __align__(16) unsigned char smem[];
  double* tb = reinterpret_cast<double*>(smem);
  unsigned char* xraw = smem + RC::tb_bytes;
  const cx<double>* twl = stage_twiddles<double, 512>(smem + RC::tw_off, tw);
  const int tid = threadIdx.x;
  const int ys = tid / TY, yl = tid % TY, xs = tid / TX, xl = tid % TX;
  // The two parts p of one (kz, r) read the two halves of the same rows.
  // Placing their blocks on one XCD halves the L2 misses but measured 10%
  // slower than the round-robin placement (14.6 against 16.1 us at 128^3).
  const int unit = blockIdx.x;
  const int p = unit & 1, r = (unit >> 1) & 1, kz = unit >> 2;
  const int mb8 = Q / 8;
  A0 += blockIdx.y * g.M;  // field f = blockIdx.y
  XR += blockIdx.y * g.M;
  // the lane's rows i = ys + SY rr, columns m = yl + TY q (channel-layout offsets)
  cx<double> yv[RY][8];
#pragma unroll
  for (int rr = 0; rr < RY; ++rr) {
    const int i = min(ys + SY * rr, NXH - 1);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int m = yl + TY * q;
      const long row = static_cast<long>(i) * mb8 + (m >> 3);
      const int off = r * 16 + p * 8 + (m & 7);
      yv[rr][q] = kz == 0 && D0 ? cx<double>{D0[row * NZT + off], 0.0}
                                : cvt<double>(A0[(row * g.nzh + kz) * NZT + off]);
    }
  }
  __syncthreads();  // twiddles
#pragma unroll
  for (int rr = 0; rr < RY; ++rr) {
    const int i = ys + SY * rr;
    if (i < NXH) {
      cx<double> wp = {1.0, 0.0};
      if (p) wp = twl[NX - 4 + i];  // exp(-2 pi i i / NX)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (p) yv[rr][q] = cmul(yv[rr][q], wp);
        if (r) yv[rr][q] = twiddle<double, RC::NY, -1>(yv[rr][q], yl + TY * q, tw_for<RC::NY>(twl));
      }
      slot_fft<double, Q, TY, -1, true>(yv[rr], xraw + ys * slot_buf_len<Q>() * 8, yl, tw_for<Q>(twl));
    }
  }
  // transpose, one real part at a time: tb[s * NXHP + i]
  cx<double> f[CX][8];
#pragma unroll
  for (int part = 0; part < 2; ++part) {
    __syncthreads();  // row transforms done (part 0) / previous part read (part 1)
#pragma unroll
    for (int rr = 0; rr < RY; ++rr) {
      const int i = ys + SY * rr;
      if (i < NXH) {
#pragma unroll
        for (int e = 0; e < 8; ++e) tb[(yl + TY * e) * NXHP + i] = part ? yv[rr][e].im : yv[rr][e].re;
      }
    }
    __syncthreads();
#pragma unroll
    for (int cc = 0; cc < CX; ++cc) {
      const int sc = min(xs + SX * cc, Q - 1);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const double t = tb[sc * NXHP + xl + TX * q];
        if (part) f[cc][q].im = t;
        else f[cc][q].re = t;
      }
    }
  }
#pragma unroll
  for (int cc = 0; cc < CX; ++cc) {
    const int sc = xs + SX * cc;
    if (sc < Q) {
      slot_fft<double, NXH, TX, -1, true>(f[cc], xraw + xs * slot_buf_len<NXH>() * 8, xl, tw_for<NXH>(twl));
      if (sum_out && kz == 0 && r == 0 && p == 0 && sc == 0 && xl == 0) {
        sum_out[0] = f[cc][0].re;
        f[cc][0] = {0.0, 0.0};
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) XR[hp_idx(kz, r, p, sc, xl + TX * e, Q, NXH)] = f[cc][e];
    }
  }
}

// ---------------------------------------------------------------- kf_plane_fwd
template <int NX, int Q, typename S, typename CV>
struct FwdCfg {
  using P = PlaneCfg<NX, Q>;
  static constexpr int QP = Q + 1;
  static constexpr int tb_bytes = align16(P::NXH * QP * static_cast<int>(sizeof(cx<S>)));
  static constexpr int xx_bytes = align16(P::SBX * slot_buf_len<P::NXH>() * P::XBYTES);  // x-slot buffers
  static constexpr int xy_bytes = align16(P::SBYH * slot_buf_len<Q>() * P::XBYTES);      // y-slot buffers
  static constexpr int tw_all = tw_bytes<double>() + tw_bytes<CV>();
  // Double buffering (two transpose buffers, separate x- and y-slot buffers)
  // lets consecutive channels overlap with one barrier each; it needs the room.
  static constexpr bool DB = 2 * tb_bytes + xx_bytes + xy_bytes + tw_all <= 160 * 1024;
  static constexpr int tb2_off = DB ? tb_bytes : 0;
  static constexpr int xx_off = DB ? 2 * tb_bytes : tb_bytes;
  static constexpr int xy_off = DB ? xx_off + xx_bytes : xx_off;
  static constexpr int tw_off = xy_off + (DB ? xy_bytes : (xx_bytes > xy_bytes ? xx_bytes : xy_bytes));
  static constexpr int smem = tw_off + tw_all;
};

// One channel in a kf_plane_fwd block: Wof(cc, e) is the weighted spectrum at
// column xs + SBX cc, k = xl + TX e (kx = 2 k + p); soff[rr][e] is the
// split-layout offset of the lane's output (row ys + SBYH rr, m = yl + TYH e)
// in a channel array.  x-inverse of the sub-plane's NX/2 points, then the
// y-inverse of every row, stored in the channel layout.  With DB, channels
// alternate between two transpose buffers and the x- and y-slot buffers are
// distinct, so the one barrier per channel suffices: a wave enters channel
// c+2 (which reuses channel c's transpose buffer) only after every wave has
// passed channel c+1's barrier, i.e. finished channel c.  Without DB a second
// barrier closes every channel.
template <typename CC, typename S, int NX, int Q, bool DB, typename WF>
__device__ __forceinline__ void plane_fwd_channel(WF&& Wof, cx<S>* tb, unsigned char* xbx, unsigned char* xby,
                                                  cx<S>* __restrict__ out,
                                                  const int (&soff)[PlaneCfg<NX, Q>::RPH][4], int r,
                                                  const cx<CC>* __restrict__ tw) {
  using P = PlaneCfg<NX, Q>;
  constexpr int QP = Q + 1;
  constexpr int XS = slot_buf_len<P::NXH>(), YS = slot_buf_len<Q>();
  const int tid = threadIdx.x;
  const int xs = tid / P::TX, xl = tid % P::TX, ys = tid / P::TYH, yl = tid % P::TYH;
#pragma unroll
  for (int cc = 0; cc < P::CPS; ++cc) {
    const int col = xs + P::SBX * cc;
    if (col < Q) {
      cx<CC> zt[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) zt[e] = cvt<CC>(Wof(cc, e));
      slot_fft<CC, P::NXH, P::TX, +1, sizeof(CC) == 8>(zt, xbx + xs * XS * P::XBYTES, xl, tw_for<P::NXH>(tw));
#pragma unroll
      for (int e = 0; e < 4; ++e) tb[(xl + P::TX * e) * QP + col] = cvt<S>(zt[e]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int rr = 0; rr < P::RPH; ++rr) {
    const int row = ys + P::SBYH * rr;
    if (row < P::NXH) {
      cx<CC> v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = cvt<CC>(tb[row * QP + yl + P::TYH * q]);
      slot_fft<CC, Q, P::TYH, +1, sizeof(CC) == 8>(v, xby + ys * YS * P::XBYTES, yl, tw_for<Q>(tw));
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = yl + P::TYH * e;
        cx<CC> b = v[e];
        if (r) b = twiddle<CC, P::NY, +1>(b, m, tw_for<P::NY>(tw));
        out[soff[rr][e]] = cvt<S>(b);
      }
    }
  }
  if constexpr (!DB) __syncthreads();
}

// Work of the forward plane kernel: (unit, channel) items, units
// u = (kz * 2 + r) * 2 + p (the sub-planes), channels in the order FP64 (eta,
// n2, mean field) then CV.  The 4 (Nz/2 + 1) units do not divide evenly over
// the CUs, so blocks take contiguous ranges of items (at most two units each)
// instead of one unit each: the last round of blocks then holds one item more
// at most.
__host__ __device__ constexpr int fwd_nf64(int nf) { return nf == 12 ? 3 : (nf == 11 ? 2 : nf); }

template <typename S, typename CV, int NX, int Q, int NF>
__global__ void __launch_bounds__(kBT) kf_plane_fwd(const cx<double>* __restrict__ XR,
                                                    const cx<double>* __restrict__ WT, cx<S>* __restrict__ SF,
                                                    const double* __restrict__ mfB, FastGeom g,
                                                    const cx<double>* __restrict__ tw64,
                                                    const cx<CV>* __restrict__ twv, int nunits) {
  using P = PlaneCfg<NX, Q>;
  using FC = FwdCfg<NX, Q, S, CV>;
  constexpr int CPS = P::CPS, NXH = P::NXH;
  constexpr int N64 = fwd_nf64(NF);
  extern __shared__ __align__(16) unsigned char smem[];
  auto tbuf = [&](int n) { return reinterpret_cast<cx<S>*>(smem + (n & 1) * FC::tb2_off); };
  unsigned char* xbx = smem + FC::xx_off;
  unsigned char* xby = smem + FC::xy_off;
  const cx<double>* tw64l = stage_twiddles<double, kBT>(smem + FC::tw_off, tw64);
  const cx<CV>* twvl = stage_twiddles<CV, kBT>(smem + FC::tw_off + tw_bytes<double>(), twv);
  const int tid = threadIdx.x;
  const int xs = tid / P::TX, xl = tid % P::TX;
  constexpr int NFP = fwd_nplane(NF);
  const long items = static_cast<long>(nunits) * NFP;
  // consecutive item ranges go to blocks of one XCD, which share its L2
  const long lb = xcd_remap(blockIdx.x, gridDim.x);
  const long i0 = items * lb / gridDim.x, i1 = items * (lb + 1) / gridDim.x;
  const double invR = 1.0 / g.R;
  __syncthreads();  // twiddles
  int nch = 0;
  // A unit split between blocks lb and lb + 1 is the last of lb's range and
  // the first of lb + 1's: odd blocks take their units in reverse order, so
  // both read the shared unit's spectrum and weights at the same time (L2).
  const int ua = static_cast<int>(i0 / NFP), ub = static_cast<int>((i1 - 1) / NFP);
  for (int uu = ua; uu <= ub && i1 > i0; ++uu) {
    const int unit = (lb & 1) ? ua + ub - uu : uu;
    const long ubase = static_cast<long>(unit) * NFP;
    const int p0 = static_cast<int>(max(i0, ubase) - ubase), p1 = static_cast<int>(min(i1, ubase + NFP) - ubase);
    const int p = unit & 1, r = (unit >> 1) & 1, kz = unit >> 2;
    int soff[P::RPH][4];  // the lane's output offsets, the same for every channel of the unit
#pragma unroll
    for (int rr = 0; rr < P::RPH; ++rr)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = min(tid / P::TYH + P::SBYH * rr, NXH - 1), m = tid % P::TYH + P::TYH * e;
        soff[rr][e] = static_cast<int>(chan_idx(kz, r, row, p, m, Q / 8, g.nzh));
      }
    // the lane's elements of the sub-plane: column xs + SBX cc, k = xl + TX e
    const long uoff = static_cast<long>(unit) * Q * NXH + xl;
    auto eoff = [&](int cc, int e) {
      const int col = xs + P::SBX * cc;
      return uoff + static_cast<long>(col < Q ? col : 0) * NXH + P::TX * e;
    };
    // FP64 spectrum and radial weights, read once per unit
    cx<double> X[CPS][4], Wr[CPS][4];
#pragma unroll
    for (int cc = 0; cc < CPS; ++cc)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        X[cc][e] = XR[eoff(cc, e)];
        Wr[cc][e] = WT[eoff(cc, e)];
      }
    const double kzv = kz * g.dkz;
    auto kxof = [&](int e) { return kvec(2 * (xl + P::TX * e) + p, NX, g.dkx); };
    auto kyof = [&](int cc) {
      const int col = xs + P::SBX * cc;
      return kvec(r + 2 * (col < Q ? col : 0), P::NY, g.dky);
    };
    // FP64 channels: eta and n2 (weights.hpp:91-108,159-162), the mean field (FP64 table,
    // mean_field.hpp:145-196)
    for (int pos = p0; pos < min(p1, N64); ++pos, ++nch) {
      const int ch = fwd_channel(NF, pos);
      plane_fwd_channel<double, S, NX, Q, FC::DB>(
          [&](int cc, int e) {
            const int col = xs + P::SBX * cc;
            if (ch == 11) return cscale(X[cc][e], col < Q ? mfB[eoff(cc, e)] : 0.0);
            return cscale(X[cc][e], ch == 0 ? Wr[cc][e].re : Wr[cc][e].im);
          },
          tbuf(nch), xbx, xby, SF + chan_store(NF, ch) * g.M, soff, r, tw64l);
    }
    if constexpr (NF >= 11) {
      if (p1 > N64) {
        // vector (i k_a w3) and tensor (aik2 k_i k_j + delta_ij w3/R) channels in CV,
        // from the spectrum and per-element weight factors narrowed to CV (weights.hpp:110-137,164-178)
        cx<CV> XC[CPS][4];
        WFac<CV> wf[CPS][4];
        CV kxc[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) kxc[e] = static_cast<CV>(kxof(e));
#pragma unroll
        for (int cc = 0; cc < CPS; ++cc)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            XC[cc][e] = cvt<CV>(X[cc][e]);
            wf[cc][e] = wfactors<CV>(Wr[cc][e], kxof(e), kyof(cc), kzv, invR);
          }
        const CV kzc = static_cast<CV>(kzv), irc = static_cast<CV>(invR);
#pragma unroll 1
        for (int pos = max(p0, N64); pos < p1; ++pos, ++nch) {
          const int ch = fwd_channel(NF, pos);
          plane_fwd_channel<CV, S, NX, Q, FC::DB>(
              [&](int cc, int e) -> cx<CV> {
                return cmul(chan_weight<CV>(ch, kxc[e], static_cast<CV>(kyof(cc)), kzc, wf[cc][e], irc), XC[cc][e]);
              },
              tbuf(nch), xbx, xby, SF + static_cast<long>(ch) * g.M, soff, r, twvl);
        }
      }
    }
  }
}

// ---------------------------------------------------------------- kg_xy_inv
// Inverse x- and y-transform of whole spectra (generic fields, no weights):
// block (unit, field f = blockIdx.y) takes sub-plane `unit` of spectrum f in
// the sub-plane order (FP64) and writes its parts in the channel layout, as
// one kf_plane_fwd channel item does.
template <typename S, int NX, int Q>
__global__ void __launch_bounds__(kBT) kg_xy_inv(const cx<double>* __restrict__ spec, cx<S>* __restrict__ out,
                                                 FastGeom g, const cx<double>* __restrict__ tw64) {
  using P = PlaneCfg<NX, Q>;
  using FC = FwdCfg<NX, Q, S, double>;
  constexpr int NXH = P::NXH;
  extern __shared__ __align__(16) unsigned char smem[];
  const cx<double>* tw64l = stage_twiddles<double, kBT>(smem + FC::tw_off, tw64);
  const int tid = threadIdx.x;
  const int xs = tid / P::TX, xl = tid % P::TX;
  const int unit = blockIdx.x;
  const int p = unit & 1, r = (unit >> 1) & 1, kz = unit >> 2;
  int soff[P::RPH][4];
#pragma unroll
  for (int rr = 0; rr < P::RPH; ++rr)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = min(tid / P::TYH + P::SBYH * rr, NXH - 1), m = tid % P::TYH + P::TYH * e;
      soff[rr][e] = static_cast<int>(chan_idx(kz, r, row, p, m, Q / 8, g.nzh));
    }
  const cx<double>* src = spec + blockIdx.y * g.M + static_cast<long>(unit) * Q * NXH + xl;
  __syncthreads();  // twiddles
  plane_fwd_channel<double, S, NX, Q, false>(
      [&](int cc, int e) { return src[static_cast<long>(xs + P::SBX * cc) * NXH + P::TX * e]; },
      reinterpret_cast<cx<S>*>(smem), smem + FC::xx_off, smem + FC::xy_off, out + blockIdx.y * g.M, soff, r,
      tw64l);
}

// ---------------------------------------------------------------- kf_plane_bwd
template <int NX, int Q, typename S, typename CV>
struct BwdCfg {
  using P = PlaneCfg<NX, Q>;
  static constexpr int NXHP = P::NXH + 1, QP = Q + 1;
  static constexpr int tb1 = align16(Q * NXHP * static_cast<int>(sizeof(cx<S>)));  // one channel transpose
  static constexpr int tbf = align16(P::NXH * QP * static_cast<int>(sizeof(cx<S>)));  // the final transpose
  static constexpr int xy_bytes = align16(P::SBYH * slot_buf_len<Q>() * P::XBYTES);
  static constexpr int xx_bytes = align16(P::SBX * slot_buf_len<P::NXH>() * P::XBYTES);
  static constexpr int tw_b = tw_bytes<CV>();
  static constexpr bool DB = 2 * tb1 + xy_bytes + xx_bytes + tw_b <= 160 * 1024;
  static constexpr int ntb = DB ? 2 : 1;
  static constexpr int xx_off = ntb * tb1 > tbf ? ntb * tb1 : tbf;
  static constexpr int xy_off = DB ? xx_off + xx_bytes : xx_off;
  static constexpr int tw_off = DB ? xy_off + xy_bytes : xx_off + (xx_bytes > xy_bytes ? xx_bytes : xy_bytes);
  static constexpr int smem = tw_off + tw_b;
};

// Work of the backward plane kernel: (unit, channel) items as in the forward
// kernel, kBwdPlaneCh channels per unit.  A block accumulates the channels of
// its range per unit and writes the inverse transform of each accumulated
// part: the part that starts at channel 0 into partial set 0, a continuation
// (the unit's remaining channels, in the next block) into set 1.  Every range
// holds at least kBwdPlaneCh items (blocks <= units), so a unit has at most
// two parts;
// bwd_unit_split() tells the consumer which units have a set-1 part.
__host__ __device__ inline int bwd_block_of_item(long item, long items, int nblk) {
  return static_cast<int>(((item + 1) * nblk - 1) / items);
}
__host__ __device__ inline bool bwd_unit_split(int unit, int nunits, int nblk) {
  const long items = static_cast<long>(nunits) * kBwdPlaneCh;
  return bwd_block_of_item(static_cast<long>(unit) * kBwdPlaneCh, items, nblk) !=
         bwd_block_of_item(static_cast<long>(unit) * kBwdPlaneCh + kBwdPlaneCh - 1, items, nblk);
}

template <typename S, typename CV, int NX, int Q>
__global__ void __launch_bounds__(kBT) kf_plane_bwd(const cx<S>* __restrict__ SB, const cx<double>* __restrict__ WT,
                                                    cx<S>* __restrict__ P, FastGeom g, const cx<CV>* __restrict__ tw,
                                                    int nunits) {
  using PC = PlaneCfg<NX, Q>;
  using BC = BwdCfg<NX, Q, S, CV>;
  constexpr int CPS = PC::CPS, NXH = PC::NXH;
  constexpr int NXHP = BC::NXHP, QP = BC::QP;
  extern __shared__ __align__(16) unsigned char smem[];
  auto tbuf = [&](int n) { return reinterpret_cast<cx<S>*>(smem + (n & 1) * (BC::DB ? BC::tb1 : 0)); };
  unsigned char* xbx = smem + BC::xx_off;
  unsigned char* xby = smem + BC::xy_off;
  const cx<CV>* twl = stage_twiddles<CV, kBT>(smem + BC::tw_off, tw);
  constexpr bool SPL = sizeof(CV) == 8;  // FP64 exchanges one real part at a time
  const int tid = threadIdx.x;
  const int xs = tid / PC::TX, xl = tid % PC::TX;
  const int ysh = tid / PC::TYH, ylh = tid % PC::TYH;
  const CV invR = static_cast<CV>(1.0 / g.R);
  const long items = static_cast<long>(nunits) * kBwdPlaneCh;
  const long lb = xcd_remap(blockIdx.x, gridDim.x);  // neighbouring ranges share an XCD's L2
  const long i0 = items * lb / gridDim.x, i1 = items * (lb + 1) / gridDim.x;
  __syncthreads();  // twiddles
  // Units in order: neighbouring blocks then read the two parts p of one
  // (kz, r), the two halves of the same channel rows, at about the same time.
  const int ua = static_cast<int>(i0 / kBwdPlaneCh), ub = static_cast<int>((i1 - 1) / kBwdPlaneCh);
  for (int unit = ua; unit <= ub && i1 > i0; ++unit) {
    const long ubase = static_cast<long>(unit) * kBwdPlaneCh;
    const int c0 = static_cast<int>(max(i0, ubase) - ubase), c1 = static_cast<int>(min(i1, ubase + kBwdPlaneCh) - ubase);
    const int p = unit & 1, r = (unit >> 1) & 1, kz = unit >> 2;
    const double kzv = kz * g.dkz;

    cx<S> nxt[PC::RPH][4];
    int loff[PC::RPH][4];  // the lane's input / output offsets, the same for every channel of the unit
#pragma unroll
    for (int rr = 0; rr < PC::RPH; ++rr) {
      const int i = ysh + PC::SBYH * rr;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        loff[rr][q] = static_cast<int>(chan_idx(kz, r, i < NXH ? i : 0, p, ylh + PC::TYH * q, Q / 8, g.nzh));
    }
    auto load_in = [&](int ci) {
      const cx<S>* in = SB + static_cast<long>(bwd_channel(ci)) * g.M;
#pragma unroll
      for (int rr = 0; rr < PC::RPH; ++rr)
#pragma unroll
        for (int q = 0; q < 4; ++q) nxt[rr][q] = in[loff[rr][q]];
    };
    load_in(c0);
    // The conj-weighted channel sum (hard_sphere.hpp:115-134) is accumulated as
    //   S = w3 A + w2 B + aik2 C,
    //   A = G_eta - i kx G_vx - i ky G_vy + (G_Txx + G_Tyy) / R,   B = G_n2,
    //   C = kx^2 G_Txx + kx ky G_Txy + kx kz G_Txz + ky^2 G_Tyy + ky kz G_Tyz,
    // since every channel weight is linear in (w3, w2, aik2) (weights.hpp:110-137,164-178): each
    // channel adds its spectrum to one or two sums with a coefficient in k only, and the radial
    // factors multiply the three sums once per unit.  Element e of the lane: k = xl + TX e
    // (kx = 2 k + p), column xs + SBX cc.
    // FP64 storage keeps one accumulator and applies each channel's full weight
    // (three FP64 sums would spill at the 128-VGPR cap of the 1024-thread block).
    constexpr bool SUMS = sizeof(CV) == 4;
    constexpr int NA = SUMS ? 1 : 0;  // extent of the arrays only one form uses
    CV kxc[4];
    cx<CV> accA[CPS][4], accB[NA * CPS + 1 - NA][4], accC[NA * CPS + 1 - NA][4];
    WFac<CV> wf[(1 - NA) * CPS + NA][4];
    int xlo = xl;  // opaque: keeps the per-element set-up inside the unit loop
    asm volatile("" : "+v"(xlo));
    const cx<double>* wu = WT + static_cast<long>(unit) * Q * NXH + xlo;
#pragma unroll
    for (int e = 0; e < 4; ++e) kxc[e] = static_cast<CV>(kvec(2 * (xlo + PC::TX * e) + p, NX, g.dkx));
#pragma unroll
    for (int cc = 0; cc < CPS; ++cc) {
      const int col = xs + PC::SBX * cc;
      const double kyd = kvec(r + 2 * (col < Q ? col : 0), PC::NY, g.dky);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        accA[cc][e] = {CV(0), CV(0)};
        if constexpr (SUMS) {
          accB[cc][e] = accC[cc][e] = {CV(0), CV(0)};
        } else {
          wf[cc][e] = wfactors<CV>(wu[static_cast<long>(col < Q ? col : 0) * NXH + PC::TX * e],
                                   kvec(2 * (xlo + PC::TX * e) + p, NX, g.dkx), kyd, kzv, 1.0 / g.R);
        }
      }
    }
    const CV kzc = static_cast<CV>(kzv);
#pragma unroll 1
    for (int ch = c0; ch < c1; ++ch) {
      cx<S>* tb = tbuf(ch);
      // y-FFT of the sub-plane's rows
      cx<CV> v[PC::RPH][4];
#pragma unroll
      for (int rr = 0; rr < PC::RPH; ++rr)
#pragma unroll
        for (int q = 0; q < 4; ++q) v[rr][q] = cvt<CV>(nxt[rr][q]);
      if (ch + 1 < c1) load_in(ch + 1);  // in flight while this channel is transformed
#pragma unroll
      for (int rr = 0; rr < PC::RPH; ++rr) {
        const int i = ysh + PC::SBYH * rr;
        if (i < NXH) {
          if (r) {
#pragma unroll
            for (int q = 0; q < 4; ++q) v[rr][q] = twiddle<CV, PC::NY, -1>(v[rr][q], ylh + PC::TYH * q, tw_for<PC::NY>(twl));
          }
          slot_fft<CV, Q, PC::TYH, -1, SPL>(v[rr], xby + ysh * slot_buf_len<Q>() * PC::XBYTES, ylh, tw_for<Q>(twl));
#pragma unroll
          for (int e = 0; e < 4; ++e) tb[(ylh + PC::TYH * e) * NXHP + i] = cvt<S>(v[rr][e]);
        }
      }
      __syncthreads();
      const int chan = bwd_channel(ch);  // the same for the whole block: uniform branches below
      const ChanForm<CV> form = chan_form<CV>(chan, true, invR);
      // x-FFT of the sub-plane's columns, accumulation into A, B, C
#pragma unroll
      for (int cc = 0; cc < CPS; ++cc) {
        const int col = xs + PC::SBX * cc;
        if (col < Q) {
          const CV kyv = static_cast<CV>(kvec(r + 2 * col, PC::NY, g.dky));
          cx<CV> f[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) f[q] = cvt<CV>(tb[col * NXHP + xl + PC::TX * q]);
          slot_fft<CV, NXH, PC::TX, -1, SPL>(f, xbx + xs * slot_buf_len<NXH>() * PC::XBYTES, xl, tw_for<NXH>(twl));
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const cx<CV> G = f[e];
            const CV kx = kxc[e];
            cx<CV>& A = accA[cc][e];
            if constexpr (!SUMS) {
              A = cadd(A, cmul(apply_form<CV>(form, kx, kyv, kzc, wf[cc][e]), G));
              continue;
            }
            cx<CV>& C = accC[cc][e];
            switch (chan) {
              case 0: A = cadd(A, G); break;                                   // eta: w3
              case 1: accB[cc][e] = cadd(accB[cc][e], G); break;               // n2: w2
              case 2: A = {A.re + kx * G.im, A.im - kx * G.re}; break;         // v_x: conj(i kx w3)
              case 3: A = {A.re + kyv * G.im, A.im - kyv * G.re}; break;       // v_y
              case 5:                                                          // T_xx
                C = {C.re + (kx * kx) * G.re, C.im + (kx * kx) * G.im};
                A = {A.re + invR * G.re, A.im + invR * G.im};
                break;
              case 6: C = {C.re + (kx * kyv) * G.re, C.im + (kx * kyv) * G.im}; break;  // T_xy
              case 7: C = {C.re + (kx * kzc) * G.re, C.im + (kx * kzc) * G.im}; break;  // T_xz
              case 8:                                                                   // T_yy
                C = {C.re + (kyv * kyv) * G.re, C.im + (kyv * kyv) * G.im};
                A = {A.re + invR * G.re, A.im + invR * G.im};
                break;
              default: C = {C.re + (kyv * kzc) * G.re, C.im + (kyv * kzc) * G.im}; break;  // T_yz
            }
          }
        }
      }
      // With DB the next channel writes the other transpose buffer and other
      // slot buffers; a wave reaches channel ch+2 (this channel's buffers) only
      // after every wave has passed channel ch+1's barrier.
      if constexpr (!BC::DB) __syncthreads();
    }
    __syncthreads();
    // FP32: S = w3 A + w2 B + aik2 C with the radial factors of the lane's elements (FP64
    // table, formed in the backward path's precision); FP64: S is the accumulator
    cx<CV> acc[CPS][4];
    if constexpr (!SUMS) {
#pragma unroll
      for (int cc = 0; cc < CPS; ++cc)
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[cc][e] = accA[cc][e];
    } else {
#pragma unroll
      for (int cc = 0; cc < CPS; ++cc) {
        const int col = xs + PC::SBX * cc;
        const double kyd = kvec(r + 2 * (col < Q ? col : 0), PC::NY, g.dky);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const WFac<CV> w = wfactors<CV>(wu[static_cast<long>(col < Q ? col : 0) * NXH + PC::TX * e],
                                          kvec(2 * (xlo + PC::TX * e) + p, NX, g.dkx), kyd, kzv, 1.0 / g.R);
          const cx<CV> a = accA[cc][e], b = accB[cc][e], c = accC[cc][e];
          acc[cc][e] = {w.w3 * a.re + w.w2 * b.re + w.aik2 * c.re, w.w3 * a.im + w.w2 * b.im + w.aik2 * c.im};
        }
      }
    }
    // x-inverse of the accumulated sub-plane, then y-inverse of every row
    cx<S>* tbf = reinterpret_cast<cx<S>*>(smem);
#pragma unroll
    for (int cc = 0; cc < CPS; ++cc) {
      const int col = xs + PC::SBX * cc;
      if (col < Q) {
        slot_fft<CV, NXH, PC::TX, +1, SPL>(acc[cc], xbx + xs * slot_buf_len<NXH>() * PC::XBYTES, xl, tw_for<NXH>(twl));
#pragma unroll
        for (int e = 0; e < 4; ++e) tbf[(xl + PC::TX * e) * QP + col] = cvt<S>(acc[cc][e]);
      }
    }
    __syncthreads();
    cx<S>* out = P + static_cast<long>(c0 > 0 ? 1 : 0) * g.M;
#pragma unroll
    for (int rr = 0; rr < PC::RPH; ++rr) {
      const int row = ysh + PC::SBYH * rr;
      if (row < NXH) {
        cx<CV> w[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = cvt<CV>(tbf[row * QP + ylh + PC::TYH * q]);
        slot_fft<CV, Q, PC::TYH, +1, SPL>(w, xby + ysh * slot_buf_len<Q>() * PC::XBYTES, ylh, tw_for<Q>(twl));
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          cx<CV> b = w[e];
          if (r) b = twiddle<CV, PC::NY, +1>(b, ylh + PC::TYH * e, tw_for<PC::NY>(twl));
          out[loff[rr][e]] = cvt<S>(b);
        }
      }
    }
    __syncthreads();  // the next part reuses the transpose and slot buffers
  }
}

}  // namespace

// ============================================================ host side
bool fast_supported(int nx, int ny, int nz, int precision) {
  auto ok = [](int n) { return n == 64 || n == 128; };
  if (!ok(ny) || !ok(nz)) return false;
  if (ok(nx)) return true;
  // x extent 256: the FP64 plane kernels of a 256 x 128 plane need more LDS than a CU has
  return nx == 256 && (precision == CDFT_PRECISION_MIXED || ny == 64);
}

FastGeom make_fast_geom(const Geom& g) {
  FastGeom f{};
  f.nx = g.nx;
  f.ny = g.ny;
  f.nz = g.nz;
  f.nzh = g.nzh;
  f.q = g.ny / 2;
  f.mb = f.q / 16;
  f.N = g.N;
  f.M = g.M;
  f.dkx = g.dkx;
  f.dky = g.dky;
  f.dkz = g.dkz;
  f.dv = g.dv;
  f.inv_n = g.inv_n;
  f.R = g.R;
  return f;
}

void fast_twiddles(void** tw64, void** tw32) {
  const double PI = 3.14159265358979323846;
  std::vector<double> t64(2 * kTwLen);
  std::vector<float> t32(2 * kTwLen);
  for (int L = 4; L <= kTwMax; L *= 2)
    for (int m = 0; m < L; ++m) {
      const int i = L - 4 + m;
      const double a = -2.0 * PI * m / L;
      t64[2 * i] = std::cos(a);
      t64[2 * i + 1] = std::sin(a);
      t32[2 * i] = static_cast<float>(t64[2 * i]);
      t32[2 * i + 1] = static_cast<float>(t64[2 * i + 1]);
    }
  if (hipMalloc(tw64, t64.size() * sizeof(double)) != hipSuccess ||
      hipMalloc(tw32, t32.size() * sizeof(float)) != hipSuccess)
    throw std::runtime_error("hipMalloc of the twiddle tables failed");
  if (hipMemcpy(*tw64, t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(*tw32, t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    throw std::runtime_error("hipMemcpy of the twiddle tables failed");
}

namespace {
__global__ __launch_bounds__(256) void k_fast_mf_layout(const cx<double>* __restrict__ spec, double* __restrict__ out,
                                                        int nx, int ny, long m) {
  const long i = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= m) return;
  const long plane = static_cast<long>(nx) * ny;
  const long kz = i / plane, rem = i - kz * plane;
  const int kx = static_cast<int>(rem / ny), ky = static_cast<int>(rem % ny);
  out[hp_idx(static_cast<int>(kz), ky & 1, kx & 1, ky >> 1, kx >> 1, ny / 2, nx / 2)] = spec[i].re;
}
}  // namespace

void launch_fast_mf_layout(const cx<double>* spec, double* out, const Geom& g, hipStream_t stream) {
  if (g.ny % 2 != 0 || g.M != static_cast<long>(g.nzh) * g.nx * g.ny)
    throw std::logic_error("fused mean-field layout needs an even y extent");
  const long m = g.M;
  k_fast_mf_layout<<<static_cast<unsigned>((m + 255) / 256), 256, 0, stream>>>(spec, out, g.nx, g.ny, m);
}

int fast_point_blocks(const FastGeom& g, bool fp64) {
  return g.nx * g.q / (fp64 ? zp_waves<double>() : zp_waves<float>());  // one line pair (x, m < NY/2) per wave
}

namespace {
template <typename T>
const T* as(const void* p) {
  return static_cast<const T*>(p);
}
// Dynamic LDS above 64 KiB has to be enabled per kernel, once per process.
// Instantiations of one kernel template share a function type, so the record
// is kept per kernel address.
template <typename K>
void set_smem(K kernel, int bytes) {
  static std::mutex mu;
  static std::vector<const void*> done;
  const void* key = reinterpret_cast<const void*>(kernel);
  std::lock_guard<std::mutex> lock(mu);
  if (std::find(done.begin(), done.end(), key) != done.end()) return;
  if (bytes > 160 * 1024) throw std::runtime_error("fused kernel needs more LDS than a CU has");
  if (hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
    throw std::runtime_error("hipFuncSetAttribute(MaxDynamicSharedMemorySize) failed");
  done.push_back(key);
}
}  // namespace

template <typename S>
void launch_fast_zfwd(const double* rho, cx<S>* out, double* D0, const FastGeom& g, const void* tw64, AliasMap amap,
                      hipStream_t st) {
  const unsigned nblk = static_cast<unsigned>((g.nx / 2) * (g.q / 8));
#define CDFT_L(NZ)                                                                                     \
  {                                                                                                    \
    auto k = kf_zfwd<S, NZ>;                                                                           \
    const int sm = zfwd_smem<NZ, S>();                                                                 \
    set_smem(k, sm);                                                                                   \
    hipLaunchKernelGGL(k, dim3(nblk), dim3(2 * NZ), sm, st, rho, out, g,                               \
                       as<cx<double>>(tw64), amap, D0);                                                \
  }
  if (g.nz == 64) CDFT_L(64) else if (g.nz == 128) CDFT_L(128) else throw std::runtime_error("fast path: nz");
#undef CDFT_L
}

template <typename S, typename CV>
void launch_fast_zinv(const cx<S>* P, double* out, const double* add, double scale, const FastGeom& g,
                      const void* twv, const double* partials, int fin_rows, double* res, hipStream_t st) {
  const unsigned nblk = static_cast<unsigned>((g.nx / 2) * (g.q / 8)) + (fin_rows > 0 ? 1u : 0u);
#define CDFT_L(NZ)                                                                                      \
  {                                                                                                     \
    auto k = kf_zinv<S, CV, NZ>;                                                                        \
    const int sm = zinv_smem<NZ, CV>();                                                                 \
    set_smem(k, sm);                                                                                    \
    hipLaunchKernelGGL(k, dim3(nblk), dim3(zinv_threads<NZ>()), sm, st, P, out, add, scale, g,          \
                       as<cx<CV>>(twv), fast_bwd_blocks(g), partials, fin_rows, res);                   \
  }
  if (g.nz == 64) CDFT_L(64) else if (g.nz == 128) CDFT_L(128) else throw std::runtime_error("fast path: nz");
#undef CDFT_L
}

template <typename S>
void launch_fast_rho_xy(const cx<S>* A0, const double* D0, void* XR, double* sum_out, const FastGeom& g,
                        const void* tw64, hipStream_t st) {
  const unsigned nblk = static_cast<unsigned>(g.nzh * 4);
#define CDFT_L(NX, Q)                                                                                 \
  {                                                                                                   \
    auto k = kf_rho_xy<S, NX, Q>;                                                                     \
    const int sm = RhoCfg<NX, Q>::smem;                                                               \
    set_smem(k, sm);                                                                                  \
    hipLaunchKernelGGL(k, dim3(nblk), dim3(512), sm, st, A0, D0, static_cast<cx<double>*>(XR), sum_out, \
                       g, as<cx<double>>(tw64));                                                      \
  }
  if (g.nx == 64 && g.q == 32) CDFT_L(64, 32)
  else if (g.nx == 64 && g.q == 64) CDFT_L(64, 64)
  else if (g.nx == 128 && g.q == 32) CDFT_L(128, 32)
  else if (g.nx == 128 && g.q == 64) CDFT_L(128, 64)
  else if (g.nx == 256 && g.q == 32) CDFT_L(256, 32)
  else if (g.nx == 256 && g.q == 64) CDFT_L(256, 64)
  else throw std::runtime_error("fast path: plane shape");
#undef CDFT_L
}

void launch_fast_wtable(void* WT, const FastGeom& g, hipStream_t st) {
  hipLaunchKernelGGL(kf_wtable, dim3(1024), dim3(256), 0, st, static_cast<cx<double>*>(WT), g);
}

template <typename S, typename CV>
void launch_fast_plane_fwd(const void* XR, const void* WT, cx<S>* SF, const double* mfB, const FastGeom& g,
                           const void* tw64, const void* twv, ChannelSet cs, hipStream_t st) {
  const int nunits = g.nzh * 4;
  const unsigned nblk = static_cast<unsigned>(std::min<long>(g.ncu, static_cast<long>(nunits) * cs.nf));
#define CDFT_L(NX, Q, NF)                                                                             \
  {                                                                                                   \
    auto k = kf_plane_fwd<S, CV, NX, Q, NF>;                                                          \
    const int sm = FwdCfg<NX, Q, S, CV>::smem;                                                        \
    set_smem(k, sm);                                                                                  \
    hipLaunchKernelGGL(k, dim3(nblk), dim3(kBT), sm, st, as<cx<double>>(XR), as<cx<double>>(WT), SF, mfB, \
                       g, as<cx<double>>(tw64), as<cx<CV>>(twv), nunits);                             \
  }
#define CDFT_NF(NX, Q)                                                                                \
  if (cs.nf == 12) CDFT_L(NX, Q, 12) else if (cs.nf == 11) CDFT_L(NX, Q, 11) else if (cs.nf == 1) CDFT_L(NX, Q, 1) \
  else throw std::runtime_error("fast path: channel set");
  if (g.nx == 64 && g.q == 32) CDFT_NF(64, 32)
  else if (g.nx == 64 && g.q == 64) CDFT_NF(64, 64)
  else if (g.nx == 128 && g.q == 32) CDFT_NF(128, 32)
  else if (g.nx == 128 && g.q == 64) CDFT_NF(128, 64)
  else if (g.nx == 256 && g.q == 32) CDFT_NF(256, 32)
  else if (g.nx == 256 && g.q == 64) CDFT_NF(256, 64)
  else throw std::runtime_error("fast path: plane shape");
#undef CDFT_NF
#undef CDFT_L
}

int fast_bwd_blocks(const FastGeom& g) { return std::min(g.ncu, g.nzh * 4); }

template <typename S, typename CV>
void launch_fast_plane_bwd(const cx<S>* SB, const void* WT, cx<S>* P, const FastGeom& g, const void* twv,
                           hipStream_t st) {
  const int nunits = g.nzh * 4;
  const unsigned nblk = static_cast<unsigned>(fast_bwd_blocks(g));
#define CDFT_L(NX, Q)                                                                                 \
  {                                                                                                   \
    auto k = kf_plane_bwd<S, CV, NX, Q>;                                                              \
    const int sm = BwdCfg<NX, Q, S, CV>::smem;                                                        \
    set_smem(k, sm);                                                                                  \
    hipLaunchKernelGGL(k, dim3(nblk), dim3(kBT), sm, st, SB, as<cx<double>>(WT), P, g, as<cx<CV>>(twv), \
                       nunits);                                                                       \
  }
  if (g.nx == 64 && g.q == 32) CDFT_L(64, 32)
  else if (g.nx == 64 && g.q == 64) CDFT_L(64, 64)
  else if (g.nx == 128 && g.q == 32) CDFT_L(128, 32)
  else if (g.nx == 128 && g.q == 64) CDFT_L(128, 64)
  else if (g.nx == 256 && g.q == 32) CDFT_L(256, 32)
  else if (g.nx == 256 && g.q == 64) CDFT_L(256, 64)
  else throw std::runtime_error("fast path: plane shape");
#undef CDFT_L
}

template <typename S, typename CV>
void launch_fast_zpoint(cx<S>* SF, const double* rho, const double* vext, double* lf, double* partials,
                        const FastGeom& g, const void* tw64, const void* twv, ChannelSet cs, FmtParams fp, double mu,
                        AliasMap amap, const double* shift_sum, DcTable dc, hipStream_t st) {
  const unsigned nblk = static_cast<unsigned>(fast_point_blocks(g, sizeof(S) == 8));
  // White-Bear II (the reference's default model) has its own instantiation
  const bool wb2 = fp.model == FMT_WHITEBEAR_II;
#define CDFT_L(NZ, NF, NB)                                                                             \
  {                                                                                                    \
    auto k = (NF >= 11 && wb2) ? kf_zpoint<S, CV, NZ, NF, NB, (NF >= 11 ? FMT_WHITEBEAR_II : kFmtRuntime)> \
                               : kf_zpoint<S, CV, NZ, NF, NB, kFmtRuntime>;                            \
    const int sm = zpoint_smem<NZ, S, CV>();                                                               \
    set_smem(k, sm);                                                                                   \
    hipLaunchKernelGGL(k, dim3(nblk), dim3(64 * zp_waves<S>()), sm, st, SF, rho, vext, lf, partials, g, \
                       as<cx<double>>(tw64), as<cx<CV>>(twv), fp, mu, amap, shift_sum, dc);            \
  }
#define CDFT_CH(NZ)                                                                                    \
  if (cs.nf == 12 && cs.nb == 11) CDFT_L(NZ, 12, 11)                                                   \
  else if (cs.nf == 12 && cs.nb == 0) CDFT_L(NZ, 12, 0)                                                \
  else if (cs.nf == 11 && cs.nb == 11) CDFT_L(NZ, 11, 11)                                              \
  else if (cs.nf == 11 && cs.nb == 0) CDFT_L(NZ, 11, 0)                                                \
  else if (cs.nf == 1 && cs.nb == 0) CDFT_L(NZ, 1, 0)                                                  \
  else if (cs.nf == 0 && cs.nb == 0) CDFT_L(NZ, 0, 0)                                                  \
  else throw std::runtime_error("fast path: channel set");
  if (g.nz == 64) CDFT_CH(64) else if (g.nz == 128) CDFT_CH(128) else throw std::runtime_error("fast path: nz");
#undef CDFT_CH
#undef CDFT_L
}

// Generic FP64 field transforms on the fused kernels (whole spectra in the
// sub-plane order, no weights): `count` fields N apart in, spectra M apart out.
template <typename I>
void launch_fields_forward(const I* in, cx<double>* spec, cx<double>* tmp, int count, const FastGeom& g,
                           const void* tw64, hipStream_t st) {
  const AliasMap id{0, 0.0, 0.0};
  const unsigned zb = static_cast<unsigned>((g.nx / 2) * (g.q / 8)), pb = static_cast<unsigned>(g.nzh * 4);
  const unsigned nf = static_cast<unsigned>(count);
#define CDFT_Z(NZ)                                                                                     \
  {                                                                                                    \
    auto k = kf_zfwd<double, NZ, I>;                                                                   \
    const int sm = zfwd_smem<NZ, double>();                                                            \
    set_smem(k, sm);                                                                                   \
    hipLaunchKernelGGL(k, dim3(zb, nf), dim3(2 * NZ), sm, st, in, tmp, g, as<cx<double>>(tw64), id,    \
                       static_cast<double*>(nullptr));                                                 \
  }
  if (g.nz == 64) CDFT_Z(64) else if (g.nz == 128) CDFT_Z(128) else throw std::runtime_error("fast path: nz");
#undef CDFT_Z
#define CDFT_L(NX, Q)                                                                                 \
  {                                                                                                   \
    auto k = kf_rho_xy<double, NX, Q>;                                                                \
    const int sm = RhoCfg<NX, Q>::smem;                                                               \
    set_smem(k, sm);                                                                                  \
    hipLaunchKernelGGL(k, dim3(pb, nf), dim3(512), sm, st, tmp, static_cast<const double*>(nullptr), spec, \
                       static_cast<double*>(nullptr), g, as<cx<double>>(tw64));                       \
  }
  if (g.nx == 64 && g.q == 32) CDFT_L(64, 32)
  else if (g.nx == 64 && g.q == 64) CDFT_L(64, 64)
  else if (g.nx == 128 && g.q == 32) CDFT_L(128, 32)
  else if (g.nx == 128 && g.q == 64) CDFT_L(128, 64)
  else if (g.nx == 256 && g.q == 32) CDFT_L(256, 32)
  else if (g.nx == 256 && g.q == 64) CDFT_L(256, 64)
  else throw std::runtime_error("fast path: plane shape");
#undef CDFT_L
}

template <typename O>
void launch_fields_inverse(const cx<double>* spec, O* out, const double* add, double scale, cx<double>* tmp,
                           int count, const FastGeom& g, const void* tw64, hipStream_t st) {
  const unsigned pb = static_cast<unsigned>(g.nzh * 4), zb = static_cast<unsigned>((g.nx / 2) * (g.q / 8));
  const unsigned nf = static_cast<unsigned>(count);
#define CDFT_L(NX, Q)                                                                                 \
  {                                                                                                   \
    auto k = kg_xy_inv<double, NX, Q>;                                                                \
    const int sm = FwdCfg<NX, Q, double, double>::smem;                                               \
    set_smem(k, sm);                                                                                  \
    hipLaunchKernelGGL(k, dim3(pb, nf), dim3(kBT), sm, st, spec, tmp, g, as<cx<double>>(tw64));       \
  }
  if (g.nx == 64 && g.q == 32) CDFT_L(64, 32)
  else if (g.nx == 64 && g.q == 64) CDFT_L(64, 64)
  else if (g.nx == 128 && g.q == 32) CDFT_L(128, 32)
  else if (g.nx == 128 && g.q == 64) CDFT_L(128, 64)
  else if (g.nx == 256 && g.q == 32) CDFT_L(256, 32)
  else if (g.nx == 256 && g.q == 64) CDFT_L(256, 64)
  else throw std::runtime_error("fast path: plane shape");
#undef CDFT_L
  // one part per sub-plane: with as many backward blocks as units no unit is split
#define CDFT_Z(NZ)                                                                                      \
  {                                                                                                     \
    auto k = kf_zinv<double, double, NZ, O>;                                                            \
    const int sm = zinv_smem<NZ, double>();                                                             \
    set_smem(k, sm);                                                                                    \
    hipLaunchKernelGGL(k, dim3(zb, nf), dim3(zinv_threads<NZ>()), sm, st, tmp, out, add, scale, g,      \
                       as<cx<double>>(tw64), static_cast<int>(pb), static_cast<const double*>(nullptr), 0, \
                       static_cast<double*>(nullptr));                                                  \
  }
  if (g.nz == 64) CDFT_Z(64) else if (g.nz == 128) CDFT_Z(128) else throw std::runtime_error("fast path: nz");
#undef CDFT_Z
}

template void launch_fields_forward<double>(const double*, cx<double>*, cx<double>*, int, const FastGeom&,
                                            const void*, hipStream_t);
template void launch_fields_forward<float>(const float*, cx<double>*, cx<double>*, int, const FastGeom&, const void*,
                                           hipStream_t);
template void launch_fields_inverse<double>(const cx<double>*, double*, const double*, double, cx<double>*, int,
                                            const FastGeom&, const void*, hipStream_t);
template void launch_fields_inverse<float>(const cx<double>*, float*, const double*, double, cx<double>*, int,
                                           const FastGeom&, const void*, hipStream_t);

#define CDFT_INST(S, CV)                                                                                         \
  template void launch_fast_zfwd<S>(const double*, cx<S>*, double*, const FastGeom&, const void*, AliasMap,      \
                                    hipStream_t);                                                                \
  template void launch_fast_rho_xy<S>(const cx<S>*, const double*, void*, double*, const FastGeom&, const void*, \
                                      hipStream_t);                                                              \
  template void launch_fast_plane_fwd<S, CV>(const void*, const void*, cx<S>*, const double*, const FastGeom&,     \
                                             const void*, const void*, ChannelSet, hipStream_t);                 \
  template void launch_fast_zpoint<S, CV>(cx<S>*, const double*, const double*, double*, double*,                 \
                                          const FastGeom&, const void*, const void*, ChannelSet, FmtParams,       \
                                          double, AliasMap, const double*, DcTable, hipStream_t);                 \
  template void launch_fast_plane_bwd<S, CV>(const cx<S>*, const void*, cx<S>*, const FastGeom&, const void*,     \
                                             hipStream_t);                                                       \
  template void launch_fast_zinv<S, CV>(const cx<S>*, double*, const double*, double, const FastGeom&,           \
                                        const void*, const double*, int, double*, hipStream_t);
CDFT_INST(float, float)
CDFT_INST(double, double)
#undef CDFT_INST

}  // namespace cdft
