// Batched line transforms of the general schedule: `nl` lines of length L
// held in one LDS buffer (line stride ls), transformed in place by a block.
//
// For the lengths listed in fft_lines, each line goes to one slot of T lanes
// of a wavefront (slot_fft.hpp), which loads it into registers, runs every
// Stockham pass there with exchanges through the line's own storage, ordered
// at wavefront level, and stores it back: the block synchronises once, at
// the end, instead of twice per radix pass (fft_inplace, which every other
// length keeps).  Lengths 3 * 2^a and 5 * 2^a start with a radix-6 or
// radix-10 pass, then radix 2; powers of two use radix 8 or 4 first.  The
// slot transforms apply every twiddle and constant as one rounded value (no
// FP32 residual).
#pragma once

#include <hip/hip_runtime.h>

#include "fft_lds.hpp"
#include "slot_fft.hpp"

namespace cdft {

// Lanes per slot of a length with a register-resident transform (fft_lines'
// list), 0 for the lengths fft_inplace takes.
__host__ __device__ constexpr int slot_lanes(int L) {
  return (L == 48 || L == 64 || L == 80) ? 8 : (L == 96 || L == 128 || L == 160) ? 16 : 0;
}

// Slots of T lanes take lines slot, slot + nslot, ...; the caller synchronises
// before the call (inputs visible), the call ends with a block barrier.
template <typename C, int L, int T, int SG>
__device__ __forceinline__ void slot_lines(cx<C>* buf, int nl, int ls, const cx<C>* __restrict__ tw) {
  constexpr int E = L / T;
  const int nslot = static_cast<int>(blockDim.x) / T;
  const int slot = static_cast<int>(threadIdx.x) / T, lane = static_cast<int>(threadIdx.x) % T;
  for (int j = slot; j < nl; j += nslot) {
    cx<C>* line = buf + static_cast<long>(j) * ls;
    cx<C> v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = line[lane + T * q];
    slot_fft<C, L, T, SG, false, false>(v, line, lane, tw);
    wave_sync();  // every lane's last exchange read is done
#pragma unroll
    for (int e = 0; e < E; ++e) line[lane + T * e] = v[e];
    wave_sync();
  }
  __syncthreads();
}

template <typename C, int L, int T>
__device__ __forceinline__ void slot_lines_sg(cx<C>* buf, int nl, int ls, int sg, const cx<C>* __restrict__ tw) {
  if (sg < 0) slot_lines<C, L, T, -1>(buf, nl, ls, tw);
  else slot_lines<C, L, T, +1>(buf, nl, ls, tw);
}

// In-place transform of nl lines (fft_inplace's contract).  The FP32 tables
// of the general plans hold the rounded twiddles in [0, L), which is all the
// slot transforms read.
template <typename C, int MAXE>
__device__ __forceinline__ void fft_lines(cx<C>* buf, int nl, int ls, const FftPlan& p, int sg, const cx<C>* __restrict__ tw) {
  switch (p.L) {
    case 48: slot_lines_sg<C, 48, 8>(buf, nl, ls, sg, tw); return;
    case 64: slot_lines_sg<C, 64, 8>(buf, nl, ls, sg, tw); return;
    case 80: slot_lines_sg<C, 80, 8>(buf, nl, ls, sg, tw); return;
    case 96: slot_lines_sg<C, 96, 16>(buf, nl, ls, sg, tw); return;
    case 128: slot_lines_sg<C, 128, 16>(buf, nl, ls, sg, tw); return;
    case 160: slot_lines_sg<C, 160, 16>(buf, nl, ls, sg, tw); return;
    default: fft_inplace<C, MAXE>(buf, nl, ls, p, sg, tw); return;
  }
}

// fft_lines for the lengths with slot transforms only (slot_lanes(p.L) != 0,
// which the caller guarantees): no LDS-pass fallback and its registers.
template <typename C>
__device__ __forceinline__ void slot_lines_only(cx<C>* buf, int nl, int ls, const FftPlan& p, int sg,
                                                const cx<C>* __restrict__ tw) {
  switch (p.L) {
    case 48: slot_lines_sg<C, 48, 8>(buf, nl, ls, sg, tw); return;
    case 64: slot_lines_sg<C, 64, 8>(buf, nl, ls, sg, tw); return;
    case 80: slot_lines_sg<C, 80, 8>(buf, nl, ls, sg, tw); return;
    case 96: slot_lines_sg<C, 96, 16>(buf, nl, ls, sg, tw); return;
    case 128: slot_lines_sg<C, 128, 16>(buf, nl, ls, sg, tw); return;
    default: slot_lines_sg<C, 160, 16>(buf, nl, ls, sg, tw); return;
  }
}

}  // namespace cdft
