// Mean-field pair potentials and lattice cell weights (init time), FP64.
//
// The pair potentials and their WCA / Barker-Henderson attractive tails follow
// include/dft/physics/potentials.hpp:53-368 of the reference; the lattice cell
// weights (zero / linear / 27-point quadratic-F rules) and the real-space
// weight table with dV/N baked in follow mean_field.hpp:41-196.  The pair
// potential and the cell rule compile for host and device; contexts build the
// table on the device (k_mf_table in context.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace cdft {

struct PairPotential {
  int kind = 0;  // 0 Lennard-Jones, 1 ten Wolde-Frenkel, 2 Wang-Ramirez-Dobnikar-Frenkel
  double sigma = 1.0, eps = 1.0, rc = -1.0, alpha = 50.0;
  double shift = 0.0, r_min = 0.0, v_min = 0.0, r_zero = 0.0, eps_eff = 0.0;

  __host__ __device__ double bare(double r) const {
    switch (kind) {
      case 0: {
        const double y2 = (sigma / r) * (sigma / r);
        const double y6 = y2 * y2 * y2;
        return 4.0 * eps * (y6 * y6 - y6);
      }
      case 1: {
        if (r < sigma) return 1e50;
        const double s = r / sigma;
        const double y = 1.0 / (s * s - 1.0);
        const double y3 = y * y * y;
        return (4.0 * eps / (alpha * alpha)) * (y3 * y3 - alpha * y3);
      }
      default: {
        if (r >= rc) return 0.0;
        const double y = sigma / r, z = rc / r;
        return eps_eff * (y * y - 1.0) * (z * z - 1.0) * (z * z - 1.0);
      }
    }
  }

  // make_lennard_jones / make_ten_wolde_frenkel / make_wang_ramirez_dobnikar_frenkel
  void finish() {
    if (kind == 0) {
      shift = rc > 0.0 ? bare(rc) : 0.0;
      r_min = std::pow(2.0, 1.0 / 6.0) * sigma;
      v_min = bare(r_min) - shift;
      r_zero = sigma * std::pow(0.5 * std::sqrt(1.0 + shift / eps) + 0.5, -1.0 / 6.0);
    } else if (kind == 1) {
      shift = rc > 0.0 ? bare(rc) : 0.0;
      r_min = sigma * std::sqrt(1.0 + std::pow(2.0 / alpha, 1.0 / 3.0));
      v_min = bare(r_min) - shift;
      r_zero = sigma * std::sqrt(1.0 + std::pow(25.0 * std::sqrt(1.0 + shift) + 25.0, -1.0 / 3.0));
    } else {
      const double q = rc / sigma;
      eps_eff = eps * 2.0 * q * q * std::pow(2.0 * (q * q - 1.0) / 3.0, -3.0);
      r_min = rc * std::pow((1.0 + 2.0 * q * q) / 3.0, -0.5);
      v_min = bare(r_min);
      shift = 0.0;
    }
  }

  // attractive tail; split 0 = WCA (flat v_min inside r_min), 1 = BH (zero inside r_zero)
  __host__ __device__ double tail(double r, int split) const {
    if (kind == 2) {
      if (r >= rc) return 0.0;
      if (split == 1) return r >= r_min ? bare(r) : 0.0;
      return r >= r_min ? bare(r) : v_min;
    }
    const double cut = rc > 0.0 ? rc : 1e30;
    if (r >= cut) return 0.0;
    if (split == 1) return r >= r_zero ? bare(r) - shift : 0.0;
    return r >= r_min ? bare(r) - shift : v_min;
  }
};

// Average of tail(r)/kT over one lattice cell at integer offset (sx,sy,sz).
__host__ __device__ inline double lattice_cell_weight(const PairPotential& pot, int split, int scheme, double kT,
                                                     double dx, long sx, long sy, long sz) {
  if (scheme == 0) {
    const double r = dx * std::sqrt(static_cast<double>(sx * sx + sy * sy + sz * sz));
    return pot.tail(r, split) / kT;
  }
  if (scheme == 3) {  // quadratic-F: offsets {-1/2, 0, 1/2}, equal weights 1/3 per axis
    double acc = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) {
          const double ox = 0.5 * (a - 1), oy = 0.5 * (b - 1), oz = 0.5 * (c - 1);
          const double rx = (sx + ox) * dx, ry = (sy + oy) * dx, rz = (sz + oz) * dx;
          acc += pot.tail(std::sqrt(rx * rx + ry * ry + rz * rz), split) / kT;
        }
    return acc / 27.0;
  }
  // linear rules and the Gauss schemes (which the reference maps to the linear rule): 8 cell corners
  double acc = 0.0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const double rx = (sx - 0.5 + a) * dx, ry = (sy - 0.5 + b) * dx, rz = (sz - 0.5 + c) * dx;
        acc += pot.tail(std::sqrt(rx * rx + ry * ry + rz * rz), split) / kT;
      }
  return acc / 8.0;
}

}  // namespace cdft
