// C ABI entry points (include/cdft.h): context management, evaluation and
// transforms.  FIRE / Minimizer and DDFT entry points live in dynamics.hip.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "api_guard.hpp"
#include "context.hpp"

namespace cdft {

thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

namespace {

// Copy a host array into a device buffer on the context stream.
void upload(Context* c, double* dst, const double* src, long n) {
  CDFT_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
}

void download(Context* c, double* dst, const double* src, long n) {
  CDFT_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  CDFT_HIP(hipStreamSynchronize(c->stream));
}

// plane-major device spectrum (kz, x, y) <-> reference half layout (x, y, kz), interleaved doubles
void spectrum_to_reference(Context* c, double* out) {
  const Geom& g = c->g;
  std::vector<double> tmp(2 * static_cast<size_t>(g.M));
  if (c->precision == CDFT_PRECISION_MIXED) {
    std::vector<float> f(2 * static_cast<size_t>(g.M));
    CDFT_HIP(hipMemcpyAsync(f.data(), c->d_S0, sizeof(float) * 2 * g.M, hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < f.size(); ++i) tmp[i] = f[i];
  } else {
    CDFT_HIP(hipMemcpyAsync(tmp.data(), c->d_S0, sizeof(double) * 2 * g.M, hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
  }
  for (int kz = 0; kz < g.nzh; ++kz)
    for (int x = 0; x < g.nx; ++x)
      for (int y = 0; y < g.ny; ++y) {
        const long s = (static_cast<long>(kz) * g.nx + x) * g.ny + y;
        const long r = kz + static_cast<long>(g.nzh) * (y + static_cast<long>(g.ny) * x);
        out[2 * r] = tmp[2 * s];
        out[2 * r + 1] = tmp[2 * s + 1];
      }
}

void reference_to_spectrum(Context* c, const double* in) {
  const Geom& g = c->g;
  std::vector<double> tmp(2 * static_cast<size_t>(g.M));
  for (int kz = 0; kz < g.nzh; ++kz)
    for (int x = 0; x < g.nx; ++x)
      for (int y = 0; y < g.ny; ++y) {
        const long s = (static_cast<long>(kz) * g.nx + x) * g.ny + y;
        const long r = kz + static_cast<long>(g.nzh) * (y + static_cast<long>(g.ny) * x);
        tmp[2 * s] = in[2 * r];
        tmp[2 * s + 1] = in[2 * r + 1];
      }
  if (c->precision == CDFT_PRECISION_MIXED) {
    std::vector<float> f(tmp.begin(), tmp.end());
    CDFT_HIP(hipMemcpyAsync(c->d_S0, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
  } else {
    CDFT_HIP(hipMemcpyAsync(c->d_S0, tmp.data(), sizeof(double) * tmp.size(), hipMemcpyHostToDevice, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
  }
}

// Staging of cdft_weighted_densities_noalloc: the n_fields * N doubles of the
// weighted densities land here between the kernels and the copy to the host.
// One copy per device, held for the life of the process.  In host code the
// array's name is not a device address: kernels and copies take the address
// resolved by scratch_on() for the current device.
constexpr size_t kScratchBytes = static_cast<size_t>(CDFT_NOALLOC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on `device`, which must be the current device;
// looked up once per device (first at context creation) and kept.
void* scratch_on(int device) {
  if (device < 0 || device >= kMaxScratchDevices) throw std::invalid_argument("device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    CDFT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HipError("hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

}  // namespace
}  // namespace cdft

using namespace cdft;

extern "C" {

const char* cdft_last_error(void) { return g_last_error.c_str(); }

int cdft_ctx_create(const cdft_grid* grid, const double* diameters, int n_species,
                    const cdft_interaction* interactions, int n_interactions, double kT,
                    const cdft_options* options, cdft_ctx** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("cdft_ctx_create: out is null");
    *out = nullptr;
    Context* c = create_context(grid, diameters, n_species, interactions, n_interactions, kT, options);
    try {
      // set-up of the no-allocation entry point's workspace on this device
      CDFT_HIP(hipSetDevice(c->device));
      (void)scratch_on(c->device);
    } catch (...) {
      delete c;
      throw;
    }
    *out = new cdft_ctx{c};
  });
}

void cdft_ctx_destroy(cdft_ctx* ctx) {
  if (!ctx) return;
  delete ctx->impl;
  delete ctx;
}

int cdft_grid_shape(const cdft_ctx* ctx, long shape[3]) {
  return guarded([&] {
    check_ctx_any(ctx);
    for (int d = 0; d < 3; ++d) shape[d] = ctx->impl->shape[d];
  });
}

int cdft_schedule(const cdft_ctx* ctx, int* schedule) {
  return guarded([&] {
    check_ctx_any(ctx);
    if (!schedule) throw std::invalid_argument("cdft_schedule: null argument");
    *schedule = ctx->impl->fast ? CDFT_SCHEDULE_FUSED : CDFT_SCHEDULE_GENERAL;
  });
}

int cdft_a_vdw(const cdft_ctx* ctx, int k, double* a_vdw) {
  return guarded([&] {
    check_ctx_any(ctx);
    if (k < 0 || k >= static_cast<int>(ctx->impl->a_vdw.size())) throw std::invalid_argument("no such interaction");
    *a_vdw = ctx->impl->a_vdw[k];
  });
}

int cdft_bulk_chemical_potential(const cdft_ctx* ctx, double rho, double* mu) {
  return guarded([&] {
    check_ctx(ctx);
    if (!(rho > 0.0)) throw std::invalid_argument("bulk density must be positive");
    *mu = ctx->impl->bulk_mu(rho);
  });
}

int cdft_total(cdft_ctx* ctx, const double* rho, const double* vext, double mu, cdft_result* result,
               double* forces) {
  return guarded([&] {
    check_ctx(ctx);
    if (!rho || !result) throw std::invalid_argument("cdft_total: rho and result are required");
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    upload(c, c->d_rho, rho, c->g.N);
    if (vext) upload(c, c->d_vext, vext, c->g.N);
    c->eval(c->d_rho, vext ? c->d_vext : nullptr, mu, forces ? c->d_force : nullptr, AliasMap{0, 0.0, 0.0});
    c->fetch_result(mu, result);
    if (forces) download(c, forces, c->d_force, c->g.N);
  });
}

int cdft_total_device(cdft_ctx* ctx, const double* d_rho, const double* d_vext, double mu, cdft_result* result,
                      double* d_forces) {
  return cdft_total_device_repeat(ctx, d_rho, d_vext, mu, 1, result, d_forces);
}

int cdft_total_device_repeat(cdft_ctx* ctx, const double* d_rho, const double* d_vext, double mu, int count,
                             cdft_result* result, double* d_forces) {
  return guarded([&] {
    check_ctx(ctx);
    if (!d_rho) throw std::invalid_argument("cdft_total_device: rho is required");
    if (count < 1) throw std::invalid_argument("count must be >= 1");
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    for (int i = 0; i < count; ++i) c->eval(d_rho, d_vext, mu, d_forces, AliasMap{0, 0.0, 0.0});
    if (result) c->fetch_result(mu, result);
    else CDFT_HIP(hipStreamSynchronize(c->stream));
  });
}

int cdft_time_total_device(cdft_ctx* ctx, const double* d_rho, const double* d_vext, double mu, int count,
                           double* d_forces, double* ms) {
  return guarded([&] {
    check_ctx(ctx);
    if (!d_rho || !ms || count < 1) throw std::invalid_argument("cdft_time_total_device: bad argument");
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    hipEvent_t e0, e1;
    CDFT_HIP(hipEventCreate(&e0));
    CDFT_HIP(hipEventCreate(&e1));
    CDFT_HIP(hipEventRecord(e0, c->stream));
    for (int i = 0; i < count; ++i) c->eval(d_rho, d_vext, mu, d_forces, AliasMap{0, 0.0, 0.0});
    CDFT_HIP(hipEventRecord(e1, c->stream));
    CDFT_HIP(hipEventSynchronize(e1));
    float t = 0.0f;
    CDFT_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}

int cdft_profile_stages(cdft_ctx* ctx, const double* d_rho, const double* d_vext, double mu, double* d_forces,
                        int reps, double* stage_ms, double* stage_bytes, int* n_stages) {
  return guarded([&] {
    check_ctx(ctx);
    if (!d_rho || !stage_ms || reps < 1) throw std::invalid_argument("cdft_profile_stages: bad argument");
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    const int ns = c->nstages();
    hipEvent_t ev[kNumStages + 1];
    for (int i = 0; i <= ns; ++i) CDFT_HIP(hipEventCreate(&ev[i]));
    for (int i = 0; i < ns; ++i) stage_ms[i] = 0.0;
    c->stage_events = ev;
    try {
      for (int r = 0; r < reps; ++r) {
        // events of skipped stages repeat the previous timestamp
        for (int i = 0; i <= ns; ++i) CDFT_HIP(hipEventRecord(ev[i], c->stream));
        c->eval(d_rho, d_vext, mu, d_forces, AliasMap{0, 0.0, 0.0});
        CDFT_HIP(hipEventSynchronize(ev[ns]));
        for (int i = 0; i < ns; ++i) {
          float t = 0.0f;
          CDFT_HIP(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
          stage_ms[i] += t / reps;
        }
      }
    } catch (...) {
      c->stage_events = nullptr;
      throw;
    }
    c->stage_events = nullptr;
    for (int i = 0; i <= ns; ++i) (void)hipEventDestroy(ev[i]);
    if (stage_bytes) c->stage_bytes(stage_bytes, d_vext != nullptr);
    if (n_stages) *n_stages = ns;
  });
}

const char* cdft_stage_name(const cdft_ctx* ctx, int i) {
  if (!ctx || !ctx->impl) return "";
  return ctx->impl->stage_name(i);
}

int cdft_weighted_densities(cdft_ctx* ctx, const double* rho, double* out) {
  return guarded([&] {
    check_ctx(ctx);
    if (!rho || !out) throw std::invalid_argument("cdft_weighted_densities: null argument");
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    const long n = c->g.N * c->cs.nf;
    double* d_out = nullptr;
    CDFT_HIP(hipMalloc(&d_out, sizeof(double) * (n > 0 ? n : 1)));
    try {
      upload(c, c->d_rho, rho, c->g.N);
      c->weighted_densities(c->d_rho, d_out);
      if (n > 0) download(c, out, d_out, n);
    } catch (...) {
      (void)hipFree(d_out);
      throw;
    }
    CDFT_HIP(hipFree(d_out));
  });
}

int cdft_weighted_densities_noalloc(cdft_ctx* ctx, const double* rho, double* out) {
  return guarded([&] {
    check_ctx(ctx);
    if (!rho || !out) throw std::invalid_argument("cdft_weighted_densities_noalloc: null argument");
    Context* c = ctx->impl;
    const long n = c->g.N * c->cs.nf;
    if (static_cast<size_t>(n) * sizeof(double) > kScratchBytes)
      throw std::invalid_argument("cdft_weighted_densities_noalloc: " + std::to_string(c->cs.nf) + " fields of " +
                                  std::to_string(c->g.N) + " points need " +
                                  std::to_string(static_cast<size_t>(n) * sizeof(double)) +
                                  " bytes, more than the " + std::to_string(kScratchBytes) +
                                  "-byte workspace");
    CDFT_HIP(hipSetDevice(c->device));
    double* d_out = static_cast<double*>(scratch_on(c->device));
    upload(c, c->d_rho, rho, c->g.N);
    c->weighted_densities(c->d_rho, d_out);
    if (n > 0) download(c, out, d_out, n);
  });
}

int cdft_fft_r2c(cdft_ctx* ctx, const double* in, double* out_half) {
  return guarded([&] {
    check_ctx(ctx);
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    upload(c, c->d_rho, in, c->g.N);
    c->r2c(c->d_rho);
    spectrum_to_reference(c, out_half);
  });
}

int cdft_fft_c2r(cdft_ctx* ctx, const double* in_half, double* out) {
  return guarded([&] {
    check_ctx(ctx);
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    reference_to_spectrum(c, in_half);
    c->c2r(c->d_force);
    download(c, out, c->d_force, c->g.N);
  });
}

}  // extern "C"
