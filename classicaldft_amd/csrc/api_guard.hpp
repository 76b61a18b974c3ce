// Exception -> status-code translation at the C ABI boundary.
#pragma once

#include <new>
#include <stdexcept>
#include <string>

#include "context.hpp"

namespace cdft {

// Runs `fn`; maps the reference's exception types onto the C status codes.
template <typename F>
int guarded(F&& fn) {
  try {
    fn();
    return CDFT_OK;
  } catch (const HipError& e) {
    set_error(e.what());
    return CDFT_EHIP;
  } catch (const Unsupported& e) {
    set_error(e.what());
    return CDFT_EUNSUPPORTED;
  } catch (const std::invalid_argument& e) {
    set_error(e.what());
    return CDFT_EINVAL;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return CDFT_ERUNTIME;
  } catch (const std::exception& e) {
    set_error(e.what());
    return CDFT_ERUNTIME;
  }
}

// Any context, one species or a mixture.
inline void check_ctx_any(const cdft_ctx* ctx) {
  if (!ctx || !ctx->impl) throw std::invalid_argument("null context");
}

// A one-species context: the algorithms (minimisation, DDFT, saddle search,
// Picard) and the one-species evaluation entry points, as in the reference
// (minimization.hpp:206-342 and dynamics.hpp act on species[0]).
inline void check_ctx(const cdft_ctx* ctx) {
  check_ctx_any(ctx);
  if (ctx->impl->n_species != 1)
    throw Unsupported("this entry point acts on one species; the context has " +
                      std::to_string(ctx->impl->n_species) + " (use cdft_total_species)");
}

}  // namespace cdft
