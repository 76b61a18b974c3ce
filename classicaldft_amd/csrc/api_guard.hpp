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

inline void check_ctx(const cdft_ctx* ctx) {
  if (!ctx || !ctx->impl) throw std::invalid_argument("null context");
}

}  // namespace cdft
