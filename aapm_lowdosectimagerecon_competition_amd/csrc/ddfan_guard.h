// ddfan_guard.h -- keeps C++ exceptions from crossing the C ABI.
//
// Every extern "C" entry point runs its body through guarded(): a host
// allocation that fails (std::bad_alloc, or a thread that cannot be started)
// returns DDFAN_ENOMEM, any other exception DDFAN_EINVAL, each with a message
// for ddfan_last_error(), instead of terminating the caller's process.
#pragma once

#include <exception>
#include <new>
#include <system_error>

#include "ddfan.h"

namespace ddf {

int fail(int code, const char* fmt, ...);

template <class F>
inline int guarded(F&& body) noexcept {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return fail(DDFAN_ENOMEM, "host allocation failed");
    } catch (const std::system_error& e) {
        return fail(DDFAN_ENOMEM, "host resources exhausted: %s", e.what());
    } catch (const std::exception& e) {
        return fail(DDFAN_EINVAL, "internal error: %s", e.what());
    } catch (...) {
        return fail(DDFAN_EINVAL, "internal error");
    }
}

}  // namespace ddf
