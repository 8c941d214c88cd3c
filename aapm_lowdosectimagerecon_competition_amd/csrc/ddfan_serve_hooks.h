// ddfan_serve_hooks.h -- the request loop's completion checks, replaceable.
//
// The worker learns that GPU work has finished, or failed, only through these
// three calls.  set_server_hooks swaps them (before the first submission), so
// a test can hand the worker an error status, as a failed kernel would, and
// check that every outstanding request then completes with DDFAN_EDEVICE
// instead of waiting for ever -- without making the device fail.
#pragma once

#include <hip/hip_runtime_api.h>

#include "ddfan.h"

namespace ddf {

struct ServerHooks {
    hipError_t (*eventQuery)(hipEvent_t) = hipEventQuery;
    hipError_t (*eventSync)(hipEvent_t) = hipEventSynchronize;
    hipError_t (*streamSync)(hipStream_t) = hipStreamSynchronize;
};

// Replaces the server's completion checks; call before its first submission.
void set_server_hooks(ddfan_server* server, const ServerHooks& hooks);

}  // namespace ddf
