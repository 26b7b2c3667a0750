// Library-internal entry points shared by fmmfft.hip and server.hip (not
// part of the C ABI; hidden from the shared object's exports).
#pragma once
#include <stdint.h>

#include "../../include/fmmfft.h"

namespace fmmfft {

// fmmfft_plan_create, optionally without the default workspace (the plan
// can then only execute through fmmfft_execute_ws / fmmfft_execute_static).
__attribute__((visibility("hidden"))) int plan_create_impl(fmmfft_plan *out, int64_t n, int32_t p_blocks,
                                                           int32_t terms, int32_t leaf, int32_t dtype,
                                                           int32_t max_batch, bool with_workspace);

}  // namespace fmmfft
