// Library-internal entry points shared by fmmfft.hip and server.hip (not
// part of the C ABI; hidden from the shared object's exports).
#pragma once
#include <stdint.h>

#include <atomic>

#include "../../include/fmmfft.h"

namespace fmmfft {

// last HIP error any library call saw (fmmfft_last_hip_error)
__attribute__((visibility("hidden"))) extern std::atomic<int> g_last_hip_error;

// fmmfft_plan_create, optionally without the default workspace (the plan
// can then only execute through fmmfft_execute_ws / fmmfft_execute_static).
__attribute__((visibility("hidden"))) int plan_create_impl(fmmfft_plan *out, int64_t n, int32_t p_blocks,
                                                           int32_t terms, int32_t leaf, int32_t dtype,
                                                           int32_t max_batch, bool with_workspace);

// fmmfft_execute_ws over count separate signals: signal i reads
// d_in_tab[i] and writes d_out_tab[i], both device tables of device
// addresses (the caller checks each address's alignment, as
// fmmfft_execute_ws does for its two pointers).  Tables and workspace must
// stay untouched until the execute has run on the stream.
__attribute__((visibility("hidden"))) int execute_table(fmmfft_plan plan, int64_t count,
                                                        const void *const *d_in_tab, void *const *d_out_tab,
                                                        int32_t dir, void *d_work, size_t work_bytes,
                                                        void *stream);

}  // namespace fmmfft
