// Hooks of the host-only HIP stand-in (fake_hip.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>

namespace fakehip {
// Run fn now, or record it if the stream is being captured.
void enqueue(hipStream_t s, std::function<void()> fn);
long live_device_allocations();
long live_streams();
long live_events();
long live_graphs();
// stream captures begun so far
long captures();
}  // namespace fakehip
