// Host-only stand-in for the HIP runtime, for an AddressSanitizer build of the library's
// host code (precond.cpp, solver.cpp, capi.cpp) on a machine without a GPU.
//
// "Device" memory is ordinary heap memory (so ASan sees every access to it), streams and
// events are heap objects (so a use after destroy or a double destroy is reported), and
// work enqueued on a stream runs at once -- or, while the stream is being captured, is
// recorded into a graph whose replay runs the recorded closures.  The kernel launchers of
// fake_kernels.cpp enqueue closures that touch exactly the buffer extents the real
// kernels read and write.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <set>
#include <vector>

#include "fake_hip.hpp"

namespace {

struct FakeStream {
  bool capturing = false;
  std::vector<std::function<void()>> rec;
};
struct FakeGraph {
  std::vector<std::function<void()>> ops;
};
struct FakeEvent {
  int recorded = 0;
};

std::set<void*> g_streams;
std::set<void*> g_events;
std::set<void*> g_graphs;
std::set<void*> g_execs;
long g_captures = 0;
long g_dev_live = 0;

[[noreturn]] void die(const char* what) {
  std::fprintf(stderr, "fake_hip: %s\n", what);
  std::abort();
}

FakeStream* stream_of(hipStream_t s) {
  if (s == nullptr) return nullptr;  // legacy default stream: never captured
  if (!g_streams.count((void*)s)) die("operation on a stream that does not exist");
  return reinterpret_cast<FakeStream*>(s);
}

}  // namespace

namespace fakehip {

void enqueue(hipStream_t s, std::function<void()> fn) {
  FakeStream* fs = stream_of(s);
  if (fs && fs->capturing) {
    fs->rec.push_back(std::move(fn));
    return;
  }
  fn();
}

long live_device_allocations() { return g_dev_live; }
long live_streams() { return (long)g_streams.size(); }
long live_events() { return (long)g_events.size(); }
long live_graphs() { return (long)(g_graphs.size() + g_execs.size()); }
long captures() { return g_captures; }

}  // namespace fakehip

extern "C" {

hipError_t hipMalloc(void** ptr, size_t size) {
  if (!ptr) return hipErrorInvalidValue;
  *ptr = std::malloc(size ? size : 1);
  if (!*ptr) return hipErrorOutOfMemory;
  std::memset(*ptr, 0xA5, size);  // device memory starts with garbage
  ++g_dev_live;
  return hipSuccess;
}

hipError_t hipFree(void* ptr) {
  if (ptr) {
    std::free(ptr);
    --g_dev_live;
  }
  return hipSuccess;
}

hipError_t hipHostMalloc(void** ptr, size_t size, unsigned int) {
  if (!ptr) return hipErrorInvalidValue;
  *ptr = std::malloc(size ? size : 1);
  return *ptr ? hipSuccess : hipErrorOutOfMemory;
}

hipError_t hipHostFree(void* ptr) {
  std::free(ptr);
  return hipSuccess;
}

hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind) {
  std::memmove(dst, src, bytes);
  return hipSuccess;
}

hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t s) {
  fakehip::enqueue(s, [=] { std::memmove(dst, src, bytes); });
  return hipSuccess;
}

hipError_t hipMemsetAsync(void* dst, int value, size_t bytes, hipStream_t s) {
  fakehip::enqueue(s, [=] { std::memset(dst, value, bytes); });
  return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  FakeStream* fs = new FakeStream();
  g_streams.insert(fs);
  *s = reinterpret_cast<hipStream_t>(fs);
  return hipSuccess;
}

hipError_t hipStreamDestroy(hipStream_t s) {
  FakeStream* fs = stream_of(s);
  if (!fs) return hipErrorInvalidHandle;
  if (fs->capturing) die("stream destroyed while capturing");
  g_streams.erase(fs);
  delete fs;
  return hipSuccess;
}

hipError_t hipStreamSynchronize(hipStream_t s) {
  FakeStream* fs = stream_of(s);
  if (fs && fs->capturing) die("synchronize of a capturing stream");
  return hipSuccess;
}

hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) {
  FakeEvent* fe = new FakeEvent();
  g_events.insert(fe);
  *e = reinterpret_cast<hipEvent_t>(fe);
  return hipSuccess;
}

static FakeEvent* event_of(hipEvent_t e) {
  if (!g_events.count((void*)e)) die("operation on an event that does not exist");
  return reinterpret_cast<FakeEvent*>(e);
}

hipError_t hipEventDestroy(hipEvent_t e) {
  FakeEvent* fe = event_of(e);
  g_events.erase(fe);
  delete fe;
  return hipSuccess;
}

hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  FakeEvent* fe = event_of(e);
  (void)stream_of(s);
  fe->recorded = 1;
  return hipSuccess;
}

hipError_t hipEventSynchronize(hipEvent_t e) {
  (void)event_of(e);
  return hipSuccess;
}

hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned int) {
  (void)stream_of(s);
  (void)event_of(e);
  return hipSuccess;
}

hipError_t hipGetDevice(int* dev) {
  *dev = 0;
  return hipSuccess;
}

// In a host-only build a device variable is its host-side shadow of the same size.
hipError_t hipGetSymbolAddress(void** ptr, const void* symbol) {
  *ptr = const_cast<void*>(symbol);
  return hipSuccess;
}

hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "fake hip error"; }

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  FakeStream* fs = stream_of(s);
  if (!fs) return hipErrorInvalidValue;  // the legacy stream cannot be captured
  if (fs->capturing) die("nested capture");
  fs->capturing = true;
  fs->rec.clear();
  ++g_captures;
  return hipSuccess;
}

hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  FakeStream* fs = stream_of(s);
  if (!fs || !fs->capturing) die("end of a capture that was not begun");
  FakeGraph* fg = new FakeGraph();
  fg->ops = std::move(fs->rec);
  fs->rec.clear();
  fs->capturing = false;
  g_graphs.insert(fg);
  *g = reinterpret_cast<hipGraph_t>(fg);
  return hipSuccess;
}

hipError_t hipGraphInstantiate(hipGraphExec_t* ex, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  if (!g_graphs.count((void*)g)) die("instantiate of a graph that does not exist");
  FakeGraph* fe = new FakeGraph(*reinterpret_cast<FakeGraph*>(g));
  g_execs.insert(fe);
  *ex = reinterpret_cast<hipGraphExec_t>(fe);
  return hipSuccess;
}

hipError_t hipGraphDestroy(hipGraph_t g) {
  if (!g_graphs.count((void*)g)) die("destroy of a graph that does not exist");
  g_graphs.erase((void*)g);
  delete reinterpret_cast<FakeGraph*>(g);
  return hipSuccess;
}

hipError_t hipGraphExecDestroy(hipGraphExec_t ex) {
  if (!g_execs.count((void*)ex)) die("destroy of a graph exec that does not exist");
  g_execs.erase((void*)ex);
  delete reinterpret_cast<FakeGraph*>(ex);
  return hipSuccess;
}

hipError_t hipGraphLaunch(hipGraphExec_t ex, hipStream_t s) {
  if (!g_execs.count((void*)ex)) die("launch of a graph exec that does not exist");
  for (const auto& op : reinterpret_cast<FakeGraph*>(ex)->ops) fakehip::enqueue(s, op);
  return hipSuccess;
}

// The real launchers are replaced by fake_kernels.cpp; the host stubs of the kernels are
// still linked in and must never run.
hipError_t hipLaunchKernel(const void*, dim3, dim3, void**, size_t, hipStream_t) {
  die("a real kernel launcher ran in the host build");
}
hipError_t __hipPushCallConfiguration(dim3, dim3, size_t, hipStream_t) {
  die("a real kernel launcher ran in the host build");
}
hipError_t __hipPopCallConfiguration(dim3*, dim3*, size_t*, hipStream_t*) {
  die("a real kernel launcher ran in the host build");
}
void** __hipRegisterFatBinary(const void*) {
  static void* handle = nullptr;
  return &handle;
}
void __hipRegisterFunction(void**, const void*, char*, const char*, unsigned int, uint3*, uint3*, dim3*,
                           dim3*, int*) {}
void __hipRegisterVar(void**, void*, char*, char*, int, size_t, int, int) {}
void __hipUnregisterFatBinary(void**) {}

}  // extern "C"
