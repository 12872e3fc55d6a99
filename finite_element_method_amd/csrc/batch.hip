// batch.hip -- stream, graph and timer helpers of the batched assembly
// (fem_batch_* in fem_host.cpp): the group launches of a batch are captured
// once into a hipGraph and replayed with one hipGraphLaunch.
#include <hip/hip_runtime.h>

#include <string>

#include "device_state.hpp"
#include "fem_device.hpp"

namespace femdev {

struct Graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
};

std::string stream_create(void*& s) {
  hipStream_t st = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  s = st;
  return "";
}

std::string event_create(void*& ev) {
  hipEvent_t e = nullptr;
  HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  ev = e;
  return "";
}

void event_destroy(void* ev) {
  if (ev) hipEventDestroy(static_cast<hipEvent_t>(ev));
}

std::string stream_wait(void* waiter, void* signaler, void* ev) {
  HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(ev), static_cast<hipStream_t>(signaler)));
  HIP_TRY(hipStreamWaitEvent(static_cast<hipStream_t>(waiter), static_cast<hipEvent_t>(ev), 0));
  return "";
}

void stream_destroy(void* s) {
  if (s) hipStreamDestroy(static_cast<hipStream_t>(s));
}

std::string stream_sync(void* s) {
  HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(s)));
  return "";
}

std::string graph_begin(void* stream) {
  HIP_TRY(hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal));
  return "";
}

std::string graph_end(void* stream, Graph*& g) {
  auto* ng = new Graph();
  hipError_t e = hipStreamEndCapture(static_cast<hipStream_t>(stream), &ng->graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&ng->exec, ng->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    graph_destroy(ng);
    return std::string("graph capture: ") + hipGetErrorString(e);
  }
  g = ng;
  return "";
}

std::string graph_launch(Graph* g, void* stream) {
  HIP_TRY(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
  return "";
}

void graph_destroy(Graph* g) {
  if (!g) return;
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  delete g;
}

Timer* timer_create() {
  auto* t = new Timer();
  if (hipEventCreate(&t->a) != hipSuccess || hipEventCreate(&t->b) != hipSuccess) {
    timer_destroy(t);
    return nullptr;
  }
  return t;
}

void timer_destroy(Timer* t) {
  if (!t) return;
  if (t->a) hipEventDestroy(t->a);
  if (t->b) hipEventDestroy(t->b);
  delete t;
}

std::string timer_start(Timer* t, void* stream) {
  HIP_TRY(hipEventRecord(t->a, static_cast<hipStream_t>(stream)));
  return "";
}

std::string timer_stop(Timer* t, void* stream, double& ms) {
  HIP_TRY(hipEventRecord(t->b, static_cast<hipStream_t>(stream)));
  HIP_TRY(hipEventSynchronize(t->b));
  float f = 0.f;
  HIP_TRY(hipEventElapsedTime(&f, t->a, t->b));
  ms = double(f);
  return "";
}

std::string timer_end(Timer* t, void* stream) {
  HIP_TRY(hipEventRecord(t->b, static_cast<hipStream_t>(stream)));
  return "";
}

std::string copy_to_host(void* dst, const void* src_device, size_t bytes) {
  if (!bytes) return "";
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(dst, src_device, bytes, hipMemcpyDeviceToHost));
  return "";
}

std::string device_arch() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
  std::string a(prop.gcnArchName);
  const size_t c = a.find(':');
  return c == std::string::npos ? a : a.substr(0, c);
}

double last_ms_between(Timer* t) {
  float f = 0.f;
  if (!t || hipEventSynchronize(t->b) != hipSuccess || hipEventElapsedTime(&f, t->a, t->b) != hipSuccess) return -1.0;
  return double(f);
}

}  // namespace femdev
