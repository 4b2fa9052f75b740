// Streams, timing events, graph capture and device facts for callers
// without PyTorch.
//
// The reference records a projector call in a CUDA graph through CuPy's
// stream capture (example/cuda/graph.py:46-51) and never needs more of the
// runtime.  These entry points give the Python drop-in (a SIRT iteration
// or a batched iteration recorded once and replayed, the batch tuner's
// timings) and C callers the same through this library alone.  Streams are
// created at the default priority.  Captured work must only launch: every
// launch of this library is a kernel (fills included), so a captured graph
// holds no memset nodes.
#include <cstring>

#include "kk_internal.h"

extern "C" int kk_device_info(int32_t* cu_count, char* name,
                              int32_t name_len) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return kk::hip_fail(e, "hipGetDevice");
  if (cu_count != nullptr) {
    int n = 0;
    e = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount,
                              dev);
    if (e != hipSuccess) return kk::hip_fail(e, "hipDeviceGetAttribute");
    *cu_count = n;
  }
  if (name != nullptr && name_len > 0) {
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return kk::hip_fail(e, "hipGetDeviceProperties");
    // the marketing name can be empty (no device-id table on the host):
    // then the architecture name, e.g. "gfx950:sramecc+:xnack-"
    const char* src = prop.name[0] != '\0' ? prop.name : prop.gcnArchName;
    std::strncpy(name, src, static_cast<size_t>(name_len) - 1);
    name[name_len - 1] = '\0';
  }
  return KK_OK;
}

extern "C" int kk_stream_create(void** stream) {
  if (stream == nullptr) return kk::fail(KK_ERR_INVALID, "kk_stream_create");
  hipStream_t s = nullptr;
  const hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) return kk::hip_fail(e, "hipStreamCreateWithFlags");
  *stream = s;
  return KK_OK;
}

extern "C" int kk_stream_destroy(void* stream) {
  if (stream == nullptr) return KK_OK;
  const hipError_t e = hipStreamDestroy(static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return kk::hip_fail(e, "hipStreamDestroy");
  return KK_OK;
}

extern "C" int kk_event_create(void** event) {
  if (event == nullptr) return kk::fail(KK_ERR_INVALID, "kk_event_create");
  hipEvent_t ev = nullptr;
  const hipError_t e = hipEventCreate(&ev);
  if (e != hipSuccess) return kk::hip_fail(e, "hipEventCreate");
  *event = ev;
  return KK_OK;
}

extern "C" int kk_event_destroy(void* event) {
  if (event == nullptr) return KK_OK;
  const hipError_t e = hipEventDestroy(static_cast<hipEvent_t>(event));
  if (e != hipSuccess) return kk::hip_fail(e, "hipEventDestroy");
  return KK_OK;
}

extern "C" int kk_event_record(void* event, void* stream) {
  if (event == nullptr) return kk::fail(KK_ERR_INVALID, "kk_event_record");
  const hipError_t e = hipEventRecord(static_cast<hipEvent_t>(event),
                                      static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return kk::hip_fail(e, "hipEventRecord");
  return KK_OK;
}

extern "C" int kk_event_synchronize(void* event) {
  if (event == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_event_synchronize");
  const hipError_t e = hipEventSynchronize(static_cast<hipEvent_t>(event));
  if (e != hipSuccess) return kk::hip_fail(e, "hipEventSynchronize");
  return KK_OK;
}

extern "C" int kk_event_elapsed_ms(void* start, void* stop, float* ms) {
  if (start == nullptr || stop == nullptr || ms == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_event_elapsed_ms");
  const hipError_t e = hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                           static_cast<hipEvent_t>(stop));
  if (e != hipSuccess) return kk::hip_fail(e, "hipEventElapsedTime");
  return KK_OK;
}

extern "C" int kk_stream_wait_event(void* stream, void* event) {
  if (event == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_stream_wait_event");
  const hipError_t e = hipStreamWaitEvent(static_cast<hipStream_t>(stream),
                                          static_cast<hipEvent_t>(event), 0);
  if (e != hipSuccess) return kk::hip_fail(e, "hipStreamWaitEvent");
  return KK_OK;
}

extern "C" int kk_graph_begin_capture(void* stream) {
  if (stream == nullptr)
    return kk::fail(KK_ERR_INVALID,
                    "kk_graph_begin_capture: the null stream cannot capture");
  const hipError_t e = hipStreamBeginCapture(
      static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return kk::hip_fail(e, "hipStreamBeginCapture");
  return KK_OK;
}

extern "C" int kk_graph_end_capture(void* stream, void** graph_exec) {
  if (stream == nullptr || graph_exec == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_graph_end_capture");
  *graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(static_cast<hipStream_t>(stream),
                                     &graph);
  if (e != hipSuccess) return kk::hip_fail(e, "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  const hipError_t e2 = hipGraphDestroy(graph);
  if (e != hipSuccess) return kk::hip_fail(e, "hipGraphInstantiate");
  if (e2 != hipSuccess) {
    hipGraphExecDestroy(exec);
    return kk::hip_fail(e2, "hipGraphDestroy");
  }
  *graph_exec = exec;
  return KK_OK;
}

extern "C" int kk_graph_launch(void* graph_exec, void* stream) {
  if (graph_exec == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_graph_launch: null graph");
  const hipError_t e = hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec),
                                      static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return kk::hip_fail(e, "hipGraphLaunch");
  return KK_OK;
}

extern "C" int kk_graph_destroy(void* graph_exec) {
  if (graph_exec == nullptr) return KK_OK;
  const hipError_t e =
      hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec));
  if (e != hipSuccess) return kk::hip_fail(e, "hipGraphExecDestroy");
  return KK_OK;
}
