// runtime.hip -- device memory, copies, streams and events behind the thin
// C-ABI layer used by the C host code.
#include "fmb_device.h"

extern "C" int fmbh_init(void) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return fmb_set_error(FMB_ERR_HIP, "no HIP device available (%s)",
                         hipGetErrorString(e));
  return FMB_OK;
}

// identity of the current device for keyed caches: architecture and CU count
extern "C" int fmbh_device_tag(char *buf, int n) {
  int dev = 0;
  FMB_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t pr;
  FMB_HIP_CHECK(hipGetDeviceProperties(&pr, dev));
  char arch[64];
  int i = 0;
  // the architecture name up to its feature suffix (gfx950:sramecc+:xnack-)
  for (; i < 63 && pr.gcnArchName[i] && pr.gcnArchName[i] != ':'; i++) arch[i] = pr.gcnArchName[i];
  arch[i] = 0;
  snprintf(buf, n, "%s/%dcu", arch, pr.multiProcessorCount);
  return FMB_OK;
}

extern "C" int fmbh_malloc(void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess)
    return fmb_set_error(FMB_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes,
                         hipGetErrorString(e));
  return FMB_OK;
}

extern "C" int fmbh_free(void *p) {
  if (p) FMB_HIP_CHECK(hipFree(p));
  return FMB_OK;
}

extern "C" int fmbh_h2d(void *dst, const void *src, size_t bytes,
                        void *stream) {
  if (!bytes) return FMB_OK;
  FMB_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                               (hipStream_t)stream));
  FMB_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return FMB_OK;
}

extern "C" int fmbh_d2h(void *dst, const void *src, size_t bytes,
                        void *stream) {
  if (!bytes) return FMB_OK;
  FMB_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost,
                               (hipStream_t)stream));
  FMB_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return FMB_OK;
}

extern "C" int fmbh_d2d(void *dst, const void *src, size_t bytes,
                        void *stream) {
  if (!bytes) return FMB_OK;
  FMB_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice,
                               (hipStream_t)stream));
  return FMB_OK;
}

__global__ void k_fill_zero(unsigned int *p, size_t nwords) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (; i < nwords; i += st) p[i] = 0u;
}

// Zero-fill with a kernel (capturable in graphs, unlike memset nodes).
extern "C" int fmbh_memset(void *dst, int v, size_t bytes, void *stream) {
  if (!bytes) return FMB_OK;
  if (v != 0 || (bytes & 3) || ((size_t)dst & 3)) {
    FMB_HIP_CHECK(hipMemsetAsync(dst, v, bytes, (hipStream_t)stream));
    return FMB_OK;
  }
  size_t nw = bytes / 4;
  int blocks = (int)((nw + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0,
                     (hipStream_t)stream, (unsigned int *)dst, nw);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

extern "C" int fmbh_sync(void *stream) {
  FMB_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return FMB_OK;
}

extern "C" int fmbh_event_create(void **ev) {
  hipEvent_t e;
  FMB_HIP_CHECK(hipEventCreate(&e));
  *ev = (void *)e;
  return FMB_OK;
}

extern "C" int fmbh_event_record(void *ev, void *stream) {
  FMB_HIP_CHECK(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
  return FMB_OK;
}

extern "C" int fmbh_event_elapsed(void *ev0, void *ev1, float *ms) {
  FMB_HIP_CHECK(hipEventSynchronize((hipEvent_t)ev1));
  FMB_HIP_CHECK(hipEventElapsedTime(ms, (hipEvent_t)ev0, (hipEvent_t)ev1));
  return FMB_OK;
}

extern "C" int fmbh_event_destroy(void *ev) {
  if (ev) FMB_HIP_CHECK(hipEventDestroy((hipEvent_t)ev));
  return FMB_OK;
}

/* streams (default priority) and graph capture of a stream's work */
extern "C" int fmbh_stream_create(void **s) {
  hipStream_t st;
  FMB_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  *s = st;
  return FMB_OK;
}

extern "C" int fmbh_stream_destroy(void *s) {
  if (s) FMB_HIP_CHECK(hipStreamDestroy((hipStream_t)s));
  return FMB_OK;
}

extern "C" int fmbh_stream_wait(void *stream, void *ev) {
  FMB_HIP_CHECK(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)ev, 0));
  return FMB_OK;
}

extern "C" int fmbh_capture_begin(void *stream) {
  FMB_HIP_CHECK(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeRelaxed));
  return FMB_OK;
}

/* ends the capture; *exec receives the instantiated graph (NULL on error,
 * the capture being closed either way) */
extern "C" int fmbh_capture_end(void *stream, void **exec) {
  hipGraph_t g = NULL;
  *exec = NULL;
  FMB_HIP_CHECK(hipStreamEndCapture((hipStream_t)stream, &g));
  hipGraphExec_t e = NULL;
  hipError_t err = hipGraphInstantiate(&e, g, NULL, NULL, 0);
  /* the graph is released whatever its instantiation gave */
  (void)hipGraphDestroy(g);
  if (err != hipSuccess)
    return fmb_set_error(FMB_ERR_HIP, "graph instantiation: %s", hipGetErrorString(err));
  *exec = e;
  return FMB_OK;
}

extern "C" int fmbh_graph_launch(void *exec, void *stream) {
  FMB_HIP_CHECK(hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream));
  return FMB_OK;
}

extern "C" int fmbh_graph_destroy(void *exec) {
  if (exec) FMB_HIP_CHECK(hipGraphExecDestroy((hipGraphExec_t)exec));
  return FMB_OK;
}

/* fork/join events (no timestamps) */
extern "C" int fmbh_event_create_sync(void **ev) {
  hipEvent_t e;
  FMB_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  *ev = e;
  return FMB_OK;
}
