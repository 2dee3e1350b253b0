// device.hip — runtime part of the C-ABI: device selection, streams,
// events, memory and copies.  Replaces the CUDA runtime call sites of the
// reference engine (src/app.cuh:288-339 allocations and streams,
// src/app.cuh:191-197/385/391 copies, src/app.cuh:421/447 stream polling,
// src/graph.h:100 pinned page buffers).  Streams are created at the default
// priority.
#include <string.h>

#include <vector>

#include "common.h"

extern "C" {

const char* scaph_error_string(int status) {
  if (status == 0) return "success";
  if (status == SCAPH_STATUS_SCRATCH_TOO_SMALL)
    return "input too large for the no-alloc entry point's fixed scratch "
           "(scaph_bfs_push_noalloc_max_bytes)";
  return hipGetErrorString(static_cast<hipError_t>(status));
}

int scaph_init(int device) {
  SCAPH_TRY(hipSetDevice(device));
  SCAPH_TRY(hipFree(nullptr));  // create the context now, not on first use
  // the no-alloc push's workspace address on this device
  return scaph_noalloc_scratch_setup(device, nullptr);
}

int scaph_device_count(int* count) {
  if (!count) return SCAPH_STATUS_INVALID;
  SCAPH_TRY(hipGetDeviceCount(count));
  return 0;
}

int scaph_get_device_props(int device, scaph_device_props* out) {
  if (!out) return SCAPH_STATUS_INVALID;
  hipDeviceProp_t p;
  SCAPH_TRY(hipGetDeviceProperties(&p, device));
  memset(out, 0, sizeof(*out));
  strncpy(out->name, p.name, sizeof(out->name) - 1);
  strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
  out->compute_units = p.multiProcessorCount;
  out->wave_size = p.warpSize;
  out->lds_per_block = p.sharedMemPerBlock;
  out->total_mem = p.totalGlobalMem;
  out->l2_bytes = p.l2CacheSize;
  out->clock_khz = p.clockRate;
  out->mem_clock_khz = p.memoryClockRate;
  out->mem_bus_width = p.memoryBusWidth;
  out->xcds = 0;
  return 0;
}

int scaph_stream_create(void** stream) {
  if (!stream) return SCAPH_STATUS_INVALID;
  hipStream_t s;
  SCAPH_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return 0;
}

int scaph_stream_destroy(void* stream) { return (int)hipStreamDestroy(as_stream(stream)); }
int scaph_stream_query(void* stream) { return (int)hipStreamQuery(as_stream(stream)); }
int scaph_stream_sync(void* stream) { return (int)hipStreamSynchronize(as_stream(stream)); }
int scaph_stream_wait_event(void* stream, void* event) {
  return (int)hipStreamWaitEvent(as_stream(stream), reinterpret_cast<hipEvent_t>(event), 0);
}

int scaph_event_create(void** event) {
  if (!event) return SCAPH_STATUS_INVALID;
  hipEvent_t e;
  SCAPH_TRY(hipEventCreate(&e));
  *event = e;
  return 0;
}
int scaph_event_destroy(void* event) { return (int)hipEventDestroy(reinterpret_cast<hipEvent_t>(event)); }
int scaph_event_record(void* event, void* stream) {
  return (int)hipEventRecord(reinterpret_cast<hipEvent_t>(event), as_stream(stream));
}
int scaph_event_query(void* event) { return (int)hipEventQuery(reinterpret_cast<hipEvent_t>(event)); }
int scaph_event_sync(void* event) { return (int)hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)); }
int scaph_event_elapsed_ms(void* start, void* stop, float* ms) {
  if (!ms) return SCAPH_STATUS_INVALID;
  return (int)hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                                  reinterpret_cast<hipEvent_t>(stop));
}
int scaph_device_sync(void) { return (int)hipDeviceSynchronize(); }
int scaph_mem_info(size_t* free_bytes, size_t* total_bytes) {
  if (!free_bytes || !total_bytes) return SCAPH_STATUS_INVALID;
  return (int)hipMemGetInfo(free_bytes, total_bytes);
}

int scaph_dev_alloc(void** ptr, size_t bytes) {
  if (!ptr) return SCAPH_STATUS_INVALID;
  *ptr = nullptr;
  if (bytes == 0) return 0;
  return (int)hipMalloc(ptr, bytes);
}
int scaph_dev_free(void* ptr) { return ptr ? (int)hipFree(ptr) : 0; }

int scaph_host_pinned_alloc(void** ptr, size_t bytes) {
  if (!ptr) return SCAPH_STATUS_INVALID;
  *ptr = nullptr;
  if (bytes == 0) return 0;
  return (int)hipHostMalloc(ptr, bytes, hipHostMallocDefault);
}
int scaph_host_pinned_free(void* ptr) { return ptr ? (int)hipHostFree(ptr) : 0; }
int scaph_host_device_pointer(const void* host_ptr, void** device_ptr) {
  if (!host_ptr || !device_ptr) return SCAPH_STATUS_INVALID;
  *device_ptr = nullptr;
  return (int)hipHostGetDevicePointer(device_ptr, const_cast<void*>(host_ptr), 0);
}

int scaph_copy_h2d_async(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return 0;
  return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream));
}
int scaph_copy_d2h_async(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return 0;
  return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream));
}
int scaph_copy_d2d_async(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return 0;
  return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream));
}

}  // extern "C"

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* dst, uint32_t value, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  // 16-byte stores for the aligned bulk
  const size_t n4 = ((reinterpret_cast<uintptr_t>(dst) & 15) == 0) ? n / 4 : 0;
  uint4* d4 = reinterpret_cast<uint4*>(dst);
  const uint4 v4 = make_uint4(value, value, value, value);
  for (size_t j = i; j < n4; j += stride) d4[j] = v4;
  for (size_t j = n4 * 4 + i; j < n; j += stride) dst[j] = value;
}

extern "C" int scaph_fill_u32(uint32_t* dst, uint32_t value, size_t n, void* stream) {
  if (n == 0) return 0;
  if (!dst) return SCAPH_STATUS_INVALID;
  size_t blocks = (n / 4 + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), dst,
                     value, n);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

// Sum of n u32 words read once each, in 16-byte loads: a kernel of known
// read bytes (4 n) against which the memory-side read counters are scaled.
__global__ void __launch_bounds__(256) k_checksum_u32(const uint32_t* src, size_t n,
                                                      unsigned long long* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t n4 = ((reinterpret_cast<uintptr_t>(src) & 15) == 0) ? n / 4 : 0;
  const uint4* s4 = reinterpret_cast<const uint4*>(src);
  unsigned long long acc = 0;
  for (size_t j = i; j < n4; j += stride) {
    const uint4 v = s4[j];
    acc += (unsigned long long)v.x + v.y + v.z + v.w;
  }
  for (size_t j = n4 * 4 + i; j < n; j += stride) acc += src[j];
  acc = wave_sum(acc);
  if (lane_id() == 0 && acc) atomicAdd(out, acc);
}

__global__ void k_zero_one_u64(unsigned long long* p) { *p = 0; }

extern "C" int scaph_checksum_u32(const uint32_t* src, size_t n, unsigned long long* out,
                                  void* stream) {
  if (!out || (n && !src)) return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_zero_one_u64, dim3(1), dim3(1), 0, st, out);
  SCAPH_LAUNCH_CHECK();
  if (n == 0) return 0;
  size_t blocks = (n / 4 + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_checksum_u32, dim3((unsigned)blocks), dim3(256), 0, st, src, n, out);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

// Copies `words` u32 from device memory into host-mapped memory, clears
// the first `clear_words` of the source, then stores `seq` into the host
// flag after a system-scope fence: the host sees the flag only once the data
// is visible.  One wave.
__global__ void __launch_bounds__(64)
k_publish(uint32_t* src, uint32_t words, uint32_t clear_words, uint32_t* dst, uint32_t* flag,
          uint32_t seq) {
  for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) {
    dst[i] = src[i];
    if (i < clear_words) src[i] = 0u;
  }
  __threadfence_system();
  __builtin_amdgcn_wave_barrier();
  if (threadIdx.x == 0)
    __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// XCC (XCD) that runs each workgroup, from the hardware register.
__global__ void k_xcc_ids(uint32_t* out, uint32_t blocks) {
  uint32_t x;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
  if (threadIdx.x == 0 && blockIdx.x < blocks) out[blockIdx.x] = x & 0xFu;
}

// XCC and the HW_ID register (CU, shader array, shader engine) of the wave
// that runs each workgroup.
__global__ void k_hw_ids(uint32_t* out, uint32_t blocks) {
  uint32_t x, h;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_HW_ID)" : "=s"(h));
  if (threadIdx.x == 0 && blockIdx.x < blocks) {
    out[2ull * blockIdx.x] = x & 0xFu;
    out[2ull * blockIdx.x + 1] = h;
  }
}

extern "C" {

int scaph_hw_ids(uint32_t* out, uint32_t blocks, void* stream) {
  if (!out || blocks == 0) return SCAPH_STATUS_INVALID;
  hipLaunchKernelGGL(k_hw_ids, dim3(blocks), dim3(64), 0, as_stream(stream), out, blocks);
  return (int)hipGetLastError();
}

int scaph_stream_create_cu_limited(void** stream, uint32_t reserved_cus) {
  if (!stream) return SCAPH_STATUS_INVALID;
  if (reserved_cus == 0) return scaph_stream_create(stream);
  int dev = 0, cus = 0;
  SCAPH_TRY(hipGetDevice(&dev));
  SCAPH_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (cus <= 0 || reserved_cus >= (uint32_t)cus) return SCAPH_STATUS_INVALID;
  // keep CUs [0, cus - reserved) of the mask; the runtime maps mask bits to
  // CUs across the XCDs
  const uint32_t keep = (uint32_t)cus - reserved_cus;
  std::vector<uint32_t> mask(((uint32_t)cus + 31) / 32, 0u);
  for (uint32_t i = 0; i < keep; ++i) mask[i / 32] |= 1u << (i % 32);
  hipStream_t s;
  SCAPH_TRY(hipExtStreamCreateWithCUMask(&s, (uint32_t)mask.size(), mask.data()));
  *stream = s;
  return 0;
}

int scaph_xcc_ids(uint32_t* out, uint32_t blocks, void* stream) {
  if (!out || blocks == 0) return SCAPH_STATUS_INVALID;
  hipLaunchKernelGGL(k_xcc_ids, dim3(blocks), dim3(64), 0, as_stream(stream), out, blocks);
  return (int)hipGetLastError();
}

int scaph_host_mapped_alloc(void** host_ptr, void** device_ptr, size_t bytes) {
  if (!host_ptr || !device_ptr || bytes == 0) return SCAPH_STATUS_INVALID;
  *host_ptr = *device_ptr = nullptr;
  SCAPH_TRY(hipHostMalloc(host_ptr, bytes, hipHostMallocMapped | hipHostMallocCoherent));
  const hipError_t e = hipHostGetDevicePointer(device_ptr, *host_ptr, 0);
  if (e != hipSuccess) {
    (void)hipHostFree(*host_ptr);  // the mapping error is the one reported
    *host_ptr = nullptr;
    return (int)e;
  }
  return 0;
}

int scaph_publish_words(uint32_t* src, uint32_t words, uint32_t clear_words, uint32_t* dst_mapped,
                        uint32_t* flag_mapped, uint32_t seq, void* stream) {
  if (!src || !dst_mapped || !flag_mapped || clear_words > words) return SCAPH_STATUS_INVALID;
  hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, as_stream(stream), src, words, clear_words,
                     dst_mapped, flag_mapped, seq);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------- hipGraph
extern "C" {

int scaph_graph_capture_begin(void* stream) {
  return (int)hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal);
}

int scaph_graph_capture_end(void* stream, void** exec) {
  if (!exec) return SCAPH_STATUS_INVALID;
  hipGraph_t g = nullptr;
  SCAPH_TRY(hipStreamEndCapture(as_stream(stream), &g));
  hipGraphExec_t ge = nullptr;
  hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return (int)e;
  *exec = ge;
  return 0;
}

int scaph_graph_launch(void* exec, void* stream) {
  return (int)hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(exec), as_stream(stream));
}

int scaph_graph_destroy(void* exec) {
  return exec ? (int)hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(exec)) : 0;
}

}  // extern "C"
