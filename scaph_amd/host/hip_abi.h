// hip_abi.h — thin C++ helpers over the C-ABI of libscaph_hip.so.  The host
// engine is compiled by the host C++ compiler and never includes a HIP
// header: every device interaction goes through include/scaph_hip.h.
#ifndef SCAPH_HOST_HIP_ABI_H
#define SCAPH_HOST_HIP_ABI_H

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "scaph_hip.h"

namespace scaph {

// Error raised by the host engine.  The CLI turns it into the reference's
// H_ERR output ("<error> in <file> at line <n>", src/gpu_engine.cuh:28-39)
// and exits; the engine C-ABI turns it into a status code.
struct Error : std::runtime_error {
  int status;
  Error(int st, const std::string& what) : std::runtime_error(what), status(st) {}
};

[[noreturn]] inline void raise_status(int st, const char* file, int line) {
  throw Error(st, std::string(scaph_error_string(st)) + " in " + file + " at line " +
                      std::to_string(line));
}

#define H_ERR(expr)                                              \
  do {                                                           \
    int _st = (expr);                                            \
    if (_st != 0) ::scaph::raise_status(_st, __FILE__, __LINE__); \
  } while (0)

// Owning device buffer.
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  ~DevBuf() { reset(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      reset();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void alloc(size_t bytes) {
    reset();
    if (bytes) H_ERR(scaph_dev_alloc(&p_, bytes));
    n_ = bytes;
  }
  void ensure(size_t bytes) {
    if (bytes > n_) alloc(bytes);
  }
  void reset() {
    if (p_) scaph_dev_free(p_);
    p_ = nullptr;
    n_ = 0;
  }
  template <typename T = void>
  T* get() const {
    return static_cast<T*>(p_);
  }
  size_t bytes() const { return n_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

// Owning pinned host buffer (hipHostMalloc; src/graph.h:100 cudaMallocHost).
class PinnedBuf {
 public:
  PinnedBuf() = default;
  ~PinnedBuf() { reset(); }
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  PinnedBuf(PinnedBuf&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  void alloc(size_t bytes) {
    reset();
    if (bytes) H_ERR(scaph_host_pinned_alloc(&p_, bytes));
    n_ = bytes;
  }
  void reset() {
    if (p_) scaph_host_pinned_free(p_);
    p_ = nullptr;
    n_ = 0;
  }
  template <typename T = void>
  T* get() const {
    return static_cast<T*>(p_);
  }
  size_t bytes() const { return n_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

// Pinned, mapped, coherent host memory that kernels write directly.
class MappedBuf {
 public:
  MappedBuf() = default;
  ~MappedBuf() { reset(); }
  MappedBuf(const MappedBuf&) = delete;
  MappedBuf& operator=(const MappedBuf&) = delete;
  void alloc(size_t bytes) {
    reset();
    H_ERR(scaph_host_mapped_alloc(&h_, &d_, bytes));
    n_ = bytes;
  }
  void reset() {
    if (h_) scaph_host_pinned_free(h_);
    h_ = d_ = nullptr;
    n_ = 0;
  }
  template <typename T = void>
  T* host() const {
    return static_cast<T*>(h_);
  }
  template <typename T = void>
  T* dev() const {
    return static_cast<T*>(d_);
  }
  size_t bytes() const { return n_; }

 private:
  void* h_ = nullptr;
  void* d_ = nullptr;
  size_t n_ = 0;
};

class Stream {
 public:
  Stream() { H_ERR(scaph_stream_create(&s_)); }
  // kernels on this stream leave `reserved_cus` CUs unused (-tricknum)
  explicit Stream(uint32_t reserved_cus) {
    H_ERR(scaph_stream_create_cu_limited(&s_, reserved_cus));
  }
  ~Stream() {
    if (s_) scaph_stream_destroy(s_);
  }
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
  void* get() const { return s_; }
  void sync() const { H_ERR(scaph_stream_sync(s_)); }
  bool idle() const {
    const int st = scaph_stream_query(s_);
    if (st == 0) return true;
    if (st == SCAPH_STATUS_NOT_READY) return false;
    raise_status(st, __FILE__, __LINE__);
  }

 private:
  void* s_ = nullptr;
};

class Event {
 public:
  Event() { H_ERR(scaph_event_create(&e_)); }
  ~Event() {
    if (e_) scaph_event_destroy(e_);
  }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
  void* get() const { return e_; }
  void record(void* stream) { H_ERR(scaph_event_record(e_, stream)); }
  void sync() { H_ERR(scaph_event_sync(e_)); }
  bool done() const {
    const int st = scaph_event_query(e_);
    if (st == 0) return true;
    if (st == SCAPH_STATUS_NOT_READY) return false;
    raise_status(st, __FILE__, __LINE__);
  }
  static float elapsed_ms(const Event& a, const Event& b) {
    float ms = 0;
    H_ERR(scaph_event_elapsed_ms(a.e_, b.e_, &ms));
    return ms;
  }

 private:
  void* e_ = nullptr;
};

}  // namespace scaph

#endif
