// Core runtime for the MI355X elliptic path: scalar types, errors, host and
// device memory, the HIP platform, the single-rank communicator and the .rc
// settings. Names and semantics follow the reference's L1 layer
// (include/types.h, include/memory.hpp, include/platform.hpp, include/comm.hpp,
// include/settings.hpp) so that code written against those headers reads the
// same; the implementation is HIP only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "libp_amd.h"

namespace libp {

typedef double dfloat;    // include/types.h:31-45
typedef int dlong;        // include/types.h:53-60
typedef long long hlong;  // include/types.h:62-69

/* Errors: LIBP_ABORT throws libp::exception (include/utils.hpp:57-76). */
class exception : public std::runtime_error {
 public:
  explicit exception(const std::string& what) : std::runtime_error(what) {}
};

#define LIBP_ABORT(msg, cond)                                                           \
  do {                                                                                  \
    if (cond) {                                                                         \
      std::stringstream _libp_ss;                                                       \
      _libp_ss << msg << " (" << __FILE__ << ":" << __LINE__ << ")";                    \
      throw ::libp::exception(_libp_ss.str());                                          \
    }                                                                                   \
  } while (0)

#define LIBP_FORCE_ABORT(msg) LIBP_ABORT(msg, true)

// Turns a nonzero status of the C ABI into an exception.
#define LIBP_CALL(expr)                                                                               \
  do {                                                                                                \
    int _st = (expr);                                                                                 \
    LIBP_ABORT(#expr << " failed: " << libp_status_string(_st) << " (" << libp_last_error() << ")", \
               _st != LIBP_OK);                                                                       \
  } while (0)

#define LIBP_HIP(expr)                                                               \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    LIBP_ABORT(#expr << " failed: " << hipGetErrorString(_e), _e != hipSuccess);     \
  } while (0)

/* Host array with shared (shallow-copy) ownership, like memory<T>
 * (include/memory.hpp:34-268). */
template <typename T>
class memory {
  std::shared_ptr<std::vector<T>> v;

 public:
  memory() {}
  explicit memory(size_t n, T init = T()) : v(std::make_shared<std::vector<T>>(n, init)) {}
  void malloc(size_t n, T init = T()) { v = std::make_shared<std::vector<T>>(n, init); }
  void free() { v.reset(); }
  size_t length() const { return v ? v->size() : 0; }
  T* ptr() { return v ? v->data() : nullptr; }
  const T* ptr() const { return v ? v->data() : nullptr; }
  T& operator[](size_t i) { return (*v)[i]; }
  const T& operator[](size_t i) const { return (*v)[i]; }
  std::vector<T>& vec() { return *v; }
  memory clone() const {
    memory m;
    if (v) m.v = std::make_shared<std::vector<T>>(*v);
    return m;
  }
};

/* Device array (hipMalloc) with shared ownership and offset views, the
 * deviceMemory<T> API of include/memory.hpp:288-474 without OCCA. */
template <typename T>
class deviceMemory {
  std::shared_ptr<void> base;
  T* p = nullptr;
  size_t n = 0;

 public:
  deviceMemory() {}
  static deviceMemory alloc(size_t count) {
    deviceMemory m;
    if (count == 0) return m;
    void* raw = nullptr;
    LIBP_HIP(hipMalloc(&raw, count * sizeof(T)));
    m.base = std::shared_ptr<void>(raw, [](void* q) { (void)hipFree(q); });
    m.p = static_cast<T*>(raw);
    m.n = count;
    return m;
  }
  // Non-owning view of caller memory (C ABI entry points).
  static deviceMemory wrap(T* ptr, size_t count) {
    deviceMemory m;
    m.p = ptr;
    m.n = count;
    return m;
  }
  bool isInitialized() const { return p != nullptr; }
  T* ptr() const { return p; }
  size_t length() const { return n; }
  deviceMemory operator+(size_t off) const {
    deviceMemory m = *this;
    m.p = p + off;
    m.n = (off <= n) ? n - off : 0;
    return m;
  }
  void copyFrom(const T* src, size_t count, size_t off, hipStream_t s) {
    if (count == 0) return;
    LIBP_ABORT("deviceMemory::copyFrom out of range", off + count > n);
    LIBP_HIP(hipMemcpyAsync(p + off, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void copyTo(T* dst, size_t count, size_t off, hipStream_t s) const {
    if (count == 0) return;
    LIBP_ABORT("deviceMemory::copyTo out of range", off + count > n);
    LIBP_HIP(hipMemcpyAsync(dst, p + off, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

/* Single-rank communicator: every collective of comm_t (include/comm.hpp) is
 * the identity on one rank. */
struct comm_t {
  int rank() const { return 0; }
  int size() const { return 1; }
  template <typename T>
  void Allreduce(T&) const {}
};

/* .rc settings: libs/core/settings.cpp. compareSetting is a substring test
 * (settings.cpp:73-75); unknown names abort. */
class settings_t {
 public:
  struct setting_t {
    std::string name, val, description;
    std::vector<std::string> options;
  };
  std::map<std::string, setting_t> settings;
  std::vector<std::string> insertOrder;

  void newSetting(const std::string& name, const std::string& val, const std::string& description = "",
                  const std::vector<std::string>& options = {});
  bool hasSetting(const std::string& name) const { return settings.count(name) > 0; }
  void changeSetting(const std::string& name, const std::string& newVal);
  std::string getSetting(const std::string& name) const;
  template <typename T>
  void getSetting(const std::string& name, T& value) const {
    std::stringstream ss(getSetting(name));
    ss >> value;
  }
  bool compareSetting(const std::string& name, const std::string& token) const;
  // settings.cpp:134-214, reading from a string instead of a file
  static std::vector<std::pair<std::string, std::string>> parseText(const std::string& text);
  void readSettingsFromText(const std::string& text);
  void report(std::ostream& os) const;
};

/* The elliptic driver's settings: the union of platformSettings_t,
 * meshSettings_t, ellipticSettings_t(+run settings) and parAlmond settings,
 * with their defaults, plus the keys this build adds. */
settings_t ellipticDefaultSettings();

/* HIP platform: one device, one stream (default priority), a reduction
 * workspace and a pinned scalar buffer. Replaces platform_t
 * (include/platform.hpp:64-233) without any kernel JIT. */
class platform_t {
 public:
  int device = 0;
  hipStream_t stream = nullptr;
  std::shared_ptr<void> state;  // owns stream + buffers
  deviceMemory<char> o_reduceWS;
  deviceMemory<double> o_scalars;  // general-purpose device scalars
  double* h_scalars = nullptr;     // pinned mirror (16 doubles), mapped for the device
  double* d_hScalars = nullptr;    // its device pointer
  hipEvent_t ev_read = nullptr;    // recorded behind each read

  platform_t() {}
  explicit platform_t(const settings_t& s);
  template <typename T>
  deviceMemory<T> malloc(size_t n) const {
    return deviceMemory<T>::alloc(n);
  }
  template <typename T>
  deviceMemory<T> malloc(const memory<T>& h) const {
    deviceMemory<T> d = deviceMemory<T>::alloc(h.length());
    d.copyFrom(h.ptr(), h.length(), 0, stream);
    return d;
  }
  template <typename T>
  deviceMemory<T> malloc(const std::vector<T>& h) const {
    deviceMemory<T> d = deviceMemory<T>::alloc(h.size());
    d.copyFrom(h.data(), h.size(), 0, stream);
    return d;
  }
  void finish() const { LIBP_HIP(hipStreamSynchronize(stream)); }
  void* s() const { return (void*)stream; }
  // Reads device scalars [0,count) of ptr into the pinned buffer: a kernel
  // writes them into the mapped buffer and the host waits on an event behind it.
  const double* readScalars(const double* dptr, int count) const;
};

}  // namespace libp
