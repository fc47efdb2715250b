// common.hip — library state, error handling and argument residency of libkukan_hip.so.
#include <cstdio>
#include <cstring>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "kukan_internal.hpp"

namespace kukan {

Lib &lib() {
  static Lib L;
  return L;
}

void hip_check(hipError_t e, const char *what, const char *file, int line) {
  if (e == hipSuccess) return;
  char buf[512];
  snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what,
           file, line);
  throw Error(100, buf);
}

void blas_check(rocblas_status s, const char *what, const char *file, int line) {
  if (s == rocblas_status_success) return;
  char buf[512];
  snprintf(buf, sizeof buf, "rocBLAS/rocSOLVER status %d in %s at %s:%d", (int)s, what, file, line);
  throw Error(101, buf);
}

void ensure_init() {
  Lib &L = lib();
  if (L.stream) return;
  int dev = L.device >= 0 ? L.device : 0;
  HIPCK(hipSetDevice(dev));
  L.device = dev;
  HIPCK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
  BLASCK(rocblas_create_handle(&L.blas));
  BLASCK(rocblas_set_stream(L.blas, L.stream));
  BLASCK(rocblas_set_pointer_mode(L.blas, rocblas_pointer_mode_host));
}

double param(const char *key, double dflt) {
  auto &p = lib().params;
  auto it = p.find(key);
  return it == p.end() ? dflt : it->second;
}

void *scratch(const std::string &tag, size_t bytes) {
  auto &m = lib().scratch;
  auto it = m.find(tag);
  if (it != m.end() && it->second.second >= bytes) return it->second.first;
  if (lib().capturing) throw Error(1, "scratch allocation inside stream capture: " + tag);
  if (it != m.end()) {
    HIPCK(hipStreamSynchronize(lib().stream));
    HIPCK(hipFree(it->second.first));
    m.erase(it);
  }
  void *p = nullptr;
  size_t b = bytes < 256 ? 256 : bytes;
  HIPCK(hipMalloc(&p, b));
  m[tag] = {p, b};
  return p;
}

void sync() { HIPCK(hipStreamSynchronize(lib().stream)); }

namespace {
struct Pending {
  std::string tag;
  hipEvent_t a, b;
  double work;
};
std::vector<Pending> &pending() {
  static std::vector<Pending> p;
  return p;
}
std::vector<hipEvent_t> &event_pool() {
  static std::vector<hipEvent_t> p;
  return p;
}
hipEvent_t get_event() {
  auto &p = event_pool();
  if (!p.empty()) {
    hipEvent_t e = p.back();
    p.pop_back();
    return e;
  }
  hipEvent_t e;
  HIPCK(hipEventCreate(&e));
  return e;
}
}  // namespace

std::map<std::string, TimingAcc> &timing_table() {
  static std::map<std::string, TimingAcc> t;
  return t;
}

Timed::Timed(const char *tag, double work) : tag_(tag), work_(work) {
  if (std::strncmp(tag, "phase_", 6) == 0) {
    roctxRangePushA(tag);
    range_ = true;
  }
  if (param("timing", 0.0) < 1.0 || lib().capturing) return;
  a_ = get_event();
  b_ = get_event();
  HIPCK(hipEventRecord(a_, lib().stream));
}

Timed::~Timed() {
  if (range_) roctxRangePop();
  if (!a_) return;
  (void)hipEventRecord(b_, lib().stream);
  pending().push_back({tag_, a_, b_, work_});
  if (pending().size() > 4096) {
    try {
      timing_collect();
    } catch (...) {
    }
  }
}

void timing_add_work(const char *tag, double work) {
  if (param("timing", 0.0) < 1.0 || lib().capturing) return;
  timing_table()[tag].work += work;
}

void timing_collect() {
  auto &p = pending();
  if (p.empty()) return;
  HIPCK(hipStreamSynchronize(lib().stream));
  for (auto &e : p) {
    float ms = 0.f;
    HIPCK(hipEventElapsedTime(&ms, e.a, e.b));
    auto &acc = timing_table()[e.tag];
    acc.ms += ms;
    acc.work += e.work;
    acc.count += 1;
    event_pool().push_back(e.a);
    event_pool().push_back(e.b);
  }
  p.clear();
}

bool is_device_ptr(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

static void *find_bound(const void *p) {
  auto &b = lib().bound;
  auto it = b.upper_bound(p);
  if (it == b.begin()) return nullptr;
  --it;
  const char *base = (const char *)it->first;
  if ((const char *)p >= base && (const char *)p < base + it->second.bytes)
    return (char *)it->second.dev + ((const char *)p - base);
  return nullptr;
}

Arr::Arr(const void *p, size_t bytes, Intent it) : host_(p), bytes_(bytes), intent_(it) {
  if (!p || bytes == 0) return;
  if (void *m = find_bound(p)) {
    dev_ = m;
    return;
  }
  if (is_device_ptr(p)) {
    dev_ = const_cast<void *>(p);
    return;
  }
  temp_ = true;
  HIPCK(hipMalloc(&dev_, bytes));
  if (it & In)
    HIPCK(hipMemcpyAsync(dev_, p, bytes, hipMemcpyHostToDevice, lib().stream));
  else
    HIPCK(hipMemsetAsync(dev_, 0, bytes, lib().stream));
}

Arr &Arr::operator=(Arr &&o) noexcept {
  host_ = o.host_; dev_ = o.dev_; bytes_ = o.bytes_; intent_ = o.intent_; temp_ = o.temp_;
  o.temp_ = false; o.dev_ = nullptr;
  return *this;
}

void Arr::finish() {
  if (temp_ && (intent_ & Out)) {
    HIPCK(hipMemcpyAsync(const_cast<void *>(host_), dev_, bytes_, hipMemcpyDeviceToHost,
                         lib().stream));
    HIPCK(hipStreamSynchronize(lib().stream));
  }
  if (temp_) {
    HIPCK(hipStreamSynchronize(lib().stream));
    (void)hipFree(dev_);
    temp_ = false;
    dev_ = nullptr;
  }
}

Arr::~Arr() {
  if (temp_) {
    (void)hipStreamSynchronize(lib().stream);
    (void)hipFree(dev_);
  }
}

std::vector<int> host_ints(const int *p, size_t n) {
  std::vector<int> v(n);
  if (n == 0 || !p) return v;
  if (void *m = find_bound(p)) {
    HIPCK(hipMemcpyAsync(v.data(), m, n * sizeof(int), hipMemcpyDeviceToHost, lib().stream));
    sync();
  } else if (is_device_ptr(p)) {
    HIPCK(hipMemcpyAsync(v.data(), p, n * sizeof(int), hipMemcpyDeviceToHost, lib().stream));
    sync();
  } else {
    memcpy(v.data(), p, n * sizeof(int));
  }
  return v;
}

std::vector<double> host_doubles(const double *p, size_t n) {
  std::vector<double> v(n);
  if (n == 0 || !p) return v;
  if (void *m = find_bound(p)) {
    HIPCK(hipMemcpyAsync(v.data(), m, n * sizeof(double), hipMemcpyDeviceToHost, lib().stream));
    sync();
  } else if (is_device_ptr(p)) {
    HIPCK(hipMemcpyAsync(v.data(), p, n * sizeof(double), hipMemcpyDeviceToHost, lib().stream));
    sync();
  } else {
    memcpy(v.data(), p, n * sizeof(double));
  }
  return v;
}

}  // namespace kukan
