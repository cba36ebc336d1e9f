// profile.hip -- named per-kernel timers with hipEvents recorded on the
// stream the kernels run on (the counterpart of PROF_rapstart/PROF_rapend,
// scalelib/src/common/scale_prof.F90:156-253).
#include <string.h>
#include <map>
#include <string>
#include <vector>
#include "shp_internal.h"

namespace shp {

struct Prof {
  struct Rec { std::string name; hipEvent_t a, b; };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<long, double>> acc;  // name -> (count, total ms)
  std::vector<std::pair<std::string, hipEvent_t>> open;  // timers started and not yet stopped (nested)
  // outermost only: timers nested inside a running one record nothing, so a
  // region is timed by two events whatever it launches
  bool outer_only = false;
  int depth = 0;
  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e; hipEventCreate(&e); return e;
  }
};

void prof_start(const Ctx& c, const char* name, hipStream_t s) {
  if (!c.prof) return;
  Prof& p = *c.prof;
  if (p.outer_only && p.depth++ > 0) return;
  hipEvent_t e = p.get();
  hipEventRecord(e, s);
  p.open.emplace_back(name, e);
}

void prof_stop(const Ctx& c, hipStream_t s) {
  if (!c.prof) return;
  Prof& p = *c.prof;
  if (p.outer_only && (p.depth == 0 || --p.depth > 0)) return;
  if (p.open.empty()) return;
  hipEvent_t b = p.get();
  hipEventRecord(b, s);
  p.pending.push_back({p.open.back().first, p.open.back().second, b});
  p.open.pop_back();
  if (p.pending.size() > 4096) prof_collect(c);
}

void prof_collect(const Ctx& c) {
  if (!c.prof) return;
  Prof& p = *c.prof;
  for (auto& r : p.pending) {
    hipEventSynchronize(r.b);
    float ms = 0.f;
    hipEventElapsedTime(&ms, r.a, r.b);
    auto& e = p.acc[r.name];
    e.first += 1;
    e.second += ms;
    p.pool.push_back(r.a);
    p.pool.push_back(r.b);
  }
  p.pending.clear();
}

Prof* prof_create() { return new Prof(); }
void prof_destroy(Prof* p) {
  if (!p) return;
  for (auto& r : p->pending) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
  for (auto& o : p->open) hipEventDestroy(o.second);
  for (auto e : p->pool) hipEventDestroy(e);
  delete p;
}
void prof_reset(Prof* p) {
  if (!p) return;
  p->acc.clear();
}
void prof_outer_only(Prof* p, bool on) {
  if (!p) return;
  p->outer_only = on;
  p->depth = 0;
}
int prof_query(Prof* p, const char* name, long* count, double* total_ms) {
  if (!p) return -1;
  auto it = p->acc.find(name);
  if (it == p->acc.end()) { *count = 0; *total_ms = 0.0; return 0; }
  *count = it->second.first;
  *total_ms = it->second.second;
  return 0;
}
int prof_names(Prof* p, char* buf, int len) {
  if (!p) return -1;
  std::string s;
  for (auto& kv : p->acc) { s += kv.first; s += "\n"; }
  int n = (int)s.size() < len - 1 ? (int)s.size() : len - 1;
  if (len > 0) { memcpy(buf, s.data(), (size_t)n); buf[n] = 0; }
  return (int)s.size();
}

// ---- HBM copy probe ----------------------------------------------------------
// A plain streaming copy of n doubles (one read and one write of 8 B each),
// 16-byte accesses, grid-stride, four independent loads in flight per lane.
// Its bytes are known exactly, so its time gives the achievable copy
// bandwidth and its FETCH_SIZE / WRITE_SIZE counters calibrate those of the
// stencil kernels.
__global__ void __launch_bounds__(256) hbm_copy_kernel(double* __restrict__ dst, const double* __restrict__ src, size_t n) {
  const size_t n2 = n / 2;
  const double2* s2 = reinterpret_cast<const double2*>(src);
  double2* d2 = reinterpret_cast<double2*>(dst);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n2; i += 4 * stride) {
    const double2 a = s2[i], b = s2[i + stride], c = s2[i + 2 * stride], e = s2[i + 3 * stride];
    d2[i] = a; d2[i + stride] = b; d2[i + 2 * stride] = c; d2[i + 3 * stride] = e;
  }
  for (; i < n2; i += stride) d2[i] = s2[i];
  if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) dst[n - 1] = src[n - 1];
}

// The same copy with non-temporal loads and stores (no L2 allocation for
// streaming data).
__global__ void __launch_bounds__(256) hbm_copy_nt_kernel(double* __restrict__ dst, const double* __restrict__ src, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const double a = __builtin_nontemporal_load(src + i), b = __builtin_nontemporal_load(src + i + stride);
    const double c = __builtin_nontemporal_load(src + i + 2 * stride), e = __builtin_nontemporal_load(src + i + 3 * stride);
    __builtin_nontemporal_store(a, dst + i);
    __builtin_nontemporal_store(b, dst + i + stride);
    __builtin_nontemporal_store(c, dst + i + 2 * stride);
    __builtin_nontemporal_store(e, dst + i + 3 * stride);
  }
  for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Read-only stream: each block sums its share of src (16-byte loads, four in
// flight per lane) and writes one partial sum, so the traffic is the read.
__global__ void __launch_bounds__(256) hbm_read_kernel(double* __restrict__ part, const double* __restrict__ src, size_t n) {
  const size_t n2 = n / 2;
  const double2* s2 = reinterpret_cast<const double2*>(src);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  for (; i + 3 * stride < n2; i += 4 * stride) {
    const double2 a = s2[i], b = s2[i + stride], c = s2[i + 2 * stride], e = s2[i + 3 * stride];
    acc += (a.x + a.y) + (b.x + b.y) + (c.x + c.y) + (e.x + e.y);
  }
  for (; i < n2; i += stride) acc += s2[i].x + s2[i].y;
  __shared__ double sh[256];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Write-only stream: fill dst with a value (16-byte stores).
__global__ void __launch_bounds__(256) hbm_fill_kernel(double* __restrict__ dst, size_t n, double v) {
  const size_t n2 = n / 2;
  double2* d2 = reinterpret_cast<double2*>(dst);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const double2 w = make_double2(v, v);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) d2[i] = w;
}

// One launch of probe `kind` over n doubles (0 copy, 1 non-temporal copy,
// 2 read-only, 3 write-only, 4 hipMemcpyAsync device to device); part needs
// ncu*8 doubles for kind 2.
hipError_t hbm_probe(const Ctx& c, int kind, double* dst, const double* src, double* part, size_t n, hipStream_t s) {
  int ncu = 256;
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c.device);
  const unsigned nb = (unsigned)ncu * 8u;
  switch (kind) {
    case 0: hipLaunchKernelGGL(hbm_copy_kernel, dim3(nb), dim3(256), 0, s, dst, src, n); break;
    case 1: hipLaunchKernelGGL(hbm_copy_nt_kernel, dim3(nb), dim3(256), 0, s, dst, src, n); break;
    case 2: hipLaunchKernelGGL(hbm_read_kernel, dim3(nb), dim3(256), 0, s, part, src, n); break;
    case 3: hipLaunchKernelGGL(hbm_fill_kernel, dim3(nb), dim3(256), 0, s, dst, n, 1.0); break;
    case 4: return hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, s);
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t hbm_copy(const Ctx& c, double* dst, const double* src, size_t n, hipStream_t s) {
  int ncu = 256;
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c.device);
  const unsigned nb = (unsigned)ncu * 8u;
  prof_start(c, "hbm_copy", s);
  hipLaunchKernelGGL(hbm_copy_kernel, dim3(nb), dim3(256), 0, s, dst, src, n);
  prof_stop(c, s);
  return hipGetLastError();
}

}  // namespace shp
