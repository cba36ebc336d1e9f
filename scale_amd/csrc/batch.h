// batch.h -- device tables of member launches for ensembles: the argument
// structs of every member, then first[nmem+1] (first logical tile of each
// member) and ni[nmem] (a per-member tile-shape word), in one allocation.
#pragma once
#include <vector>

#include "common.h"
#include "shp_internal.h"

namespace shp {

template <typename A>
inline hipError_t batch_upload(const std::vector<A>& args, const std::vector<unsigned>& first,
                               const std::vector<unsigned>& ni, BatchBuf& b) {
  const size_t na = args.size() * sizeof(A), nf = first.size() * sizeof(unsigned), nn = ni.size() * sizeof(unsigned);
  const size_t off_f = (na + 255) / 256 * 256, off_n = off_f + (nf + 255) / 256 * 256;
  batch_free(b);
  hipError_t e = hipMalloc(&b.dev, off_n + nn + 256);
  if (e != hipSuccess) { b.dev = nullptr; return e; }
  char* p = (char*)b.dev;
  if ((e = hipMemcpy(p, args.data(), na, hipMemcpyHostToDevice)) != hipSuccess) return e;
  if ((e = hipMemcpy(p + off_f, first.data(), nf, hipMemcpyHostToDevice)) != hipSuccess) return e;
  if ((e = hipMemcpy(p + off_n, ni.data(), nn, hipMemcpyHostToDevice)) != hipSuccess) return e;
  b.nmem = (int)args.size();
  b.total = first.back();
  return hipSuccess;
}

template <typename A>
inline void batch_tables(const BatchBuf& b, const A*& tab, const unsigned*& first, const unsigned*& ni) {
  const size_t na = (size_t)b.nmem * sizeof(A), nf = ((size_t)b.nmem + 1) * sizeof(unsigned);
  const size_t off_f = (na + 255) / 256 * 256, off_n = off_f + (nf + 255) / 256 * 256;
  tab = (const A*)b.dev;
  first = (const unsigned*)((const char*)b.dev + off_f);
  ni = (const unsigned*)((const char*)b.dev + off_n);
}

// grid of a batched launch: one block per tile, or a persistent grid of a
// multiple of 8 blocks (each XCD then walks one range of logical tiles)
inline unsigned batch_grid(unsigned total, int blocks_per_cu, int ncu) {
  if (blocks_per_cu <= 0) return total;
  unsigned g = (unsigned)(blocks_per_cu * ncu);
  g = g / 8 * 8;
  if (g < 8) g = 8;
  return g < total ? g : total;
}

// the member of logical tile lt: the last m with first[m] <= lt
__device__ __forceinline__ int batch_member(const unsigned* __restrict__ first, int nmem, unsigned lt) {
  int lo = 0, hi = nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (first[mid] <= lt) lo = mid; else hi = mid - 1;
  }
  return lo;
}

}  // namespace shp
