// hash_api.cpp -- the spatial hashing service (include/aqflip_hash.h): the FLIP
// step's hash, stable counting sort and cell ranges on caller-provided records.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/aqflip_hash.h"
#include "aqflip_internal.h"
#include "flip_view.h"

using namespace aqf;

namespace {

aq_status hfail(aq_status s, const std::string& m) {
  set_last_error(m);
  return s;
}

#define HHIP(expr)                                                                            \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return hfail(e_ == hipErrorOutOfMemory ? AQ_ENOMEM : AQ_EHIP,                           \
                   std::string(#expr) + ": " + hipGetErrorString(e_));                        \
  } while (0)

#define HTRY(expr)               \
  do {                           \
    aq_status s_ = (expr);       \
    if (s_ != AQ_OK) return s_;  \
  } while (0)

}  // namespace

struct aq_hash {
  Geom G{};
  hipStream_t stream = nullptr;
  Bins B{};
  long long cap = 0;  // particles the per-particle buffers hold
  // staging for the host entry point
  float *d_rec = nullptr, *d_sorted = nullptr;
  int *d_start = nullptr, *d_end = nullptr, *d_shash = nullptr;
  size_t rec_cap = 0;
  std::vector<void*> fixed, per_particle, staging;

  ~aq_hash() {
    if (stream) hipStreamSynchronize(stream);
    for (void* p : staging) hipFree(p);
    for (void* p : per_particle) hipFree(p);
    for (void* p : fixed) hipFree(p);
    if (stream) hipStreamDestroy(stream);
  }
  template <typename T>
  aq_status alloc(T** p, size_t n, std::vector<void*>& list) {
    HHIP(hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)));
    list.push_back(*p);
    return AQ_OK;
  }
  aq_status ensure(long long n) {
    if (n <= cap) return AQ_OK;
    HHIP(hipStreamSynchronize(stream));
    for (void* p : per_particle) hipFree(p);
    per_particle.clear();
    HTRY(alloc(&B.hash, (size_t)n, per_particle));
    HTRY(alloc(&B.rank, (size_t)n, per_particle));
    HTRY(alloc(&B.perm, (size_t)n, per_particle));
    cap = n;
    return AQ_OK;
  }
  aq_status ensure_staging(long long n, int stride) {
    const size_t need = (size_t)std::max<long long>(n, 1) * stride;
    if (need <= rec_cap) return AQ_OK;
    HHIP(hipStreamSynchronize(stream));
    for (void* p : staging) hipFree(p);
    staging.clear();
    HTRY(alloc(&d_rec, need, staging));
    HTRY(alloc(&d_sorted, need, staging));
    HTRY(alloc(&d_shash, (size_t)std::max<long long>(n, 1), staging));
    rec_cap = need;
    return AQ_OK;
  }

  // The whole pipeline on device records, in order on stream s; outputs are device
  // pointers (may be null).
  aq_status run(const float* rec, long long n, int stride, int* perm, int* shash, int* start, int* end,
                float* sorted, hipStream_t s) {
    HTRY(ensure(n));
    const int ncells = G.ncells_hash;
    launch_fill_u32(B.count, 0u, ncells, s);
    if (n > 0) launch_hash_records(G, rec, n, stride, B, s);
    launch_scan(G, B, s);
    if (n > 0) launch_scatter_segsort(G, n, B, s);
    launch_hash_outputs(G, n, stride, B, rec, start, end, shash, sorted, s);
    if (perm && n > 0) HHIP(hipMemcpyAsync(perm, B.perm, n * sizeof(int), hipMemcpyDeviceToDevice, s));
    HHIP(hipGetLastError());
    return AQ_OK;
  }
};

namespace {

aq_status check_args(const aq_hash* h, const void* rec, long long n, int stride) {
  if (!h) return hfail(AQ_EINVAL, "null handle");
  if (n < 0 || n > 0x7fffffffLL) return hfail(AQ_EINVAL, "particle count must be in [0, 2^31)");
  if (n > 0 && !rec) return hfail(AQ_EINVAL, "null records");
  if (stride < 3) return hfail(AQ_EINVAL, "stride must be >= 3 floats");
  return AQ_OK;
}

// cell_count: entries of the caller's cell_start / cell_end arrays; 0 means the grid's
// cells. The reference clears cellCount entries (Fluid/Fluid_3D_common.cuh:47-48),
// which the mesher sets above the hashed grid's cells (Rendering/Renderer3D/Mesher.cuh:155).
aq_status check_cell_count(const aq_hash* h, long long cell_count) {
  if (cell_count != 0 && cell_count < h->G.ncells_hash)
    return hfail(AQ_EINVAL, "cell_count must be 0 or at least size_x*size_y*size_z");
  return AQ_OK;
}

aq_status check_grid(int sx, int sy, int sz, float cell_size) {
  if (sx < 1 || sy < 1 || sz < 1) return hfail(AQ_EINVAL, "grid sizes must be >= 1");
  if ((long long)sx * sy * sz >= 0x7fffffffLL) return hfail(AQ_EINVAL, "grid too large");
  if (!(cell_size > 0) || !std::isfinite(cell_size)) return hfail(AQ_EINVAL, "cell_size must be positive");
  return AQ_OK;
}

Geom hash_geom(int sx, int sy, int sz, float cell_size) {
  Geom G{};
  G.X = sx, G.Y = sy, G.Z = sz;
  G.ncells_hash = sx * sy * sz;
  G.h = cell_size;
  G.inv_h = 1.0 / (double)cell_size;
  G.divZ = make_fastdiv((uint32_t)sz);
  G.divYZ = make_fastdiv((uint32_t)(sy * sz));
  return G;
}

// ---- the entry point that allocates nothing (aq_hash_sort_device_noalloc)
// Its bins live in one file-scope device array: per-cell counts, offsets and multi-run
// flags, the scan partials, and per-record hash, rank and permutation, for at most kStaticCells cells
// and kStaticRecords records (69 MB on every GPU the library runs on, held until the
// process ends). In host code `scratch` names a host-side stand-in, so the array's
// device address is looked up once per device with hipGetSymbolAddress and only that
// address is used from then on.
constexpr long long kStaticCells = 1LL << 21;    // 128^3
constexpr long long kStaticRecords = 1LL << 22;  // 4,194,304
constexpr long long kStaticParts = (kStaticCells + 4095) / 4096 + 1;
constexpr long long align64(long long v) { return (v + 63) / 64 * 64; }  // ints, 256 bytes
constexpr long long kOffCount = 0;
constexpr long long kOffOffsets = kOffCount + align64(kStaticCells);
constexpr long long kOffParts = kOffOffsets + align64(kStaticCells + 1);
constexpr long long kOffHash = kOffParts + align64(kStaticParts);
constexpr long long kOffRank = kOffHash + align64(kStaticRecords);
constexpr long long kOffPerm = kOffRank + align64(kStaticRecords);
constexpr long long kOffMulti = kOffPerm + align64(kStaticRecords);  // kStaticCells bytes
constexpr long long kStaticInts = kOffMulti + align64(kStaticCells / 4);

}  // namespace

static __device__ __align__(256) int scratch[kStaticInts];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
int* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on the first call
// made with that device current).
aq_status static_scratch(int** out) {
  int dev = -1;
  HHIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return hfail(AQ_EINVAL, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HHIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return hfail(AQ_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = static_cast<int*>(p);
  }
  *out = g_scratch_ptr[dev];
  return AQ_OK;
}

}  // namespace

extern "C" {

aq_status aq_hash_create(int32_t sx, int32_t sy, int32_t sz, float cell_size, int64_t capacity,
                         aq_hash** out) {
  if (!out) return hfail(AQ_EINVAL, "null output");
  *out = nullptr;
  HTRY(check_grid(sx, sy, sz, cell_size));
  if (capacity < 0) return hfail(AQ_EINVAL, "negative capacity");
  aq_hash* h = new aq_hash;
  Geom& G = h->G;
  G = hash_geom(sx, sy, sz, cell_size);
  auto bail = [&](aq_status s) {
    delete h;
    return s;
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(hfail(AQ_EHIP, "stream creation failed"));
  const int nch = G.ncells_hash;
  aq_status st = h->alloc(&h->B.count, (size_t)nch, h->fixed);
  if (st == AQ_OK) st = h->alloc(&h->B.offsets, (size_t)nch + 1, h->fixed);
  if (st == AQ_OK) st = h->alloc(&h->B.partials, (size_t)(nch + 4095) / 4096 + 1, h->fixed);
  if (st == AQ_OK) st = h->alloc(&h->B.multi, (size_t)nch, h->fixed);
  // the multi-run flags start at zero; k_segsort clears each flag it reads
  if (st == AQ_OK && hipMemsetAsync(h->B.multi, 0, (size_t)nch, h->stream) != hipSuccess)
    st = hfail(AQ_EHIP, "clearing the bins failed");
  if (st == AQ_OK) st = h->alloc(&h->d_start, (size_t)nch, h->fixed);
  if (st == AQ_OK) st = h->alloc(&h->d_end, (size_t)nch, h->fixed);
  if (st == AQ_OK) st = h->ensure(std::max<int64_t>(capacity, 1));
  // the clear must be complete before a sort queued on a caller's stream reads the flags
  if (st == AQ_OK && hipStreamSynchronize(h->stream) != hipSuccess) st = hfail(AQ_EHIP, "stream synchronisation failed");
  if (st != AQ_OK) return bail(st);
  *out = h;
  return AQ_OK;
}

void aq_hash_destroy(aq_hash* h) { delete h; }

aq_status aq_hash_sort_host(aq_hash* h, const float* rec, int64_t n, int32_t stride, int32_t* perm,
                            int32_t* shash, int32_t* start, int32_t* end, int64_t cell_count, float* sorted) {
  HTRY(check_args(h, rec, n, stride));
  HTRY(check_cell_count(h, cell_count));
  for (long long i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a)
      if (!std::isfinite(rec[i * stride + a])) return hfail(AQ_EINVAL, "non-finite particle position");
  HTRY(h->ensure_staging(n, stride));
  const size_t bytes = (size_t)n * stride * sizeof(float);
  if (n > 0) HHIP(hipMemcpyAsync(h->d_rec, rec, bytes, hipMemcpyHostToDevice, h->stream));
  HTRY(h->run(h->d_rec, n, stride, nullptr, shash ? h->d_shash : nullptr, start ? h->d_start : nullptr,
              end ? h->d_end : nullptr, sorted ? h->d_sorted : nullptr, h->stream));
  const long long ncells = h->G.ncells_hash;
  const size_t cb = (size_t)ncells * sizeof(int);
  if (perm && n > 0) HHIP(hipMemcpyAsync(perm, h->B.perm, n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  if (shash && n > 0) HHIP(hipMemcpyAsync(shash, h->d_shash, n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  if (start) HHIP(hipMemcpyAsync(start, h->d_start, cb, hipMemcpyDeviceToHost, h->stream));
  if (end) HHIP(hipMemcpyAsync(end, h->d_end, cb, hipMemcpyDeviceToHost, h->stream));
  if (sorted && n > 0) HHIP(hipMemcpyAsync(sorted, h->d_sorted, bytes, hipMemcpyDeviceToHost, h->stream));
  HHIP(hipStreamSynchronize(h->stream));
  if (cell_count > ncells) {
    if (start) std::fill(start + ncells, start + cell_count, -1);
    if (end) std::fill(end + ncells, end + cell_count, -1);
  }
  return AQ_OK;
}

aq_status aq_hash_sort_device(aq_hash* h, const float* d_rec, int64_t n, int32_t stride, int32_t* d_perm,
                              int32_t* d_shash, int32_t* d_start, int32_t* d_end, int64_t cell_count,
                              float* d_sorted, void* stream) {
  HTRY(check_args(h, d_rec, n, stride));
  HTRY(check_cell_count(h, cell_count));
  if (d_sorted && d_sorted == d_rec) return hfail(AQ_EINVAL, "sorted records must not alias the input");
  // The kernels run on the caller's stream, after the work already queued there (the
  // reference's performSpatialHashing2 runs in the order of the default stream). The
  // handle's own stream last ran a host-array call, which completed before it returned.
  const hipStream_t s = (hipStream_t)stream;
  HTRY(h->run(d_rec, n, stride, d_perm, d_shash, d_start, d_end, d_sorted, s));
  const long long ncells = h->G.ncells_hash;
  if (cell_count > ncells) {
    if (d_start) launch_fill_u32(d_start + ncells, 0xffffffffu, cell_count - ncells, s);
    if (d_end) launch_fill_u32(d_end + ncells, 0xffffffffu, cell_count - ncells, s);
    HHIP(hipGetLastError());
  }
  HHIP(hipStreamSynchronize(s));
  return AQ_OK;
}

aq_status aq_hash_sort_device_noalloc(int32_t sx, int32_t sy, int32_t sz, float cell_size,
                                      const float* d_rec, int64_t n, int32_t stride, int32_t* d_perm,
                                      int32_t* d_shash, int32_t* d_start, int32_t* d_end,
                                      int64_t cell_count, float* d_sorted, void* stream) {
  HTRY(check_grid(sx, sy, sz, cell_size));
  const Geom G = hash_geom(sx, sy, sz, cell_size);
  if (G.ncells_hash > kStaticCells)
    return hfail(AQ_EINVAL, "aq_hash_sort_device_noalloc: size_x*size_y*size_z exceeds 2097152 cells");
  if (n < 0 || n > kStaticRecords)
    return hfail(AQ_EINVAL, "aq_hash_sort_device_noalloc: record count must be in [0, 4194304]");
  if (n > 0 && !d_rec) return hfail(AQ_EINVAL, "null records");
  if (stride < 3) return hfail(AQ_EINVAL, "stride must be >= 3 floats");
  if (cell_count != 0 && cell_count < G.ncells_hash)
    return hfail(AQ_EINVAL, "cell_count must be 0 or at least size_x*size_y*size_z");
  if (d_sorted && d_sorted == d_rec) return hfail(AQ_EINVAL, "sorted records must not alias the input");
  int* base = nullptr;
  HTRY(static_scratch(&base));
  Bins B{};
  B.count = base + kOffCount;
  B.offsets = base + kOffOffsets;
  B.partials = base + kOffParts;
  B.hash = base + kOffHash;
  B.rank = base + kOffRank;
  B.perm = base + kOffPerm;
  B.multi = reinterpret_cast<uint8_t*>(base + kOffMulti);
  const hipStream_t s = (hipStream_t)stream;
  // The same kernels as aq_hash_sort_device; the counts and the multi-run flags start
  // from fill kernels (a previous call that failed part-way may have left flags set).
  launch_fill_u32(B.count, 0u, G.ncells_hash, s);
  launch_fill_u32(B.multi, 0u, (G.ncells_hash + 3) / 4, s);
  if (n > 0) launch_hash_records(G, d_rec, n, stride, B, s);
  launch_scan(G, B, s);
  if (n > 0) launch_scatter_segsort(G, n, B, s);
  launch_hash_outputs(G, n, stride, B, d_rec, d_start, d_end, d_shash, d_sorted, s);
  if (d_perm && n > 0) HHIP(hipMemcpyAsync(d_perm, B.perm, n * sizeof(int), hipMemcpyDeviceToDevice, s));
  if (cell_count > G.ncells_hash) {
    if (d_start) launch_fill_u32(d_start + G.ncells_hash, 0xffffffffu, cell_count - G.ncells_hash, s);
    if (d_end) launch_fill_u32(d_end + G.ncells_hash, 0xffffffffu, cell_count - G.ncells_hash, s);
  }
  HHIP(hipGetLastError());
  HHIP(hipStreamSynchronize(s));
  return AQ_OK;
}

}  // extern "C"
