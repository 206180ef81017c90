// F(Q) of one structure without any device allocation (pyiid_fq_noalloc).
//
// Every intermediate of the F-only path lives in `scratch`, a file-scope
// device array that each GPU keeps until the process ends:
//   * the packed atoms;
//   * the positions in element order;
//   * the work items;
//   * the normalisation;
//   * the per-item partial sums.
// The work list is built on the device from the element indices.  The
// partial sums run in passes of at most kPartBytes, so the array's size does
// not depend on the atom count.
//
// In host code the array's name is a stand-in, not a device address.  Host
// code reaches the array only through the address that hipGetSymbolAddress
// returns on each device, looked up once per device.  Nothing here
// allocates device memory or waits for the stream.  Calls that overlap on
// one GPU race on the array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "../../include/pyiid_hip.h"
#include "internal.h"

using namespace pyiid;

namespace {

constexpr int kMaxAtoms = PYIID_NOALLOC_MAX_ATOMS;
constexpr int kMaxElem = PYIID_NOALLOC_MAX_ELEMENTS;
constexpr int kMaxQ = PYIID_NOALLOC_MAX_BINS;
constexpr int kMaxItems = 65536;
constexpr int kMaxTiles = kMaxAtoms / 64 + kMaxElem;
constexpr size_t kPartBytes = size_t(32) << 20;  // per-item F partials of one pass
constexpr int kMaxDevices = 64;
constexpr int kPrepThreads = 1024;
constexpr int kPrepWaves = kPrepThreads / 64;

constexpr size_t up256(size_t x) { return (x + 255) & ~size_t(255); }

// Byte offsets into the scratch array.
constexpr size_t kOffHdr = 0;        // Hdr
constexpr size_t kOffCounts = 256;   // int64 counts[kMaxElem], then the atom count
constexpr size_t kOffPos = kOffCounts + up256(sizeof(int64_t) * (kMaxElem + 1));
constexpr size_t kOffPerm = kOffPos + sizeof(AtomRec) * kMaxAtoms;
constexpr size_t kOffItems = kOffPerm + up256(sizeof(int32_t) * kMaxAtoms);
constexpr size_t kOffNorm = kOffItems + up256(sizeof(Item) * kMaxItems);  // fsu, fso, gs
constexpr size_t kOffAcc = kOffNorm + 3 * sizeof(double) * kMaxQ;
constexpr size_t kOffPart = kOffAcc + sizeof(double) * kMaxQ;
constexpr size_t kScratchBytes = kOffPart + kPartBytes;
static_assert(kScratchBytes <= (size_t(PYIID_NOALLOC_SCRATCH_MIB) << 20),
              "scratch layout exceeds the size the header states");

struct Hdr {
  int n_items;  // items written by prep_kernel
  int bad;      // bit 0: an element index outside [0, n_elem); bit 1: item list overflow
};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Builds everything the F-only kernel reads, in one block:
//   * pos[i]: AtomRec of atom i, its element index in w;
//   * perm[p]: the atoms in element order (stable), the element runs starting
//     at start[e];
//   * counts[e]: the atoms of element e, then the atom count;
//   * items: per i tile of at most 64 positions of one element, j ranges
//     over the positions below the tile's end, cut every jc positions and at
//     element boundaries (split = the tile's element, pad = the range's);
//     the list is padded with empty items up to item_cap.
// This is the same work list as the context's unordered-pair F items, with
// element order alone as the position order.  An out-of-range element index
// reads row 0 and sets Hdr.bad.
__global__ __launch_bounds__(kPrepThreads) void prep_kernel(const float* __restrict__ posf,
                                                            const int32_t* __restrict__ elem,
                                                            int n, int n_elem, int jc,
                                                            int item_cap,
                                                            unsigned char* __restrict__ s) {
  __shared__ int cnt[kMaxElem], start[kMaxElem + 1], run[kMaxElem], tfirst[kMaxElem + 1];
  __shared__ int wcnt[kPrepWaves][kMaxElem];
  __shared__ int tcount[kMaxTiles + 1];
  __shared__ int bad;
  Hdr* hdr = reinterpret_cast<Hdr*>(s + kOffHdr);
  int64_t* counts = reinterpret_cast<int64_t*>(s + kOffCounts);
  AtomRec* pos = reinterpret_cast<AtomRec*>(s + kOffPos);
  int32_t* perm = reinterpret_cast<int32_t*>(s + kOffPerm);
  Item* items = reinterpret_cast<Item*>(s + kOffItems);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  auto element = [&](int i) {
    const int e = elem ? elem[i] : 0;
    return (e < 0 || e >= n_elem) ? -1 : e;
  };

  if (tid < kMaxElem) cnt[tid] = 0;
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int i = tid; i < n; i += kPrepThreads) {
    int e = element(i);
    if (e < 0) {
      bad = 1;
      e = 0;
    }
    atomicAdd(&cnt[e], 1);
    pos[i] = AtomRec{static_cast<double>(posf[3 * i]), static_cast<double>(posf[3 * i + 1]),
                     static_cast<double>(posf[3 * i + 2]), static_cast<double>(e)};
  }
  __syncthreads();
  if (tid == 0) {
    start[0] = 0;
    tfirst[0] = 0;
    for (int e = 0; e < n_elem; ++e) {
      start[e + 1] = start[e] + cnt[e];
      tfirst[e + 1] = tfirst[e] + (cnt[e] + 63) / 64;
      run[e] = 0;
      counts[e] = cnt[e];
    }
    counts[kMaxElem] = n;
  }
  __syncthreads();

  // Stable counting sort, kPrepThreads atoms per round.  Within a wave, the
  // lanes of one element rank themselves by ballot; the waves' counts per
  // element are then scanned in wave order.
  for (int base = 0; base < n; base += kPrepThreads) {
    for (int k = tid; k < kPrepWaves * kMaxElem; k += kPrepThreads) (&wcnt[0][0])[k] = 0;
    __syncthreads();
    const int i = base + tid;
    const bool valid = i < n;
    int e = 0;
    if (valid) e = max(element(i), 0);
    int rank = 0;
    uint64_t todo = __ballot(valid);
    while (todo != 0) {  // wave-uniform: one pass per distinct element in the wave
      const int leader = __builtin_ctzll(todo);
      const int le = __shfl(e, leader);
      const uint64_t same = __ballot(valid && e == le);
      if (valid && e == le) rank = __popcll(same & ((1ull << lane) - 1ull));
      if (lane == leader) wcnt[wave][le] = __popcll(same);
      todo &= ~same;
    }
    __syncthreads();
    if (tid < n_elem) {
      int off = run[tid];
      for (int w = 0; w < kPrepWaves; ++w) {
        const int t = wcnt[w][tid];
        wcnt[w][tid] = off;
        off += t;
      }
      run[tid] = off;
    }
    __syncthreads();
    if (valid) perm[start[e] + wcnt[wave][e] + rank] = i;
    __syncthreads();
  }

  // Work items: count per tile, scan, write.
  const int n_tiles = tfirst[n_elem];
  auto tile_span = [&](int t, int& e, int& ib, int& ie) {
    e = 0;
    while (t >= tfirst[e + 1]) ++e;
    ib = start[e] + 64 * (t - tfirst[e]);
    ie = min(start[e + 1], ib + 64);
  };
  for (int t = tid; t < n_tiles; t += kPrepThreads) {
    int e, ib, ie;
    tile_span(t, e, ib, ie);
    int c = 0;
    for (int f = 0; f < n_elem; ++f)
      if (start[f] < ie && cnt[f] > 0) c += (min(start[f + 1], ie) - start[f] + jc - 1) / jc;
    tcount[t] = c;
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int t = 0; t < n_tiles; ++t) {
      const int c = tcount[t];
      tcount[t] = acc;
      acc += c;
    }
    tcount[n_tiles] = acc;
    if (acc > item_cap) bad |= 2;
    hdr->n_items = acc > item_cap ? 0 : acc;
    hdr->bad = bad;
  }
  __syncthreads();
  const int total = tcount[n_tiles] > item_cap ? 0 : tcount[n_tiles];
  if (total > 0) {
    for (int t = tid; t < n_tiles; t += kPrepThreads) {
      int e, ib, ie;
      tile_span(t, e, ib, ie);
      int k = tcount[t];
      for (int f = 0; f < n_elem; ++f) {
        const int fe = min(start[f + 1], ie);
        for (int jb = start[f]; jb < fe; jb += jc)
          items[k++] = Item{ib, ie, jb, min(jb + jc, fe), e, 0, t, f};
      }
    }
  }
  for (int k = total + tid; k < item_cap; k += kPrepThreads) items[k] = Item{0, 0, 0, 0, 0, 0, 0, 0};
}

// One pass over the partials of items [base, base + cap): acc[q] = (first ?
// 0 : acc[q]) + their sum in a fixed order (16 slices, added in slice
// order).  Items at or beyond the list's length are skipped.  The last pass
// writes fq[q] = float32(fsu[q] acc[q]), or NaN in every bin when Hdr.bad is
// set.  A block covers 64 bins.
__global__ __launch_bounds__(1024) void accumulate_kernel(const unsigned char* __restrict__ s_ro,
                                                          double* __restrict__ acc,
                                                          const double* __restrict__ part,
                                                          int base, int cap, int nq, int first,
                                                          int last, float* __restrict__ fq) {
  __shared__ double sh[16][65];
  const Hdr hdr = *reinterpret_cast<const Hdr*>(s_ro + kOffHdr);
  const double* fsu = reinterpret_cast<const double*>(s_ro + kOffNorm);
  const int qo = threadIdx.x & 63, sl = threadIdx.x >> 6;
  const int q = blockIdx.x * 64 + qo;
  const int cnt = max(0, min(cap, hdr.n_items - base));
  double a = 0.0;
  if (q < nq)
    for (int p = sl; p < cnt; p += 16) a += part[(size_t)p * nq + q];
  sh[sl][qo] = a;
  __syncthreads();
  if (sl != 0 || q >= nq) return;
  double t = first ? 0.0 : acc[q];
#pragma unroll
  for (int k = 0; k < 16; ++k) t += sh[k][qo];
  acc[q] = t;
  if (last) fq[q] = hdr.bad ? __builtin_nanf("") : static_cast<float>(t * fsu[q]);
}

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of the scratch array on `device` (current), looked up once.
int scratch_on(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(PYIID_E_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                       hipGetErrorString(e));
    g_scratch[device] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch[device]);
  return PYIID_OK;
}

// Restores the caller's current device.
struct DeviceGuard {
  int prev = -1;
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

#define NA_HIP(expr)                                                                       \
  do {                                                                                     \
    const hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                                  \
      return set_error(PYIID_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                                \
  } while (0)

}  // namespace

extern "C" int pyiid_fq_noalloc(int device, void* stream, const float* pos, const int32_t* elem,
                                int64_t n, const float* table, int32_t n_elem, int32_t n_q,
                                double qbin, float* fq) {
  if (n < 0 || n > kMaxAtoms)
    return set_error(PYIID_E_INVALID, "n = %lld atoms is outside [0, %d], the no-allocation limit",
                     static_cast<long long>(n), kMaxAtoms);
  if (n_elem < 1 || n_elem > kMaxElem)
    return set_error(PYIID_E_INVALID,
                     "n_elem = %d is outside [1, %d], the no-allocation limit", n_elem, kMaxElem);
  if (n_q < 1 || n_q > kMaxQ)
    return set_error(PYIID_E_INVALID, "n_q = %d is outside [1, %d], the no-allocation limit",
                     n_q, kMaxQ);
  if (!(qbin > 0.0)) return set_error(PYIID_E_INVALID, "qbin must be positive");
  if (!table || !fq || (n > 0 && !pos)) return set_error(PYIID_E_INVALID, "null pointer argument");
  int n_dev = 0;
  if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev < 1)
    return set_error(PYIID_E_NODEVICE, "no HIP device");
  if (device < 0 || device >= std::min(n_dev, kMaxDevices))
    return set_error(PYIID_E_INVALID, "device %d outside [0, %d)", device,
                     std::min(n_dev, kMaxDevices));
  DeviceGuard guard;
  NA_HIP(hipGetDevice(&guard.prev));
  NA_HIP(hipSetDevice(device));
  unsigned char* s = nullptr;
  {
    const int rc = scratch_on(device, &s);
    if (rc != PYIID_OK) return rc;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int ni = static_cast<int>(n);
  const bool weighted = n_elem > 1;
  const int jc = 256 * std::max(1, ni / (256 * 16));
  // Upper bound on the work list.  There are at most (n + 63 n_elem) / 64
  // tiles.  Tile t ends at a position ie <= min(n, 64 (t + 1)), and its j
  // ranges cut the positions below ie into per-element runs of at most jc:
  // at most (ie + (jc - 1) n_elem) / jc of them.
  const int64_t tiles = (n + 63 * (int64_t)n_elem) / 64;
  int64_t item_cap = 0;
  for (int64_t t = 0; t < tiles; ++t)
    item_cap += (std::min<int64_t>(n, 64 * (t + 1)) + (int64_t)(jc - 1) * n_elem) / jc;
  if (item_cap > kMaxItems)
    return set_error(PYIID_E_INVALID, "work list bound %lld exceeds %d items",
                     static_cast<long long>(item_cap), kMaxItems);
  hipLaunchKernelGGL(prep_kernel, dim3(1), dim3(kPrepThreads), 0, st, pos, elem, ni, n_elem, jc,
                     static_cast<int>(item_cap), s);
  NA_HIP(hipGetLastError());
  int64_t* counts = reinterpret_cast<int64_t*>(s + kOffCounts);
  double* fsu = reinterpret_cast<double*>(s + kOffNorm);
  NA_HIP(launch_norm(st, table, counts, n_elem, n_q, counts + kMaxElem, 1, weighted, nullptr, fsu,
                     fsu + kMaxQ, fsu + 2 * kMaxQ));
  const int per_pass = static_cast<int>(kPartBytes / (sizeof(double) * n_q));
  const int passes = std::max<int>(1, static_cast<int>((item_cap + per_pass - 1) / per_pass));
  const double dq = static_cast<double>(static_cast<float>(qbin));
  const AtomRec* prec = reinterpret_cast<const AtomRec*>(s + kOffPos);
  const int32_t* perm = reinterpret_cast<const int32_t*>(s + kOffPerm);
  const Item* items = reinterpret_cast<const Item*>(s + kOffItems);
  double* acc = reinterpret_cast<double*>(s + kOffAcc);
  double* part = reinterpret_cast<double*>(s + kOffPart);
  for (int p = 0; p < passes; ++p) {
    const int base = p * per_pass;
    const int cnt = static_cast<int>(std::min<int64_t>(per_pass, item_cap - base));
    NA_HIP(launch_fq(st, kFqM, false, prec, perm, table, n_q, dq, items + base, cnt, weighted,
                     part, nullptr));
    hipLaunchKernelGGL(accumulate_kernel, dim3((n_q + 63) / 64), dim3(1024), 0, st, s, acc, part,
                       base, per_pass, n_q, p == 0 ? 1 : 0, p == passes - 1 ? 1 : 0, fq);
    NA_HIP(hipGetLastError());
  }
  return PYIID_OK;
}
