// C ABI of the gfx950 scattering backend (declared in include/pyiid_hip.h).
//
// A context holds a batch of B independent structures ("chains") stored back
// to back (B = 1 for the single-structure calls), the per-element form-factor
// tables, the F(Q) -> G(r) transform, the work-item lists of the Debye
// kernels, workspaces, and the device-resident sampler state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

#include "../../include/pyiid_hip.h"
#include "internal.h"

using namespace pyiid;

namespace {

// Launch shapes used without the autotuner (pyiid_set_autotune(ctx, 0)):
// target wave counts for the fused kernel's j split and the contracted-force
// kernel's Q groups.
constexpr int64_t kDefaultFusedWaves = 32768;
// Longest j range of a fused-kernel item (one FP32 run of the gradient sums).
constexpr int64_t kMaxJRun = 2048;
constexpr int64_t kDefaultPfWaves = 8192;

thread_local char g_err[1024] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(PYIID_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                  \
  } while (0)

#define CHK(expr)                    \
  do {                               \
    int rc_ = (expr);                \
    if (rc_ != PYIID_OK) return rc_; \
  } while (0)

// Counts device reallocations; a captured graph records it and is captured
// again when any workspace it may reference has moved since.
std::atomic<uint64_t> g_alloc_epoch{1};

// Device buffer that only grows.
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  int reserve(size_t want) {
    if (want <= bytes) return PYIID_OK;
    g_alloc_epoch.fetch_add(1);
    if (p) {
      (void)hipFree(p);
      p = nullptr;
      bytes = 0;
    }
    if (want == 0) return PYIID_OK;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(PYIID_E_NOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    bytes = want;
    return PYIID_OK;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

struct Table {
  Buf tab;  // float [n_elem][nq]
  int n_elem = 0, nq = 0;
  bool set = false;
  // per-chain normalisation [chain][nq] (recomputed when table or composition changes)
  Buf na, fsu, fso, gs;
  uint64_t norm_stamp = ~0ull;
};

// A work-item list and its ranges, cached against the structure.
struct ItemSet {
  Buf items, ranges;
  Buf perm;  // int32 [n]: atom index per position (F-only items of a weighted table)
  int n_items = 0, n_ranges = 0, split = 0;
  uint64_t stamp = ~0ull;
  int key = -1;
};

// Work list of the symmetric F + gradient kernel: tiles of at most 64
// positions in order_atoms order, items (I group x J range of k tiles each)
// in decreasing order of pair count, and every tile's slot list.
struct SymSet {
  Buf items, tiles, entries, tile_el, perm, franges;
  int n_items = 0, n_tiles = 0;
  int64_t n_slots = 0;
  uint64_t stamp = ~0ull;
  int key = -1;
};

// One captured proposal and the parameters it was captured with.
struct GraphSlot {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int leap = -1;
  double eps = 0.0, beta = 0.0;
  uint64_t seed = 0, stamp = 0, epoch = 0;
  int64_t cap = 0;
};

struct Hmc {
  bool set = false;
  bool fcur_stale = false;  // fcur predates surrogate-driven moves
  int slot = 0, potential = 0;
  double qbin = 0.0, conv = 1.0;
  Buf mass, mom, fcur, fnew, pos_save, kin0, ecur, ehist, acchist, counter, chains, gobs;
  int64_t hist_cap = 0;
  GraphSlot exact_g, sur_g;
};

// Learned integrator state (network, neighbour lists, output-layer fit).
struct Sur {
  bool set = false, w3_set = false;
  float rbf[4] = {0.f, 0.f, 0.f, 0.f};  // rc, mu0, dmu, eta
  float skin = 0.5f;
  int max_nbr = 0;
  Buf net, w3, onehot, nbr, cnt, d, cbuf, eatom, fsur, phi, gpart, gout, viol, viol_count,
      atom_chain, tiles64, tiles256, chunks, granges, efeat, offs, efe;
  int n_t64 = 0, n_t256 = 0, n_chunks = 0;
  uint64_t stamp = ~0ull;
  std::vector<double> acc;  // per chain: accumulated Phi Phi^T, Phi y, y.y (FP64)
  std::vector<double> r2;   // fit quality of the last training sample per chain
  // energy-difference rows: weight (A^-2) and the previous training state's
  // feature energies (B x 64) and exact energies (B)
  double energy_weight = 0.0;
  bool eprev_set = false;
  std::vector<double> eprev, Eprev;
};

}  // namespace

struct pyiid_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;

  // structure: n_chain chains, atoms [off[b], off[b+1])
  int64_t n = 0;
  int n_chain = 1;
  std::vector<int64_t> off{0, 0};
  bool pos_set = false;
  Buf posf;  // float [n][3] staging
  Buf elem;  // int32 [n]
  Buf pos;   // AtomRec [n]
  bool pos_dirty = true;
  std::vector<int32_t> elem_h;
  int max_elem = 0;
  Buf counts_d, natoms_d;
  uint64_t struct_stamp = 1;  // bumps when sizes / offsets / elements change

  Table tab[2];
  uint64_t tab_stamp[2] = {1, 1};

  // transform
  Buf tmat;  // float [n_r][nq]
  int t_nr = 0, t_nq = 0;
  bool t_set = false;

  ItemSet fq_items;     // unordered pairs, ranges per chain
  ItemSet fused_items;  // ordered pairs with j splits, ranges per i tile
  ItemSet pair_items;   // tile pairs (I, J <= I) of the contracted-force kernel
  SymSet sym;           // unordered tile pairs of the symmetric F + gradient kernel
  Buf pair_entries;     // (work, side) lists per tile
  int pf_groups = 1, pf_span = 32;

  // launch shapes (set by the autotuner per problem shape)
  int grad_m = kGradM;                     // Q bins per wave of the fused kernel
  int fq_m = kFqM;                         // Q bins per wave of the F-only kernel
  int fq_m_fixed = 0;                      // set by pyiid_set_fq_bins_per_wave (0: autotuned)
  int grad_m_fixed = 0;                    // set by pyiid_set_grad_bins_per_wave (0: autotuned)
  int64_t fused_waves = kDefaultFusedWaves;  // target wave count that sets the fused j split
  int sym_k = 0;  // tiles per I group / J range of the symmetric kernel (0: ordered kernel)
  int sym_k_fixed = 0;  // set by pyiid_set_grad_tiles_per_item (0: autotuned)
  int sym_m = 16;       // Q bins per wave of the symmetric kernel (16 or 12)
  int sym_m_fixed = 0;  // set by pyiid_set_grad_sym_bins_per_wave (0: autotuned)
  int grad_kernel = PYIID_GRAD_AUTO;  // set by pyiid_set_grad_kernel
  int64_t pf_waves = kDefaultPfWaves;        // target wave count that sets the force Q groups
  bool autotune = true;
  bool compensated = false;  // compensated FP32 sums in the F and F + gradient kernels
  std::map<std::vector<int64_t>, std::vector<int64_t>> tuned;

  // workspaces
  Buf gpart, fpart, fq64, fq32, g64, io_a, io_b, io_c, w, out5, u, uq, gobs, pfpart;
  Buf formcnt;  // uint64 [2]: F-only steps in the three-term / Reinsch form (while timing)

  Hmc hmc;
  Sur sur;
  bool capturing = false;

  // timing: event pairs recorded around launches, resolved when queried
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  double t_ms[5] = {0, 0, 0, 0, 0};
  int64_t t_n[5] = {0, 0, 0, 0, 0};
};

namespace {

int ensure_device(pyiid_ctx* c) {
  HIPCHK(hipSetDevice(c->device));
  return PYIID_OK;
}

int check_ctx(pyiid_ctx* c) {
  if (!c) return fail(PYIID_E_INVALID, "null context");
  return ensure_device(c);
}

// Copy `bytes` from a user pointer (host or device per flags) into dst (device).
int upload(pyiid_ctx* c, void* dst, const void* src, size_t bytes, int flags) {
  if (bytes == 0) return PYIID_OK;
  HIPCHK(hipMemcpyAsync(dst, src, bytes,
                        (flags & PYIID_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        c->stream));
  return PYIID_OK;
}

int download(pyiid_ctx* c, void* dst, const void* src, size_t bytes, int flags) {
  if (bytes == 0) return PYIID_OK;
  HIPCHK(hipMemcpyAsync(dst, src, bytes,
                        (flags & PYIID_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                        c->stream));
  return PYIID_OK;
}

int finish(pyiid_ctx* c, int flags) {
  if (!(flags & PYIID_DEVICE)) HIPCHK(hipStreamSynchronize(c->stream));
  return PYIID_OK;
}

__global__ void pack_kernel(const float* __restrict__ p, const int32_t* __restrict__ e, int64_t n,
                            AtomRec* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  AtomRec a;
  a.x = p[3 * i + 0];
  a.y = p[3 * i + 1];
  a.z = p[3 * i + 2];
  a.w = static_cast<double>(e[i]);
  out[i] = a;
}

__global__ void unpack_kernel(const AtomRec* __restrict__ a, int64_t n, float* __restrict__ p) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  p[3 * i + 0] = static_cast<float>(a[i].x);
  p[3 * i + 1] = static_cast<float>(a[i].y);
  p[3 * i + 2] = static_cast<float>(a[i].z);
}

int ensure_packed(pyiid_ctx* c) {
  if (!c->pos_set) return fail(PYIID_E_STATE, "positions not set");
  if (!c->pos_dirty || c->n == 0) return PYIID_OK;
  const int64_t n = c->n;
  hipLaunchKernelGGL(pack_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     c->stream, c->posf.as<float>(), c->elem.as<int32_t>(), n,
                     c->pos.as<AtomRec>());
  HIPCHK(hipGetLastError());
  c->pos_dirty = false;
  return PYIID_OK;
}

int check_table(pyiid_ctx* c, int slot) {
  if (slot != PYIID_SLOT_FQ && slot != PYIID_SLOT_PDF)
    return fail(PYIID_E_INVALID, "slot %d is not 0 (F(Q) scatter) or 1 (PDF scatter)", slot);
  Table& t = c->tab[slot];
  if (!t.set) return fail(PYIID_E_STATE, "scatter table for slot %d not set", slot);
  if (c->max_elem >= t.n_elem)
    return fail(PYIID_E_INVALID, "element index %d out of range for a %d-row scatter table",
                c->max_elem, t.n_elem);
  return PYIID_OK;
}

bool weighted_table(const Table& t) { return t.n_elem > 1; }

int ensure_norm(pyiid_ctx* c, int slot) {
  Table& t = c->tab[slot];
  const uint64_t stamp = c->tab_stamp[slot] * 1000003ull + c->struct_stamp;
  if (t.norm_stamp == stamp) return PYIID_OK;
  if (c->capturing) return fail(PYIID_E_STATE, "normalisation changed during graph capture");
  const int B = c->n_chain;
  const size_t b = sizeof(double) * (size_t)std::max(1, t.nq) * B;
  CHK(t.na.reserve(b));
  CHK(t.fsu.reserve(b));
  CHK(t.fso.reserve(b));
  CHK(t.gs.reserve(b));
  CHK(c->counts_d.reserve(sizeof(int64_t) * (size_t)std::max(1, t.n_elem) * B));
  CHK(c->natoms_d.reserve(sizeof(int64_t) * B));
  std::vector<int64_t> cnt((size_t)t.n_elem * B, 0), na(B);
  for (int ch = 0; ch < B; ++ch) {
    na[ch] = c->off[ch + 1] - c->off[ch];
    for (int64_t i = c->off[ch]; i < c->off[ch + 1]; ++i) {
      const int e = c->elem_h.empty() ? 0 : c->elem_h[i];
      if (e < t.n_elem) cnt[(size_t)ch * t.n_elem + e]++;
    }
  }
  HIPCHK(hipMemcpyAsync(c->counts_d.p, cnt.data(), sizeof(int64_t) * cnt.size(),
                        hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->natoms_d.p, na.data(), sizeof(int64_t) * B, hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(launch_norm(c->stream, t.tab.as<float>(), c->counts_d.as<int64_t>(), t.n_elem, t.nq,
                     c->natoms_d.as<int64_t>(), B, weighted_table(t), t.na.as<double>(),
                     t.fsu.as<double>(), t.fso.as<double>(), t.gs.as<double>()));
  // cnt / na are host temporaries: wait for the copies
  HIPCHK(hipStreamSynchronize(c->stream));
  t.norm_stamp = stamp;
  return PYIID_OK;
}

int upload_items(pyiid_ctx*, ItemSet& s, const std::vector<Item>& v, const std::vector<Range>& r) {
  CHK(s.items.reserve(sizeof(Item) * std::max<size_t>(1, v.size())));
  CHK(s.ranges.reserve(sizeof(Range) * std::max<size_t>(1, r.size())));
  if (!v.empty())
    HIPCHK(hipMemcpy(s.items.p, v.data(), sizeof(Item) * v.size(), hipMemcpyHostToDevice));
  if (!r.empty())
    HIPCHK(hipMemcpy(s.ranges.p, r.data(), sizeof(Range) * r.size(), hipMemcpyHostToDevice));
  s.n_items = static_cast<int>(v.size());
  s.n_ranges = static_cast<int>(r.size());
  return PYIID_OK;
}

// Order of every chain's atoms for the unordered-pair work lists, as
// positions -> atom indices (perm): element order for a weighted table, then
// the Z-order (Morton code over the chain's bounding box) of the positions
// when the lists are built, ties in atom order.  A 64-atom tile in Z-order is
// compact in space, so the distances from one j atom to its atoms span a
// narrow range and fewer F-only waves meet a step near a multiple of pi (the
// slower Reinsch form).  seg_end[p]: end of the element run holding position
// p (the chain's end when unweighted).
int order_atoms(pyiid_ctx* c, bool weighted, std::vector<int32_t>& perm,
                std::vector<int64_t>& seg_end) {
  perm.resize(static_cast<size_t>(c->n));
  seg_end.resize(static_cast<size_t>(c->n));
  for (int64_t a = 0; a < c->n; ++a) perm[a] = static_cast<int32_t>(a);
  if (c->n == 0) return PYIID_OK;
  auto el = [&](int32_t a) { return c->elem_h.empty() ? 0 : c->elem_h[a]; };
  std::vector<AtomRec> h(static_cast<size_t>(c->n));
  HIPCHK(hipMemcpyAsync(h.data(), c->pos.p, sizeof(AtomRec) * h.size(), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  auto spread = [](uint32_t x) {  // 10 bits -> every third of 30
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
  };
  std::vector<uint32_t> zkey(static_cast<size_t>(c->n), 0u);
  for (int b = 0; b < c->n_chain; ++b) {
    const int64_t o = c->off[b], e = c->off[b + 1];
    if (e <= o) continue;
    double lo[3] = {h[o].x, h[o].y, h[o].z}, hi[3] = {h[o].x, h[o].y, h[o].z};
    for (int64_t a = o; a < e; ++a) {
      const double p3[3] = {h[a].x, h[a].y, h[a].z};
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], p3[k]);
        hi[k] = std::max(hi[k], p3[k]);
      }
    }
    const double span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-30});
    for (int64_t a = o; a < e; ++a) {
      const double p3[3] = {h[a].x, h[a].y, h[a].z};
      uint32_t q[3];
      for (int k = 0; k < 3; ++k)
        q[k] = static_cast<uint32_t>(std::min(1023.0, std::max(0.0, (p3[k] - lo[k]) / span * 1023.0)));
      zkey[a] = spread(q[0]) | (spread(q[1]) << 1) | (spread(q[2]) << 2);
    }
    std::stable_sort(perm.begin() + o, perm.begin() + e, [&](int32_t x, int32_t y) {
      const int ex = weighted ? el(x) : 0, ey = weighted ? el(y) : 0;
      return ex != ey ? ex < ey : zkey[x] < zkey[y];
    });
    seg_end[e - 1] = e;
    for (int64_t p = e - 2; p >= o; --p)
      seg_end[p] = (!weighted || el(perm[p]) == el(perm[p + 1])) ? seg_end[p + 1] : p + 1;
  }
  return PYIID_OK;
}

int upload_perm(pyiid_ctx*, ItemSet& s, const std::vector<int32_t>& perm) {
  CHK(s.perm.reserve(sizeof(int32_t) * std::max<size_t>(1, perm.size())));
  if (!perm.empty())
    HIPCHK(hipMemcpy(s.perm.p, perm.data(), sizeof(int32_t) * perm.size(), hipMemcpyHostToDevice));
  return PYIID_OK;
}

// Unordered-pair items: per chain, per i tile of at most 64 atoms, j ranges
// of the chain below the tile's end.  Ranges: one per chain.  Item indices
// are positions in perm (order_atoms); for a weighted table no i tile or j
// range crosses an element boundary, item.split is the i tile's element and
// item.pad the j range's, so that the kernel sums plain sin(Q r)/r and
// weights each bin's total by f_i(Q) f_j(Q) once.
int ensure_fq_items(pyiid_ctx* c, bool weighted) {
  ItemSet& s = c->fq_items;
  const int key = weighted ? 1 : 0;
  if (s.stamp == c->struct_stamp && s.key == key) return PYIID_OK;
  if (c->capturing) return fail(PYIID_E_STATE, "structure changed during graph capture");
  std::vector<int32_t> perm;
  std::vector<int64_t> seg_end;
  CHK(order_atoms(c, weighted, perm, seg_end));
  auto el = [&](int32_t a) { return c->elem_h.empty() ? 0 : c->elem_h[a]; };
  std::vector<Item> v;
  std::vector<Range> r;
  int tile = 0;
  for (int b = 0; b < c->n_chain; ++b) {
    const int64_t o = c->off[b], nb = c->off[b + 1] - o;
    const int begin = static_cast<int>(v.size());
    const int64_t jc = 256 * std::max<int64_t>(1, nb / (256 * 16));
    for (int64_t ib = o, ie = o; ib < o + nb; ib = ie, ++tile) {
      ie = std::min(seg_end[ib], ib + 64);
      const int ei = weighted ? el(perm[ib]) : 0;
      for (int64_t jb = o; jb < ie;) {
        const int64_t je = std::min({ie, jb + jc, seg_end[jb]});
        const int ej = weighted ? el(perm[jb]) : 0;
        v.push_back(Item{static_cast<int>(ib), static_cast<int>(ie), static_cast<int>(jb),
                         static_cast<int>(je), ei, b, tile, ej});
        jb = je;
      }
    }
    r.push_back(Range{begin, static_cast<int>(v.size()), static_cast<int>(o),
                      static_cast<int>(o + nb)});
  }
  CHK(upload_items(c, s, v, r));
  CHK(upload_perm(c, s, perm));
  s.stamp = c->struct_stamp;
  s.key = key;
  return PYIID_OK;
}

// Ordered-pair items: per chain, per i tile, the same j ranges covering the
// chain: `split` chunks, and for a weighted table the chunks cut further at
// element boundaries of the chain's atoms in element order (j indices are
// positions in perm; item.pad is the range's element).  Every tile has as
// many items as the chain with the most j ranges (padded with empty ranges)
// so that the gradient partial buffer [split][n][3][nq], one plane per j
// range, is fully written.  Ranges: one per i tile (for the force reduction).
int ensure_fused_items(pyiid_ctx* c, int nq, bool weighted) {
  ItemSet& s = c->fused_items;
  int64_t n_tiles = 0, max_nb = 0;
  for (int b = 0; b < c->n_chain; ++b) {
    const int64_t nb = c->off[b + 1] - c->off[b];
    n_tiles += (nb + 63) / 64;
    max_nb = std::max(max_nb, nb);
  }
  const int64_t base = std::max<int64_t>(1, n_tiles * ((nq + c->grad_m - 1) / c->grad_m));
  int64_t split = (c->fused_waves + base - 1) / base;
  // The gradient sums run in FP32 over an item's j range and in FP64 across
  // ranges, and their rounding error grows with the FP32 run.  Runs of at
  // most kMaxJRun atoms hold the gradient of a 40,000-atom Pt/Ni alloy to
  // 1.1e-6 of its largest entry, against 8.2e-6 with two runs of 20,000, and
  // take no longer (the fused kernel's time does not depend on the split).
  // (compensated sums do not need it: their FP32 rounding acts on one j tile)
  if (!c->compensated) split = std::max<int64_t>(split, (max_nb + kMaxJRun - 1) / kMaxJRun);
  split = std::min<int64_t>(split, 32);
  split = std::min<int64_t>(split, std::max<int64_t>(1, max_nb / 64));
  split = std::max<int64_t>(1, split);
  const int key = static_cast<int>(split) * 2 + (weighted ? 1 : 0);
  if (s.stamp == c->struct_stamp && s.key == key) return PYIID_OK;
  if (c->capturing) return fail(PYIID_E_STATE, "structure changed during graph capture");
  std::vector<int32_t> perm(static_cast<size_t>(c->n));
  for (int64_t a = 0; a < c->n; ++a) perm[a] = static_cast<int32_t>(a);
  auto el = [&](int32_t a) { return c->elem_h.empty() ? 0 : c->elem_h[a]; };
  struct JRange {
    int64_t jb, je;
    int elem;
  };
  std::vector<std::vector<JRange>> jr(static_cast<size_t>(c->n_chain));
  size_t planes = 1;
  for (int b = 0; b < c->n_chain; ++b) {
    const int64_t o = c->off[b], nb = c->off[b + 1] - o;
    const int64_t jc = (nb + split - 1) / split;
    if (weighted)
      std::stable_sort(perm.begin() + o, perm.begin() + o + nb,
                       [&](int32_t x, int32_t y) { return el(x) < el(y); });
    for (int64_t sp = 0; sp < split; ++sp) {
      const int64_t jb = std::min(o + nb, o + sp * jc);
      const int64_t je = std::min(o + nb, jb + jc);
      if (!weighted || jb == je) {
        jr[b].push_back(JRange{jb, je, 0});
        continue;
      }
      for (int64_t p = jb; p < je;) {
        int64_t q = p + 1;
        while (q < je && el(perm[q]) == el(perm[p])) ++q;
        jr[b].push_back(JRange{p, q, el(perm[p])});
        p = q;
      }
    }
    planes = std::max(planes, jr[b].size());
  }
  std::vector<Item> v;
  std::vector<Range> r;
  int tile = 0;
  for (int b = 0; b < c->n_chain; ++b) {
    const int64_t o = c->off[b], nb = c->off[b + 1] - o;
    for (int64_t ib = o; ib < o + nb; ib += 64, ++tile) {
      const int64_t ie = std::min(o + nb, ib + 64);
      const int begin = static_cast<int>(v.size());
      for (size_t sp = 0; sp < planes; ++sp) {
        const JRange x = sp < jr[b].size() ? jr[b][sp] : JRange{o + nb, o + nb, 0};
        v.push_back(Item{static_cast<int>(ib), static_cast<int>(ie), static_cast<int>(x.jb),
                         static_cast<int>(x.je), static_cast<int>(sp), b, tile, x.elem});
      }
      r.push_back(Range{begin, static_cast<int>(v.size()), static_cast<int>(ib),
                        static_cast<int>(ie)});
    }
  }
  // ranges[n_ranges]: every item (the F of a single structure)
  const int n_tile_ranges = static_cast<int>(r.size());
  r.push_back(Range{0, static_cast<int>(v.size()), 0, static_cast<int>(c->n)});
  CHK(upload_items(c, s, v, r));
  CHK(s.perm.reserve(sizeof(int32_t) * std::max<size_t>(1, perm.size())));
  if (!perm.empty())
    HIPCHK(hipMemcpy(s.perm.p, perm.data(), sizeof(int32_t) * perm.size(), hipMemcpyHostToDevice));
  s.n_ranges = n_tile_ranges;
  s.split = static_cast<int>(planes);
  s.key = key;
  s.stamp = c->struct_stamp;
  return PYIID_OK;
}

// Slot buffer limit of the symmetric kernel (8 GiB of the 288 GB); a shape
// that needs more for every group size runs the ordered kernel.
constexpr size_t kSymSlotBudget = static_cast<size_t>(8) << 30;
// Tiles per I group / J range tried by the autotuner.  A J range of at most
// 32 tiles keeps every FP32 run of the gradient sums within kMaxJRun atoms.
constexpr int kSymGroups[] = {4, 8, 16, 32};

struct SymPlan {
  std::vector<int32_t> perm;
  std::vector<Range> tiles;  // {slot list begin, end, position begin, end}
  std::vector<int> tile_el;
  std::vector<int> chain_tiles{0};  // tiles of chain b: [chain_tiles[b], chain_tiles[b + 1])
  std::vector<SymItem> items;
  std::vector<int> entries;
  int64_t n_slots = 0;
};

// Tiles of the symmetric kernel: runs of at most 64 positions of one chain
// (and, for a weighted table, of one element) in order_atoms order.
int plan_sym_tiles(pyiid_ctx* c, bool weighted, SymPlan& pl) {
  std::vector<int64_t> seg_end;
  CHK(order_atoms(c, weighted, pl.perm, seg_end));
  auto el = [&](int64_t p) { return weighted && !c->elem_h.empty() ? c->elem_h[pl.perm[p]] : 0; };
  pl.tiles.clear();
  pl.tile_el.clear();
  pl.chain_tiles.assign(1, 0);
  for (int b = 0; b < c->n_chain; ++b) {
    for (int64_t p = c->off[b]; p < c->off[b + 1];) {
      const int64_t q = std::min(p + 64, seg_end[p]);
      pl.tiles.push_back(Range{0, 0, static_cast<int>(p), static_cast<int>(q)});
      pl.tile_el.push_back(el(p));
      p = q;
    }
    pl.chain_tiles.push_back(static_cast<int>(pl.tiles.size()));
  }
  return PYIID_OK;
}

// Items for k tiles per I group and per J range (J ranges also end at element
// changes), their slots, and every tile's slot list.  Items are ordered by
// decreasing pair count so that the longest start first.
int plan_sym_items(SymPlan& pl, int k) {
  pl.items.clear();
  pl.entries.clear();
  std::vector<double> cost;
  std::vector<std::vector<int>> per_tile(pl.tiles.size());
  int64_t ns = 0;
  for (size_t b = 0; b + 1 < pl.chain_tiles.size(); ++b) {
    const int t0 = pl.chain_tiles[b], t1 = pl.chain_tiles[b + 1];
    std::vector<std::pair<int, int>> jr;
    for (int t = t0; t < t1;) {
      int u = t + 1;
      while (u < t1 && u - t < k && pl.tile_el[u] == pl.tile_el[t]) ++u;
      jr.emplace_back(t, u);
      t = u;
    }
    for (int g0 = t0; g0 < t1; g0 += k) {
      const int g1 = std::min(t1, g0 + k);
      for (const auto& r : jr) {
        if (r.first >= g1) break;
        if (ns + 2 * k > INT32_MAX) return fail(PYIID_E_INVALID, "too many atom pairs");
        SymItem it{g0, g1, r.first, r.second, 0, 0, 0, 0};
        const int i_first = std::max(g0, r.first);
        it.islot0 = static_cast<int>(ns);
        double w = 0.0;
        for (int ti = i_first; ti < g1; ++ti) {
          per_tile[ti].push_back(static_cast<int>(ns++));
          w += std::min(r.second, ti + 1) - r.first - (ti < r.second ? 0.5 : 0.0);
        }
        it.jslot0 = static_cast<int>(ns);
        for (int tj = r.first; tj < std::min(r.second, g1); ++tj)
          per_tile[tj].push_back(static_cast<int>(ns++));
        pl.items.push_back(it);
        cost.push_back(w);
      }
    }
  }
  std::vector<int> ord(pl.items.size());
  for (size_t i = 0; i < ord.size(); ++i) ord[i] = static_cast<int>(i);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  std::vector<SymItem> sorted(pl.items.size());
  for (size_t i = 0; i < ord.size(); ++i) sorted[i] = pl.items[ord[i]];
  pl.items.swap(sorted);
  for (size_t t = 0; t < pl.tiles.size(); ++t) {
    pl.tiles[t].begin = static_cast<int>(pl.entries.size());
    pl.entries.insert(pl.entries.end(), per_tile[t].begin(), per_tile[t].end());
    pl.tiles[t].end = static_cast<int>(pl.entries.size());
  }
  pl.n_slots = ns;
  return PYIID_OK;
}

size_t sym_slot_bytes(int64_t n_slots, int nq) {
  return sizeof(float) * static_cast<size_t>(std::max<int64_t>(1, n_slots)) * 3 * 64 *
         static_cast<size_t>(std::max(1, nq));
}

template <typename T>
int upload_vec(Buf& b, const std::vector<T>& v) {
  CHK(b.reserve(sizeof(T) * std::max<size_t>(1, v.size())));
  if (!v.empty()) HIPCHK(hipMemcpy(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return PYIID_OK;
}

int ensure_sym_items(pyiid_ctx* c, bool weighted, int k) {
  SymSet& s = c->sym;
  const int key = k * 2 + (weighted ? 1 : 0);
  if (s.stamp == c->struct_stamp && s.key == key) return PYIID_OK;
  if (c->capturing) return fail(PYIID_E_STATE, "structure changed during graph capture");
  SymPlan pl;
  CHK(plan_sym_tiles(c, weighted, pl));
  CHK(plan_sym_items(pl, k));
  CHK(upload_vec(s.items, pl.items));
  CHK(upload_vec(s.tiles, pl.tiles));
  CHK(upload_vec(s.entries, pl.entries));
  CHK(upload_vec(s.tile_el, pl.tile_el));
  CHK(upload_vec(s.perm, pl.perm));
  const std::vector<Range> fr{Range{0, static_cast<int>(pl.items.size()), 0, static_cast<int>(c->n)}};
  CHK(upload_vec(s.franges, fr));
  s.n_items = static_cast<int>(pl.items.size());
  s.n_tiles = static_cast<int>(pl.tiles.size());
  s.n_slots = pl.n_slots;
  s.key = key;
  s.stamp = c->struct_stamp;
  return PYIID_OK;
}

// Group sizes of the symmetric kernel whose slots fit kSymSlotBudget.
int sym_candidates(pyiid_ctx* c, int nq, bool weighted, std::vector<int>& out) {
  out.clear();
  SymPlan pl;
  CHK(plan_sym_tiles(c, weighted, pl));
  for (int k : kSymGroups) {
    CHK(plan_sym_items(pl, k));
    if (sym_slot_bytes(pl.n_slots, nq) <= kSymSlotBudget) out.push_back(k);
  }
  return PYIID_OK;
}

hipEvent_t take_event(pyiid_ctx* c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Resolve recorded event pairs into the per-kind totals.
int resolve_timing(pyiid_ctx* c) {
  for (auto& pe : c->ev_pending) {
    HIPCHK(hipEventSynchronize(pe.second.second));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
    c->t_ms[pe.first] += ms;
    c->t_n[pe.first] += 1;
    c->ev_pool.push_back(pe.second.first);
    c->ev_pool.push_back(pe.second.second);
  }
  c->ev_pending.clear();
  return PYIID_OK;
}

// Brackets one kernel launch with HIP events on the launch stream (no
// synchronisation; the pair is resolved by pyiid_kernel_time).  Inactive
// while a graph is being captured.
struct Timed {
  pyiid_ctx* c;
  int kind;
  hipEvent_t e0 = nullptr;
  Timed(pyiid_ctx* c_, int k) : c(c_), kind(k) {
    if (c->timing && !c->capturing) {
      e0 = take_event(c);
      if (e0) (void)hipEventRecord(e0, c->stream);
    }
  }
  int done() {
    if (!e0) return PYIID_OK;
    hipEvent_t e1 = take_event(c);
    if (!e1) return fail(PYIID_E_HIP, "hipEventCreate failed");
    HIPCHK(hipEventRecord(e1, c->stream));
    c->ev_pending.push_back({kind, {e0, e1}});
    if (c->ev_pending.size() > 4096) return resolve_timing(c);
    return PYIID_OK;
  }
};

int64_t min_chain_atoms(pyiid_ctx* c) {
  int64_t m = INT64_MAX;
  for (int b = 0; b < c->n_chain; ++b) m = std::min(m, c->off[b + 1] - c->off[b]);
  return c->n_chain ? m : 0;
}

int tune_fq(pyiid_ctx* c, int slot, double qbin);

// counted: add the launch's three-term / Reinsch steps to the context's
// counts while timing is on (not for the tuner's trial launches)
hipError_t launch_fq_items(pyiid_ctx* c, const Table& t, double qbin, bool counted = false) {
  unsigned long long* forms =
      counted && c->timing && !c->capturing ? c->formcnt.as<unsigned long long>() : nullptr;
  return launch_fq(c->stream, c->fq_m, c->compensated, c->pos.as<AtomRec>(),
                   c->fq_items.perm.as<int32_t>(),
                   t.tab.as<float>(), t.nq, static_cast<double>(static_cast<float>(qbin)),
                   c->fq_items.items.as<Item>(), c->fq_items.n_items, weighted_table(t),
                   c->fpart.as<double>(), forms);
}

// F(Q) of every chain (full grid) into c->fq64 [B][nq] (double) and c->fq32.
int compute_fq(pyiid_ctx* c, int slot, double qbin) {
  Table& t = c->tab[slot];
  const int nq = t.nq, B = c->n_chain;
  CHK(c->fq64.reserve(sizeof(double) * (size_t)std::max(1, nq) * B));
  CHK(c->fq32.reserve(sizeof(float) * (size_t)std::max(1, nq) * B));
  CHK(ensure_packed(c));
  CHK(ensure_norm(c, slot));
  CHK(ensure_fq_items(c, weighted_table(t)));
  CHK(c->fpart.reserve(sizeof(double) * (size_t)std::max(1, c->fq_items.n_items) * nq));
  CHK(tune_fq(c, slot, qbin));
  Timed tm(c, 1);
  HIPCHK(launch_fq_items(c, t, qbin, true));
  CHK(tm.done());
  // chains without pairs have empty ranges: their F is 0
  HIPCHK(launch_reduce_ranges(c->stream, c->fpart.as<double>(), c->fq_items.ranges.as<Range>(), B,
                              nq, t.fsu.as<double>(), c->fq64.as<double>(), c->fq32.as<float>()));
  return PYIID_OK;
}

// Fused kernel over all chains: gradient partials and F partials.
int tune_fused(pyiid_ctx* c, int slot, double qbin);

int run_fused(pyiid_ctx* c, int slot, double qbin, bool tune = true) {
  Table& t = c->tab[slot];
  const int nq = t.nq;
  CHK(ensure_packed(c));
  CHK(ensure_norm(c, slot));
  if (tune) CHK(tune_fused(c, slot, qbin));
  CHK(ensure_fused_items(c, nq, weighted_table(t)));
  const ItemSet& s = c->fused_items;
  CHK(c->fpart.reserve(sizeof(double) * (size_t)std::max(1, s.n_items) * nq));
  CHK(c->gpart.reserve(sizeof(float) * (size_t)s.split * std::max<int64_t>(1, c->n) * 3 * nq));
  const double dq = static_cast<double>(static_cast<float>(qbin));
  Timed tm(c, 0);
  HIPCHK(launch_fused_grad(c->stream, c->grad_m, c->compensated, c->pos.as<AtomRec>(),
                           s.perm.as<int32_t>(),
                           static_cast<int>(c->n), t.tab.as<float>(), nq, dq, s.items.as<Item>(),
                           s.n_items, weighted_table(t), c->gpart.as<float>(),
                           c->fpart.as<double>()));
  CHK(tm.done());
  return PYIID_OK;
}

// Symmetric kernel over the single chain: gradient slots and F partials.
int run_sym(pyiid_ctx* c, int slot, double qbin) {
  Table& t = c->tab[slot];
  const int nq = t.nq;
  CHK(ensure_packed(c));
  CHK(ensure_norm(c, slot));
  CHK(ensure_sym_items(c, weighted_table(t), c->sym_k));
  const SymSet& s = c->sym;
  CHK(c->fpart.reserve(sizeof(double) * (size_t)std::max(1, s.n_items) * nq));
  CHK(c->gpart.reserve(sym_slot_bytes(s.n_slots, nq)));
  const double dq = static_cast<double>(static_cast<float>(qbin));
  Timed tm(c, 0);
  HIPCHK(launch_sym_grad(c->stream, c->sym_m, c->pos.as<AtomRec>(), s.perm.as<int32_t>(),
                         s.tiles.as<Range>(), s.tile_el.as<int>(), t.tab.as<float>(), nq, dq,
                         s.items.as<SymItem>(), s.n_items, weighted_table(t), c->gpart.as<float>(),
                         c->fpart.as<double>()));
  CHK(tm.done());
  return PYIID_OK;
}

hipError_t reduce_sym(pyiid_ctx* c, const Table& t, void* out, bool out_f64) {
  return launch_reduce_slots(c->stream, c->gpart.as<float>(), c->sym.entries.as<int>(),
                             c->sym.tiles.as<Range>(), c->sym.n_tiles, c->sym.perm.as<int32_t>(),
                             t.nq, t.gs.as<double>(), out, out_f64);
}

// Tile pairs (I, J <= I) of every chain for the contracted-force kernel, the
// Q grouping (enough waves to fill the device, LDS within its limit) and the
// (work, side) entry list of every tile for the deterministic reduction.
// Tiles are runs of at most 64 positions in order_atoms order; for a weighted
// table they do not cross element boundaries and item.split = e_I | e_J << 16.
int ensure_pair_items(pyiid_ctx* c, int nq, bool weighted) {
  ItemSet& s = c->pair_items;
  int64_t n_items = 0;
  for (int b = 0; b < c->n_chain; ++b) {
    const int64_t t = (c->off[b + 1] - c->off[b] + 63) / 64;
    n_items += t * (t + 1) / 2;
  }
  const int n_chunks = std::max(1, (nq + 31) / 32);
  int64_t groups = (c->pf_waves + std::max<int64_t>(1, n_items) - 1) / std::max<int64_t>(1, n_items);
  groups = std::max<int64_t>(1, std::min<int64_t>(groups, n_chunks));
  int span = 0;
  for (;; ++groups) {
    const int per = static_cast<int>((n_chunks + groups - 1) / groups);
    span = per * 32;
    if (sizeof(float) * 2 * kWaves * static_cast<size_t>(span) <= kPairForceLds) break;
    if (groups >= n_chunks) return fail(PYIID_E_INVALID, "contracted-force tables exceed LDS");
  }
  groups = (n_chunks + span / 32 - 1) / (span / 32);
  const int key = static_cast<int>(groups * 100000 + span) * 2 + (weighted ? 1 : 0);
  if (s.stamp == c->struct_stamp && s.key == key) return PYIID_OK;
  if (c->capturing) return fail(PYIID_E_STATE, "structure changed during graph capture");
  std::vector<int32_t> perm;
  std::vector<int64_t> seg_end;
  CHK(order_atoms(c, weighted, perm, seg_end));
  auto el = [&](int64_t p) { return weighted && !c->elem_h.empty() ? c->elem_h[perm[p]] : 0; };
  std::vector<Item> v;
  std::vector<std::vector<int>> per_tile;
  std::vector<Range> tiles;
  int tile0 = 0;
  for (int b = 0; b < c->n_chain; ++b) {
    const int64_t o = c->off[b], e = c->off[b + 1];
    std::vector<std::pair<int, int>> run;  // the chain's tiles [lo, hi)
    for (int64_t p = o; p < e;) {
      const int64_t q = std::min(p + 64, seg_end[p]);
      run.emplace_back(static_cast<int>(p), static_cast<int>(q));
      tiles.push_back(Range{0, 0, static_cast<int>(p), static_cast<int>(q)});
      per_tile.emplace_back();
      p = q;
    }
    const int nt = static_cast<int>(run.size());
    for (int ti = 0; ti < nt; ++ti)
      for (int tj = 0; tj <= ti; ++tj) {
        const int k = static_cast<int>(v.size());
        const int es = el(run[ti].first) | (el(run[tj].first) << 16);
        v.push_back(Item{run[ti].first, run[ti].second, run[tj].first, run[tj].second, es, b,
                         tile0 + ti, tile0 + tj});
        for (int g = 0; g < groups; ++g) {
          const int work = k * static_cast<int>(groups) + g;
          per_tile[tile0 + ti].push_back(2 * work);
          per_tile[tile0 + tj].push_back(2 * work + 1);
        }
      }
    tile0 += nt;
  }
  if (static_cast<int64_t>(v.size()) * groups > INT32_MAX / 2)
    return fail(PYIID_E_INVALID, "too many atom pairs");
  std::vector<int> entries;
  for (size_t t = 0; t < tiles.size(); ++t) {
    tiles[t].begin = static_cast<int>(entries.size());
    entries.insert(entries.end(), per_tile[t].begin(), per_tile[t].end());
    tiles[t].end = static_cast<int>(entries.size());
  }
  CHK(upload_items(c, s, v, tiles));
  CHK(upload_perm(c, s, perm));
  CHK(c->pair_entries.reserve(sizeof(int) * std::max<size_t>(1, entries.size())));
  if (!entries.empty())
    HIPCHK(hipMemcpy(c->pair_entries.p, entries.data(), sizeof(int) * entries.size(),
                     hipMemcpyHostToDevice));
  c->pf_groups = static_cast<int>(groups);
  c->pf_span = span;
  s.key = key;
  s.stamp = c->struct_stamp;
  return PYIID_OK;
}

int launch_contract(pyiid_ctx* c, int slot, double qbin) {
  Table& t = c->tab[slot];
  const int nq = t.nq;
  const double dq = static_cast<double>(static_cast<float>(qbin));
  CHK(ensure_pair_items(c, nq, weighted_table(t)));
  const ItemSet& s = c->pair_items;
  CHK(c->pfpart.reserve(sizeof(float) * 384 * static_cast<size_t>(std::max(1, s.n_items)) *
                        c->pf_groups));
  HIPCHK(launch_pair_force(c->stream, c->pos.as<AtomRec>(), s.perm.as<int32_t>(),
                           t.tab.as<float>(), nq, dq, s.items.as<Item>(), s.n_items, c->pf_groups,
                           c->pf_span, weighted_table(t), c->u.as<float>(), c->uq.as<float>(),
                           c->pfpart.as<float>()));
  return PYIID_OK;
}

// Times one launch of `fn` on the context stream with HIP events (ms).
template <typename F>
int time_launch(pyiid_ctx* c, F fn, double* ms) {
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) return fail(PYIID_E_HIP, "hipEventCreate failed");
  HIPCHK(hipEventRecord(a, c->stream));
  const int rc = fn();
  HIPCHK(hipEventRecord(b, c->stream));
  HIPCHK(hipEventSynchronize(b));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, a, b));
  c->ev_pool.push_back(a);
  c->ev_pool.push_back(b);
  *ms = t;
  return rc;
}

// Autotuner: the first call for a problem shape runs every candidate launch
// shape once to warm it and once timed, keeps the fastest, and remembers it
// for the shape.  Never runs while a graph is captured (defaults apply).
int tune_contract(pyiid_ctx* c, int slot, double qbin) {
  if (!c->autotune) {
    c->pf_waves = kDefaultPfWaves;
    return PYIID_OK;
  }
  Table& t = c->tab[slot];
  const std::vector<int64_t> key{1, c->n, c->n_chain, t.nq, weighted_table(t) ? 1 : 0,
                                 static_cast<int64_t>(c->struct_stamp)};
  auto hit = c->tuned.find(key);
  if (hit != c->tuned.end()) {
    c->pf_waves = hit->second[0];
    return PYIID_OK;
  }
  // a shape first met under graph capture runs the default shape
  c->pf_waves = kDefaultPfWaves;
  if (c->capturing) return PYIID_OK;
  double best = 1e300;
  int64_t best_w = c->pf_waves;
  for (int64_t w : {2048LL, 8192LL, 32768LL}) {
    c->pf_waves = w;
    double ms = 0.0;
    CHK(launch_contract(c, slot, qbin));
    CHK(time_launch(c, [&]() { return launch_contract(c, slot, qbin); }, &ms));
    if (ms < best) {
      best = ms;
      best_w = w;
    }
  }
  c->pf_waves = best_w;
  c->tuned[key] = {best_w};
  return PYIID_OK;
}

// F-only kernel: Q bins per wave.  A block spans 4 waves of m bins and every
// span repeats the per-pair setup, so the best m depends on nq (how full the
// last span is) and on the item shapes.
int tune_fq(pyiid_ctx* c, int slot, double qbin) {
  Table& t = c->tab[slot];
  const std::vector<int64_t> key{2, c->n, c->n_chain, t.nq, weighted_table(t) ? 1 : 0,
                                 static_cast<int64_t>(c->struct_stamp)};
  if (c->compensated) {  // one shape
    c->fq_m = kFqM;
    return PYIID_OK;
  }
  if (c->fq_m_fixed || !c->autotune) {
    c->fq_m = c->fq_m_fixed ? c->fq_m_fixed : kFqM;
    return PYIID_OK;
  }
  auto hit = c->tuned.find(key);
  if (hit != c->tuned.end()) {
    c->fq_m = static_cast<int>(hit->second[0]);
    return PYIID_OK;
  }
  c->fq_m = kFqM;
  if (c->capturing) return PYIID_OK;
  double best = 1e300;
  int best_m = kFqM;
  for (int m : {64, 72, 80, 88, 96, 112}) {
    c->fq_m = m;
    double ms = 0.0;
    HIPCHK(launch_fq_items(c, t, qbin));
    CHK(time_launch(c, [&]() {
      HIPCHK(launch_fq_items(c, t, qbin));
      return PYIID_OK;
    }, &ms));
    if (ms < best) {
      best = ms;
      best_m = m;
    }
  }
  c->fq_m = best_m;
  c->tuned[key] = {best_m};
  return PYIID_OK;
}

int tune_fused(pyiid_ctx* c, int slot, double qbin) {
  // compensated sums: 16 bins per wave only; the wave target is still tuned
  const std::vector<int> shapes = c->compensated ? std::vector<int>{16} : std::vector<int>{32, 16};
  if (c->grad_m_fixed || !c->autotune) {
    c->grad_m = c->compensated ? 16 : (c->grad_m_fixed ? c->grad_m_fixed : kGradM);
    c->fused_waves = kDefaultFusedWaves;
    return PYIID_OK;
  }
  Table& t = c->tab[slot];
  const std::vector<int64_t> key{c->compensated ? 3 : 0, c->n, c->n_chain, t.nq,
                                 weighted_table(t) ? 1 : 0,
                                 static_cast<int64_t>(c->struct_stamp)};
  auto hit = c->tuned.find(key);
  if (hit != c->tuned.end()) {
    c->grad_m = static_cast<int>(hit->second[0]);
    c->fused_waves = hit->second[1];
    return PYIID_OK;
  }
  // a shape first met under graph capture runs the default shape
  c->grad_m = shapes[0];
  c->fused_waves = kDefaultFusedWaves;
  if (c->capturing) return PYIID_OK;
  double best = 1e300;
  int best_m = c->grad_m;
  int64_t best_w = c->fused_waves;
  // a candidate is timed with the reduction of its split partials (the
  // split changes both), twice after a warm-up launch; the faster counts
  CHK(c->io_b.reserve(sizeof(float) * static_cast<size_t>(c->n) * 3 * std::max(1, t.nq)));
  auto trial = [&]() {
    CHK(run_fused(c, slot, qbin, false));
    HIPCHK(launch_reduce_grad(c->stream, c->gpart.as<float>(), c->fused_items.split, c->n, t.nq,
                              t.gs.as<double>(), c->io_b.p, false));
    return PYIID_OK;
  };
  for (int m : shapes)
    for (int64_t w : {16384LL, 32768LL, 65536LL}) {
      c->grad_m = m;
      c->fused_waves = w;
      double ms = 0.0, ms2 = 0.0;
      CHK(trial());
      CHK(time_launch(c, trial, &ms));
      CHK(time_launch(c, trial, &ms2));
      ms = std::min(ms, ms2);
      if (ms < best) {
        best = ms;
        best_m = m;
        best_w = w;
      }
    }
  c->grad_m = best_m;
  c->fused_waves = best_w;
  c->tuned[key] = {best_m, best_w};
  return PYIID_OK;
}

// Kernel of the F + gradient path: the symmetric kernel at the fastest group
// size and shape, or the ordered kernel (its own shapes tuned by tune_fused)
// when that is faster, when compensated sums are on, when its bins per wave
// are fixed or when the autotuner is off.
int tune_grad(pyiid_ctx* c, int slot, double qbin) {
  Table& t = c->tab[slot];
  const bool weighted = weighted_table(t);
  if (c->compensated || c->grad_m_fixed || c->grad_kernel == PYIID_GRAD_ORDERED) {
    c->sym_k = 0;
    return tune_fused(c, slot, qbin);
  }
  if (c->sym_k_fixed) {
    c->sym_k = c->sym_k_fixed;
    c->sym_m = c->sym_m_fixed ? c->sym_m_fixed : 16;
    return PYIID_OK;
  }
  const bool sym_only = c->grad_kernel == PYIID_GRAD_SYMMETRIC;
  // without the autotuner, or first met under graph capture: default shapes
  if (!c->autotune || c->capturing) {
    if (!sym_only) {
      c->sym_k = 0;
      return tune_fused(c, slot, qbin);
    }
    c->sym_k = 16;
    c->sym_m = c->sym_m_fixed ? c->sym_m_fixed : 16;
    return PYIID_OK;
  }
  const std::vector<int64_t> key{5, c->n, c->n_chain, t.nq, weighted ? 1 : 0,
                                 static_cast<int64_t>(c->struct_stamp), c->sym_m_fixed,
                                 c->grad_kernel};
  auto hit = c->tuned.find(key);
  if (hit != c->tuned.end()) {
    c->sym_k = static_cast<int>(hit->second[0]);
    c->sym_m = static_cast<int>(hit->second[1]);
    return c->sym_k ? PYIID_OK : tune_fused(c, slot, qbin);
  }
  std::vector<int> cand;
  CHK(sym_candidates(c, t.nq, weighted, cand));
  if (sym_only && cand.empty()) cand.push_back(kSymGroups[3]);
  CHK(c->io_b.reserve(sizeof(float) * static_cast<size_t>(c->n) * 3 * std::max(1, t.nq)));
  // every candidate is timed twice (the faster counts) with the reduction of
  // its partials; the ordered kernel at the shape its own tuner chose
  auto time2 = [&](auto trial, double* best_ms) {
    double a = 0.0, b = 0.0;
    CHK(trial());
    CHK(time_launch(c, trial, &a));
    CHK(time_launch(c, trial, &b));
    *best_ms = std::min(a, b);
    return PYIID_OK;
  };
  double best = 1e300;
  int best_k = 0, best_m = 16;
  if (!sym_only) {
    c->sym_k = 0;
    CHK(tune_fused(c, slot, qbin));
    auto trial = [&]() {
      CHK(run_fused(c, slot, qbin, false));
      HIPCHK(launch_reduce_grad(c->stream, c->gpart.as<float>(), c->fused_items.split, c->n,
                                t.nq, t.gs.as<double>(), c->io_b.p, false));
      return PYIID_OK;
    };
    CHK(time2(trial, &best));
  }
  const std::vector<int> ms_shapes =
      c->sym_m_fixed ? std::vector<int>{c->sym_m_fixed} : std::vector<int>{16, 12};
  for (int k : cand)
    for (int m : ms_shapes) {
      c->sym_k = k;
      c->sym_m = m;
      auto trial = [&]() {
        CHK(run_sym(c, slot, qbin));
        HIPCHK(reduce_sym(c, t, c->io_b.p, false));
        return PYIID_OK;
      };
      double ms = 0.0;
      CHK(time2(trial, &ms));
      if (ms < best) {
        best = ms;
        best_k = k;
        best_m = m;
      }
    }
  c->sym_k = best_k;
  c->sym_m = best_m;
  c->tuned[key] = {best_k, best_m};
  // (the ordered kernel's shape is cached by tune_fused)
  return best_k ? PYIID_OK : tune_fused(c, slot, qbin);
}

// forces (n x 3 doubles) = scale * sum_q u[chain][q] grad(i, c, q), by the
// unordered-pair contracted-force kernel.
int contract_forces(pyiid_ctx* c, int slot, double qbin, double scale, double* out) {
  Table& t = c->tab[slot];
  const int nq = t.nq, B = c->n_chain;
  CHK(ensure_packed(c));
  CHK(ensure_norm(c, slot));
  CHK(c->uq.reserve(sizeof(float) * static_cast<size_t>(std::max(1, nq)) * B));
  HIPCHK(launch_uq(c->stream, c->u.as<float>(), B, nq,
                   static_cast<double>(static_cast<float>(qbin)), c->uq.as<float>()));
  CHK(tune_contract(c, slot, qbin));
  Timed tm(c, 4);
  CHK(launch_contract(c, slot, qbin));
  CHK(tm.done());
  HIPCHK(launch_reduce_pair_force(c->stream, c->pfpart.as<float>(), c->pair_entries.as<int>(),
                                  c->pair_items.ranges.as<Range>(), c->pair_items.n_ranges,
                                  c->pair_items.perm.as<int32_t>(), scale, out));
  return PYIID_OK;
}

int gemm_t(pyiid_ctx* c, const void* a, bool a64, int64_t rows, double* out) {
  Timed tm(c, 2);
  HIPCHK(launch_gemm_nt(c->stream, a, a64, rows, c->t_nq, c->tmat.as<float>(), c->t_nr, out,
                        nullptr));
  return tm.done();
}

// Device-only energy and forces of every chain against gobs_dev [B][n_r]:
// out5 [B][5], w [B][n_r]; forces (n x 3 doubles) = force_scale * dP/dG .
// gradG in the reference convention (gradG from -1/2 dF/dq).  No host
// synchronisation, no allocation once warm: safe inside graph capture.
int energy_forces_device(pyiid_ctx* c, int slot, double qbin, int potential,
                         const double* gobs_dev, double force_scale, double* forces_dev) {
  const int nq = c->t_nq, nr = c->t_nr, B = c->n_chain;
  CHK(compute_fq(c, slot, qbin));
  CHK(c->g64.reserve(sizeof(double) * (size_t)nr * B));
  CHK(gemm_t(c, c->fq32.p, false, B, c->g64.as<double>()));
  CHK(c->out5.reserve(sizeof(double) * 5 * B));
  CHK(c->w.reserve(sizeof(double) * (size_t)nr * B));
  HIPCHK(launch_potential(c->stream, potential, gobs_dev, c->g64.as<double>(), nr, B,
                          c->out5.as<double>(), c->w.as<double>()));
  if (!forces_dev) return PYIID_OK;
  CHK(c->u.reserve(sizeof(float) * (size_t)nq * B));
  HIPCHK(launch_adjoint(c->stream, c->tmat.as<float>(), nr, nq, c->w.as<double>(),
                        c->tab[slot].gs.as<double>(), B, c->u.as<float>()));
  return contract_forces(c, slot, qbin, force_scale, forces_dev);
}

int check_pdf_setup(pyiid_ctx* c, int slot, int potential) {
  if (!c->pos_set) return fail(PYIID_E_STATE, "positions not set");
  CHK(check_table(c, slot));
  if (!c->t_set) return fail(PYIID_E_STATE, "transform not set");
  if (c->tab[slot].nq != c->t_nq)
    return fail(PYIID_E_INVALID, "table has %d Q bins, transform expects %d", c->tab[slot].nq,
                c->t_nq);
  if (potential != PYIID_POT_RW && potential != PYIID_POT_CHI_SQ)
    return fail(PYIID_E_INVALID, "unknown potential %d", potential);
  return PYIID_OK;
}

void destroy_slot(GraphSlot& g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  g.exec = nullptr;
  g.graph = nullptr;
  g.leap = -1;
}

void destroy_graph(Hmc& h) {
  destroy_slot(h.exact_g);
  destroy_slot(h.sur_g);
}

int set_structure(pyiid_ctx* c, const float* pos, int64_t n, const std::vector<int64_t>& off,
                  int flags) {
  if (n < 0 || n > 2000000000LL)
    return fail(PYIID_E_INVALID, "atom count %lld out of range", (long long)n);
  if (n > 0 && !pos) return fail(PYIID_E_INVALID, "null positions");
  const bool same = (n == c->n) && (off == c->off);
  if (!same) {
    // the composition changes: elements reset to 0
    CHK(c->elem.reserve(sizeof(int32_t) * std::max<int64_t>(1, n)));
    c->elem_h.assign(static_cast<size_t>(n), 0);
    if (n > 0)
      HIPCHK(hipMemcpyAsync(c->elem.p, c->elem_h.data(), sizeof(int32_t) * n,
                            hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->max_elem = 0;
    c->n = n;
    c->off = off;
    c->n_chain = static_cast<int>(off.size()) - 1;
    c->struct_stamp++;
    c->hmc.set = false;
    destroy_graph(c->hmc);
  }
  CHK(c->posf.reserve(sizeof(float) * 3 * std::max<int64_t>(1, n)));
  CHK(c->pos.reserve(sizeof(AtomRec) * std::max<int64_t>(1, n)));
  CHK(upload(c, c->posf.p, pos, sizeof(float) * 3 * n, flags));
  c->pos_set = true;
  c->pos_dirty = true;
  CHK(ensure_packed(c));
  return finish(c, flags);
}

// One proposal of the batched sampler (fixed launch sequence).
int hmc_proposal(pyiid_ctx* c, int n_leapfrog, double eps, double beta, uint64_t seed) {
  Hmc& h = c->hmc;
  const int B = c->n_chain;
  const double fscale = 2.0 * h.conv;  // exact -dE/dq from the reference-convention contraction
  HIPCHK(launch_hmc_begin(c->stream, c->pos.as<AtomRec>(), h.pos_save.as<double>(),
                          h.mom.as<double>(), h.fcur.as<double>(), h.mass.as<double>(),
                          h.chains.as<HmcChain>(), B, beta, eps, seed, h.counter.as<int64_t>(),
                          h.kin0.as<double>()));
  for (int l = 1; l <= n_leapfrog; ++l) {
    HIPCHK(launch_hmc_drift(c->stream, c->pos.as<AtomRec>(), h.mom.as<double>(),
                            h.mass.as<double>(), c->n, eps));
    CHK(energy_forces_device(c, h.slot, h.qbin, h.potential, h.gobs.as<double>(), fscale,
                             h.fnew.as<double>()));
    HIPCHK(launch_hmc_kick(c->stream, h.mom.as<double>(), h.fnew.as<double>(), c->n,
                           l < n_leapfrog ? eps : 0.5 * eps));
  }
  HIPCHK(launch_hmc_end(c->stream, c->pos.as<AtomRec>(), h.pos_save.as<double>(),
                        h.mom.as<double>(), h.mass.as<double>(), h.fcur.as<double>(),
                        h.fnew.as<double>(), h.ecur.as<double>(), c->out5.as<double>(), h.conv,
                        h.chains.as<HmcChain>(), B, beta, seed, h.counter.as<int64_t>(),
                        h.kin0.as<double>(), h.ehist.as<double>(), h.acchist.as<int32_t>(),
                        h.hist_cap, nullptr, nullptr));
  return PYIID_OK;
}

// ---- learned integrator ----

constexpr int kGramChunk = 2048;  // flat force columns per Gram partial
// learned-force proposals after which a block's acceptance is first checked
constexpr int kSurAcceptProbe = 3;

uint16_t host_bf16(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

// Tiles, chunk lists and per-atom buffers of the surrogate for the current
// batch layout.  The output-layer fit restarts when the layout changes.
int ensure_sur_layout(pyiid_ctx* c) {
  Sur& s = c->sur;
  if (s.stamp == c->struct_stamp) return PYIID_OK;
  if (c->capturing) return fail(PYIID_E_STATE, "structure changed during graph capture");
  const int B = c->n_chain;
  const int64_t n = c->n;
  std::vector<Item> t64, t256, chunks;
  std::vector<Range> gr(B);
  std::vector<int32_t> ach(static_cast<size_t>(std::max<int64_t>(1, n)), 0);
  int slot = 0;
  for (int b = 0; b < B; ++b) {
    const int o = static_cast<int>(c->off[b]), e = static_cast<int>(c->off[b + 1]);
    for (int ib = o; ib < e; ib += 64) t64.push_back(Item{ib, std::min(e, ib + 64), o, e, 0, b, 0, 0});
    for (int ib = o; ib < e; ib += 256)
      t256.push_back(Item{ib, std::min(e, ib + 256), o, e, 0, b, 0, 0});
    for (int i = o; i < e; ++i) ach[i] = b;
    gr[b].begin = slot;
    for (int64_t c0 = 3LL * o; c0 < 3LL * e; c0 += kGramChunk)
      chunks.push_back(Item{0, 0, static_cast<int>(c0),
                            static_cast<int>(std::min<int64_t>(3LL * e, c0 + kGramChunk)), slot++,
                            b, 0, 0});
    gr[b].end = slot;
    gr[b].i_base = b;
    gr[b].i_end = b;
  }
  const int64_t n1 = std::max<int64_t>(1, n);
  CHK(s.tiles64.reserve(sizeof(Item) * std::max<size_t>(1, t64.size())));
  CHK(s.tiles256.reserve(sizeof(Item) * std::max<size_t>(1, t256.size())));
  CHK(s.chunks.reserve(sizeof(Item) * std::max<size_t>(1, chunks.size())));
  CHK(s.granges.reserve(sizeof(Range) * B));
  CHK(s.atom_chain.reserve(sizeof(int32_t) * n1));
  CHK(s.nbr.reserve(sizeof(int32_t) * static_cast<size_t>(s.max_nbr) * n1));
  CHK(s.cnt.reserve(sizeof(int32_t) * n1));
  CHK(s.d.reserve(sizeof(float) * kSurRbf * n1));
  CHK(s.cbuf.reserve(sizeof(float) * kSurRbf * n1));
  CHK(s.eatom.reserve(sizeof(float) * n1));
  CHK(s.fsur.reserve(sizeof(double) * 3 * n1));
  CHK(s.gpart.reserve(sizeof(float) * kSurGram * std::max<size_t>(1, chunks.size())));
  CHK(s.gout.reserve(sizeof(double) * kSurGram * B));
  CHK(s.viol.reserve(sizeof(int32_t) * B));
  CHK(s.viol_count.reserve(sizeof(int32_t) * B));
  CHK(s.w3.reserve(sizeof(uint16_t) * kSurHidden * B));
  CHK(s.efeat.reserve(sizeof(float) * kSurHidden * n1));
  CHK(s.efe.reserve(sizeof(double) * kSurHidden * B));
  CHK(upload_vec(s.offs, c->off));
  if (!t64.empty())
    HIPCHK(hipMemcpy(s.tiles64.p, t64.data(), sizeof(Item) * t64.size(), hipMemcpyHostToDevice));
  if (!t256.empty())
    HIPCHK(hipMemcpy(s.tiles256.p, t256.data(), sizeof(Item) * t256.size(), hipMemcpyHostToDevice));
  if (!chunks.empty())
    HIPCHK(hipMemcpy(s.chunks.p, chunks.data(), sizeof(Item) * chunks.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s.granges.p, gr.data(), sizeof(Range) * B, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s.atom_chain.p, ach.data(), sizeof(int32_t) * n1, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(s.viol.p, 0, sizeof(int32_t) * B));
  HIPCHK(hipMemset(s.viol_count.p, 0, sizeof(int32_t) * B));
  s.n_t64 = static_cast<int>(t64.size());
  s.n_t256 = static_cast<int>(t256.size());
  s.n_chunks = static_cast<int>(chunks.size());
  s.acc.assign(static_cast<size_t>(kSurGram) * B, 0.0);
  s.r2.assign(B, 0.0);
  s.w3_set = false;
  s.eprev_set = false;
  s.stamp = c->struct_stamp;
  return PYIID_OK;
}

int sur_list(pyiid_ctx* c) {
  Sur& s = c->sur;
  HIPCHK(launch_nbr_build(c->stream, c->pos.as<AtomRec>(), static_cast<int>(c->n),
                          s.tiles256.as<Item>(), s.n_t256, s.rbf[0] + s.skin, s.max_nbr,
                          s.nbr.as<int>(), s.cnt.as<int>(), s.viol.as<int>()));
  return PYIID_OK;
}

// Surrogate forces (or one feature-force column) from the current list.
int sur_eval(pyiid_ctx* c, const uint16_t* w3, int w3_stride, double* force, float* phi,
             float* eatom, bool descriptors = true) {
  Sur& s = c->sur;
  const int n = static_cast<int>(c->n);
  if (descriptors)
    HIPCHK(launch_descriptors(c->stream, c->pos.as<AtomRec>(), n, s.nbr.as<int>(),
                              s.cnt.as<int>(), s.rbf, s.d.as<float>()));
  HIPCHK(launch_mlp(c->stream, s.d.as<float>(), n, s.tiles64.as<Item>(), s.n_t64, s.net.p, w3,
                    w3_stride, s.cbuf.as<float>(), eatom));
  HIPCHK(launch_sur_force(c->stream, c->pos.as<AtomRec>(), n, s.nbr.as<int>(), s.cnt.as<int>(),
                          s.rbf, s.cbuf.as<float>(), force, phi));
  return PYIID_OK;
}

// Raises if a neighbour list overflowed (clears the flags).
int sur_check_lists(pyiid_ctx* c) {
  Sur& s = c->sur;
  std::vector<int32_t> v(c->n_chain);
  HIPCHK(hipMemcpyAsync(v.data(), s.viol.p, sizeof(int32_t) * v.size(), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int b = 0; b < c->n_chain; ++b)
    if (v[b]) {
      HIPCHK(hipMemset(s.viol.p, 0, sizeof(int32_t) * c->n_chain));
      return fail(PYIID_E_INVALID,
                  "chain %d: an atom has more than %d neighbours within %g A (raise max_nbr)", b,
                  s.max_nbr, static_cast<double>(s.rbf[0] + s.skin));
    }
  return PYIID_OK;
}

// Feature forces Phi[k][3i + c] = -d/dx_ic sum_a h2_k(D_a) for the 64
// output units (the forces of the network with w3 = e_k), and the feature
// energies efe[b][k] = sum_{a in chain b} h2_k(D_a) (its chain energies).
int sur_features(pyiid_ctx* c) {
  Sur& s = c->sur;
  const int64_t n3 = 3 * c->n;
  CHK(s.phi.reserve(sizeof(float) * kSurHidden * static_cast<size_t>(std::max<int64_t>(1, n3))));
  CHK(sur_list(c));
  CHK(sur_check_lists(c));
  HIPCHK(launch_descriptors(c->stream, c->pos.as<AtomRec>(), static_cast<int>(c->n),
                            s.nbr.as<int>(), s.cnt.as<int>(), s.rbf, s.d.as<float>()));
  for (int k = 0; k < kSurHidden; ++k)
    CHK(sur_eval(c, s.onehot.as<uint16_t>() + kSurHidden * k, 0, nullptr,
                 s.phi.as<float>() + k * n3, s.efeat.as<float>() + k * c->n, false));
  HIPCHK(launch_chain_feature_energy(c->stream, s.efeat.as<float>(), c->n, s.offs.as<int64_t>(),
                                     c->n_chain, s.efe.as<double>()));
  return PYIID_OK;
}

// Solves (A + lam I) w = r in place by Cholesky (FP64); false if not SPD.
bool cholesky_solve(std::vector<double>& a, std::vector<double>& r, int n) {
  for (int j = 0; j < n; ++j) {
    double d = a[j * n + j];
    for (int k = 0; k < j; ++k) d -= a[j * n + k] * a[j * n + k];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    a[j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      double v = a[i * n + j];
      for (int k = 0; k < j; ++k) v -= a[i * n + k] * a[j * n + k];
      a[i * n + j] = v / d;
    }
  }
  for (int i = 0; i < n; ++i) {
    double v = r[i];
    for (int k = 0; k < i; ++k) v -= a[i * n + k] * r[k];
    r[i] = v / a[i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double v = r[i];
    for (int k = i + 1; k < n; ++k) v -= a[k * n + i] * r[k];
    r[i] = v / a[i * n + i];
  }
  return true;
}

// One training step at the current state: exact energy and forces (they also
// refresh the sampler's state), feature forces and energies, the chains'
// normal equations added to the running ones (decayed by `forget`), ridge
// solve for w3.  The normal equations hold the force rows Phi w = F of this
// state and, from the second state on, one energy row per chain,
// sqrt(lambda) (e - e_prev) w = sqrt(lambda) (E - E_prev): the exact energy
// change since the previous training state, which fixes the network's
// energy up to the constant it cannot see.
int sur_train_step(pyiid_ctx* c, double forget, double ridge) {
  Hmc& h = c->hmc;
  Sur& s = c->sur;
  const int B = c->n_chain;
  CHK(ensure_sur_layout(c));
  CHK(energy_forces_device(c, h.slot, h.qbin, h.potential, h.gobs.as<double>(), 2.0 * h.conv,
                           h.fcur.as<double>()));
  HIPCHK(launch_energy_from_out5(c->stream, c->out5.as<double>(), B, h.conv, h.ecur.as<double>()));
  h.fcur_stale = false;
  CHK(sur_features(c));
  HIPCHK(launch_gram(c->stream, s.phi.as<float>(), 3 * c->n, h.fcur.as<double>(),
                     s.chunks.as<Item>(), s.n_chunks, s.gpart.as<float>(), s.granges.as<Range>(),
                     B, s.gout.as<double>()));
  std::vector<double> g(static_cast<size_t>(kSurGram) * B);
  std::vector<double> ef(static_cast<size_t>(kSurHidden) * B), en(B);
  HIPCHK(hipMemcpyAsync(g.data(), s.gout.p, sizeof(double) * g.size(), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipMemcpyAsync(ef.data(), s.efe.p, sizeof(double) * ef.size(), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipMemcpyAsync(en.data(), h.ecur.p, sizeof(double) * B, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  constexpr int H = kSurHidden;
  const bool erow = s.eprev_set && s.energy_weight > 0.0;
  std::vector<uint16_t> w3h(static_cast<size_t>(H) * B);
  std::vector<double> a(H * H), w(H);
  for (int b = 0; b < B; ++b) {
    double* acc = &s.acc[static_cast<size_t>(b) * kSurGram];
    const double* gn = &g[static_cast<size_t>(b) * kSurGram];
    for (int e = 0; e < kSurGram; ++e) acc[e] = forget * acc[e] + gn[e];
    if (erow) {
      const double* e1 = &ef[static_cast<size_t>(b) * H];
      const double* e0 = &s.eprev[static_cast<size_t>(b) * H];
      const double dE = en[b] - s.Eprev[b], lam = s.energy_weight;
      for (int k = 0; k < H; ++k) {
        const double dk = e1[k] - e0[k];
        for (int l = 0; l < H; ++l) acc[k * H + l] += lam * dk * (e1[l] - e0[l]);
        acc[H * H + k] += lam * dk * dE;
      }
      acc[H * H + H] += lam * dE * dE;
    }
    double tr = 0.0;
    for (int k = 0; k < H; ++k) tr += acc[k * H + k];
    std::fill(w.begin(), w.end(), 0.0);
    if (tr > 0.0) {
      double lam = std::max(ridge, 1e-12) * tr / H;
      for (int attempt = 0; attempt < 8; ++attempt, lam *= 10.0) {
        for (int e = 0; e < H * H; ++e) a[e] = acc[e];
        for (int k = 0; k < H; ++k) a[k * H + k] += lam;
        for (int k = 0; k < H; ++k) w[k] = acc[H * H + k];
        if (cholesky_solve(a, w, H)) break;
        std::fill(w.begin(), w.end(), 0.0);
      }
    }
    // fit quality on this sample: 1 - |y - Phi w|^2 / |y|^2
    const double yy = gn[H * H + H];
    double res = yy;
    for (int k = 0; k < H; ++k) {
      res -= 2.0 * w[k] * gn[H * H + k];
      double t = 0.0;
      for (int l = 0; l < H; ++l) t += gn[k * H + l] * w[l];
      res += w[k] * t;
    }
    s.r2[b] = yy > 0.0 ? 1.0 - res / yy : 0.0;
    for (int k = 0; k < H; ++k) w3h[static_cast<size_t>(b) * H + k] = host_bf16(static_cast<float>(w[k]));
  }
  HIPCHK(hipMemcpy(s.w3.p, w3h.data(), sizeof(uint16_t) * w3h.size(), hipMemcpyHostToDevice));
  s.w3_set = true;
  s.eprev.swap(ef);
  s.Eprev.swap(en);
  s.eprev_set = true;
  return PYIID_OK;
}

// One proposal driven by surrogate forces; the exact energy of the end point
// decides acceptance.  Fixed launch sequence (capturable).
int sur_proposal(pyiid_ctx* c, int n_leapfrog, double eps, double beta, uint64_t seed) {
  Hmc& h = c->hmc;
  Sur& s = c->sur;
  const int B = c->n_chain;
  const uint16_t* w3 = s.w3.as<uint16_t>();
  CHK(sur_list(c));
  CHK(sur_eval(c, w3, kSurHidden, s.fsur.as<double>(), nullptr, nullptr));
  HIPCHK(launch_hmc_begin(c->stream, c->pos.as<AtomRec>(), h.pos_save.as<double>(),
                          h.mom.as<double>(), s.fsur.as<double>(), h.mass.as<double>(),
                          h.chains.as<HmcChain>(), B, beta, eps, seed, h.counter.as<int64_t>(),
                          h.kin0.as<double>()));
  for (int l = 1; l <= n_leapfrog; ++l) {
    HIPCHK(launch_hmc_drift(c->stream, c->pos.as<AtomRec>(), h.mom.as<double>(),
                            h.mass.as<double>(), c->n, eps));
    HIPCHK(launch_disp_check(c->stream, c->pos.as<AtomRec>(), h.pos_save.as<double>(),
                             s.atom_chain.as<int>(), c->n, 0.5 * s.skin, s.viol.as<int>()));
    CHK(sur_eval(c, w3, kSurHidden, s.fsur.as<double>(), nullptr, nullptr));
    HIPCHK(launch_hmc_kick(c->stream, h.mom.as<double>(), s.fsur.as<double>(), c->n,
                           l < n_leapfrog ? eps : 0.5 * eps));
  }
  CHK(energy_forces_device(c, h.slot, h.qbin, h.potential, h.gobs.as<double>(), 0.0, nullptr));
  HIPCHK(launch_hmc_end(c->stream, c->pos.as<AtomRec>(), h.pos_save.as<double>(),
                        h.mom.as<double>(), h.mass.as<double>(), nullptr, nullptr,
                        h.ecur.as<double>(), c->out5.as<double>(), h.conv,
                        h.chains.as<HmcChain>(), B, beta, seed, h.counter.as<int64_t>(),
                        h.kin0.as<double>(), h.ehist.as<double>(), h.acchist.as<int32_t>(),
                        h.hist_cap, s.viol.as<int>(), s.viol_count.as<int>()));
  h.fcur_stale = true;
  return PYIID_OK;
}

// Runs one proposal of either kind, from a captured graph when asked: the
// first call with new parameters (or after any workspace moved) runs the
// proposal directly, then captures it for the following calls.
int run_proposal(pyiid_ctx* c, bool surrogate, bool use_graph, int n_leapfrog, double eps,
                 double beta, uint64_t seed) {
  Hmc& h = c->hmc;
  auto body = [&]() {
    return surrogate ? sur_proposal(c, n_leapfrog, eps, beta, seed)
                     : hmc_proposal(c, n_leapfrog, eps, beta, seed);
  };
  if (!use_graph) return body();
  GraphSlot& g = surrogate ? h.sur_g : h.exact_g;
  const bool same = g.exec && g.leap == n_leapfrog && g.eps == eps && g.beta == beta &&
                    g.seed == seed && g.stamp == c->struct_stamp && g.cap == h.hist_cap &&
                    g.epoch == g_alloc_epoch.load();
  if (same) {
    HIPCHK(hipGraphLaunch(g.exec, c->stream));
    if (surrogate) h.fcur_stale = true;
    return PYIID_OK;
  }
  destroy_slot(g);
  CHK(body());  // this proposal; it also sizes every workspace
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  const int rc = body();  // recorded, not executed
  c->capturing = false;
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &gr);
  if (rc != PYIID_OK) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  HIPCHK(e);
  g.graph = gr;
  HIPCHK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
  g.leap = n_leapfrog;
  g.eps = eps;
  g.beta = beta;
  g.seed = seed;
  g.stamp = c->struct_stamp;
  g.cap = h.hist_cap;
  g.epoch = g_alloc_epoch.load();
  return PYIID_OK;
}

// Exact energy and forces of the current state when surrogate moves made
// the sampler's cached ones stale.
int refresh_exact(pyiid_ctx* c) {
  Hmc& h = c->hmc;
  if (!h.fcur_stale) return PYIID_OK;
  CHK(energy_forces_device(c, h.slot, h.qbin, h.potential, h.gobs.as<double>(), 2.0 * h.conv,
                           h.fcur.as<double>()));
  HIPCHK(launch_energy_from_out5(c->stream, c->out5.as<double>(), c->n_chain, h.conv,
                                 h.ecur.as<double>()));
  h.fcur_stale = false;
  return PYIID_OK;
}

int prepare_history(pyiid_ctx* c, int n_proposals) {
  Hmc& h = c->hmc;
  const int B = c->n_chain;
  if (n_proposals > h.hist_cap) {
    destroy_graph(h);  // graphs record the history buffers' addresses
    CHK(h.ehist.reserve(sizeof(double) * (size_t)n_proposals * B));
    CHK(h.acchist.reserve(sizeof(int32_t) * (size_t)n_proposals * B));
    h.hist_cap = n_proposals;
  }
  const int64_t zero = 0;
  HIPCHK(hipMemcpyAsync(h.counter.p, &zero, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PYIID_OK;
}

// Per-proposal completion times (ms since the run's start) from events
// recorded on the stream after each proposal.
struct RunClock {
  pyiid_ctx* c;
  hipEvent_t start = nullptr;
  std::vector<hipEvent_t> ev;
  explicit RunClock(pyiid_ctx* c_, bool on) : c(c_) {
    if (on && (start = take_event(c))) (void)hipEventRecord(start, c->stream);
  }
  int mark() {
    if (!start) return PYIID_OK;
    hipEvent_t e = take_event(c);
    if (!e) return fail(PYIID_E_HIP, "hipEventCreate failed");
    HIPCHK(hipEventRecord(e, c->stream));
    ev.push_back(e);
    return PYIID_OK;
  }
  int collect(double* out) {
    if (!start) return PYIID_OK;
    for (size_t k = 0; k < ev.size(); ++k) {
      HIPCHK(hipEventSynchronize(ev[k]));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, start, ev[k]));
      if (out) out[k] = ms;
    }
    return PYIID_OK;
  }
  ~RunClock() {
    if (start) c->ev_pool.push_back(start);
    for (hipEvent_t e : ev) c->ev_pool.push_back(e);
  }
};

}  // namespace

int pyiid::set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

extern "C" {

int pyiid_abi_version(void) { return PYIID_ABI_VERSION; }

const char* pyiid_last_error(void) { return g_err; }

int pyiid_device_count(int* count) {
  if (!count) return fail(PYIID_E_INVALID, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return PYIID_OK;
}

int pyiid_create(int device, pyiid_ctx** out) {
  if (!out) return fail(PYIID_E_INVALID, "null output pointer");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(PYIID_E_NODEVICE, "no HIP device is visible");
  if (device < 0 || device >= n)
    return fail(PYIID_E_INVALID, "device %d out of range [0, %d)", device, n);
  HIPCHK(hipSetDevice(device));
  pyiid_ctx* c = new pyiid_ctx();
  c->device = device;
  hipError_t e = hipStreamCreate(&c->own_stream);
  if (e != hipSuccess) {
    delete c;
    return fail(PYIID_E_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
  }
  c->stream = c->own_stream;
  *out = c;
  return PYIID_OK;
}

int pyiid_destroy(pyiid_ctx* c) {
  if (!c) return PYIID_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  destroy_graph(c->hmc);
  for (Buf* b : {&c->posf, &c->elem, &c->pos, &c->counts_d, &c->natoms_d, &c->tmat, &c->gpart,
                 &c->fpart, &c->fq64, &c->fq32, &c->g64, &c->io_a, &c->io_b,
                 &c->io_c, &c->w, &c->out5, &c->u, &c->gobs, &c->fq_items.items,
                 &c->fq_items.ranges, &c->fq_items.perm, &c->fused_items.items, &c->fused_items.ranges,
                 &c->fused_items.perm, &c->hmc.mass,
                 &c->hmc.mom, &c->hmc.fcur, &c->hmc.fnew, &c->hmc.pos_save, &c->hmc.kin0,
                 &c->hmc.ecur, &c->hmc.ehist, &c->hmc.acchist, &c->hmc.counter, &c->hmc.chains,
                 &c->hmc.gobs, &c->sur.net, &c->sur.w3, &c->sur.onehot, &c->sur.nbr, &c->sur.cnt,
                 &c->sur.d, &c->sur.cbuf, &c->sur.eatom, &c->sur.fsur, &c->sur.phi,
                 &c->sur.gpart, &c->sur.gout, &c->sur.viol, &c->sur.viol_count,
                 &c->sur.atom_chain, &c->sur.tiles64, &c->sur.tiles256, &c->sur.chunks,
                 &c->sur.granges, &c->pair_items.items, &c->pair_items.ranges, &c->pair_items.perm,
                 &c->pair_entries, &c->uq, &c->pfpart})
    b->release();
  for (Table& t : c->tab) {
    t.tab.release();
    t.na.release();
    t.fsu.release();
    t.fso.release();
    t.gs.release();
  }
  for (auto& pe : c->ev_pending) {
    (void)hipEventDestroy(pe.second.first);
    (void)hipEventDestroy(pe.second.second);
  }
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return PYIID_OK;
}

int pyiid_set_stream(pyiid_ctx* c, void* stream) {
  CHK(check_ctx(c));
  c->stream = stream ? static_cast<hipStream_t>(stream) : c->own_stream;
  destroy_graph(c->hmc);
  return PYIID_OK;
}

int pyiid_synchronize(pyiid_ctx* c) {
  CHK(check_ctx(c));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PYIID_OK;
}

int pyiid_enable_timing(pyiid_ctx* c, int enable) {
  CHK(check_ctx(c));
  CHK(resolve_timing(c));
  c->timing = enable != 0;
  for (int k = 0; k < 5; ++k) {
    c->t_ms[k] = 0.0;
    c->t_n[k] = 0;
  }
  if (c->timing) {
    CHK(c->formcnt.reserve(2 * sizeof(unsigned long long)));
    HIPCHK(hipMemsetAsync(c->formcnt.p, 0, 2 * sizeof(unsigned long long), c->stream));
  }
  return PYIID_OK;
}

int pyiid_fq_form_counts(pyiid_ctx* c, int64_t* three_term, int64_t* reinsch) {
  CHK(check_ctx(c));
  unsigned long long h[2] = {0, 0};
  if (c->formcnt.p) {
    HIPCHK(hipMemcpyAsync(h, c->formcnt.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (three_term) *three_term = static_cast<int64_t>(h[0]);
  if (reinsch) *reinsch = static_cast<int64_t>(h[1]);
  return PYIID_OK;
}

// The launch-shape setters below drop captured proposals when the setting
// changes: a graph records the shapes it was captured with.
int pyiid_set_autotune(pyiid_ctx* c, int enable) {
  CHK(check_ctx(c));
  if (c->autotune != (enable != 0)) destroy_graph(c->hmc);
  c->autotune = enable != 0;
  return PYIID_OK;
}

int pyiid_set_precision(pyiid_ctx* c, int mode) {
  CHK(check_ctx(c));
  if (mode != PYIID_PRECISION_DEFAULT && mode != PYIID_PRECISION_COMPENSATED)
    return fail(PYIID_E_INVALID, "precision mode %d is not %d or %d", mode,
                PYIID_PRECISION_DEFAULT, PYIID_PRECISION_COMPENSATED);
  const bool comp = mode == PYIID_PRECISION_COMPENSATED;
  if (c->compensated != comp) destroy_graph(c->hmc);
  c->compensated = comp;
  return PYIID_OK;
}

int pyiid_set_fq_bins_per_wave(pyiid_ctx* c, int m) {
  CHK(check_ctx(c));
  if (m != 0 && m != 64 && m != 72 && m != 80 && m != 88 && m != 96 && m != 112)
    return fail(PYIID_E_INVALID, "bins per wave %d is not 0, 64, 72, 80, 88, 96 or 112", m);
  if (c->fq_m_fixed != m) destroy_graph(c->hmc);
  c->fq_m_fixed = m;
  return PYIID_OK;
}

int pyiid_set_grad_bins_per_wave(pyiid_ctx* c, int m) {
  CHK(check_ctx(c));
  if (m != 0 && m != 16 && m != 32)
    return fail(PYIID_E_INVALID, "bins per wave %d is not 0, 16 or 32", m);
  if (c->grad_m_fixed != m) destroy_graph(c->hmc);
  c->grad_m_fixed = m;
  return PYIID_OK;
}

int pyiid_set_grad_tiles_per_item(pyiid_ctx* c, int k) {
  CHK(check_ctx(c));
  if (k != 0 && k != 4 && k != 8 && k != 16 && k != 32)
    return fail(PYIID_E_INVALID, "tiles per item %d is not 0, 4, 8, 16 or 32", k);
  c->sym_k_fixed = k;
  return PYIID_OK;
}

int pyiid_set_grad_kernel(pyiid_ctx* c, int kernel) {
  CHK(check_ctx(c));
  if (kernel != PYIID_GRAD_AUTO && kernel != PYIID_GRAD_ORDERED && kernel != PYIID_GRAD_SYMMETRIC)
    return fail(PYIID_E_INVALID, "F + gradient kernel %d is not 0, 1 or 2", kernel);
  c->grad_kernel = kernel;
  return PYIID_OK;
}

int pyiid_grad_shape(pyiid_ctx* c, int* symmetric, int* bins_per_wave, int* group) {
  CHK(check_ctx(c));
  if (!symmetric || !bins_per_wave || !group) return fail(PYIID_E_INVALID, "null output");
  *symmetric = c->sym_k ? 1 : 0;
  *bins_per_wave = c->sym_k ? c->sym_m : c->grad_m;
  *group = c->sym_k ? c->sym_k : c->fused_items.split;
  return PYIID_OK;
}

int pyiid_set_grad_sym_bins_per_wave(pyiid_ctx* c, int m) {
  CHK(check_ctx(c));
  if (m != 0 && m != 12 && m != 16)
    return fail(PYIID_E_INVALID, "symmetric bins per wave %d is not 0, 12 or 16", m);
  c->sym_m_fixed = m;
  return PYIID_OK;
}

int pyiid_kernel_time(pyiid_ctx* c, int kind, double* total_ms, int64_t* launches) {
  CHK(check_ctx(c));
  if (kind < 0 || kind > 4) return fail(PYIID_E_INVALID, "kernel kind %d not in [0, 4]", kind);
  CHK(resolve_timing(c));
  if (total_ms) *total_ms = c->t_ms[kind];
  if (launches) *launches = c->t_n[kind];
  return PYIID_OK;
}

int pyiid_set_positions(pyiid_ctx* c, const float* pos, int64_t n, int flags) {
  CHK(check_ctx(c));
  std::vector<int64_t> off{0, n};
  return set_structure(c, pos, n, off, flags);
}

int pyiid_set_batch(pyiid_ctx* c, const float* pos, int64_t n_total, const int64_t* offsets,
                    int32_t n_chain, int flags) {
  CHK(check_ctx(c));
  if (n_chain < 1 || !offsets) return fail(PYIID_E_INVALID, "need n_chain >= 1 and offsets");
  std::vector<int64_t> off(offsets, offsets + n_chain + 1);
  if (off[0] != 0 || off[n_chain] != n_total)
    return fail(PYIID_E_INVALID, "offsets must run from 0 to n_total");
  for (int b = 0; b < n_chain; ++b)
    if (off[b + 1] < off[b]) return fail(PYIID_E_INVALID, "offsets must be non-decreasing");
  return set_structure(c, pos, n_total, off, flags);
}

int pyiid_get_positions(pyiid_ctx* c, float* pos, int flags) {
  CHK(check_ctx(c));
  if (!c->pos_set) return fail(PYIID_E_STATE, "positions not set");
  if (!pos) return fail(PYIID_E_INVALID, "null output");
  CHK(ensure_packed(c));
  if (c->n == 0) return PYIID_OK;
  float* dst = pos;
  if (!(flags & PYIID_DEVICE)) {
    CHK(c->io_a.reserve(sizeof(float) * 3 * c->n));
    dst = c->io_a.as<float>();
  }
  hipLaunchKernelGGL(unpack_kernel, dim3(static_cast<unsigned>((c->n + 255) / 256)), dim3(256), 0,
                     c->stream, c->pos.as<AtomRec>(), c->n, dst);
  HIPCHK(hipGetLastError());
  if (!(flags & PYIID_DEVICE)) CHK(download(c, pos, dst, sizeof(float) * 3 * c->n, flags));
  return finish(c, flags);
}

int pyiid_set_elements(pyiid_ctx* c, const int32_t* elem, int64_t n, int flags) {
  CHK(check_ctx(c));
  if (!c->pos_set) return fail(PYIID_E_STATE, "set positions before elements");
  if (n != c->n)
    return fail(PYIID_E_INVALID, "element count %lld != atom count %lld", (long long)n,
                (long long)c->n);
  std::vector<int32_t> h(static_cast<size_t>(n));
  if (n > 0) {
    if (!elem) return fail(PYIID_E_INVALID, "null elements");
    if (flags & PYIID_DEVICE) {
      HIPCHK(hipMemcpyAsync(h.data(), elem, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    } else {
      std::memcpy(h.data(), elem, sizeof(int32_t) * n);
    }
  }
  int mx = 0;
  for (int32_t e : h) {
    if (e < 0 || e >= 1024) return fail(PYIID_E_INVALID, "element index %d out of range [0, 1024)", e);
    mx = std::max(mx, static_cast<int>(e));
  }
  c->elem_h = h;
  c->max_elem = mx;
  if (n > 0)
    HIPCHK(hipMemcpyAsync(c->elem.p, h.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->struct_stamp++;
  c->hmc.set = false;
  destroy_graph(c->hmc);
  c->pos_dirty = true;
  CHK(ensure_packed(c));
  return finish(c, flags);
}

int pyiid_set_scatter(pyiid_ctx* c, int slot, const float* table, int32_t n_elem, int32_t n_q,
                      int flags) {
  CHK(check_ctx(c));
  if (slot != PYIID_SLOT_FQ && slot != PYIID_SLOT_PDF)
    return fail(PYIID_E_INVALID, "slot %d is not 0 or 1", slot);
  if (n_elem < 1 || n_q < 1 || !table)
    return fail(PYIID_E_INVALID, "scatter table needs n_elem >= 1, n_q >= 1 (got %d, %d)", n_elem,
                n_q);
  Table& t = c->tab[slot];
  CHK(t.tab.reserve(sizeof(float) * (size_t)n_elem * n_q));
  CHK(upload(c, t.tab.p, table, sizeof(float) * (size_t)n_elem * n_q, flags));
  t.n_elem = n_elem;
  t.nq = n_q;
  t.set = true;
  c->tab_stamp[slot]++;
  c->hmc.set = false;
  destroy_graph(c->hmc);
  return finish(c, flags);
}

int pyiid_fq(pyiid_ctx* c, int slot, double qbin, float* fq, int flags) {
  CHK(check_ctx(c));
  if (!c->pos_set) return fail(PYIID_E_STATE, "positions not set");
  CHK(check_table(c, slot));
  if (!(qbin > 0.0)) return fail(PYIID_E_INVALID, "qbin must be positive");
  if (!fq) return fail(PYIID_E_INVALID, "null output");
  const int nq = c->tab[slot].nq;
  const size_t cnt = (size_t)nq * c->n_chain;
  CHK(compute_fq(c, slot, qbin));
  if (flags & PYIID_DEVICE) {
    HIPCHK(hipMemcpyAsync(fq, c->fq32.p, sizeof(float) * cnt, hipMemcpyDeviceToDevice, c->stream));
  } else {
    CHK(download(c, fq, c->fq32.p, sizeof(float) * cnt, flags));
  }
  return finish(c, flags);
}

int pyiid_grad_fq(pyiid_ctx* c, int slot, double qbin, void* grad, float* fq, int flags) {
  CHK(check_ctx(c));
  if (!c->pos_set) return fail(PYIID_E_STATE, "positions not set");
  CHK(check_table(c, slot));
  if (!(qbin > 0.0)) return fail(PYIID_E_INVALID, "qbin must be positive");
  if (!grad) return fail(PYIID_E_INVALID, "null gradient output");
  if (c->n_chain != 1) return fail(PYIID_E_INVALID, "pyiid_grad_fq takes a single structure");
  const int nq = c->tab[slot].nq;
  const int64_t n = c->n;
  const size_t esz = (flags & PYIID_F64) ? sizeof(double) : sizeof(float);
  const size_t gbytes = esz * (size_t)n * 3 * nq;
  if (n < 2) {
    if (flags & PYIID_DEVICE) {
      if (gbytes) HIPCHK(hipMemsetAsync(grad, 0, gbytes, c->stream));
      if (fq) HIPCHK(hipMemsetAsync(fq, 0, sizeof(float) * nq, c->stream));
    } else {
      if (gbytes) std::memset(grad, 0, gbytes);
      if (fq) std::memset(fq, 0, sizeof(float) * nq);
    }
    return finish(c, flags);
  }
  CHK(ensure_packed(c));
  CHK(ensure_norm(c, slot));
  CHK(tune_grad(c, slot, qbin));
  const bool sym = c->sym_k != 0;
  if (sym)
    CHK(run_sym(c, slot, qbin));
  else
    CHK(run_fused(c, slot, qbin, false));
  Table& t = c->tab[slot];
  void* gdst = grad;
  if (!(flags & PYIID_DEVICE)) {
    CHK(c->io_c.reserve(gbytes));
    gdst = c->io_c.p;
  }
  if (sym)
    HIPCHK(reduce_sym(c, t, gdst, (flags & PYIID_F64) != 0));
  else
    HIPCHK(launch_reduce_grad(c->stream, c->gpart.as<float>(), c->fused_items.split, n, nq,
                              t.gs.as<double>(), gdst, (flags & PYIID_F64) != 0));
  if (!(flags & PYIID_DEVICE)) CHK(download(c, grad, gdst, gbytes, flags));
  if (fq) {
    // F from all items of the single chain (for the ordered kernel, the range
    // stored after the tile ranges)
    CHK(c->io_a.reserve(sizeof(float) * nq));
    float* fdst = (flags & PYIID_DEVICE) ? fq : c->io_a.as<float>();
    const Range* fr = sym ? c->sym.franges.as<Range>()
                          : c->fused_items.ranges.as<Range>() + c->fused_items.n_ranges;
    HIPCHK(launch_reduce_ranges(c->stream, c->fpart.as<double>(), fr, 1, nq, t.fsu.as<double>(),
                                nullptr, fdst));
    if (!(flags & PYIID_DEVICE)) CHK(download(c, fq, fdst, sizeof(float) * nq, flags));
  }
  return finish(c, flags);
}

int pyiid_pair_geometry(pyiid_ctx* c, int64_t k_offset, int64_t k_count, float* d, float* r,
                        int flags) {
  CHK(check_ctx(c));
  CHK(ensure_packed(c));
  if (c->n_chain != 1) return fail(PYIID_E_INVALID, "pyiid_pair_geometry takes a single structure");
  const int64_t kmax = c->n * (c->n - 1) / 2;
  if (k_offset < 0 || k_count < 0 || k_offset + k_count > kmax)
    return fail(PYIID_E_INVALID, "pair range [%lld, %lld) outside [0, %lld)", (long long)k_offset,
                (long long)(k_offset + k_count), (long long)kmax);
  if (k_count == 0) return PYIID_OK;
  float* dd = d;
  float* rr = r;
  if (!(flags & PYIID_DEVICE)) {
    CHK(c->io_c.reserve(sizeof(float) * 4 * (size_t)k_count));
    dd = d ? c->io_c.as<float>() : nullptr;
    rr = r ? c->io_c.as<float>() + 3 * k_count : nullptr;
  }
  Timed tm(c, 3);
  HIPCHK(launch_pair_geometry(c->stream, c->pos.as<AtomRec>(), k_offset, k_count, dd, rr));
  CHK(tm.done());
  if (!(flags & PYIID_DEVICE)) {
    if (d) CHK(download(c, d, dd, sizeof(float) * 3 * k_count, flags));
    if (r) CHK(download(c, r, rr, sizeof(float) * k_count, flags));
  }
  return finish(c, flags);
}

int pyiid_normalization(pyiid_ctx* c, int slot, double* na, int flags) {
  CHK(check_ctx(c));
  if (!c->pos_set) return fail(PYIID_E_STATE, "positions not set");
  CHK(check_table(c, slot));
  if (!na) return fail(PYIID_E_INVALID, "null output");
  CHK(ensure_norm(c, slot));
  CHK(download(c, na, c->tab[slot].na.p, sizeof(double) * c->tab[slot].nq * c->n_chain, flags));
  return finish(c, flags);
}

int pyiid_set_transform(pyiid_ctx* c, double rstep, double qstep, const double* rgrid,
                        int32_t n_r, double qmin, int32_t n_q) {
  CHK(check_ctx(c));
  if (!(rstep > 0.0) || !(qstep > 0.0) || n_r < 1 || n_q < 1 || !rgrid || !(qmin >= 0.0))
    return fail(PYIID_E_INVALID, "invalid transform parameters");
  // get_pdf_at_qmin: pad F to nfromdr = ceil(pi/rstep/qstep) (master_kernel.py:65-70),
  // npad2 = 2 * 2^ceil(log2(round(qmin/qstep) + len)) (master_kernel.py:151-155),
  // drpad = pi/(npad2 qstep) (:74), r -> x = r/(2 drpad), linear interpolation (:77-84),
  // G = 2 * lerp, gpad[j] = (qstep/pi) sum_m F_m sin(2 pi j m / npad2).
  const int64_t nfromdr = static_cast<int64_t>(std::ceil(M_PI / rstep / qstep));
  const int64_t length = std::max<int64_t>(n_q, nfromdr);
  const int64_t npad1 = static_cast<int64_t>(std::nearbyint(qmin / qstep)) + length;
  const int64_t npad2 =
      (int64_t(1) << static_cast<int>(std::ceil(std::log(static_cast<double>(npad1)) / std::log(2.0)))) * 2;
  const double drpad = M_PI / (static_cast<double>(npad2) * qstep);
  const int q_zero = static_cast<int>(std::ceil(qmin / qstep));
  std::vector<float> tm(static_cast<size_t>(n_r) * n_q);
  const double two_pi_over = 2.0 * M_PI / static_cast<double>(npad2);
  for (int t = 0; t < n_r; ++t) {
    const double x = rgrid[t] / drpad / 2.0;
    if (!(x >= 0.0) || x + 1.0 >= static_cast<double>(npad2))
      return fail(PYIID_E_INVALID, "r grid value %g outside the transform grid", rgrid[t]);
    const int64_t lo = static_cast<int64_t>(x);
    const double w = x - static_cast<double>(lo);
    for (int m = 0; m < n_q; ++m) {
      double v = 0.0;
      if (m >= q_zero) {
        const double a = static_cast<double>((lo * m) % npad2) * two_pi_over;
        const double b = static_cast<double>(((lo + 1) * m) % npad2) * two_pi_over;
        v = (2.0 * qstep / M_PI) * ((1.0 - w) * std::sin(a) + w * std::sin(b));
      }
      tm[(size_t)t * n_q + m] = static_cast<float>(v);
    }
  }
  CHK(c->tmat.reserve(sizeof(float) * tm.size()));
  HIPCHK(hipMemcpy(c->tmat.p, tm.data(), sizeof(float) * tm.size(), hipMemcpyHostToDevice));
  c->t_nr = n_r;
  c->t_nq = n_q;
  c->t_set = true;
  c->hmc.set = false;
  destroy_graph(c->hmc);
  return PYIID_OK;
}

int pyiid_pdf(pyiid_ctx* c, const float* fq, double* gr, int flags) {
  CHK(check_ctx(c));
  if (!c->t_set) return fail(PYIID_E_STATE, "transform not set");
  if (!fq || !gr) return fail(PYIID_E_INVALID, "null pointer");
  const float* src = fq;
  if (!(flags & PYIID_DEVICE)) {
    CHK(c->io_a.reserve(sizeof(float) * c->t_nq));
    CHK(upload(c, c->io_a.p, fq, sizeof(float) * c->t_nq, flags));
    src = c->io_a.as<float>();
  }
  double* dst = gr;
  if (!(flags & PYIID_DEVICE)) {
    CHK(c->io_c.reserve(sizeof(double) * c->t_nr));
    dst = c->io_c.as<double>();
  }
  CHK(gemm_t(c, src, false, 1, dst));
  if (!(flags & PYIID_DEVICE)) CHK(download(c, gr, dst, sizeof(double) * c->t_nr, flags));
  return finish(c, flags);
}

int pyiid_grad_pdf(pyiid_ctx* c, const void* grad_fq, int64_t rows, double* grad_gr, int flags) {
  CHK(check_ctx(c));
  if (!c->t_set) return fail(PYIID_E_STATE, "transform not set");
  if (rows < 0 || (rows > 0 && (!grad_fq || !grad_gr))) return fail(PYIID_E_INVALID, "bad arguments");
  if (rows == 0) return PYIID_OK;
  const bool a64 = (flags & PYIID_F64) != 0;
  const size_t in_bytes = (a64 ? sizeof(double) : sizeof(float)) * (size_t)rows * c->t_nq;
  const size_t out_bytes = sizeof(double) * (size_t)rows * c->t_nr;
  const void* src = grad_fq;
  double* dst = grad_gr;
  if (!(flags & PYIID_DEVICE)) {
    CHK(c->io_a.reserve(in_bytes));
    CHK(upload(c, c->io_a.p, grad_fq, in_bytes, flags));
    src = c->io_a.p;
    CHK(c->io_c.reserve(out_bytes));
    dst = c->io_c.as<double>();
  }
  CHK(gemm_t(c, src, a64, rows, dst));
  if (!(flags & PYIID_DEVICE)) CHK(download(c, grad_gr, dst, out_bytes, flags));
  return finish(c, flags);
}

int pyiid_potential(pyiid_ctx* c, int potential, const double* gobs, const double* gcalc,
                    int32_t n, double* value, double* scale, int flags) {
  CHK(check_ctx(c));
  if (potential != PYIID_POT_RW && potential != PYIID_POT_CHI_SQ)
    return fail(PYIID_E_INVALID, "unknown potential %d", potential);
  if (n < 1 || !gobs || !gcalc) return fail(PYIID_E_INVALID, "bad arguments");
  CHK(c->io_a.reserve(sizeof(double) * n));
  CHK(c->io_b.reserve(sizeof(double) * n));
  CHK(c->w.reserve(sizeof(double) * n));
  CHK(c->out5.reserve(sizeof(double) * 5));
  CHK(upload(c, c->io_a.p, gcalc, sizeof(double) * n, flags));
  CHK(upload(c, c->io_b.p, gobs, sizeof(double) * n, flags));
  HIPCHK(launch_potential(c->stream, potential, c->io_b.as<double>(), c->io_a.as<double>(), n, 1,
                          c->out5.as<double>(), c->w.as<double>()));
  double out5[5];
  HIPCHK(hipMemcpyAsync(out5, c->out5.p, sizeof(out5), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (value) *value = out5[0];
  if (scale) *scale = out5[1];
  return PYIID_OK;
}

int pyiid_potential_grad(pyiid_ctx* c, int potential, const double* grad_gcalc, int64_t rows,
                         const double* gobs, const double* gcalc, int32_t n, double* grad,
                         int flags) {
  CHK(check_ctx(c));
  if (rows < 0 || (rows > 0 && (!grad_gcalc || !grad)))
    return fail(PYIID_E_INVALID, "bad arguments");
  CHK(pyiid_potential(c, potential, gobs, gcalc, n, nullptr, nullptr, flags));
  if (rows == 0) return PYIID_OK;
  const double* x = grad_gcalc;
  double* y = grad;
  if (!(flags & PYIID_DEVICE)) {
    CHK(c->io_c.reserve(sizeof(double) * (size_t)rows * n));
    CHK(upload(c, c->io_c.p, grad_gcalc, sizeof(double) * (size_t)rows * n, flags));
    x = c->io_c.as<double>();
    CHK(c->io_a.reserve(sizeof(double) * rows));
    y = c->io_a.as<double>();
  }
  HIPCHK(launch_gemv_rows(c->stream, x, rows, n, c->w.as<double>(), y));
  if (!(flags & PYIID_DEVICE)) CHK(download(c, grad, y, sizeof(double) * rows, flags));
  return finish(c, flags);
}

int pyiid_pdf_energy_forces(pyiid_ctx* c, int slot, double qbin, const double* gobs, int potential,
                            double conv, double* energy, double* forces, int flags) {
  CHK(check_ctx(c));
  CHK(check_pdf_setup(c, slot, potential));
  if (!gobs) return fail(PYIID_E_INVALID, "null target");
  const int B = c->n_chain, nr = c->t_nr;
  CHK(c->gobs.reserve(sizeof(double) * (size_t)nr * B));
  CHK(upload(c, c->gobs.p, gobs, sizeof(double) * (size_t)nr * B, flags));
  double* fdst = nullptr;
  if (forces) {
    if (flags & PYIID_DEVICE) {
      fdst = forces;
    } else {
      CHK(c->io_c.reserve(sizeof(double) * 3 * std::max<int64_t>(1, c->n)));
      fdst = c->io_c.as<double>();
    }
  }
  CHK(energy_forces_device(c, slot, qbin, potential, c->gobs.as<double>(), conv, fdst));
  if (energy) {
    std::vector<double> o5(5 * (size_t)B);
    HIPCHK(hipMemcpyAsync(o5.data(), c->out5.p, sizeof(double) * 5 * B, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int b = 0; b < B; ++b) energy[b] = o5[5 * (size_t)b] * conv;
  }
  if (forces && !(flags & PYIID_DEVICE))
    CHK(download(c, forces, fdst, sizeof(double) * 3 * c->n, flags));
  return finish(c, flags);
}

int pyiid_fq_energy_forces(pyiid_ctx* c, int slot, double qbin, int32_t q_first,
                           const double* gobs, int potential, double conv, double* energy,
                           double* forces, int flags) {
  CHK(check_ctx(c));
  if (!c->pos_set) return fail(PYIID_E_STATE, "positions not set");
  CHK(check_table(c, slot));
  if (c->n_chain != 1) return fail(PYIID_E_INVALID, "pyiid_fq_energy_forces takes a single structure");
  if (potential != PYIID_POT_RW && potential != PYIID_POT_CHI_SQ)
    return fail(PYIID_E_INVALID, "unknown potential %d", potential);
  const int nq = c->tab[slot].nq;
  if (q_first < 0 || q_first >= nq) return fail(PYIID_E_INVALID, "q_first %d outside [0, %d)", q_first, nq);
  if (!gobs) return fail(PYIID_E_INVALID, "null target");
  const int ng = nq - q_first;
  CHK(compute_fq(c, slot, qbin));
  CHK(c->gobs.reserve(sizeof(double) * ng));
  CHK(upload(c, c->gobs.p, gobs, sizeof(double) * ng, flags));
  CHK(c->out5.reserve(sizeof(double) * 5));
  CHK(c->w.reserve(sizeof(double) * ng));
  HIPCHK(launch_potential(c->stream, potential, c->gobs.as<double>(),
                          c->fq64.as<double>() + q_first, ng, 1, c->out5.as<double>(),
                          c->w.as<double>()));
  double out5[5];
  HIPCHK(hipMemcpyAsync(out5, c->out5.p, sizeof(out5), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (energy) *energy = out5[0] * conv;
  if (!forces) return PYIID_OK;
  if (c->n < 2) {
    if (flags & PYIID_DEVICE) HIPCHK(hipMemsetAsync(forces, 0, sizeof(double) * 3 * c->n, c->stream));
    else std::memset(forces, 0, sizeof(double) * 3 * c->n);
    return finish(c, flags);
  }
  CHK(c->u.reserve(sizeof(float) * nq));
  HIPCHK(launch_adjoint_fq(c->stream, nq, q_first, c->w.as<double>(), c->tab[slot].gs.as<double>(),
                           c->u.as<float>()));
  CHK(c->io_c.reserve(sizeof(double) * 3 * c->n));
  double* fdst = (flags & PYIID_DEVICE) ? forces : c->io_c.as<double>();
  CHK(contract_forces(c, slot, qbin, conv, fdst));
  if (!(flags & PYIID_DEVICE)) CHK(download(c, forces, fdst, sizeof(double) * 3 * c->n, flags));
  return finish(c, flags);
}

int pyiid_hmc_setup(pyiid_ctx* c, int slot, double qbin, const double* gobs, int potential,
                    double conv, const double* masses, int flags) {
  CHK(check_ctx(c));
  CHK(check_pdf_setup(c, slot, potential));
  if (!gobs || !masses) return fail(PYIID_E_INVALID, "null target or masses");
  if (min_chain_atoms(c) < 2) return fail(PYIID_E_INVALID, "every chain needs at least two atoms");
  Hmc& h = c->hmc;
  destroy_graph(h);
  const int B = c->n_chain, nr = c->t_nr;
  const int64_t n = c->n;
  CHK(h.mass.reserve(sizeof(double) * n));
  CHK(h.mom.reserve(sizeof(double) * 3 * n));
  CHK(h.fcur.reserve(sizeof(double) * 3 * n));
  CHK(h.fnew.reserve(sizeof(double) * 3 * n));
  CHK(h.pos_save.reserve(sizeof(double) * 3 * n));
  CHK(h.kin0.reserve(sizeof(double) * B));
  CHK(h.ecur.reserve(sizeof(double) * B));
  CHK(h.counter.reserve(sizeof(int64_t)));
  CHK(h.chains.reserve(sizeof(HmcChain) * B));
  CHK(h.gobs.reserve(sizeof(double) * (size_t)nr * B));
  CHK(upload(c, h.mass.p, masses, sizeof(double) * n, flags));
  CHK(upload(c, h.gobs.p, gobs, sizeof(double) * (size_t)nr * B, flags));
  std::vector<HmcChain> ch(B);
  for (int b = 0; b < B; ++b)
    ch[b] = HmcChain{static_cast<int>(c->off[b]), static_cast<int>(c->off[b + 1])};
  HIPCHK(hipMemcpyAsync(h.chains.p, ch.data(), sizeof(HmcChain) * B, hipMemcpyHostToDevice,
                        c->stream));
  h.slot = slot;
  h.qbin = qbin;
  h.potential = potential;
  h.conv = conv;
  // initial energy and exact forces of the current state
  CHK(energy_forces_device(c, slot, qbin, potential, h.gobs.as<double>(), 2.0 * conv,
                           h.fcur.as<double>()));
  HIPCHK(launch_energy_from_out5(c->stream, c->out5.as<double>(), B, conv, h.ecur.as<double>()));
  HIPCHK(hipStreamSynchronize(c->stream));  // `ch` is a host temporary
  h.set = true;
  c->sur.eprev_set = false;  // a new energy function: no energy difference to the last fit
  h.fcur_stale = false;
  c->sur.stamp = ~0ull;  // a new target: the surrogate fit starts again
  return PYIID_OK;
}

int pyiid_hmc_run(pyiid_ctx* c, int n_proposals, int n_leapfrog, double step, double beta,
                  uint64_t seed, int use_graph, double* energies, int32_t* accepted,
                  double* times_ms, int flags) {
  CHK(check_ctx(c));
  Hmc& h = c->hmc;
  if (!h.set) return fail(PYIID_E_STATE, "call pyiid_hmc_setup first");
  if (n_proposals < 0 || n_leapfrog < 1 || !(beta > 0.0))
    return fail(PYIID_E_INVALID, "need n_proposals >= 0, n_leapfrog >= 1, beta > 0");
  const int B = c->n_chain;
  if (n_proposals == 0) return PYIID_OK;
  CHK(prepare_history(c, n_proposals));
  CHK(refresh_exact(c));
  RunClock clock(c, times_ms != nullptr);
  for (int p = 0; p < n_proposals; ++p) {
    CHK(run_proposal(c, false, use_graph != 0, n_leapfrog, step, beta, seed));
    CHK(clock.mark());
  }
  CHK(clock.collect(times_ms));
  if (energies)
    CHK(download(c, energies, h.ehist.p, sizeof(double) * (size_t)n_proposals * B, flags));
  if (accepted)
    CHK(download(c, accepted, h.acchist.p, sizeof(int32_t) * (size_t)n_proposals * B, flags));
  return finish(c, flags);
}

int pyiid_sur_setup(pyiid_ctx* c, const float* params, const float* w1, const float* b1,
                    const float* w2, const float* b2, const float* d_mean,
                    const float* d_inv_scale, int32_t max_nbr) {
  CHK(check_ctx(c));
  if (!params || !w1 || !b1 || !w2 || !b2 || !d_mean || !d_inv_scale)
    return fail(PYIID_E_INVALID, "null network parameter");
  const float rc = params[0], skin = params[1], mu0 = params[2], dmu = params[3],
              eta = params[4];
  if (!(rc > 0.f) || !(skin > 0.f) || !(dmu > 0.f) || !(eta > 0.f) || !std::isfinite(mu0))
    return fail(PYIID_E_INVALID, "need rc, skin, dmu, eta > 0");
  if (max_nbr < 1 || max_nbr > 4096) return fail(PYIID_E_INVALID, "max_nbr %d not in [1, 4096]", max_nbr);
  Sur& s = c->sur;
  destroy_graph(c->hmc);
  std::vector<unsigned char> blob(kSurNetBytes);
  unsigned char* p = blob.data();
  for (int e = 0; e < 16 * 64; ++e, p += 2) {
    const uint16_t v = host_bf16(w1[e]);
    std::memcpy(p, &v, 2);
  }
  for (int e = 0; e < 64 * 64; ++e, p += 2) {
    const uint16_t v = host_bf16(w2[e]);
    std::memcpy(p, &v, 2);
  }
  std::memcpy(p, b1, 64 * 4), p += 256;
  std::memcpy(p, b2, 64 * 4), p += 256;
  std::memcpy(p, d_mean, 16 * 4), p += 64;
  std::memcpy(p, d_inv_scale, 16 * 4), p += 64;
  CHK(s.net.reserve(kSurNetBytes));
  HIPCHK(hipMemcpy(s.net.p, blob.data(), kSurNetBytes, hipMemcpyHostToDevice));
  std::vector<uint16_t> onehot(kSurHidden * kSurHidden, 0);
  for (int k = 0; k < kSurHidden; ++k) onehot[k * kSurHidden + k] = host_bf16(1.0f);
  CHK(s.onehot.reserve(sizeof(uint16_t) * onehot.size()));
  HIPCHK(hipMemcpy(s.onehot.p, onehot.data(), sizeof(uint16_t) * onehot.size(),
                   hipMemcpyHostToDevice));
  s.rbf[0] = rc;
  s.rbf[1] = mu0;
  s.rbf[2] = dmu;
  s.rbf[3] = eta;
  s.skin = skin;
  s.max_nbr = max_nbr;
  s.stamp = ~0ull;  // buffers are sized with max_nbr
  s.set = true;
  return PYIID_OK;
}

static int sur_ready(pyiid_ctx* c) {
  CHK(check_ctx(c));
  if (!c->sur.set) return fail(PYIID_E_STATE, "call pyiid_sur_setup first");
  if (!c->pos_set) return fail(PYIID_E_STATE, "positions not set");
  CHK(ensure_packed(c));
  return ensure_sur_layout(c);
}

int pyiid_sur_set_output(pyiid_ctx* c, const float* w3, int flags) {
  CHK(sur_ready(c));
  if (!w3) return fail(PYIID_E_INVALID, "null w3");
  std::vector<float> w(static_cast<size_t>(kSurHidden) * c->n_chain);
  if (flags & PYIID_DEVICE) {
    HIPCHK(hipMemcpy(w.data(), w3, sizeof(float) * w.size(), hipMemcpyDeviceToHost));
  } else {
    std::memcpy(w.data(), w3, sizeof(float) * w.size());
  }
  std::vector<uint16_t> h(w.size());
  for (size_t e = 0; e < w.size(); ++e) h[e] = host_bf16(w[e]);
  HIPCHK(hipMemcpy(c->sur.w3.p, h.data(), sizeof(uint16_t) * h.size(), hipMemcpyHostToDevice));
  c->sur.w3_set = true;
  return PYIID_OK;
}

int pyiid_sur_descriptors(pyiid_ctx* c, float* d, int flags) {
  CHK(sur_ready(c));
  if (!d) return fail(PYIID_E_INVALID, "null output");
  Sur& s = c->sur;
  CHK(sur_list(c));
  CHK(sur_check_lists(c));
  HIPCHK(launch_descriptors(c->stream, c->pos.as<AtomRec>(), static_cast<int>(c->n),
                            s.nbr.as<int>(), s.cnt.as<int>(), s.rbf, s.d.as<float>()));
  if (flags & PYIID_DEVICE) {
    HIPCHK(hipMemcpyAsync(d, s.d.p, sizeof(float) * kSurRbf * c->n, hipMemcpyDeviceToDevice,
                          c->stream));
  } else {
    CHK(download(c, d, s.d.p, sizeof(float) * kSurRbf * c->n, flags));
  }
  return finish(c, flags);
}

int pyiid_sur_forces(pyiid_ctx* c, double* forces, float* atom_energy, int flags) {
  CHK(sur_ready(c));
  Sur& s = c->sur;
  if (!s.w3_set) return fail(PYIID_E_STATE, "surrogate output layer not set (train or set it)");
  CHK(sur_list(c));
  CHK(sur_check_lists(c));
  CHK(sur_eval(c, s.w3.as<uint16_t>(), kSurHidden, s.fsur.as<double>(), nullptr,
               s.eatom.as<float>()));
  if (forces) CHK(download(c, forces, s.fsur.p, sizeof(double) * 3 * c->n, flags));
  if (atom_energy) CHK(download(c, atom_energy, s.eatom.p, sizeof(float) * c->n, flags));
  return finish(c, flags);
}

int pyiid_sur_features(pyiid_ctx* c, float* phi, int flags) {
  CHK(sur_ready(c));
  if (!phi) return fail(PYIID_E_INVALID, "null output");
  CHK(sur_features(c));
  CHK(download(c, phi, c->sur.phi.p, sizeof(float) * kSurHidden * 3 * c->n, flags));
  return finish(c, flags);
}

int pyiid_sur_feature_energies(pyiid_ctx* c, double* e, int flags) {
  CHK(sur_ready(c));
  if (!e) return fail(PYIID_E_INVALID, "null output");
  CHK(sur_features(c));
  CHK(download(c, e, c->sur.efe.p, sizeof(double) * kSurHidden * c->n_chain, flags));
  return finish(c, flags);
}

int pyiid_sur_get_output(pyiid_ctx* c, float* w3, int flags) {
  CHK(sur_ready(c));
  if (!w3) return fail(PYIID_E_INVALID, "null output");
  if (!c->sur.w3_set) return fail(PYIID_E_STATE, "no output layer has been fitted or set");
  std::vector<uint16_t> h(static_cast<size_t>(kSurHidden) * c->n_chain);
  HIPCHK(hipMemcpy(h.data(), c->sur.w3.p, sizeof(uint16_t) * h.size(), hipMemcpyDeviceToHost));
  std::vector<float> w(h.size());
  for (size_t e = 0; e < h.size(); ++e) {
    const uint32_t u = static_cast<uint32_t>(h[e]) << 16;
    std::memcpy(&w[e], &u, 4);
  }
  if (flags & PYIID_DEVICE)
    HIPCHK(hipMemcpy(w3, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
  else
    std::memcpy(w3, w.data(), sizeof(float) * w.size());
  return PYIID_OK;
}

int pyiid_sur_set_energy_weight(pyiid_ctx* c, double weight) {
  CHK(check_ctx(c));
  if (!(weight >= 0.0) || !std::isfinite(weight))
    return fail(PYIID_E_INVALID, "energy weight must be finite and >= 0");
  c->sur.energy_weight = weight;
  return PYIID_OK;
}

int pyiid_sur_train(pyiid_ctx* c, double forget, double ridge, double* r2, int flags) {
  CHK(sur_ready(c));
  if (!c->hmc.set) return fail(PYIID_E_STATE, "call pyiid_hmc_setup first (targets)");
  if (!(forget >= 0.0 && forget <= 1.0) || !(ridge >= 0.0))
    return fail(PYIID_E_INVALID, "need 0 <= forget <= 1 and ridge >= 0");
  CHK(sur_train_step(c, forget, ridge));
  if (r2) std::memcpy(r2, c->sur.r2.data(), sizeof(double) * c->n_chain);
  (void)flags;
  return PYIID_OK;
}

int pyiid_sur_hmc_run(pyiid_ctx* c, int n_proposals, int n_leapfrog, double step, double beta,
                      uint64_t seed, int use_graph, int train_every, double forget, double ridge,
                      double min_r2, double min_acceptance, double* energies, int32_t* accepted,
                      int32_t* exact_mode, double* times_ms, int32_t* violations, int flags) {
  CHK(sur_ready(c));
  Hmc& h = c->hmc;
  Sur& s = c->sur;
  if (!h.set) return fail(PYIID_E_STATE, "call pyiid_hmc_setup first");
  if (n_proposals < 0 || n_leapfrog < 1 || !(beta > 0.0) || train_every < 0)
    return fail(PYIID_E_INVALID,
                "need n_proposals >= 0, n_leapfrog >= 1, beta > 0, train_every >= 0");
  if (!(forget >= 0.0 && forget <= 1.0) || !(ridge >= 0.0))
    return fail(PYIID_E_INVALID, "need 0 <= forget <= 1 and ridge >= 0");
  if (min_chain_atoms(c) < 2) return fail(PYIID_E_INVALID, "every chain needs at least two atoms");
  const int B = c->n_chain;
  if (n_proposals == 0) return PYIID_OK;
  CHK(prepare_history(c, n_proposals));
  RunClock clock(c, times_ms != nullptr);
  bool exact_block = false;
  int block_start = 0;
  std::vector<int32_t> acc_h;
  // true when proposals [block_start, p) were accepted less often than min_acceptance
  auto low_acceptance = [&](int p, bool* low) -> int {
    acc_h.resize(static_cast<size_t>(p - block_start) * B);
    HIPCHK(hipMemcpyAsync(acc_h.data(), h.acchist.as<int32_t>() + (size_t)block_start * B,
                          sizeof(int32_t) * acc_h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int64_t n_acc = 0;
    for (int32_t a : acc_h) n_acc += a;
    *low = static_cast<double>(n_acc) < min_acceptance * static_cast<double>(acc_h.size());
    return PYIID_OK;
  };
  for (int p = 0; p < n_proposals; ++p) {
    if (!s.w3_set || (train_every > 0 && p % train_every == 0)) {
      // acceptance of the learned-force block that just ended
      bool low_acc = false;
      if (min_acceptance > 0.0 && p > block_start && !exact_block) CHK(low_acceptance(p, &low_acc));
      CHK(sur_train_step(c, forget, ridge));
      double mean = 0.0;
      for (int b = 0; b < B; ++b) mean += s.r2[b];
      // poor fit, or learned proposals mostly rejected: exact forces until the next fit
      exact_block = mean / B < min_r2 || low_acc;
      block_start = p;
    } else if (min_acceptance > 0.0 && !exact_block && train_every > kSurAcceptProbe &&
               p - block_start == kSurAcceptProbe) {
      // the block's first learned proposals already fall short: exact forces for the rest of it
      CHK(low_acceptance(p, &exact_block));
    }
    if (exact_block) CHK(refresh_exact(c));
    CHK(run_proposal(c, !exact_block, use_graph != 0, n_leapfrog, step, beta, seed));
    if (exact_mode) exact_mode[p] = exact_block ? 1 : 0;
    CHK(clock.mark());
  }
  CHK(clock.collect(times_ms));
  if (energies)
    CHK(download(c, energies, h.ehist.p, sizeof(double) * (size_t)n_proposals * B, flags));
  if (accepted)
    CHK(download(c, accepted, h.acchist.p, sizeof(int32_t) * (size_t)n_proposals * B, flags));
  if (violations) CHK(download(c, violations, s.viol_count.p, sizeof(int32_t) * B, flags));
  return finish(c, flags);
}

}  // extern "C"
