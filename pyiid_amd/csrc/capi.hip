// C ABI of the gfx950 scattering backend (declared in include/pyiid_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <utility>
#include <vector>

#include "../../include/pyiid_hip.h"
#include "internal.h"

using namespace pyiid;

namespace {

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

#define CHK(expr)              \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != PYIID_OK) return rc_; \
  } while (0)

// Device buffer that only grows.
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  int reserve(size_t want) {
    if (want <= bytes) return PYIID_OK;
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
  // derived normalisation (recomputed when the table or the composition changes)
  Buf na, fsu, fso, gs;  // double [nq]
  uint64_t norm_stamp = ~0ull;
};

}  // namespace

struct pyiid_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;

  int64_t n = 0;
  bool pos_set = false;
  Buf posf;   // float [n][3] staging
  Buf elem;   // int32 [n]
  Buf pos;    // AtomRec [n]
  bool pos_dirty = true;
  std::vector<int64_t> counts{0};
  int max_elem = 0;
  Buf counts_d;
  uint64_t comp_stamp = 1;  // bumps when positions count / elements change

  Table tab[2];
  uint64_t tab_stamp[2] = {1, 1};

  // transform
  Buf tmat;  // float [n_r][nq]
  int t_nr = 0, t_nq = 0;
  bool t_set = false;

  // F-only work items
  Buf items;
  int n_items = 0;
  int64_t items_n = -1;

  // workspaces
  Buf gpart, fpart, forcepart, fq64, g64, io_a, io_b, io_c, small, u;

  // timing: event pairs recorded around launches, resolved when queried
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> ev_pool;   // free events
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  double t_ms[4] = {0, 0, 0, 0};
  int64_t t_n[4] = {0, 0, 0, 0};
};

namespace {

int ensure_device(pyiid_ctx* c) {
  HIPCHK(hipSetDevice(c->device));
  return PYIID_OK;
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
  if (t.n_elem > 1 && t.n_elem > kMaxElem)
    return fail(PYIID_E_INVALID, "at most %d elements are supported (got %d)", kMaxElem,
                t.n_elem);
  return PYIID_OK;
}

bool weighted_table(const Table& t) { return t.n_elem > 1; }

int ensure_norm(pyiid_ctx* c, int slot) {
  Table& t = c->tab[slot];
  const uint64_t stamp = c->tab_stamp[slot] * 1000003ull + c->comp_stamp;
  if (t.norm_stamp == stamp) return PYIID_OK;
  const size_t b = sizeof(double) * std::max(1, t.nq);
  CHK(t.na.reserve(b));
  CHK(t.fsu.reserve(b));
  CHK(t.fso.reserve(b));
  CHK(t.gs.reserve(b));
  CHK(c->counts_d.reserve(sizeof(int64_t) * std::max<size_t>(1, t.n_elem)));
  std::vector<int64_t> cnt(t.n_elem, 0);
  for (size_t e = 0; e < c->counts.size() && e < cnt.size(); ++e) cnt[e] = c->counts[e];
  HIPCHK(hipMemcpyAsync(c->counts_d.p, cnt.data(), sizeof(int64_t) * t.n_elem,
                        hipMemcpyHostToDevice, c->stream));
  HIPCHK(launch_norm(c->stream, t.tab.as<float>(), c->counts_d.as<int64_t>(), t.n_elem, t.nq,
                     c->n, weighted_table(t), t.na.as<double>(), t.fsu.as<double>(),
                     t.fso.as<double>(), t.gs.as<double>()));
  // counts is a host temporary: wait for the copy before it goes out of scope
  HIPCHK(hipStreamSynchronize(c->stream));
  t.norm_stamp = stamp;
  return PYIID_OK;
}

int ensure_items(pyiid_ctx* c) {
  if (c->items_n == c->n) return PYIID_OK;
  const int64_t n = c->n;
  const int n_itile = static_cast<int>((n + 63) / 64);
  int64_t jc = 256 * std::max<int64_t>(1, n / (256 * 16));
  std::vector<FqItem> v;
  for (int it = 0; it < n_itile; ++it) {
    const int64_t jmax = std::min<int64_t>(n, (int64_t)it * 64 + 64);
    for (int64_t jb = 0; jb < jmax; jb += jc)
      v.push_back(FqItem{it, static_cast<int>(jb), static_cast<int>(std::min(jmax, jb + jc)), 0});
  }
  CHK(c->items.reserve(sizeof(FqItem) * std::max<size_t>(1, v.size())));
  HIPCHK(hipMemcpy(c->items.p, v.data(), sizeof(FqItem) * v.size(), hipMemcpyHostToDevice));
  c->n_items = static_cast<int>(v.size());
  c->items_n = n;
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
// synchronisation; the pair is resolved by pyiid_kernel_time).
struct Timed {
  pyiid_ctx* c;
  int kind;
  hipEvent_t e0 = nullptr;
  Timed(pyiid_ctx* c_, int k) : c(c_), kind(k) {
    if (c->timing) {
      e0 = take_event(c);
      if (e0) (void)hipEventRecord(e0, c->stream);
    }
  }
  int done() {
    if (!c->timing || !e0) return PYIID_OK;
    hipEvent_t e1 = take_event(c);
    if (!e1) return fail(PYIID_E_HIP, "hipEventCreate failed");
    HIPCHK(hipEventRecord(e1, c->stream));
    c->ev_pending.push_back({kind, {e0, e1}});
    if (c->ev_pending.size() > 4096) return resolve_timing(c);
    return PYIID_OK;
  }
};

// Number of j splits of the fused kernel: enough waves to fill 256 CUs many
// times over, each split covering at least 64 atoms.
int choose_split(int64_t n, int nq) {
  const int64_t base = ((n + 63) / 64) * ((nq + kGradM - 1) / kGradM);
  int64_t s = (32768 + base - 1) / std::max<int64_t>(1, base);
  s = std::min<int64_t>(s, 32);
  s = std::min<int64_t>(s, std::max<int64_t>(1, n / 64));
  return static_cast<int>(std::max<int64_t>(1, s));
}

// F(Q) (full grid) into c->fq64 (double, nq) using the F-only kernel.
int compute_fq(pyiid_ctx* c, int slot, double qbin) {
  Table& t = c->tab[slot];
  const int nq = t.nq;
  CHK(c->fq64.reserve(sizeof(double) * std::max(1, nq)));
  if (c->n < 2) {
    HIPCHK(hipMemsetAsync(c->fq64.p, 0, sizeof(double) * std::max(1, nq), c->stream));
    return PYIID_OK;
  }
  CHK(ensure_packed(c));
  CHK(ensure_norm(c, slot));
  CHK(ensure_items(c));
  CHK(c->fpart.reserve(sizeof(double) * (size_t)c->n_items * nq));
  const double dq = static_cast<double>(static_cast<float>(qbin));
  Timed tm(c, 1);
  HIPCHK(launch_fq(c->stream, c->pos.as<AtomRec>(), static_cast<int>(c->n), t.tab.as<float>(),
                   t.n_elem, nq, dq, c->items.as<FqItem>(), c->n_items, weighted_table(t),
                   c->fpart.as<double>()));
  CHK(tm.done());
  HIPCHK(launch_reduce_rows(c->stream, c->fpart.as<double>(), c->n_items, nq, t.fsu.as<double>(),
                            c->fq64.as<double>(), nullptr));
  return PYIID_OK;
}

// Fused kernel: gradient partials (or contracted forces when u != null) and F.
int run_fused(pyiid_ctx* c, int slot, double qbin, const float* u, int* n_split_out,
              int* n_part_force) {
  Table& t = c->tab[slot];
  const int nq = t.nq;
  const int64_t n = c->n;
  CHK(ensure_packed(c));
  CHK(ensure_norm(c, slot));
  const int S = choose_split(n, nq);
  const int j_chunk = static_cast<int>((n + S - 1) / S);
  const int n_itile = static_cast<int>((n + 63) / 64);
  const int n_qgroup = (nq + kWaves * kGradM - 1) / (kWaves * kGradM);
  CHK(c->fpart.reserve(sizeof(double) * (size_t)S * n_itile * nq));
  if (u) {
    CHK(c->forcepart.reserve(sizeof(float) * (size_t)S * n_qgroup * kWaves * n * 3));
  } else {
    CHK(c->gpart.reserve(sizeof(float) * (size_t)S * n * 3 * nq));
  }
  const double dq = static_cast<double>(static_cast<float>(qbin));
  Timed tm(c, 0);
  HIPCHK(launch_fused_grad(c->stream, c->pos.as<AtomRec>(), static_cast<int>(n), t.tab.as<float>(),
                           t.n_elem, nq, dq, S, j_chunk, weighted_table(t), c->gpart.as<float>(),
                           c->fpart.as<double>(), u, c->forcepart.as<float>()));
  CHK(tm.done());
  *n_split_out = S;
  if (n_part_force) *n_part_force = S * n_qgroup * kWaves;
  return PYIID_OK;
}

int gemm_t(pyiid_ctx* c, const void* a, bool a64, int64_t rows, double* out) {
  Timed tm(c, 2);
  HIPCHK(launch_gemm_nt(c->stream, a, a64, rows, c->t_nq, c->tmat.as<float>(), c->t_nr, out,
                        nullptr));
  return tm.done();
}

int check_ctx(pyiid_ctx* c) {
  if (!c) return fail(PYIID_E_INVALID, "null context");
  return ensure_device(c);
}

// Energy + forces given gcalc (device, n_g doubles) and a u-builder.
int potential_on_device(pyiid_ctx* c, int potential, const double* gobs_user, int n_g,
                        const double* gcalc_dev, int flags, double* out5_host) {
  CHK(c->io_b.reserve(sizeof(double) * n_g));  // gobs
  CHK(c->io_c.reserve(sizeof(double) * n_g));  // w
  CHK(c->small.reserve(sizeof(double) * 8));
  CHK(upload(c, c->io_b.p, gobs_user, sizeof(double) * n_g, flags));
  HIPCHK(launch_potential(c->stream, potential, c->io_b.as<double>(), gcalc_dev, n_g,
                          c->small.as<double>(), c->io_c.as<double>()));
  HIPCHK(hipMemcpyAsync(out5_host, c->small.p, sizeof(double) * 5, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PYIID_OK;
}

int forces_from_u(pyiid_ctx* c, int slot, double qbin, double conv, double* forces, int flags) {
  int S = 0, n_part = 0;
  CHK(run_fused(c, slot, qbin, c->u.as<float>(), &S, &n_part));
  CHK(c->io_a.reserve(sizeof(double) * c->n * 3));
  double* dst = (flags & PYIID_DEVICE) ? forces : c->io_a.as<double>();
  HIPCHK(launch_reduce_force(c->stream, c->forcepart.as<float>(), n_part, c->n, conv, dst));
  if (!(flags & PYIID_DEVICE)) CHK(download(c, forces, dst, sizeof(double) * c->n * 3, flags));
  return PYIID_OK;
}

}  // namespace

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
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return fail(PYIID_E_HIP, "hipEventCreate failed");
  }
  *out = c;
  return PYIID_OK;
}

int pyiid_destroy(pyiid_ctx* c) {
  if (!c) return PYIID_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (Buf* b : {&c->posf, &c->elem, &c->pos, &c->counts_d, &c->tmat, &c->items, &c->gpart,
                 &c->fpart, &c->forcepart, &c->fq64, &c->g64, &c->io_a, &c->io_b, &c->io_c,
                 &c->small, &c->u})
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
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return PYIID_OK;
}

int pyiid_set_stream(pyiid_ctx* c, void* stream) {
  CHK(check_ctx(c));
  c->stream = stream ? static_cast<hipStream_t>(stream) : c->own_stream;
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
  for (int k = 0; k < 4; ++k) {
    c->t_ms[k] = 0.0;
    c->t_n[k] = 0;
  }
  return PYIID_OK;
}

int pyiid_kernel_time(pyiid_ctx* c, int kind, double* total_ms, int64_t* launches) {
  CHK(check_ctx(c));
  if (kind < 0 || kind > 3) return fail(PYIID_E_INVALID, "kernel kind %d not in [0, 3]", kind);
  CHK(resolve_timing(c));
  if (total_ms) *total_ms = c->t_ms[kind];
  if (launches) *launches = c->t_n[kind];
  return PYIID_OK;
}

int pyiid_set_positions(pyiid_ctx* c, const float* pos, int64_t n, int flags) {
  CHK(check_ctx(c));
  if (n < 0 || n > 2000000000LL) return fail(PYIID_E_INVALID, "atom count %lld out of range", (long long)n);
  if (n > 0 && !pos) return fail(PYIID_E_INVALID, "null positions");
  if (n != c->n) {
    // composition changes with N: reset elements to all-zero
    CHK(c->elem.reserve(sizeof(int32_t) * std::max<int64_t>(1, n)));
    if (n > 0) HIPCHK(hipMemsetAsync(c->elem.p, 0, sizeof(int32_t) * n, c->stream));
    c->counts.assign(1, n);
    c->max_elem = 0;
    c->n = n;
    c->comp_stamp++;
  }
  CHK(c->posf.reserve(sizeof(float) * 3 * std::max<int64_t>(1, n)));
  CHK(c->pos.reserve(sizeof(AtomRec) * std::max<int64_t>(1, n)));
  CHK(upload(c, c->posf.p, pos, sizeof(float) * 3 * n, flags));
  c->pos_set = true;
  c->pos_dirty = true;
  CHK(ensure_packed(c));
  return finish(c, flags);
}

int pyiid_set_elements(pyiid_ctx* c, const int32_t* elem, int64_t n, int flags) {
  CHK(check_ctx(c));
  if (!c->pos_set) return fail(PYIID_E_STATE, "set positions before elements");
  if (n != c->n) return fail(PYIID_E_INVALID, "element count %lld != atom count %lld", (long long)n, (long long)c->n);
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
  std::vector<int64_t> cnt(mx + 1, 0);
  for (int32_t e : h) cnt[e]++;
  c->counts = cnt;
  c->max_elem = mx;
  if (n > 0) HIPCHK(hipMemcpyAsync(c->elem.p, h.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->comp_stamp++;
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
    return fail(PYIID_E_INVALID, "scatter table needs n_elem >= 1, n_q >= 1 (got %d, %d)", n_elem, n_q);
  Table& t = c->tab[slot];
  CHK(t.tab.reserve(sizeof(float) * (size_t)n_elem * n_q));
  CHK(upload(c, t.tab.p, table, sizeof(float) * (size_t)n_elem * n_q, flags));
  t.n_elem = n_elem;
  t.nq = n_q;
  t.set = true;
  c->tab_stamp[slot]++;
  return finish(c, flags);
}

int pyiid_fq(pyiid_ctx* c, int slot, double qbin, float* fq, int flags) {
  CHK(check_ctx(c));
  if (!c->pos_set) return fail(PYIID_E_STATE, "positions not set");
  CHK(check_table(c, slot));
  if (!(qbin > 0.0)) return fail(PYIID_E_INVALID, "qbin must be positive");
  if (!fq) return fail(PYIID_E_INVALID, "null output");
  const int nq = c->tab[slot].nq;
  CHK(compute_fq(c, slot, qbin));
  CHK(c->io_a.reserve(sizeof(float) * nq));
  float* dst = (flags & PYIID_DEVICE) ? fq : c->io_a.as<float>();
  HIPCHK(launch_f64_to_f32(c->stream, c->fq64.as<double>(), nq, dst));
  if (!(flags & PYIID_DEVICE)) CHK(download(c, fq, dst, sizeof(float) * nq, flags));
  return finish(c, flags);
}

int pyiid_grad_fq(pyiid_ctx* c, int slot, double qbin, void* grad, float* fq, int flags) {
  CHK(check_ctx(c));
  if (!c->pos_set) return fail(PYIID_E_STATE, "positions not set");
  CHK(check_table(c, slot));
  if (!(qbin > 0.0)) return fail(PYIID_E_INVALID, "qbin must be positive");
  if (!grad) return fail(PYIID_E_INVALID, "null gradient output");
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
  int S = 0;
  CHK(run_fused(c, slot, qbin, nullptr, &S, nullptr));
  Table& t = c->tab[slot];
  void* gdst = grad;
  if (!(flags & PYIID_DEVICE)) {
    CHK(c->g64.reserve(gbytes));
    gdst = c->g64.p;
  }
  HIPCHK(launch_reduce_grad(c->stream, c->gpart.as<float>(), S, n, nq, t.gs.as<double>(), gdst,
                            (flags & PYIID_F64) != 0));
  if (!(flags & PYIID_DEVICE)) CHK(download(c, grad, gdst, gbytes, flags));
  if (fq) {
    CHK(c->io_a.reserve(sizeof(float) * nq));
    float* fdst = (flags & PYIID_DEVICE) ? fq : c->io_a.as<float>();
    const int n_itile = static_cast<int>((n + 63) / 64);
    HIPCHK(launch_reduce_rows(c->stream, c->fpart.as<double>(), S * n_itile, nq,
                              t.fso.as<double>(), nullptr, fdst));
    if (!(flags & PYIID_DEVICE)) CHK(download(c, fq, fdst, sizeof(float) * nq, flags));
  }
  return finish(c, flags);
}

int pyiid_pair_geometry(pyiid_ctx* c, int64_t k_offset, int64_t k_count, float* d, float* r,
                        int flags) {
  CHK(check_ctx(c));
  CHK(ensure_packed(c));
  const int64_t kmax = c->n * (c->n - 1) / 2;
  if (k_offset < 0 || k_count < 0 || k_offset + k_count > kmax)
    return fail(PYIID_E_INVALID, "pair range [%lld, %lld) outside [0, %lld)", (long long)k_offset,
                (long long)(k_offset + k_count), (long long)kmax);
  if (k_count == 0) return PYIID_OK;
  float* dd = d;
  float* rr = r;
  if (!(flags & PYIID_DEVICE)) {
    CHK(c->g64.reserve(sizeof(float) * 4 * (size_t)k_count));
    dd = d ? c->g64.as<float>() : nullptr;
    rr = r ? c->g64.as<float>() + 3 * k_count : nullptr;
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
  CHK(download(c, na, c->tab[slot].na.p, sizeof(double) * c->tab[slot].nq, flags));
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
  const int64_t npad2 = (int64_t(1) << static_cast<int>(std::ceil(std::log(static_cast<double>(npad1)) / std::log(2.0)))) * 2;
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
    CHK(c->g64.reserve(sizeof(double) * c->t_nr));
    dst = c->g64.as<double>();
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
    CHK(c->g64.reserve(out_bytes));
    dst = c->g64.as<double>();
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
  CHK(upload(c, c->io_a.p, gcalc, sizeof(double) * n, flags));
  double out5[5];
  CHK(potential_on_device(c, potential, gobs, n, c->io_a.as<double>(), flags, out5));
  if (value) *value = out5[0];
  if (scale) *scale = out5[1];
  return PYIID_OK;
}

int pyiid_potential_grad(pyiid_ctx* c, int potential, const double* grad_gcalc, int64_t rows,
                         const double* gobs, const double* gcalc, int32_t n, double* grad,
                         int flags) {
  CHK(check_ctx(c));
  if (potential != PYIID_POT_RW && potential != PYIID_POT_CHI_SQ)
    return fail(PYIID_E_INVALID, "unknown potential %d", potential);
  if (n < 1 || rows < 0 || !gobs || !gcalc || (rows > 0 && (!grad_gcalc || !grad)))
    return fail(PYIID_E_INVALID, "bad arguments");
  CHK(c->io_a.reserve(sizeof(double) * n));
  CHK(upload(c, c->io_a.p, gcalc, sizeof(double) * n, flags));
  double out5[5];
  CHK(potential_on_device(c, potential, gobs, n, c->io_a.as<double>(), flags, out5));
  if (rows == 0) return PYIID_OK;
  const double* x = grad_gcalc;
  double* y = grad;
  if (!(flags & PYIID_DEVICE)) {
    CHK(c->g64.reserve(sizeof(double) * (size_t)rows * n));
    CHK(upload(c, c->g64.p, grad_gcalc, sizeof(double) * (size_t)rows * n, flags));
    x = c->g64.as<double>();
    CHK(c->io_a.reserve(sizeof(double) * rows));
    y = c->io_a.as<double>();
  }
  HIPCHK(launch_gemv_rows(c->stream, x, rows, n, c->io_c.as<double>(), y));
  if (!(flags & PYIID_DEVICE)) CHK(download(c, grad, y, sizeof(double) * rows, flags));
  return finish(c, flags);
}

int pyiid_pdf_energy_forces(pyiid_ctx* c, int slot, double qbin, const double* gobs, int potential,
                            double conv, double* energy, double* forces, int flags) {
  CHK(check_ctx(c));
  if (!c->pos_set) return fail(PYIID_E_STATE, "positions not set");
  CHK(check_table(c, slot));
  if (!c->t_set) return fail(PYIID_E_STATE, "transform not set");
  if (c->tab[slot].nq != c->t_nq)
    return fail(PYIID_E_INVALID, "table has %d Q bins, transform expects %d", c->tab[slot].nq, c->t_nq);
  if (potential != PYIID_POT_RW && potential != PYIID_POT_CHI_SQ)
    return fail(PYIID_E_INVALID, "unknown potential %d", potential);
  if (!gobs) return fail(PYIID_E_INVALID, "null target");
  const int nq = c->t_nq, nr = c->t_nr;
  CHK(compute_fq(c, slot, qbin));
  CHK(c->io_a.reserve(sizeof(float) * nq));
  HIPCHK(launch_f64_to_f32(c->stream, c->fq64.as<double>(), nq, c->io_a.as<float>()));
  CHK(c->g64.reserve(sizeof(double) * nr));
  CHK(gemm_t(c, c->io_a.p, false, 1, c->g64.as<double>()));
  double out5[5];
  CHK(potential_on_device(c, potential, gobs, nr, c->g64.as<double>(), flags, out5));
  if (energy) *energy = out5[0] * conv;
  if (!forces) return PYIID_OK;
  if (c->n < 2) {
    if (flags & PYIID_DEVICE) HIPCHK(hipMemsetAsync(forces, 0, sizeof(double) * 3 * c->n, c->stream));
    else std::memset(forces, 0, sizeof(double) * 3 * c->n);
    return finish(c, flags);
  }
  CHK(ensure_norm(c, slot));
  CHK(c->u.reserve(sizeof(float) * nq));
  HIPCHK(launch_adjoint(c->stream, c->tmat.as<float>(), nr, nq, c->io_c.as<double>(),
                        c->tab[slot].gs.as<double>(), c->u.as<float>()));
  CHK(forces_from_u(c, slot, qbin, conv, forces, flags));
  return finish(c, flags);
}

int pyiid_fq_energy_forces(pyiid_ctx* c, int slot, double qbin, int32_t q_first,
                           const double* gobs, int potential, double conv, double* energy,
                           double* forces, int flags) {
  CHK(check_ctx(c));
  if (!c->pos_set) return fail(PYIID_E_STATE, "positions not set");
  CHK(check_table(c, slot));
  if (potential != PYIID_POT_RW && potential != PYIID_POT_CHI_SQ)
    return fail(PYIID_E_INVALID, "unknown potential %d", potential);
  const int nq = c->tab[slot].nq;
  if (q_first < 0 || q_first >= nq) return fail(PYIID_E_INVALID, "q_first %d outside [0, %d)", q_first, nq);
  if (!gobs) return fail(PYIID_E_INVALID, "null target");
  const int ng = nq - q_first;
  CHK(compute_fq(c, slot, qbin));
  double out5[5];
  CHK(potential_on_device(c, potential, gobs, ng, c->fq64.as<double>() + q_first, flags, out5));
  if (energy) *energy = out5[0] * conv;
  if (!forces) return PYIID_OK;
  if (c->n < 2) {
    if (flags & PYIID_DEVICE) HIPCHK(hipMemsetAsync(forces, 0, sizeof(double) * 3 * c->n, c->stream));
    else std::memset(forces, 0, sizeof(double) * 3 * c->n);
    return finish(c, flags);
  }
  CHK(ensure_norm(c, slot));
  CHK(c->u.reserve(sizeof(float) * nq));
  HIPCHK(launch_adjoint_fq(c->stream, nq, q_first, c->io_c.as<double>(),
                           c->tab[slot].gs.as<double>(), c->u.as<float>()));
  CHK(forces_from_u(c, slot, qbin, conv, forces, flags));
  return finish(c, flags);
}

}  // extern "C"
