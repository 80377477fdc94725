// Host side of the HIP engine: plans, root-of-unity tables, launch sequences
// and the extern "C" ABI declared in include/bluecrunch_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bluecrunch_gpu.h"
#include "launch.h"

using namespace bcg;

namespace {

// ------------------------------------------------------------------ state

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

struct RootsDev {
  cd *lo = nullptr;
  cd *hi = nullptr;
  int sh = 0;
  unsigned long long four_n = 0;
};

struct Pending {
  int kind;
  hipEvent_t a, b;
};

struct Engine {
  bool ready = false;
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t cap = nullptr;                     // graph capture of batches
  // side streams of a batch: one-pass length classes run beside each other
  // (forked from and joined back into the batch's stream by events)
  static constexpr int NAUX = 3;
  hipStream_t aux[NAUX] = {nullptr, nullptr, nullptr};
  hipEvent_t fork_ev = nullptr, join_ev[NAUX] = {nullptr, nullptr, nullptr};
  std::map<unsigned long long, RootsDev> roots;  // keyed by N
  std::map<int, cd *> ltab;                      // e^{2 pi i t/L}, keyed by L
  DevBuf bufA, bufB, wA, wB, wC, blk, margin, zbuf, desc, counters;
  bool prof = false;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> free_events;
  unsigned prof_kinds = ~0u;  // kinds bracketed by events while profiling
  long prof_n[BCG_PROF_KINDS] = {0};
  double prof_ms[BCG_PROF_KINDS] = {0};
};

Engine g;
bcg_stats g_stats;  // see bcg_stats_get
std::recursive_mutex g_lock;

void note_product(unsigned long long points, int bits, double margin) {
  g_stats.products++;
  if (margin > g_stats.worst_margin) g_stats.worst_margin = margin;
  if (points > g_stats.max_points) {
    g_stats.max_points = points;
    g_stats.bits_at_max = bits;
    g_stats.margin_at_max = margin;
  } else if (points == g_stats.max_points && margin > g_stats.margin_at_max) {
    g_stats.margin_at_max = margin;
  }
}
thread_local std::string g_err;

int fail(int code, const char *what, hipError_t e = hipSuccess) {
  char buf[512];
  if (e != hipSuccess)
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  else
    snprintf(buf, sizeof buf, "%s", what);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                      \
  do {                                                 \
    hipError_t e_ = (x);                               \
    if (e_ != hipSuccess) return fail(BCG_E_HIP, #x, e_); \
  } while (0)

void ep_trim(size_t keep = 0);  // idle blocks of the e pipeline's pool

// True while bcg_mul_dev_noalloc runs: the engine's buffers are then slices
// of the static scratch (their sizes are limits), host transfers bypass the
// pinned ring, and nothing is allocated.
bool g_fixed = false;

int ensure(DevBuf &b, size_t bytes) {
  if (b.bytes >= bytes) return 0;
  if (g_fixed) return fail(BCG_E_ARG, "product too large for the allocation-free scratch");
  if (b.p) {
    HIPCHK(hipStreamSynchronize(g.stream));
    HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes < 256 ? 256 : bytes;
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) {  // give the pool's idle blocks back and try once more
    (void)hipGetLastError();
    ep_trim();
    e = hipMalloc(&b.p, want);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(BCG_E_NOMEM, "device allocation failed", e);
  }
  b.bytes = want;
  return 0;
}

// Pinned host staging for the library's own host<->device transfers
// (descriptors, lengths, margins, small operands): a pageable copy goes
// through a driver bounce buffer and can cost milliseconds on a host range the
// driver has not seen; through this ring it is one DMA. Regions are handed out
// in order; when the ring wraps, the device is drained first, so no queued
// transfer still reads a region that is handed out again.
struct HostRing {
  char *p = nullptr;
  size_t bytes = 0, pos = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
};
HostRing g_ring;

int ring_take(size_t n, char **out) {
  n = (n + 255) & ~(size_t)255;
  if (g_ring.pos + n > g_ring.bytes) {
    if (g_ring.p) HIPCHK(hipDeviceSynchronize());
    g_ring.pos = 0;
    if (n > g_ring.bytes) {
      if (g_ring.p) HIPCHK(hipHostFree(g_ring.p));
      g_ring.p = nullptr;
      g_ring.bytes = 0;
      const size_t want = n > ((size_t)16 << 20) ? n : ((size_t)16 << 20);
      HIPCHK(hipHostMalloc((void **)&g_ring.p, want, hipHostMallocDefault));
      g_ring.bytes = want;
    }
  }
  *out = g_ring.p + g_ring.pos;
  g_ring.pos += n;
  return 0;
}

// Host copy of n bytes spread over four threads (page-aligned parts): the
// first-touch page faults of a fresh destination and the copy itself run on
// several cores.
void par_memcpy(void *dst, const void *src, size_t n) {
  constexpr int NT = 4;
  if (n < ((size_t)1 << 20)) {
    memcpy(dst, src, n);
    return;
  }
  const size_t part = (n / NT + 4095) & ~(size_t)4095;
  std::thread th[NT - 1];
  for (int k = 1; k < NT; k++) {
    const size_t a0 = k * part < n ? k * part : n, a1 = (k + 1) * part < n ? (k + 1) * part : n;
    auto piece = [=] { memcpy((char *)dst + a0, (const char *)src + a0, a1 - a0); };
    try {
      th[k - 1] = std::thread(piece);
    } catch (...) {  // no thread to be had: this part on the calling thread
      piece();
    }
  }
  memcpy(dst, src, part < n ? part : n);
  for (auto &t : th)
    if (t.joinable()) t.join();
}

constexpr size_t RING_CHUNK = (size_t)8 << 20;

int ring_events() {
  for (int i = 0; i < 2; i++)
    if (!g_ring.ev[i]) HIPCHK(hipEventCreateWithFlags(&g_ring.ev[i], hipEventDisableTiming));
  return 0;
}

// Host to device, queued on s (the host data is copied before returning).
// Large copies go in 8 MiB chunks through two ring regions: chunk i is copied
// into region i & 1 once the DMA of chunk i - 2 has left it, so the host copy
// of one chunk overlaps the DMA of the one before.
int h2d(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (!bytes) return 0;
  if (g_fixed) {  // from the caller's memory, complete on return
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
  }
  constexpr size_t CH = RING_CHUNK;
  char *st;
  if (bytes <= CH) {
    if (int rc = ring_take(bytes, &st)) return rc;
    memcpy(st, src, bytes);
    HIPCHK(hipMemcpyAsync(dst, st, bytes, hipMemcpyHostToDevice, s));
    return 0;
  }
  if (int rc = ring_events()) return rc;
  if (int rc = ring_take(2 * CH, &st)) return rc;
  const size_t nch = (bytes + CH - 1) / CH;
  for (size_t i = 0; i < nch; i++) {
    const size_t n = bytes - i * CH < CH ? bytes - i * CH : CH;
    if (i >= 2) HIPCHK(hipEventSynchronize(g_ring.ev[i & 1]));
    par_memcpy(st + (i & 1) * CH, (const char *)src + i * CH, n);
    HIPCHK(hipMemcpyAsync((char *)dst + i * CH, st + (i & 1) * CH, n, hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(g_ring.ev[i & 1], s));
  }
  return 0;
}

// Device to host after the work queued on s; returns when dst holds the data.
// Large copies go in 8 MiB chunks through two ring regions, the host copy of
// one chunk overlapping the DMA of the next.
int d2h(void *dst, const void *src, size_t bytes, hipStream_t s) {
  constexpr size_t CH = RING_CHUNK;
  if (g_fixed) {
    if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
  }
  if (bytes <= CH) {
    char *st;
    if (int rc = ring_take(bytes ? bytes : 1, &st)) return rc;
    if (bytes) HIPCHK(hipMemcpyAsync(st, src, bytes, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    memcpy(dst, st, bytes);
    return 0;
  }
  if (int rc = ring_events()) return rc;
  char *st;
  if (int rc = ring_take(2 * CH, &st)) return rc;
  const size_t nch = (bytes + CH - 1) / CH;
  for (size_t i = 0; i <= nch; i++) {
    if (i < nch) {  // chunk i into region i & 1 (chunk i - 2 left it at step i - 1)
      const size_t n = bytes - i * CH < CH ? bytes - i * CH : CH;
      HIPCHK(hipMemcpyAsync(st + (i & 1) * CH, (const char *)src + i * CH, n, hipMemcpyDeviceToHost, s));
      HIPCHK(hipEventRecord(g_ring.ev[i & 1], s));
    }
    if (i >= 1) {  // chunk i - 1 out of its region while chunk i is in flight
      const size_t j = i - 1, n = bytes - j * CH < CH ? bytes - j * CH : CH;
      HIPCHK(hipEventSynchronize(g_ring.ev[j & 1]));
      par_memcpy((char *)dst + j * CH, st + (j & 1) * CH, n);
    }
  }
  return 0;
}

int prepare_noalloc();  // tables and the scratch address of bcg_mul_dev_noalloc

int init_locked(int device) {
  if (g.ready) return 0;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) return fail(BCG_E_NODEV, "no HIP device", e);
  if (device < 0 || device >= n) return fail(BCG_E_ARG, "device index out of range");
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&g.cap, hipStreamNonBlocking));
  for (int i = 0; i < Engine::NAUX; i++) {
    HIPCHK(hipStreamCreateWithFlags(&g.aux[i], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&g.join_ev[i], hipEventDisableTiming));
  }
  HIPCHK(hipEventCreateWithFlags(&g.fork_ev, hipEventDisableTiming));
  g.device = device;
  g.ready = true;
  if (int rc = ensure(g.margin, sizeof(unsigned int) * MARGIN_SLOTS * MARGIN_STRIDE)) return rc;
  // every kernel family's code object now rather than inside a first product
  // (a failure here only defers the load to the first launch)
  for (auto f : {preload_mul1, preload_fwd_pow2, preload_fwd_three, preload_inv_pow2, preload_inv_three, preload_mid,
                 preload_keep})
    if (f()) (void)hipGetLastError();
  return prepare_noalloc();
}

#define ENTER()                                   \
  std::lock_guard<std::recursive_mutex> lk_(g_lock); \
  do {                                            \
    if (int rc_ = init_locked(0)) return rc_;     \
    (void)hipSetDevice(g.device);                 \
  } while (0)

hipStream_t pick(void *s) { return s ? (hipStream_t)s : g.stream; }

// ----------------------------------------------------------------- timing

hipEvent_t get_event() {
  if (!g.free_events.empty()) {
    hipEvent_t e = g.free_events.back();
    g.free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

struct ProfScope {
  int kind;
  hipStream_t s;
  hipEvent_t a = nullptr;
  ProfScope(int k, hipStream_t st) : kind(k), s(st) {
    if (g.prof && ((g.prof_kinds >> k) & 1u)) {
      a = get_event();
      (void)hipEventRecord(a, s);
    }
  }
  ~ProfScope() {
    if (g.prof && a) {
      hipEvent_t b = get_event();
      (void)hipEventRecord(b, s);
      g.pending.push_back({kind, a, b});
    }
  }
};

void prof_collect() {
  for (auto &p : g.pending) {
    (void)hipEventSynchronize(p.b);
    float ms = 0;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      g.prof_n[p.kind]++;
      g.prof_ms[p.kind] += ms;
    }
    g.free_events.push_back(p.a);
    g.free_events.push_back(p.b);
  }
  g.pending.clear();
}

// ----------------------------------------------------------------- tables

// cos/sin of 2 pi num/den in long double (64-bit significand), reduced to the
// first octant so the long double argument stays small and exact-ish.
void root_ld(unsigned long long num, unsigned long long den, double *re, double *im) {
  // e^{2 pi i num / den}; reduce num/den to [0, 1/8] with symmetries.
  num %= den;
  // work with x = num/den in [0,1)
  long double two_pi = 6.283185307179586476925286766559005768L;
  unsigned long long n8 = num * 8;  // compare against den multiples
  int oct = (int)(n8 / den);        // 0..7
  // angle = 2 pi num/den = oct * pi/4 + r, r in [0, pi/4)
  long double r = two_pi * (long double)(n8 - (unsigned long long)oct * den) / (8.0L * (long double)den);
  long double c = cosl(r), s = sinl(r);
  long double h = 0.70710678118654752440084436210484903928L;
  long double cr, ci;
  switch (oct) {
    case 0: cr = c; ci = s; break;
    case 1: cr = h * (c - s); ci = h * (c + s); break;
    case 2: cr = -s; ci = c; break;
    case 3: cr = -h * (c + s); ci = h * (c - s); break;
    case 4: cr = -c; ci = -s; break;
    case 5: cr = -h * (c - s); ci = -h * (c + s); break;
    case 6: cr = s; ci = -c; break;
    default: cr = h * (c + s); ci = -h * (c - s); break;
  }
  *re = (double)cr;
  *im = (double)ci;
}

int upload(cd **dst, const std::vector<double> &host) {
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, host.size() * sizeof(double));
  if (e != hipSuccess) return fail(BCG_E_NOMEM, "table allocation failed", e);
  HIPCHK(hipMemcpy(p, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice));
  *dst = (cd *)p;
  return 0;
}

// Table-free fields of Roots for 4N = four_n: 4N = 8 q8, q8 = 2^qa or 3 * 2^qa.
void roots_direct(unsigned long long four_n, Roots *r) {
  r->four_n = four_n;
  r->q8 = (four_n % 8 == 0) ? four_n / 8 : 0;
  r->q3 = (r->q8 && r->q8 % 3 == 0) ? 1 : 0;
  unsigned long long q = r->q3 ? r->q8 / 3 : r->q8;
  r->qa = 0;
  while (q > 1) {
    q >>= 1;
    r->qa++;
  }
  const long double quarter_pi = 0.785398163397448309615660845819875721L;
  r->step = r->q8 ? (double)(quarter_pi / (long double)r->q8) : 0.0;
}

// lo[u] = Omega^u (u < nlo), hi[u] = Omega^{u 2^sh} (u < nhi), from omega_direct.
__global__ void k_fill_roots(Roots r, cd *lo, unsigned long long nlo, cd *hi, unsigned long long nhi) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nlo)
    lo[i] = omega_direct(r, i);
  else if (i < nlo + nhi)
    hi[i - nlo] = omega_direct(r, (i - nlo) << r.sh);
}

// Root tables for transform length N, built once on the device (N even, so
// omega_direct's octant split applies; the tables agree with correctly rounded
// values within its 1.8e-16).
int get_roots(unsigned long long N, Roots *out) {
  auto it = g.roots.find(N);
  if (it == g.roots.end() && g_fixed) return fail(BCG_E_ARG, "root table not prepared at initialisation");
  if (it == g.roots.end()) {
    RootsDev r;
    r.four_n = 4 * N;
    int lg = 0;
    while ((1ull << lg) < r.four_n) lg++;
    r.sh = (lg + 1) / 2;
    const unsigned long long nlo = 1ull << r.sh;
    const unsigned long long nhi = (r.four_n + nlo - 1) >> r.sh;
    Roots d;
    memset(&d, 0, sizeof d);
    roots_direct(r.four_n, &d);
    d.sh = r.sh;
    if (N % 2 || !d.q8) {  // (no such length class) host tables, long double
      std::vector<double> lo(2 * nlo), hi(2 * nhi);
      for (unsigned long long u = 0; u < nlo; u++) root_ld(u, r.four_n, &lo[2 * u], &lo[2 * u + 1]);
      for (unsigned long long u = 0; u < nhi; u++) root_ld(u << r.sh, r.four_n, &hi[2 * u], &hi[2 * u + 1]);
      if (int rc = upload(&r.lo, lo)) return rc;
      if (int rc = upload(&r.hi, hi)) return rc;
    } else {
      void *plo = nullptr, *phi = nullptr;
      hipError_t e = hipMalloc(&plo, nlo * sizeof(cd));
      if (e == hipSuccess) e = hipMalloc(&phi, nhi * sizeof(cd));
      if (e != hipSuccess) {
        if (plo) (void)hipFree(plo);
        return fail(BCG_E_NOMEM, "table allocation failed", e);
      }
      r.lo = (cd *)plo;
      r.hi = (cd *)phi;
      const unsigned long long n = nlo + nhi;
      hipLaunchKernelGGL(k_fill_roots, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g.stream, d, r.lo, nlo, r.hi,
                         nhi);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(g.stream));  // ready before any stream uses them
    }
    it = g.roots.emplace(N, r).first;
  }
  out->lo = it->second.lo;
  out->hi = it->second.hi;
  out->sh = it->second.sh;
  roots_direct(it->second.four_n, out);
  return 0;
}

int get_ltab(int L, const cd **out) {
  auto it = g.ltab.find(L);
  if (it == g.ltab.end() && g_fixed) return fail(BCG_E_ARG, "pass table not prepared at initialisation");
  if (it == g.ltab.end()) {
    int n = L / 2 > 0 ? L / 2 : 1;
    std::vector<double> t(2 * n);
    for (int i = 0; i < n; i++) root_ld((unsigned long long)i, (unsigned long long)L, &t[2 * i], &t[2 * i + 1]);
    cd *p = nullptr;
    if (int rc = upload(&p, t)) return rc;
    it = g.ltab.emplace(L, p).first;
  }
  *out = it->second;
  return 0;
}

// ------------------------------------------------------------------ plans

bool is_three(unsigned long long n) { return n % 3 == 0; }

int grid_for(unsigned long long n);

// Smallest 2^k or 3*2^k >= n (fft/fft.c:22-41 length classes), at least 64.
unsigned long long class_ceil(unsigned long long n) {
  unsigned long long b = 32;
  for (;;) {
    if (2 * b >= n) return 2 * b;
    if (3 * b >= n) return 3 * b;
    b *= 2;
  }
}

int ilog2u(unsigned long long v) { return v ? 63 - __builtin_clzll(v) : 0; }

// Largest digit size whose worst rounding error stays well below 1/4 at
// transform length N (unsigned digits, right-angle convolution). Lengths of
// the form 2^k and 3 * 2^k (every length class) are memoised: batch planning
// asks for them once per candidate digit size of every product.
int bmax_for(unsigned long long N) {
  static int memo[2][64];
  const int k = ilog2u(N);
  const int t = N == (1ull << k) ? 0 : (k >= 1 && N == (3ull << (k - 1)) ? 1 : -1);
  if (t >= 0 && memo[t][k]) return memo[t][k];
  double lg = std::log2((double)N);
  double b = (50.5 - lg - std::log2(lg)) / 2.0;
  int bi = (int)std::floor(b);
  if (bi > 20) bi = 20;
  if (bi < 4) bi = 4;
  if (t >= 0) memo[t][k] = bi;
  return bi;
}

struct Plan {
  unsigned long long N = 0;
  int bits = 0;
  int P = 0;
  int L[3] = {0, 0, 0};
  int C[3] = {1, 1, 1};
  unsigned long long M[3] = {1, 1, 1};
  unsigned long long dA = 0, dB = 0;
};

unsigned long long digits_of(unsigned long long words, int bits) {
  return (32ull * words + (unsigned long long)bits - 1) / (unsigned long long)bits;
}

int tile_cols(int L, unsigned long long M) {
  int c = 8192 / L;
  if (is_three((unsigned long long)L)) c = 6144 / L;
  if (c > 16) c = 16;
  if (c < 1) c = 1;
  while ((unsigned long long)c > M) c >>= 1;
  return c;
}

bool fill_strides(Plan &p) {
  unsigned long long prod = 1;
  for (int i = 0; i < p.P; i++) prod *= (unsigned long long)p.L[i];
  if (prod != p.N) return false;
  unsigned long long m = p.N;
  for (int i = 0; i < p.P; i++) {
    m /= (unsigned long long)p.L[i];
    p.M[i] = m;
    p.C[i] = (i == p.P - 1) ? 1 : tile_cols(p.L[i], m);
  }
  return true;
}

bool is_pow2(unsigned long long v) { return v && !(v & (v - 1)); }

// Tile width of inverse pass i: a middle pass (i > 0) of 256-point tiles runs
// 32 columns wide (its two-stage split tile still fits two workgroups per CU,
// and 512-byte row runs measured faster than 256-byte ones); every other pass
// uses the plan's width, the forward one included.
int inv_cols(const Plan &p, int i) {
  if (i > 0 && i < p.P - 1 && p.L[i] == 256 && p.C[i] == 16 && p.M[i] % 32 == 0) return 32;
  return p.C[i];
}

// Adds p and, for each strided pass with 16-column tiles, the variant with
// 8-column tiles (half the LDS per workgroup: more workgroups per CU).
void push_with_tiles(std::vector<Plan> &out, const Plan &p) {
  const int strided = p.P - 1;
  for (int mask = 0; mask < (1 << strided); mask++) {
    Plan q = p;
    bool ok = true;
    for (int i = 0; i < strided; i++)
      if (mask & (1 << i)) {
        if (q.C[i] != 16) ok = false;
        q.C[i] = 8;
      }
    if (ok) out.push_back(q);
  }
}

// Every pass split of N the kernels support: the innermost (contiguous) pass
// carries the 3*2^k factor and is <= 8192 (<= 6144); outer passes are powers
// of two in [64, 8192]; at most three passes. Ordered by decreasing inner
// pass, two-pass before three-pass; each split first with its widest tiles.
std::vector<Plan> factor_candidates(unsigned long long N) {
  std::vector<Plan> out;
  const bool three = is_three(N);
  const unsigned long long MAXIN = three ? 6144 : 8192;
  if (N <= MAXIN) {
    Plan p;
    p.N = N;
    p.P = 1;
    p.L[0] = (int)N;
    fill_strides(p);
    out.push_back(p);
    // the largest one-pass class can also run as two passes (64 x 128), which
    // the length-class tuner weighs against the one-workgroup kernel
    if (N < 8192) return out;
  }
  for (unsigned long long inner = MAXIN; inner >= (three ? 192ull : 128ull); inner /= 2) {
    if (N % inner) continue;
    unsigned long long rest = N / inner;
    if (!is_pow2(rest)) continue;
    if (rest >= 64 && rest <= 8192) {
      Plan p;
      p.N = N;
      p.P = 2;
      p.L[0] = (int)rest;
      p.L[1] = (int)inner;
      if (fill_strides(p)) push_with_tiles(out, p);
    }
    for (unsigned long long l1 = 64; l1 <= 8192; l1 *= 2) {
      unsigned long long l2 = rest / l1;
      if (l1 * l2 != rest || l2 < 64 || l2 > 8192) continue;
      Plan p;
      p.N = N;
      p.P = 3;
      p.L[0] = (int)l1;
      p.L[1] = (int)l2;
      p.L[2] = (int)inner;
      if (fill_strides(p)) push_with_tiles(out, p);
    }
  }
  return out;
}

// Tuned choice per transform length: index into factor_candidates(N), for
// single products (bcg_autotune, bcg_autotune_classes) and for the products of
// batched calls (bcg_autotune_classes), which run every pass over a whole
// length class at once and may prefer another split.
std::map<unsigned long long, int> g_tuned;
std::map<unsigned long long, int> g_tuned_batch;

// Untuned choice: a two-pass split whose strided pass still moves >= 8
// columns (128-byte runs) per row; otherwise the three-pass split with the
// largest inner pass <= 4096 and the shortest first pass, so every strided
// pass runs 16 columns wide (narrow strided tiles reach about half the
// bandwidth of 16-column ones on MI355X).
int default_candidate(const std::vector<Plan> &c) {
  for (size_t i = 0; i < c.size(); i++)
    if (c[i].P <= 2 && (c[i].P == 1 || c[i].C[0] >= 8)) return (int)i;
  int best = -1;
  for (size_t i = 0; i < c.size(); i++) {
    if (c[i].P != 3 || c[i].L[2] > 4096) continue;
    if (best < 0 || c[i].L[2] > c[best].L[2] || (c[i].L[2] == c[best].L[2] && c[i].L[0] < c[best].L[0]))
      best = (int)i;
  }
  return best < 0 ? 0 : best;
}

bool factor(unsigned long long N, Plan &p) {
  std::vector<Plan> c = factor_candidates(N);
  if (c.empty()) return false;
  int idx = default_candidate(c);
  auto it = g_tuned.find(N);
  if (it != g_tuned.end() && it->second >= 0 && it->second < (int)c.size()) idx = it->second;
  int bits = p.bits;
  unsigned long long dA = p.dA, dB = p.dB;
  p = c[idx];
  p.bits = bits;
  p.dA = dA;
  p.dB = dB;
  return true;
}

// Split cache: factor_candidates is only built once per transform length.
std::map<unsigned long long, std::vector<Plan>> g_cands;

bool factor_cached(unsigned long long N, Plan &p, bool batched = false) {
  auto it = g_cands.find(N);
  if (it == g_cands.end()) it = g_cands.emplace(N, factor_candidates(N)).first;
  const std::vector<Plan> &c = it->second;
  if (c.empty()) return false;
  int idx = default_candidate(c);
  auto t = g_tuned.find(N);
  if (t != g_tuned.end() && t->second >= 0 && t->second < (int)c.size()) idx = t->second;
  if (batched) {
    auto tb = g_tuned_batch.find(N);
    if (tb != g_tuned_batch.end() && tb->second >= 0 && tb->second < (int)c.size()) idx = tb->second;
  }
  int bits = p.bits;
  unsigned long long dA = p.dA, dB = p.dB;
  p = c[idx];
  p.bits = bits;
  p.dA = dA;
  p.dB = dB;
  return true;
}

// Digit sizes planned above the rounding bound b_max(N) (bcg_plan_digit_bias;
// 0 in normal use).
int g_digit_bias = 0;

int make_plan(unsigned long long AL, unsigned long long BL, int force_bits, Plan &best, bool batched = false) {
  best = Plan();
  unsigned long long bestN = 0;
  int bestb = 0;
  // N grows as b falls, so the first feasible b gives the shortest length;
  // smaller b keeping that length are preferred (more margin)
  for (int b = 20; b >= 4; b--) {
    if (force_bits > 0 && b != force_bits) continue;
    unsigned long long dA = digits_of(AL, b), dB = digits_of(BL, b);
    unsigned long long need = dA + dB - 1;
    unsigned long long N = class_ceil((need + 1) / 2);
    if (bestN && N > bestN) break;
    if (force_bits <= 0 && b > bmax_for(N) + g_digit_bias) continue;
    if (bestN == 0 || N < bestN || (N == bestN && b < bestb)) {
      bestN = N;
      bestb = b;
    }
  }
  if (bestN == 0) return fail(BCG_E_ARG, "no transform plan for this size");
  best.N = bestN;
  best.bits = bestb;
  best.dA = digits_of(AL, bestb);
  best.dB = digits_of(BL, bestb);
  if (!factor_cached(bestN, best, batched)) return fail(BCG_E_ARG, "no pass split for this length");
  return 0;
}

// --------------------------------------------------------------- launching

#define LAUNCH(kern, grid, block, stream, ...) \
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, stream, __VA_ARGS__)

// grid.y > 1 only for batched launches (PassArgs::desc set).
#define LAUNCH2(kern, gx, gy, block, stream, ...) \
  hipLaunchKernelGGL(kern, dim3(gx, gy), dim3(block), 0, stream, __VA_ARGS__)

int coef_bits(const Plan &p) {
  unsigned long long m = p.dA < p.dB ? p.dA : p.dB;
  int cb = 2 * p.bits + ilog2u(m + 1) + 2;
  return cb > 63 ? 63 : cb;
}

// Fill the common pass arguments for pass i of plan p. The packing pass
// (i = 0 of a multi-pass plan) slices digits assuming M b and N b are
// multiples of 32.
int pass_args(const Plan &p, int i, PassArgs &a) {
  if (p.P > 1 && i == 0 && (p.M[0] % 32 || p.N % 32)) return fail(BCG_E_ARG, "packing pass needs 32 | M and 32 | N");
  // the packing pass addresses each half of an operand with 32-bit byte offsets
  if (p.P > 1 && i == 0 && p.N * (unsigned long long)p.bits / 8 >= 0xfffff000ull)
    return fail(BCG_E_ARG, "transform too long for the packing pass");
  a.N = p.N;
  a.M = p.M[i];
  a.S = (unsigned long long)p.L[i] * p.M[i];
  a.tw_mul = 4ull * p.N / a.S;
  if (int rc = get_ltab(p.L[i], &a.tl)) return rc;
  if (int rc = get_roots(p.N, &a.roots)) return rc;
  a.bits = p.bits;
  a.inv_n = 1.0 / (double)p.N;
  a.fast_round = coef_bits(p) <= 51 ? 1 : 0;
  a.ntiles = (int)(p.N / ((unsigned long long)p.L[i] * p.C[i]));
  return 0;
}

// Carry resolve: coefficients -> CL words.
int run_carry(const long long *coef, unsigned long long ncoef, int bits, int cbits, uint32_t *out,
              unsigned long long CL, hipStream_t s, const uint32_t *add = nullptr, unsigned long long addl = 0,
              unsigned cM = 0, unsigned crows = 0) {
  if (CL == 0) return 0;
  const unsigned long long wpb = (unsigned long long)CT * CWPT;
  unsigned long long nblk = (CL + wpb - 1) / wpb;
  if (int rc = ensure(g.blk, (nblk + 15) & ~15ull)) return rc;
  CarryArgs c;
  memset(&c, 0, sizeof c);
  c.coef = coef;
  c.ncoef = ncoef;
  c.bits = bits;
  c.cbits = cbits;
  c.out = out;
  c.CL = CL;
  c.blk_fn = (unsigned char *)g.blk.p;
  c.nblk = nblk;
  c.add = add;
  c.addl = addl;
  c.coef_T = cM ? 1 : 0;
  c.cM = cM;
  c.crows = crows;
  c.wpb = (unsigned)wpb;
  c.lds_slots = carry_slots(bits, cbits);
  {
    ProfScope ps(BCG_K_CARRY, s);
    hipLaunchKernelGGL(k_carry_local, dim3((unsigned)nblk), dim3(CT), c.lds_slots * sizeof(long long), s, c);
    LAUNCH(k_carry_fix, 1, 1024, s, c);
    LAUNCH(k_carry_apply, (unsigned)nblk, 256, s, c);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// Carry resolve of plan p's last pass output (group sums or coefficients in
// bufB) into CL words (+ addend D).
int group_words(const Plan &p);
int group_stride(int K);
bool coef_transposed(const Plan &p);
int carry_of_plan(const Plan &p, uint32_t *out, unsigned long long CL, hipStream_t s, const uint32_t *add,
                  unsigned long long addl) {
  const int K = group_words(p);
  if (!K)
    return run_carry((const long long *)g.bufB.p, 2 * p.N, p.bits, coef_bits(p), out, CL, s, add, addl,
                     coef_transposed(p) ? (unsigned)p.M[0] : 0u, 2u * (unsigned)p.L[0]);
  if (CL == 0) return 0;
  const unsigned long long wpb = (unsigned long long)GCT * GWPT;
  unsigned long long nblk = (CL + wpb - 1) / wpb;
  if (int rc = ensure(g.blk, (nblk + 15) & ~15ull)) return rc;
  CarryArgs c;
  memset(&c, 0, sizeof c);
  c.wpb = (unsigned)wpb;
  c.ncoef = 2 * p.N;
  c.bits = p.bits;
  c.cbits = coef_bits(p);
  c.out = out;
  c.CL = CL;
  c.blk_fn = (unsigned char *)g.blk.p;
  c.nblk = nblk;
  c.add = add;
  c.addl = addl;
  c.ybuf = (const uint32_t *)g.bufB.p;
  c.gC = p.C[0];
  c.ng = (long long)(c.ncoef / (unsigned long long)c.gC);
  c.gK = K;
  c.gKs = group_stride(K);
  // LDS: the groups under one block's GCT * GWPT words
  size_t glds = ((size_t)(32 * GCT * GWPT) / (size_t)(p.C[0] * p.bits) + CARRY_GCOV + 2) * c.gKs;
  glds = (glds + glds / 32 + 4) * 4;  // one pad word per 32
  {
    ProfScope ps(BCG_K_CARRY, s);
    // groups covering one word: ceil(K / (G b / 32)) when group starts are
    // word-aligned, else the bound of group_words
    const int gb = p.C[0] * p.bits;
    const int gcov = gb % 32 == 0 ? (K + gb / 32 - 1) / (gb / 32) : (32 * K + gb - 1) / gb + 1;
    if (gcov <= 2)
      hipLaunchKernelGGL(k_carry_groups<2>, dim3((unsigned)nblk), dim3(GCT), glds, s, c);
    else if (gcov == 3)
      hipLaunchKernelGGL(k_carry_groups<3>, dim3((unsigned)nblk), dim3(GCT), glds, s, c);
    else
      hipLaunchKernelGGL(k_carry_groups<CARRY_GCOV>, dim3((unsigned)nblk), dim3(GCT), glds, s, c);
    LAUNCH(k_carry_fix, 1, 1024, s, c);
    LAUNCH(k_carry_apply, (unsigned)nblk, 256, s, c);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int zero_margin(hipStream_t s) {
  const int n = MARGIN_SLOTS * MARGIN_STRIDE;
  LAUNCH(k_zero_u32, (n + 255) / 256, 256, s, (unsigned int *)g.margin.p, n);
  return 0;
}

// Deferred margin checks (bcg_defer_begin / bcg_defer_end): the products of
// the thread that opened the scope each write their margin slots into a set
// of their own and are not read back or re-run one by one (no host
// synchronisation per product); the sets are read together when they run
// out and at the end of the scope, which returns the worst margin. Products
// whose output aliases an input, and all products of other threads, check
// at once as usual.
constexpr int DEFER_SETS = 64;
constexpr size_t MARGIN_SET_BYTES = sizeof(unsigned int) * MARGIN_SLOTS * MARGIN_STRIDE;
struct DeferScope {
  bool open = false;
  bool in_product = false;  // a deferred product is running: its margin is not read
  std::thread::id owner;
  DevBuf sets;
  int n = 0;  // sets in use
  unsigned long long points[DEFER_SETS];
  int bits[DEFER_SETS];
  double worst = 0.0;
};
DeferScope g_defer;

int read_margin(hipStream_t s, double *m) {
  if (g_defer.in_product) {  // read with the scope's other sets later
    *m = 0.0;
    return 0;
  }
  unsigned int slots[MARGIN_SLOTS * MARGIN_STRIDE];
  if (int rc = d2h(slots, g.margin.p, sizeof slots, s)) return rc;
  unsigned int bits = 0;
  for (int i = 0; i < MARGIN_SLOTS; i++) bits = slots[i * MARGIN_STRIDE] > bits ? slots[i * MARGIN_STRIDE] : bits;
  float f;
  memcpy(&f, &bits, sizeof f);
  *m = (double)f;
  return 0;
}

// Group mode: the last inverse pass of a multi-pass plan with 8- or
// 16-column tiles writes, per row segment of C digits, the exact sum
// sum_i c_i 2^{b i} as K words instead of C 64-bit coefficients (about a
// quarter of the bytes), and the carry reads those sums. Returns K, or 0 when
// the plan keeps plain coefficients; group_stride rounds K up to 4 words.
int group_stride(int K) { return (K + 3) & ~3; }
// Words per exact group sum of G digits of b bits (coefficients below 2^cb),
// 0 when the group-sum carry does not cover that shape.
int group_words_for(int G, int b, int cb) {
  if ((G != 8 && G != 16) || cb > 51) return 0;
  int off_max = 0;
  for (int g = 0; g < 32; g++) {
    int o = (G * b * g) & 31;
    off_max = o > off_max ? o : off_max;
  }
  const int K = (off_max + b * (G - 1) + cb + 31) / 32;
  // groups covering one word: at most ceil(32 K / (G b)) + 1
  const int cover = (32 * K + G * b - 1) / (G * b) + 1;
  if (cover > CARRY_GCOV || group_stride(K) > 8) return 0;
  return K;
}

int group_words(const Plan &p) { return p.P < 2 ? 0 : group_words_for(p.C[0], p.bits, coef_bits(p)); }

// Coefficients of plan p are written transposed when the last pass has
// narrow column tiles (see unpack_point); the carry reads them back that way.
bool coef_transposed(const Plan &p) { return p.P >= 2 && p.C[0] < 8; }

// Chunked middle passes of three-pass products: the second forward pass, the
// middle pass and the second inverse pass only touch blocks of S1 = L[1] M[1]
// consecutive points, so they run one group of J blocks at a time, all three
// launches on a group before the next; a group's points (both operands) total
// at most g_chunk_bytes and stay in the memory-side cache (256 MB on MI355X)
// from one launch to the next (cache-keeping kernels). 0 turns chunking off.
// With two streams, odd groups run on a side stream beside the even ones, so
// the end of one group's launch overlaps the start of the next group's (two
// groups in flight: 128 MB each measured best at 2^26 points, 2.50 ms against
// 2.56 ms unchunked; one stream loses the cache gain to launch boundaries).
size_t g_chunk_bytes = 128ull << 20;
int g_chunk_streams = 2;

struct ChunkPlan {
  fwd_fn fwd = nullptr, inv = nullptr;
  mid_fn mid = nullptr;
  unsigned long long S1 = 0, nb = 0, J = 0;  // block points, blocks, blocks per group
};

// True when plan p runs chunked: three passes, cache-keeping kernels for its
// tiles, and more than one group.
bool chunk_plan(const Plan &p, bool square, ChunkPlan &c) {
  if (p.P != 3 || !g_chunk_bytes) return false;
  c.fwd = pick_fwd_keep(p.L[1], p.C[1]);
  c.inv = pick_inv_keep(p.L[1], inv_cols(p, 1));
  c.mid = pick_mid_keep(p.L[2]);
  if (!c.fwd || !c.inv || !c.mid) return false;
  c.S1 = (unsigned long long)p.L[1] * p.M[1];
  c.nb = p.N / c.S1;
  const unsigned long long per_block = (square ? 1ull : 2ull) * c.S1 * sizeof(cd);
  c.J = g_chunk_bytes / per_block;
  if (c.J < 1) c.J = 1;
  return c.J < c.nb;
}

int run_chunk(const Plan &p, const ChunkPlan &ch, const PassArgs &f, const PassArgs &m, const PassArgs &v, int ci,
              cd *da, cd *db, bool square, unsigned long long b0, hipStream_t s);

int run_chunked(const Plan &p, const ChunkPlan &ch, const PassArgs &a, cd *da, cd *db, bool square, hipStream_t s) {
  PassArgs f = a, m = a, v = a;
  if (int rc = pass_args(p, 1, f)) return rc;
  if (int rc = pass_args(p, 2, m)) return rc;
  if (int rc = pass_args(p, 1, v)) return rc;
  const int ci = inv_cols(p, 1);
  const unsigned long long groups = (ch.nb + ch.J - 1) / ch.J;
  const int ns = (int)std::min<unsigned long long>((unsigned long long)g_chunk_streams, groups);
  if (ns > 1) {
    HIPCHK(hipEventRecord(g.fork_ev, s));
    for (int i = 0; i < ns - 1; i++) HIPCHK(hipStreamWaitEvent(g.aux[i], g.fork_ev, 0));
  }
  int rc = 0;
  for (unsigned long long b0 = 0, k = 0; b0 < ch.nb && !rc; b0 += ch.J, k++) {
    const int q = (int)(k % (unsigned long long)ns);
    rc = run_chunk(p, ch, f, m, v, ci, da, db, square, b0, q ? g.aux[q - 1] : s);
  }
  // the side streams rejoin s (also before an error returns)
  for (int i = 0; i < ns - 1; i++) {
    HIPCHK(hipEventRecord(g.join_ev[i], g.aux[i]));
    HIPCHK(hipStreamWaitEvent(s, g.join_ev[i], 0));
  }
  return rc;
}

// One group of the chunked middle passes: blocks [b0, b0 + J) on stream s.
int run_chunk(const Plan &p, const ChunkPlan &ch, const PassArgs &f, const PassArgs &m, const PassArgs &v, int ci,
              cd *da, cd *db, bool square, unsigned long long b0, hipStream_t s) {
  const unsigned long long jc = std::min(ch.J, ch.nb - b0);
  const size_t off = (size_t)(b0 * ch.S1);
  {
    PassArgs x = f;
    x.data = da + off;
    if (!square) {
      x.pair = 1;
      x.data_b = db + off;
    }
    x.ntiles = (int)(jc * (p.M[1] / (unsigned long long)p.C[1]));
    ProfScope ps(BCG_K_FWD, s);
    if (ch.fwd(false, x.ntiles, square ? 1 : 2, s, x)) return fail(BCG_E_ARG, "kernel rejected this pass");
    HIPCHK(hipGetLastError());
  }
  {
    PassArgs y = m;
    y.data = da + off;
    y.data_b = db + off;
    ProfScope ps(BCG_K_MID, s);
    if (ch.mid(square, false, (int)(jc * ch.S1 / (unsigned long long)p.L[2]), 1, s, y))
      return fail(BCG_E_ARG, "kernel rejected this pass");
    HIPCHK(hipGetLastError());
  }
  {
    PassArgs z = v;
    z.data = da + off;
    z.ntiles = (int)(jc * (p.M[1] / (unsigned long long)ci));
    ProfScope ps(BCG_K_INV, s);
    if (ch.inv(false, z.ntiles, 1, s, z)) return fail(BCG_E_ARG, "kernel rejected this pass");
    HIPCHK(hipGetLastError());
  }
  return 0;
}

// Run the transforms of plan p up to the rounded coefficients (in bufB).
int run_transforms(const Plan &p, const uint32_t *A, unsigned long long AL, const uint32_t *B,
                   unsigned long long BL, bool square, hipStream_t s) {
  const size_t arr = (size_t)p.N * sizeof(cd);
  if (int rc = ensure(g.bufA, arr)) return rc;
  if (int rc = ensure(g.bufB, arr)) return rc;  // B spectrum, then coefficients (2N x 8 bytes)
  cd *da = (cd *)g.bufA.p;
  cd *db = (cd *)g.bufB.p;
  long long *coef = (long long *)g.bufB.p;
  if (int rc = zero_margin(s)) return rc;
  PassArgs a;
  memset(&a, 0, sizeof a);
  a.wa = A;
  a.al = AL;
  a.wb = B;
  a.bl = BL;
  a.margin = (unsigned int *)g.margin.p;
  a.margin_mask = MARGIN_SLOTS - 1;
  if (p.P == 1) {
    if (int rc = pass_args(p, 0, a)) return rc;
    a.data = da;
    a.data_b = db;
    a.coef = coef;
    mid_fn f = pick_mid(p.L[0]);
    if (!f) return fail(BCG_E_ARG, "no kernel for this length");
    ProfScope ps(BCG_K_MID, s);
    if (f(square, true, 1, 1, s, a)) return fail(BCG_E_ARG, "kernel rejected this pass");
    HIPCHK(hipGetLastError());
    return 0;
  }
  ChunkPlan ch;
  const bool chunked = chunk_plan(p, square, ch);
  // forward outer passes for A (and B)
  // (A and B share each launch: grid.y = 2)
  for (int i = 0; i < p.P - 1; i++) {
    if (chunked && i == 1) break;  // with the middle pass, chunk by chunk
    PassArgs b = a;
    if (int rc = pass_args(p, i, b)) return rc;
    b.data = da;
    b.wa = A;
    b.al = AL;
    {
      // the packing pass of operands that leave the upper digit half empty
      // (at most N b / 32 words each) runs the kernel without upper digits
      const unsigned long long half_w = (p.N * (unsigned long long)p.bits) >> 5;
      b.lower = (i == 0 && AL <= half_w && (square || BL <= half_w)) ? 1 : 0;
    }
    if (!square) {
      b.pair = 1;
      b.data_b = db;
      b.wb = B;
      b.bl = BL;
    }
    fwd_fn f = pick_fwd(p.L[i], p.C[i]);
    if (!f) return fail(BCG_E_ARG, "no forward kernel for this tile");
    ProfScope ps(i == 0 ? BCG_K_FWD_PACK : BCG_K_FWD, s);
    if (f(i == 0, b.ntiles, square ? 1 : 2, s, b)) return fail(BCG_E_ARG, "kernel rejected this pass");
    HIPCHK(hipGetLastError());
  }
  // middle: last forward pass + pointwise + first inverse pass, in place in A
  if (chunked) {
    ProfScope ps(BCG_K_MIDDLE, s);
    if (int rc = run_chunked(p, ch, a, da, db, square, s)) return rc;
  } else {
    PassArgs b = a;
    int i = p.P - 1;
    if (int rc = pass_args(p, i, b)) return rc;
    b.data = da;
    b.data_b = db;
    mid_fn f = pick_mid(p.L[i]);
    if (!f) return fail(BCG_E_ARG, "no middle kernel for this length");
    ProfScope ps(BCG_K_MID, s);
    if (f(square, false, (int)(p.N / (unsigned long long)p.L[i]), 1, s, b)) return fail(BCG_E_ARG, "kernel rejected this pass");
    HIPCHK(hipGetLastError());
  }
  // inverse outer passes, the outermost one rounding into coefficients
  for (int i = p.P - 2; i >= 0; i--) {
    if (chunked && i == 1) continue;  // ran chunk by chunk above
    PassArgs b = a;
    if (int rc = pass_args(p, i, b)) return rc;
    b.data = da;
    b.coef = coef;
    b.coef_T = (i == 0 && coef_transposed(p)) ? 1 : 0;
    if (i == 0 && group_words(p)) {
      b.ybuf = (uint32_t *)g.bufB.p;
      b.ygK = group_stride(group_words(p));
    }
    const int ci = inv_cols(p, i);
    b.ntiles = (int)(p.N / ((unsigned long long)p.L[i] * (unsigned long long)ci));
    fwd_fn f = pick_inv(p.L[i], ci);
    if (!f) return fail(BCG_E_ARG, "no inverse kernel for this tile");
    ProfScope ps(i == 0 ? BCG_K_INV_UNPACK : BCG_K_INV, s);
    if (f(i == 0, b.ntiles, 1, s, b)) return fail(BCG_E_ARG, "kernel rejected this pass");
    HIPCHK(hipGetLastError());
  }
  return 0;
}

__global__ void k_basecase(const uint32_t *A, unsigned long long AL, const uint32_t *B, unsigned long long BL,
                           uint32_t *C, unsigned long long CL, const uint32_t *D, unsigned long long DL) {
  // column sums of 32x32-bit products split into low and high halves
  extern __shared__ unsigned long long colsum[];
  const unsigned long long W = AL + BL;
  for (unsigned long long w = threadIdx.x; w < W; w += blockDim.x) {
    unsigned long long lo = 0, hi = 0;
    unsigned long long i0 = w >= BL ? w - BL + 1 : 0;
    for (unsigned long long i = i0; i < AL && i <= w; i++) {
      unsigned long long pr = (unsigned long long)A[i] * B[w - i];
      lo += pr & 0xffffffffull;
      hi += pr >> 32;
    }
    colsum[2 * w] = lo;
    colsum[2 * w + 1] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long carry = 0;  // < 2^40
    for (unsigned long long w = 0; w < CL; w++) {
      unsigned long long v = carry;
      if (w < W) v += colsum[2 * w];
      if (w >= 1 && w - 1 < W) v += colsum[2 * (w - 1) + 1];
      if (D && w < DL) v += D[w];
      C[w] = (uint32_t)v;
      carry = v >> 32;
    }
  }
}

int basecase_dev(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL,
                 hipStream_t s, const uint32_t *D = nullptr, size_t DL = 0) {
  if (AL + BL > 4000) return fail(BCG_E_ARG, "basecase limited to AL + BL <= 4000 words");
  size_t sh = (AL + BL + 1) * 2 * sizeof(unsigned long long);
  // an output overlapping an input (in-place products) goes through a scratch;
  // otherwise the kernel writes C directly
  const uintptr_t c0 = (uintptr_t)C, c1 = c0 + CL * sizeof(uint32_t);
  auto overlaps = [&](const uint32_t *x, size_t n) {
    return x && n && (uintptr_t)x < c1 && c0 < (uintptr_t)(x + n);
  };
  const bool via_scratch = overlaps(A, AL) || overlaps(B, BL) || overlaps(D, DL);
  uint32_t *dst = C;
  if (via_scratch) {
    if (int rc = ensure(g.zbuf, CL * sizeof(uint32_t) + 64)) return rc;
    dst = (uint32_t *)g.zbuf.p;
  }
  {
    ProfScope ps(BCG_K_OTHER, s);
    hipLaunchKernelGGL(k_basecase, dim3(1), dim3(256), sh, s, A, (unsigned long long)AL, B,
                       (unsigned long long)BL, dst, (unsigned long long)CL, D, (unsigned long long)DL);
  }
  HIPCHK(hipGetLastError());
  if (via_scratch) HIPCHK(hipMemcpyAsync(C, dst, CL * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  return 0;
}

__global__ void k_copy_words(uint32_t *dst, const uint32_t *src, unsigned long long n, unsigned long long srcn) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long st = (unsigned long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) dst[i] = i < srcn ? src[i] : 0u;
}

// Plain streaming copy of n 16-byte vectors (grid-stride).
__global__ __launch_bounds__(256) void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src,
                                                unsigned long long n) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long st = (unsigned long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) dst[i] = src[i];
}

int fill_zero(uint32_t *C, size_t CL, hipStream_t s) {
  if (!CL) return 0;
  int grid = (int)((CL + 255) / 256);
  if (grid > 4096) grid = 4096;
  LAUNCH(k_fill_u32, grid, 256, s, C, (unsigned long long)CL, 0u);
  HIPCHK(hipGetLastError());
  return 0;
}

// Workgroups resident per CU for a one-pass class kernel (LDS bound).
int mul1_grid(int L, int count) {
  size_t lds = (size_t)(L + L / 16) * 16 + 256;
  int per_cu = (int)(160 * 1024 / lds);
  if (per_cu < 1) per_cu = 1;
  if (per_cu > 8) per_cu = 8;
  int grid = 256 * per_cu;
  return count < grid ? count : grid;
}

__global__ void k_zero_i32(int *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Launch the one-pass products desc[first .. first+count) of length class L
// (descriptors and margin slots already on the device).
int run_class(int L, const MulDesc *ddesc, int count, unsigned int *dmargins, int *dcounter, hipStream_t s) {
  ClassArgs ca;
  ca.desc = ddesc;
  ca.count = count;
  if (int rc = get_roots((unsigned long long)L, &ca.roots)) return rc;
  if (int rc = get_ltab(L, &ca.tl)) return rc;
  ca.margins = dmargins;
  ca.next = dcounter;
  mul1_fn f = pick_mul1(L);
  if (!f) return fail(BCG_E_ARG, "no one-pass kernel for this length");
  ProfScope ps(BCG_K_MID, s);
  f(mul1_grid(L, count), s, ca);
  HIPCHK(hipGetLastError());
  return 0;
}

// One product whose plan is a single pass: the whole product in one workgroup.
int mul_onepass(const Plan &p, uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL,
                hipStream_t s, double *margin, const uint32_t *D = nullptr, size_t DL = 0) {
  if (int rc = ensure(g.desc, sizeof(MulDesc))) return rc;
  if (int rc = ensure(g.counters, 64)) return rc;
  MulDesc d;
  d.a = A;
  d.b = B;
  d.c = C;
  d.d = D;
  d.al = AL;
  d.bl = BL;
  d.cl = CL;
  d.dl = D ? DL : 0;
  d.bits = p.bits;
  d.cbits = coef_bits(p);
  if (int rc = h2d(g.desc.p, &d, sizeof d, s)) return rc;
  if (int rc = zero_margin(s)) return rc;
  LAUNCH(k_zero_i32, 1, 64, s, (int *)g.counters.p, 1);
  if (int rc = run_class(p.L[0], (const MulDesc *)g.desc.p, 1, (unsigned int *)g.margin.p, (int *)g.counters.p, s))
    return rc;
  return read_margin(s, margin);
}

// Batched multi-pass products of one plan (same N and pass split): every pass
// is one launch over (tiles x products); product p owns arrays p*N in the
// batch buffers and margin slot p. Descriptors already on the device.
// Scratch of batched launches: descriptors, margins, class counters, the
// batched transform arrays and carry block functions. The plain batch call
// uses the engine's set; a captured batch owns one (its graph keeps the
// addresses).
struct BatchBufs {
  DevBuf tdesc, mdesc, cdesc, margins, counters, tA, tB, blk;
  void release() {
    for (DevBuf *b : {&tdesc, &mdesc, &cdesc, &margins, &counters, &tA, &tB, &blk}) {
      if (b->p) (void)hipFree(b->p);
      b->p = nullptr;
      b->bytes = 0;
    }
  }
};

BatchBufs g_bb;  // scratch of the plain batch call

int run_multipass_class(const Plan &p, const MulDesc *ddesc, const std::vector<MulDesc> &hdesc,
                        unsigned int *dmargins, hipStream_t s, BatchBufs &bb) {
  const int count = (int)hdesc.size();
  const size_t arr = (size_t)p.N * sizeof(cd) * (size_t)count;
  if (int rc = ensure(bb.tA, arr)) return rc;
  if (int rc = ensure(bb.tB, arr)) return rc;
  cd *da = (cd *)bb.tA.p;
  cd *db = (cd *)bb.tB.p;
  PassArgs a;
  memset(&a, 0, sizeof a);
  a.desc = ddesc;
  a.margin = dmargins;
  int fast = 1;  // every product of the class has coefficients below 2^51
  for (const MulDesc &d : hdesc) fast = fast && d.cbits <= 51;
  // both operands' forward passes in one launch each (grid.y = 2 count) while
  // grid.y allows it, else one launch per operand
  const bool both = 2 * count <= 65535;
  for (int op = 0; op < (both ? 1 : 2); op++) {
    for (int i = 0; i < p.P - 1; i++) {
      PassArgs b = a;
      if (int rc = pass_args(p, i, b)) return rc;
      b.data = op == 0 ? da : db;
      b.operand = op;
      if (both) {
        b.data_b = db;
        b.pcount = count;
      }
      fwd_fn f = pick_fwd(p.L[i], p.C[i]);
      if (!f) return fail(BCG_E_ARG, "no forward kernel for this tile");
      ProfScope ps(i == 0 ? BCG_K_FWD_PACK : BCG_K_FWD, s);
      if (f(i == 0, b.ntiles, both ? 2 * count : count, s, b)) return fail(BCG_E_ARG, "kernel rejected this pass");
      HIPCHK(hipGetLastError());
    }
  }
  {
    PassArgs b = a;
    int i = p.P - 1;
    if (int rc = pass_args(p, i, b)) return rc;
    b.data = da;
    b.data_b = db;
    mid_fn f = pick_mid(p.L[i]);
    if (!f) return fail(BCG_E_ARG, "no middle kernel for this length");
    ProfScope ps(BCG_K_MID, s);
    if (f(false, false, (int)(p.N / (unsigned long long)p.L[i]), count, s, b)) return fail(BCG_E_ARG, "kernel rejected this pass");
    HIPCHK(hipGetLastError());
  }
  // group mode when every product of the class has exact group sums: the last
  // inverse pass writes them (one stride for the class, the widest product's)
  // into the B buffer and the group-sum carry resolves them per product
  bool groups = fast && p.P >= 2 && (p.C[0] == 8 || p.C[0] == 16);
  int gks = 0, gcov = 0, bmin = 64;
  for (const MulDesc &d : hdesc) {
    if (!d.gk) groups = false;
    gks = std::max(gks, group_stride(d.gk));
    bmin = std::min(bmin, d.bits);
    const int gb = p.C[0] * d.bits;
    const int cv = gb % 32 == 0 ? (d.gk + gb / 32 - 1) / (gb / 32) : (32 * d.gk + gb - 1) / gb + 1;
    gcov = std::max(gcov, cv);
  }
  const unsigned long long ng = groups ? 2 * p.N / (unsigned long long)p.C[0] : 0;
  const unsigned long long ystride = ng * (unsigned long long)gks;
  for (int i = p.P - 2; i >= 0; i--) {
    PassArgs b = a;
    if (int rc = pass_args(p, i, b)) return rc;
    b.data = da;
    b.coef = (long long *)db;
    b.fast_round = fast;
    if (i == 0 && groups) {
      b.coef = nullptr;
      b.ybuf = (uint32_t *)db;
      b.ygK = gks;
      b.ystride = ystride;
    }
    const int ci = inv_cols(p, i);
    b.ntiles = (int)(p.N / ((unsigned long long)p.L[i] * (unsigned long long)ci));
    fwd_fn f = pick_inv(p.L[i], ci);
    if (!f) return fail(BCG_E_ARG, "no inverse kernel for this tile");
    ProfScope ps(i == 0 ? BCG_K_INV_UNPACK : BCG_K_INV, s);
    if (f(i == 0, b.ntiles, count, s, b)) return fail(BCG_E_ARG, "kernel rejected this pass");
    HIPCHK(hipGetLastError());
  }
  // carries, one product per grid row
  const unsigned long long wpb = groups ? (unsigned long long)GCT * GWPT : (unsigned long long)CT * CWPT;
  unsigned long long nblk_max = 1;
  for (const MulDesc &d : hdesc) {
    unsigned long long nb = (d.cl + wpb - 1) / wpb;
    if (nb > nblk_max) nblk_max = nb;
  }
  nblk_max = (nblk_max + 15) & ~15ull;  // 16-byte aligned rows of block functions
  if (int rc = ensure(bb.blk, nblk_max * (size_t)count)) return rc;
  CarryArgs c;
  memset(&c, 0, sizeof c);
  c.desc = ddesc;
  c.coef_stride = 2 * p.N;
  c.coef = (const long long *)db;
  c.ncoef = 2 * p.N;
  c.blk_fn = (unsigned char *)bb.blk.p;
  c.nblk = nblk_max;
  c.wpb = (unsigned)wpb;
  {
    int bmin = 64, cbmax = 0;
    for (const MulDesc &d : hdesc) {
      bmin = d.bits < bmin ? d.bits : bmin;
      cbmax = d.cbits > cbmax ? d.cbits : cbmax;
    }
    c.lds_slots = carry_slots(bmin, cbmax);
  }
  if (groups) {
    c.coef = nullptr;
    c.ybuf = (const uint32_t *)db;
    c.ystride = ystride;
    c.gC = p.C[0];
    c.ng = (long long)ng;
    c.gK = gks;  // per product from its descriptor
    c.gKs = gks;
    // LDS: the groups under one block's GCT * GWPT words at the smallest digit size
    size_t glds = ((size_t)(32 * GCT * GWPT) / (size_t)(p.C[0] * bmin) + CARRY_GCOV + 2) * (size_t)gks;
    glds = (glds + glds / 32 + 4) * 4;  // one pad word per 32
    ProfScope ps(BCG_K_CARRY, s);
    const dim3 grid((unsigned)nblk_max, (unsigned)count);
    if (gcov <= 2)
      hipLaunchKernelGGL(k_carry_groups<2>, grid, dim3(GCT), glds, s, c);
    else if (gcov == 3)
      hipLaunchKernelGGL(k_carry_groups<3>, grid, dim3(GCT), glds, s, c);
    else
      hipLaunchKernelGGL(k_carry_groups<CARRY_GCOV>, grid, dim3(GCT), glds, s, c);
    LAUNCH(k_carry_fix, count, 1024, s, c);
    LAUNCH2(k_carry_apply, (unsigned)nblk_max, (unsigned)count, 256, s, c);
  } else {
    ProfScope ps(BCG_K_CARRY, s);
    hipLaunchKernelGGL(k_carry_local, dim3((unsigned)nblk_max, (unsigned)count), dim3(CT),
                       c.lds_slots * sizeof(long long), s, c);
    LAUNCH(k_carry_fix, count, 1024, s, c);
    LAUNCH2(k_carry_apply, (unsigned)nblk_max, (unsigned)count, 256, s, c);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// Reads the margin sets of the open deferral scope (after the device is
// idle), records their products in the statistics and frees the sets.
int defer_flush() {
  DeferScope &d = g_defer;
  if (!d.n) return 0;
  HIPCHK(hipDeviceSynchronize());
  std::vector<unsigned int> w((size_t)d.n * MARGIN_SLOTS * MARGIN_STRIDE);
  HIPCHK(hipMemcpy(w.data(), d.sets.p, w.size() * sizeof(unsigned int), hipMemcpyDeviceToHost));
  for (int i = 0; i < d.n; i++) {
    unsigned int b = 0;
    for (int k = 0; k < MARGIN_SLOTS; k++) b = std::max(b, w[((size_t)i * MARGIN_SLOTS + k) * MARGIN_STRIDE]);
    float f;
    memcpy(&f, &b, sizeof f);
    d.worst = std::max(d.worst, (double)f);
    note_product(d.points[i], d.bits[i], (double)f);
  }
  d.n = 0;
  return 0;
}

int mul_dev_body(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL, int force_bits,
                 hipStream_t s, bcg_mul_info *info, const uint32_t *D, size_t DL);

int mul_dev_locked(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL,
                   int force_bits, hipStream_t s, bcg_mul_info *info, const uint32_t *D = nullptr, size_t DL = 0) {
  DeferScope &d = g_defer;
  if (!d.open || d.in_product || g_fixed || std::this_thread::get_id() != d.owner || C == A || C == B)
    return mul_dev_body(C, CL, A, AL, B, BL, force_bits, s, info, D, DL);
  // a deferred product: margin slots of its own set, no read-back
  if (d.n == DEFER_SETS)
    if (int rc = defer_flush()) return rc;
  void *keep = g.margin.p;
  g.margin.p = (unsigned char *)d.sets.p + (size_t)d.n * MARGIN_SET_BYTES;
  d.in_product = true;
  bcg_mul_info inf;
  memset(&inf, 0, sizeof inf);
  const int rc = mul_dev_body(C, CL, A, AL, B, BL, force_bits, s, &inf, D, DL);
  d.in_product = false;
  g.margin.p = keep;
  if (rc) return rc;
  if (inf.points) {  // an FFT product: its set holds its margin (schoolbook products have none)
    d.points[d.n] = inf.points;
    d.bits[d.n] = inf.bits;
    d.n++;
  }
  if (info) *info = inf;
  return 0;
}

int mul_dev_body(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL, int force_bits,
                 hipStream_t s, bcg_mul_info *info, const uint32_t *D, size_t DL) {
  bcg_mul_info inf;
  memset(&inf, 0, sizeof inf);
  bool square = (A == B && AL == BL);
  inf.square = square ? 1 : 0;
  if (!D) DL = 0;
  if (D && C == D) return fail(BCG_E_ARG, "the addend may not alias the output");
  if (AL == 0 || BL == 0 || CL == 0) {
    if (info) *info = inf;
    if (int rc = fill_zero(C, CL, s)) return rc;
    if (DL && CL) {
      LAUNCH(k_copy_words, grid_for(CL), 256, s, C, (const uint32_t *)D, (unsigned long long)CL,
             (unsigned long long)DL);
      HIPCHK(hipGetLastError());
    }
    return 0;
  }
  if (AL + BL < 100 && force_bits <= 0) {
    if (info) *info = inf;
    return basecase_dev(C, CL, A, AL, B, BL, s, D, DL);
  }
  Plan p;
  if (int rc = make_plan(AL, BL, force_bits, p)) return rc;
  bool alias = (C == A || C == B);
  for (;;) {
    if (p.P == 1) {
      // whole product in one workgroup; aliasing outputs go through a scratch so a
      // margin retry still sees the inputs
      double m = 0;
      uint32_t *dst = C;
      if (alias) {
        if (int rc = ensure(g.zbuf, CL * sizeof(uint32_t) + 64)) return rc;
        dst = (uint32_t *)g.zbuf.p;
      }
      if (int rc = mul_onepass(p, dst, CL, A, AL, B, BL, s, &m, D, DL)) return rc;
      inf.bits = p.bits;
      inf.passes = 1;
      inf.points = (size_t)p.N;
      inf.margin = m;
      if (m >= 0.25 && p.bits > 4) {
        Plan q;
        if (int rc = make_plan(AL, BL, p.bits - 1, q)) return rc;
        p = q;
        inf.retries++;
        g_stats.retries++;
        continue;
      }
      if (m >= 0.25) return fail(BCG_E_MARGIN, "rounding margin >= 0.25 at the smallest digit size");
      if (alias) HIPCHK(hipMemcpyAsync(C, dst, CL * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
      break;
    }
    if (int rc = run_transforms(p, A, AL, B, BL, square, s)) return rc;
    double m = 0;
    if (alias) {
      // the carry pass overwrites the inputs: check the margin first
      if (int rc = read_margin(s, &m)) return rc;
    }
    if (!alias) {
      if (int rc = carry_of_plan(p, C, CL, s, D, DL)) return rc;
      if (int rc = read_margin(s, &m)) return rc;
    }
    inf.bits = p.bits;
    inf.passes = p.P;
    inf.points = (size_t)p.N;
    inf.margin = m;
    inf.group_digits = group_words(p) ? p.C[0] : 0;
    inf.group_stride = group_words(p) ? group_stride(group_words(p)) : 0;
    {
      ChunkPlan ch;
      inf.chunks = chunk_plan(p, square, ch) ? (int)((ch.nb + ch.J - 1) / ch.J) : 0;
    }
    if (m >= 0.25) {
      if (p.bits <= 4) {
        if (info) *info = inf;
        return fail(BCG_E_MARGIN, "rounding margin >= 0.25 at the smallest digit size");
      }
      Plan q;
      if (int rc = make_plan(AL, BL, p.bits - 1, q)) return rc;
      p = q;
      inf.retries++;
      g_stats.retries++;
      continue;
    }
    if (alias) {
      if (int rc = carry_of_plan(p, C, CL, s, D, DL)) return rc;
    }
    break;
  }
  if (!g_defer.in_product) note_product(inf.points, inf.bits, inf.margin);  // deferred: when its set is read
  if (info) *info = inf;
  return 0;
}

int defer_begin() {
  DeferScope &d = g_defer;
  if (d.open) return fail(BCG_E_ARG, "a deferral scope is already open");
  if (int rc = ensure(d.sets, DEFER_SETS * MARGIN_SET_BYTES)) return rc;
  d.open = true;
  d.owner = std::this_thread::get_id();
  d.n = 0;
  d.worst = 0.0;
  return 0;
}

int defer_end(double *worst) {
  DeferScope &d = g_defer;
  if (!d.open || std::this_thread::get_id() != d.owner) return fail(BCG_E_ARG, "no deferral scope open on this thread");
  const int rc = defer_flush();
  d.open = false;
  d.n = 0;
  if (worst) *worst = d.worst;
  return rc;
}

// ------------------------------------------------- reference-order transforms

__global__ void k_small_dft(cd *z, int L, int inv, int three, int lg) {
  // naive DFT for L < 64, in the reference order; one workgroup
  __shared__ cd in[64];
  int t = threadIdx.x;
  if (t < L) in[t] = z[t];
  __syncthreads();
  if (t < L) {
    cd acc = make_double2(0, 0);
    if (!inv) {
      // output position t holds frequency k(t)
      unsigned k = three ? brev((unsigned)t / 3u, lg) + ((unsigned)t % 3u) * (unsigned)(L / 3) : brev((unsigned)t, lg);
      for (int j = 0; j < L; j++) {
        double sn, cs;
        sincospi(2.0 * (double)(((unsigned long long)j * k) % (unsigned long long)L) / (double)L, &sn, &cs);
        acc = cadd(acc, cmul(in[j], make_double2(cs, sn)));
      }
    } else {
      for (int r = 0; r < L; r++) {
        unsigned k = three ? brev((unsigned)r / 3u, lg) + ((unsigned)r % 3u) * (unsigned)(L / 3) : brev((unsigned)r, lg);
        double sn, cs;
        sincospi(2.0 * (double)(((unsigned long long)t * k) % (unsigned long long)L) / (double)L, &sn, &cs);
        acc = cadd(acc, cmulc(in[r], make_double2(cs, sn)));
      }
    }
    z[t] = acc;
  }
}

bool valid_length(unsigned long long n) {
  if (n < 2) return false;
  unsigned long long p = is_three(n) ? n / 3 : n;
  return (p & (p - 1)) == 0;
}

int transform_dev(double *z, size_t length, bool inverse, hipStream_t s) {
  if (!valid_length(length)) return fail(BCG_E_ARG, "transform length must be 2^k or 3*2^k");
  if (length < 64) {
    int three = is_three(length) ? 1 : 0;
    int lg = ilog2u(three ? length / 3 : length);
    ProfScope ps(BCG_K_OTHER, s);
    LAUNCH(k_small_dft, 1, 64, s, (cd *)z, (int)length, inverse ? 1 : 0, three, lg);
    HIPCHK(hipGetLastError());
    return 0;
  }
  Plan p;
  p.N = length;
  if (!factor(length, p)) return fail(BCG_E_ARG, "no transform plan for this length");
  for (int i = 0; i < p.P; i++)
    if (i == p.P - 1) p.C[i] = 1;
  PassArgs a;
  memset(&a, 0, sizeof a);
  if (!inverse) {
    for (int i = 0; i < p.P; i++) {
      PassArgs b = a;
      if (int rc = pass_args(p, i, b)) return rc;
      b.data = (cd *)z;
      fwd_fn f = pick_fwd(p.L[i], p.C[i]);
      if (!f) return fail(BCG_E_ARG, "no forward kernel for this tile");
      ProfScope ps(BCG_K_FWD, s);
      if (f(false, b.ntiles, 1, s, b)) return fail(BCG_E_ARG, "kernel rejected this pass");
      HIPCHK(hipGetLastError());
    }
  } else {
    for (int i = p.P - 1; i >= 0; i--) {
      PassArgs b = a;
      if (int rc = pass_args(p, i, b)) return rc;
      b.data = (cd *)z;
      fwd_fn f = pick_inv(p.L[i], p.C[i]);
      if (!f) return fail(BCG_E_ARG, "no inverse kernel for this tile");
      ProfScope ps(BCG_K_INV, s);
      if (f(false, b.ntiles, 1, s, b)) return fail(BCG_E_ARG, "kernel rejected this pass");
      HIPCHK(hipGetLastError());
    }
  }
  return 0;
}

// ----------------------------------------------------------- truncated FFT
//
// tft_forward(T, len, k): the first len outputs (bit-reversed order) of the
// length-2^k forward transform of T[0..len) zero-padded; the rest zeroed.
// tft_inverse(T, len, k): given those len outputs of a signal supported on
// [0, len), returns 2^k times the signal in T[0..len), zeros after
// (fft/tft.c entry points; semantics pinned by test/test_tft.c).
//
// The inverse recursion on n = 2^k follows the first DIF level: outputs in the
// first half are the half-length transform of u_j = x_j + x_{j+n/2}, outputs in
// the second half that of v_j = (x_j - x_{j+n/2}) w^j (w = e^{2 pi i/n}).
//  * len <= n/2: x_{j+n/2} = 0, so u = x and the problem halves.
//  * len > n/2: u follows from the full first half; for j >= len - n/2 the
//    upper input is zero, so v_j = u_j w^j is known there; removing those known
//    inputs from the second-half outputs leaves a half-size inverse TFT for the
//    remaining v_j, and x_j, x_{j+n/2} = (u_j +- v_j w^{-j}) / 2.

__global__ void k_scale(cd *z, unsigned long long n, double s) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long st = (unsigned long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) z[i] = make_double2(z[i].x * s, z[i].y * s);
}

__global__ void k_zero_range(cd *z, unsigned long long lo, unsigned long long hi) {
  unsigned long long i = lo + (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long st = (unsigned long long)gridDim.x * blockDim.x;
  for (; i < hi; i += st) z[i] = make_double2(0.0, 0.0);
}

// vk[j] = 0 for j < l2, u[j] * w^j for j >= l2 (w = e^{2 pi i / n}, j < n/2)
__global__ void k_tft_known(const cd *u, cd *vk, unsigned long long half, unsigned long long l2, unsigned long long n) {
  unsigned long long j = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long st = (unsigned long long)gridDim.x * blockDim.x;
  for (; j < half; j += st) {
    if (j < l2) {
      vk[j] = make_double2(0.0, 0.0);
    } else {
      double sn, cs;
      sincospi(2.0 * (double)j / (double)n, &sn, &cs);
      vk[j] = cmul(u[j], make_double2(cs, sn));
    }
  }
}

// r[p] = y[p] - f[p], p < l2
__global__ void k_sub(cd *r, const cd *y, const cd *f, unsigned long long l2) {
  unsigned long long p = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long st = (unsigned long long)gridDim.x * blockDim.x;
  for (; p < l2; p += st) r[p] = csub(y[p], f[p]);
}

// x_j, x_{j+n/2} from u_j and v_j (j < l2); x_j = u_j for l2 <= j < half
__global__ void k_tft_combine(cd *x, const cd *u, const cd *v, unsigned long long half, unsigned long long l2,
                              unsigned long long n) {
  unsigned long long j = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long st = (unsigned long long)gridDim.x * blockDim.x;
  for (; j < half; j += st) {
    if (j < l2) {
      double sn, cs;
      sincospi(2.0 * (double)j / (double)n, &sn, &cs);
      cd t = cmulc(v[j], make_double2(cs, sn));
      x[j] = make_double2(0.5 * (u[j].x + t.x), 0.5 * (u[j].y + t.y));
      x[j + half] = make_double2(0.5 * (u[j].x - t.x), 0.5 * (u[j].y - t.y));
    } else {
      x[j] = u[j];
    }
  }
}

__global__ void k_pointwise(cd *a, const cd *b, unsigned long long n) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long st = (unsigned long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) a[i] = cmul(a[i], b[i]);
}

__global__ void k_int_to_fft(cd *z, unsigned long long length, const uint32_t *A, unsigned long long AL, int bits,
                             unsigned long long pts) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long st = (unsigned long long)gridDim.x * blockDim.x;
  unsigned long long n = length > pts ? length : pts;
  for (; i < n; i += st) z[i] = make_double2(i < pts ? digit_at(A, AL, i, bits) : 0.0, 0.0);
}

// Points at or beyond `avail` (the caller's array length) read as zero.
__global__ void k_round_points(const cd *z, unsigned long long pts, unsigned long long avail, double scale,
                               long long *coef, unsigned int *margin) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long st = (unsigned long long)gridDim.x * blockDim.x;
  double worst = 0;
  for (; i < pts; i += st) {
    double f = i < avail ? z[i].x * scale : 0.0;
    worst = fmax(worst, fabs(f - rint(f)));
    long long r = (long long)(f + 0.5);
    coef[i] = r > 0 ? r : 0;
  }
  margin_reduce(margin, worst, MARGIN_SLOTS - 1);
}

int grid_for(unsigned long long n) {
  unsigned long long g2 = (n + 255) / 256;
  if (g2 > 8192) g2 = 8192;
  if (g2 < 1) g2 = 1;
  return (int)g2;
}

// True-scale inverse TFT: y[0..len) -> x[0..len) (x supported on [0, len)),
// transform length n = 2^k; y and x may be the same device array.
int tft_inverse_rec(cd *y, unsigned long long len, unsigned long long n, cd *x, hipStream_t s) {
  if (len == 0) return 0;
  if (n == 1) {
    if (x != y) HIPCHK(hipMemcpyAsync(x, y, sizeof(cd), hipMemcpyDeviceToDevice, s));
    return 0;
  }
  if (len == n) {
    if (x != y) HIPCHK(hipMemcpyAsync(x, y, n * sizeof(cd), hipMemcpyDeviceToDevice, s));
    if (int rc = transform_dev((double *)x, n, true, s)) return rc;
    LAUNCH(k_scale, grid_for(n), 256, s, x, n, 1.0 / (double)n);
    HIPCHK(hipGetLastError());
    return 0;
  }
  const unsigned long long half = n / 2;
  if (len <= half) return tft_inverse_rec(y, len, half, x, s);
  const unsigned long long l2 = len - half;
  void *mem = nullptr;
  hipError_t e = hipMalloc(&mem, 4 * half * sizeof(cd));
  if (e != hipSuccess) return fail(BCG_E_NOMEM, "tft scratch", e);
  cd *u = (cd *)mem, *vk = u + half, *r = vk + half, *v = r + half;
  int rc = 0;
  do {
    // u = inverse of the full first half (true scale)
    if ((e = hipMemcpyAsync(u, y, half * sizeof(cd), hipMemcpyDeviceToDevice, s)) != hipSuccess) { rc = fail(BCG_E_HIP, "copy", e); break; }
    if ((rc = transform_dev((double *)u, half, true, s))) break;
    LAUNCH(k_scale, grid_for(half), 256, s, u, half, 1.0 / (double)half);
    // known upper inputs of the second half, their outputs removed from y
    LAUNCH(k_tft_known, grid_for(half), 256, s, u, vk, half, l2, n);
    if ((rc = transform_dev((double *)vk, half, false, s))) break;
    LAUNCH(k_sub, grid_for(l2), 256, s, r, y + half, vk, l2);
    if ((rc = tft_inverse_rec(r, l2, half, v, s))) break;
    LAUNCH(k_tft_combine, grid_for(half), 256, s, x, u, v, half, l2, n);
    if ((e = hipGetLastError()) != hipSuccess) { rc = fail(BCG_E_HIP, "tft kernels", e); break; }
    if ((e = hipStreamSynchronize(s)) != hipSuccess) { rc = fail(BCG_E_HIP, "sync", e); break; }
  } while (0);
  (void)hipStreamSynchronize(s);
  (void)hipFree(mem);
  return rc;
}

__global__ void k_fill_hash(uint32_t *p, unsigned long long n, unsigned long long seed) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long st = (unsigned long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) {
    unsigned long long z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    p[i] = (uint32_t)((z >> 32) % 1000000000ull) | (i + 1 == n ? 1u : 0u);
  }
}

#include "e_pipeline.inc"
#include "bigfloat_dev.inc"

}  // namespace

// ======================================================== allocation-free

namespace {

// Scratch of bcg_mul_dev_noalloc: the engine's per-product buffers as fixed
// slices of one static device array, for products whose transform has at
// most NOALLOC_MAX_POINTS points: two transform arrays (16 bytes a point),
// the margin slots, carry block functions, one product descriptor, counters,
// and the staging of an aliased output (one-pass and schoolbook products).
constexpr unsigned long long NOALLOC_MAX_POINTS = 1ull << 20;
constexpr size_t NA_ARR = (size_t)NOALLOC_MAX_POINTS * sizeof(cd);
constexpr size_t NA_MARGIN = (size_t)MARGIN_SLOTS * MARGIN_STRIDE * sizeof(unsigned int);
constexpr size_t NA_BLK = (size_t)64 << 10;
constexpr size_t NA_DESC = 256;
constexpr size_t NA_COUNTERS = 256;
constexpr size_t NA_ZBUF = (size_t)64 << 10;
constexpr size_t NA_BYTES = 2 * NA_ARR + NA_MARGIN + NA_BLK + NA_DESC + NA_COUNTERS + NA_ZBUF;

}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {

// Device address of `scratch` per device (0: not looked up). Host code uses
// these, never the symbol itself (on the host it is only a stand-in).
constexpr int NA_MAX_DEVICES = 64;
unsigned char *g_scratch_ptr[NA_MAX_DEVICES] = {nullptr};

// At initialisation, with the device current: the scratch's address, and
// every table and pass split a product up to NOALLOC_MAX_POINTS points can
// use (root tables per length class, per-pass tables per pass length), so that
// the allocation-free entry point finds them built.
int prepare_noalloc() {
  if (g.device < 0 || g.device >= NA_MAX_DEVICES) return fail(BCG_E_ARG, "device index beyond the scratch table");
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return fail(BCG_E_HIP, "scratch address", e);
  g_scratch_ptr[g.device] = (unsigned char *)ptr;
  for (unsigned long long b = 32; 2 * b <= NOALLOC_MAX_POINTS; b *= 2) {
    for (unsigned long long N : {2 * b, 3 * b}) {
      if (N > NOALLOC_MAX_POINTS) continue;
      Plan p;
      if (!factor_cached(N, p)) continue;
      Roots r;
      if (int rc = get_roots(N, &r)) return rc;
      for (const Plan &c : g_cands[N])
        for (int i = 0; i < c.P; i++) {
          const cd *t = nullptr;
          if (int rc = get_ltab(c.L[i], &t)) return rc;
        }
    }
  }
  return 0;
}

// The engine's buffers pointed at the scratch for one call (under the engine
// lock), restored afterwards.
struct ScratchScope {
  DevBuf *bufs[7] = {&g.bufA, &g.bufB, &g.margin, &g.blk, &g.desc, &g.counters, &g.zbuf};
  DevBuf saved[7];
  explicit ScratchScope(unsigned char *base) {
    const size_t sizes[7] = {NA_ARR, NA_ARR, NA_MARGIN, NA_BLK, NA_DESC, NA_COUNTERS, NA_ZBUF};
    size_t off = 0;
    for (int i = 0; i < 7; i++) {
      saved[i] = *bufs[i];
      bufs[i]->p = base + off;
      bufs[i]->bytes = sizes[i];
      off += sizes[i];
    }
    g_fixed = true;
  }
  ~ScratchScope() {
    for (int i = 0; i < 7; i++) *bufs[i] = saved[i];
    g_fixed = false;
  }
};

}  // namespace

// =================================================================== C ABI

extern "C" {

int bcg_init(int device) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  return init_locked(device);
}

int bcg_shutdown(void) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  if (!g.ready) return 0;
  (void)hipStreamSynchronize(g.stream);
  g_defer.open = false;
  g_defer.n = 0;
  for (DevBuf *b : {&g.bufA, &g.bufB, &g.wA, &g.wB, &g.wC, &g.blk, &g.margin, &g.zbuf, &g.desc, &g.counters,
                    &g_defer.sets}) {
    if (b->p) (void)hipFree(b->p);
    b->p = nullptr;
    b->bytes = 0;
  }
  for (auto &r : g.roots) {
    (void)hipFree(r.second.lo);
    (void)hipFree(r.second.hi);
  }
  g.roots.clear();
  for (auto &t : g.ltab) (void)hipFree(t.second);
  g.ltab.clear();
  prof_collect();
  for (auto e : g.free_events) (void)hipEventDestroy(e);
  g.free_events.clear();
  for (int i = 0; i < Engine::NAUX; i++) {
    if (g.aux[i]) (void)hipStreamDestroy(g.aux[i]);
    if (g.join_ev[i]) (void)hipEventDestroy(g.join_ev[i]);
    g.aux[i] = nullptr;
    g.join_ev[i] = nullptr;
  }
  if (g.fork_ev) (void)hipEventDestroy(g.fork_ev);
  g.fork_ev = nullptr;
  (void)hipStreamDestroy(g.stream);
  (void)hipStreamDestroy(g.cap);
  g_bb.release();
  ep_trim();
  if (g_ring.p) (void)hipHostFree(g_ring.p);
  for (auto &e : g_ring.ev)
    if (e) (void)hipEventDestroy(e);
  g_ring = HostRing();
  g.stream = nullptr;
  g.cap = nullptr;
  g.ready = false;
  return 0;
}

const char *bcg_last_error(void) { return g_err.c_str(); }

void *bcg_stream(void) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  if (init_locked(0)) return nullptr;
  return (void *)g.stream;
}

int bcg_sync(void) {
  ENTER();
  HIPCHK(hipStreamSynchronize(g.stream));
  return 0;
}

int bcg_malloc(void **ptr, size_t bytes) {
  ENTER();
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess) return fail(BCG_E_NOMEM, "hipMalloc", e);
  return 0;
}

int bcg_free(void *ptr) {
  ENTER();
  if (ptr) HIPCHK(hipFree(ptr));
  return 0;
}

int bcg_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  ENTER();
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  return 0;
}

int bcg_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  ENTER();
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  return 0;
}

int bcg_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  ENTER();
  if (bytes % 4) return fail(BCG_E_ARG, "d2d copy size must be a multiple of 4 bytes");
  if (!bytes) return 0;
  if (!dst || !src) return fail(BCG_E_ARG, "null pointer");
  if (bytes % 16 == 0 && ((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0) {
    const unsigned long long n = bytes / 16;
    unsigned long long grid = (n + 255) / 256;
    if (grid > 16384) grid = 16384;
    LAUNCH(k_copy16, (int)grid, 256, g.stream, (uint4 *)dst, (const uint4 *)src, n);
  } else {
    const unsigned long long n = bytes / 4;
    LAUNCH(k_copy_words, grid_for(n), 256, g.stream, (uint32_t *)dst, (const uint32_t *)src, n, n);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(g.stream));
  return 0;
}

int bcg_mul_dev_bits(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL, int bits,
                     void *stream, bcg_mul_info *info) {
  ENTER();
  if ((AL && !A) || (BL && !B) || (CL && !C)) return fail(BCG_E_ARG, "null pointer");
  if (bits < 0 || bits > 24) return fail(BCG_E_ARG, "bits out of range");
  return mul_dev_locked(C, CL, A, AL, B, BL, bits, pick(stream), info);
}

int bcg_mul_dev(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL, void *stream,
                bcg_mul_info *info) {
  return bcg_mul_dev_bits(C, CL, A, AL, B, BL, 0, stream, info);
}

int bcg_mul_dev_noalloc(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL,
                        void *stream, bcg_mul_info *info) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  if (!g.ready) return fail(BCG_E_ARG, "call bcg_init before bcg_mul_dev_noalloc");
  HIPCHK(hipSetDevice(g.device));
  unsigned char *base = g_scratch_ptr[g.device];
  if (!base) return fail(BCG_E_HIP, "allocation-free scratch not set up");
  if ((AL && !A) || (BL && !B) || (CL && !C)) return fail(BCG_E_ARG, "null pointer");
  if (AL && BL && AL + BL >= 100) {
    Plan p;
    if (int rc = make_plan(AL, BL, 0, p)) return rc;
    if (p.N > NOALLOC_MAX_POINTS) return fail(BCG_E_ARG, "transform longer than the allocation-free limit (2^20 points)");
  }
  ScratchScope scope(base);
  return mul_dev_locked(C, CL, A, AL, B, BL, 0, pick(stream), info);
}

int bcg_mul_host(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL,
                 bcg_mul_info *info) {
  ENTER();
  if ((AL && !A) || (BL && !B) || (CL && !C)) return fail(BCG_E_ARG, "null pointer");
  hipStream_t s = g.stream;
  bool square = (A == B && AL == BL);
  if (int rc = ensure(g.wA, AL * 4 + 4)) return rc;
  if (!square)
    if (int rc = ensure(g.wB, BL * 4 + 4)) return rc;
  if (int rc = ensure(g.wC, CL * 4 + 4)) return rc;
  uint32_t *dA = (uint32_t *)g.wA.p;
  uint32_t *dB = square ? dA : (uint32_t *)g.wB.p;
  uint32_t *dC = (uint32_t *)g.wC.p;
  // through the pinned staging ring: C may alias A or B, and the operands are
  // in staging before the product is written back
  if (int rc = h2d(dA, A, AL * 4, s)) return rc;
  if (!square)
    if (int rc = h2d(dB, B, BL * 4, s)) return rc;
  if (int rc = mul_dev_locked(dC, CL, dA, AL, dB, BL, 0, s, info)) return rc;
  return d2h(C, dC, CL * 4, s);
}

int bcg_basecase_dev(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL,
                     void *stream) {
  ENTER();
  if (AL == 0 || BL == 0) return fill_zero(C, CL, pick(stream));
  return basecase_dev(C, CL, A, AL, B, BL, pick(stream));
}

int bcg_basecase_host(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL) {
  ENTER();
  hipStream_t s = g.stream;
  if (int rc = ensure(g.wA, AL * 4 + 4)) return rc;
  if (int rc = ensure(g.wB, BL * 4 + 4)) return rc;
  if (int rc = ensure(g.wC, CL * 4 + 4)) return rc;
  if (AL) HIPCHK(hipMemcpyAsync(g.wA.p, A, AL * 4, hipMemcpyHostToDevice, s));
  if (BL) HIPCHK(hipMemcpyAsync(g.wB.p, B, BL * 4, hipMemcpyHostToDevice, s));
  int rc = (AL == 0 || BL == 0) ? fill_zero((uint32_t *)g.wC.p, CL, s)
                                : basecase_dev((uint32_t *)g.wC.p, CL, (uint32_t *)g.wA.p, AL, (uint32_t *)g.wB.p, BL, s);
  if (rc) return rc;
  if (CL) HIPCHK(hipMemcpyAsync(C, g.wC.p, CL * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

}  // extern "C"

namespace {

// One batch call, planned on the host: tiny products (one wave each),
// one-pass products grouped by length class (one persistent k_mul1 launch per
// class), multi-pass products grouped by transform length (one launch per
// pass, the product in grid.y), and the products that go alone (aliasing or
// empty operands).
struct BatchPlan {
  std::vector<MulDesc> tiny, mdesc, cdesc;
  std::vector<std::pair<Plan, std::pair<size_t, size_t>>> mgroups;  // plan, (first, count) in mdesc
  std::vector<std::pair<int, std::pair<size_t, size_t>>> cgroups;   // L, (first, count) in cdesc
  std::vector<size_t> morder, corder, single;                        // product index of each entry
  std::vector<Plan> plans;
  // the call's arguments (kept for the products re-run alone)
  uint32_t *C = nullptr;
  const uint32_t *A = nullptr, *B = nullptr, *D = nullptr;
  std::vector<size_t> c_off, cl, a_off, al, b_off, bl, d_off, dl;
};

int batch_plan(BatchPlan &bp, uint32_t *C, const uint32_t *A, const uint32_t *B, const uint32_t *D,
               const size_t *c_off, const size_t *cl, const size_t *a_off, const size_t *al, const size_t *b_off,
               const size_t *bl, const size_t *d_off, const size_t *dl, size_t count) {
  bp = BatchPlan();
  bp.C = C;
  bp.A = A;
  bp.B = B;
  bp.D = D;
  bp.c_off.assign(c_off, c_off + count);
  bp.cl.assign(cl, cl + count);
  bp.a_off.assign(a_off, a_off + count);
  bp.al.assign(al, al + count);
  bp.b_off.assign(b_off, b_off + count);
  bp.bl.assign(bl, bl + count);
  bp.d_off.assign(count, 0);
  bp.dl.assign(count, 0);
  if (D && dl) {
    for (size_t i = 0; i < count; i++) {
      bp.dl[i] = dl[i];
      bp.d_off[i] = d_off ? d_off[i] : 0;
    }
  }
  bp.plans.assign(count, Plan());
  auto desc_of = [&](size_t i, const Plan *pl) {
    MulDesc d;
    d.a = A + a_off[i];
    d.b = B + b_off[i];
    d.c = C + c_off[i];
    d.d = bp.dl[i] ? D + bp.d_off[i] : nullptr;
    d.al = al[i];
    d.bl = bl[i];
    d.cl = cl[i];
    d.dl = bp.dl[i];
    d.bits = pl ? pl->bits : 0;
    d.cbits = pl ? coef_bits(*pl) : 0;
    d.gk = (pl && pl->P >= 2) ? group_words_for(pl->C[0], d.bits, d.cbits) : 0;
    return d;
  };
  std::map<int, std::vector<size_t>> classes;
  std::map<unsigned long long, std::vector<size_t>> mclasses;
  // the products of one tree level mostly share their lengths: plan each
  // length pair once
  std::map<std::pair<size_t, size_t>, Plan> plan_of;
  for (size_t i = 0; i < count; i++) {
    const uint32_t *pa = A + a_off[i];
    const uint32_t *pb = B + b_off[i];
    uint32_t *pc = C + c_off[i];
    const uint32_t *pd = bp.dl[i] ? D + bp.d_off[i] : nullptr;
    bool alias = (pc == pa || pc == pb);
    if (pd && pd == pc) return fail(BCG_E_ARG, "an addend may not alias its output");
    if (al[i] == 0 || bl[i] == 0 || cl[i] == 0) {
      bp.single.push_back(i);
      continue;
    }
    if (al[i] + bl[i] < 100) {
      if (alias || cl[i] > 4096)
        bp.single.push_back(i);
      else
        bp.tiny.push_back(desc_of(i, nullptr));
      continue;
    }
    auto pit = plan_of.find({al[i], bl[i]});
    if (pit == plan_of.end()) {
      Plan pl;
      if (int rc = make_plan(al[i], bl[i], 0, pl, true)) return rc;
      pit = plan_of.emplace(std::make_pair(al[i], bl[i]), pl).first;
    }
    bp.plans[i] = pit->second;
    if (alias)
      bp.single.push_back(i);
    else if (bp.plans[i].P == 1)
      classes[bp.plans[i].L[0]].push_back(i);
    else
      mclasses[bp.plans[i].N].push_back(i);
  }
  for (auto &kv : mclasses) {
    const size_t first = bp.mdesc.size();
    for (size_t i : kv.second) {
      bp.mdesc.push_back(desc_of(i, &bp.plans[i]));
      bp.morder.push_back(i);
    }
    bp.mgroups.push_back({bp.plans[kv.second[0]], {first, kv.second.size()}});  // same N and pass split
  }
  for (auto &kv : classes) {
    const size_t first = bp.cdesc.size();
    for (size_t i : kv.second) {
      bp.cdesc.push_back(desc_of(i, &bp.plans[i]));
      bp.corder.push_back(i);
    }
    bp.cgroups.push_back({kv.first, {first, kv.second.size()}});
  }
  return 0;
}

// Descriptors to the device and scratch sized (no allocation afterwards).
int batch_upload(const BatchPlan &bp, BatchBufs &bb, hipStream_t s) {
  const size_t nm = bp.mdesc.size(), nc = bp.cdesc.size();
  if (int rc = ensure(bb.tdesc, bp.tiny.size() * sizeof(MulDesc))) return rc;
  if (int rc = ensure(bb.mdesc, nm * sizeof(MulDesc))) return rc;
  if (int rc = ensure(bb.cdesc, nc * sizeof(MulDesc))) return rc;
  if (int rc = ensure(bb.margins, (nm + nc) * sizeof(unsigned int))) return rc;
  if (int rc = ensure(bb.counters, (bp.cgroups.size() + 1) * sizeof(int))) return rc;
  size_t arr = 0, blk = 16;
  for (auto &gr : bp.mgroups) {
    const Plan &p = gr.first;
    arr = std::max(arr, (size_t)p.N * sizeof(cd) * gr.second.second);
    const unsigned long long wpb = (unsigned long long)CT * CWPT;
    unsigned long long nb = 1;
    for (size_t k = 0; k < gr.second.second; k++)
      nb = std::max(nb, (unsigned long long)((bp.mdesc[gr.second.first + k].cl + wpb - 1) / wpb));
    blk = std::max(blk, (size_t)((nb + 15) & ~15ull) * gr.second.second);
  }
  if (arr) {
    if (int rc = ensure(bb.tA, arr)) return rc;
    if (int rc = ensure(bb.tB, arr)) return rc;
    if (int rc = ensure(bb.blk, blk)) return rc;
  }
  if (!bp.tiny.empty())
    if (int rc = h2d(bb.tdesc.p, bp.tiny.data(), bp.tiny.size() * sizeof(MulDesc), s)) return rc;
  if (nm)
    if (int rc = h2d(bb.mdesc.p, bp.mdesc.data(), nm * sizeof(MulDesc), s)) return rc;
  if (nc)
    if (int rc = h2d(bb.cdesc.p, bp.cdesc.data(), nc * sizeof(MulDesc), s)) return rc;
  return 0;
}

// Every batched launch of the plan, kernels only (capturable).
int batch_launch(const BatchPlan &bp, BatchBufs &bb, hipStream_t s) {
  const size_t nm = bp.mdesc.size(), nc = bp.cdesc.size();
  unsigned int *mm = (unsigned int *)bb.margins.p, *cm = mm + nm;
  if (nm + nc) LAUNCH(k_zero_u32, (int)((nm + nc + 255) / 256), 256, s, mm, (int)(nm + nc));
  if (!bp.cgroups.empty()) LAUNCH(k_zero_i32, 1, 64, s, (int *)bb.counters.p, (int)bp.cgroups.size());
  HIPCHK(hipGetLastError());
  // one-pass length classes: independent persistent launches, spread over the
  // side streams (largest class first), beside the multi-pass groups on s, so
  // that small classes do not leave the GPU idle one after another; the tiny
  // products go to a side stream too; joined back into s at the end. The
  // multi-pass groups, the longest chain of dependent launches, are issued
  // first (4096-product replay 1.27-1.28 -> 1.26 ms; issuing the largest
  // one-pass classes ahead of them measured 1.30 ms).
  const bool fork = !bp.cgroups.empty() || (!bp.tiny.empty() && !bp.mgroups.empty());
  if (fork) {
    HIPCHK(hipEventRecord(g.fork_ev, s));
    for (int i = 0; i < Engine::NAUX; i++) HIPCHK(hipStreamWaitEvent(g.aux[i], g.fork_ev, 0));
  }
  int rc = 0;
  for (auto &gr : bp.mgroups) {
    const size_t first = gr.second.first, n = gr.second.second;
    std::vector<MulDesc> hd(bp.mdesc.begin() + first, bp.mdesc.begin() + first + n);
    rc = run_multipass_class(gr.first, (const MulDesc *)bb.mdesc.p + first, hd, mm + first, s, bb);
    if (rc) break;
  }
  if (!rc && fork) {
    std::vector<int> order(bp.cgroups.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = (int)k;
    std::sort(order.begin(), order.end(), [&](int x, int y) {
      const double wx = (double)bp.cgroups[x].second.second * bp.cgroups[x].first;
      const double wy = (double)bp.cgroups[y].second.second * bp.cgroups[y].first;
      return wx > wy;
    });
    for (size_t k = 0; k < order.size() && !rc; k++) {
      const int ci = order[k];
      const auto &gr = bp.cgroups[ci];
      const size_t first = gr.second.first, n = gr.second.second;
      rc = run_class(gr.first, (const MulDesc *)bb.cdesc.p + first, (int)n, cm + first, (int *)bb.counters.p + ci,
                     g.aux[k % Engine::NAUX]);
    }
  }
  if (!rc && !bp.tiny.empty()) {
    hipStream_t ts = fork ? g.aux[bp.cgroups.size() % Engine::NAUX] : s;
    ProfScope ps(BCG_K_OTHER, ts);
    LAUNCH(k_basecase_batch, (int)bp.tiny.size(), 64, ts, (const MulDesc *)bb.tdesc.p, (int)bp.tiny.size());
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail(BCG_E_HIP, "k_basecase_batch", e);
  }
  if (fork) {  // the side streams rejoin s (also before an error returns)
    for (int i = 0; i < Engine::NAUX; i++) {
      HIPCHK(hipEventRecord(g.join_ev[i], g.aux[i]));
      HIPCHK(hipStreamWaitEvent(s, g.join_ev[i], 0));
    }
  }
  return rc;
}

// Margins back; products at >= 0.25 re-run alone with b - 1; the products
// that go alone run now.
int batch_finish(const BatchPlan &bp, BatchBufs &bb, hipStream_t s, double *worst_margin) {
  double worst = 0;
  const size_t nm = bp.mdesc.size(), nc = bp.cdesc.size();
  std::vector<size_t> retry;
  if (nm + nc) {
    std::vector<unsigned int> mb(nm + nc);
    if (int rc = d2h(mb.data(), bb.margins.p, (nm + nc) * sizeof(unsigned int), s)) return rc;
    for (size_t k = 0; k < nm + nc; k++) {
      float f;
      memcpy(&f, &mb[k], sizeof f);
      const size_t i = k < nm ? bp.morder[k] : bp.corder[k - nm];
      if (f >= 0.25f) {
        retry.push_back(i);
      } else {
        if (f > worst) worst = f;
        note_product(bp.plans[i].N, bp.plans[i].bits, f);
      }
    }
  }
  auto one = [&](size_t i, int force_bits) -> int {
    bcg_mul_info inf;
    const uint32_t *pd = bp.dl[i] ? bp.D + bp.d_off[i] : nullptr;
    if (int rc = mul_dev_locked(bp.C + bp.c_off[i], bp.cl[i], bp.A + bp.a_off[i], bp.al[i], bp.B + bp.b_off[i],
                                bp.bl[i], force_bits, s, &inf, pd, pd ? bp.dl[i] : 0))
      return rc;
    if (inf.margin > worst) worst = inf.margin;
    return 0;
  };
  for (size_t i : bp.single)
    if (int rc = one(i, 0)) return rc;
  for (size_t i : retry) {
    int fb = bp.plans[i].bits - 1;
    if (fb < 4) return fail(BCG_E_MARGIN, "rounding margin >= 0.25 at the smallest digit size");
    g_stats.retries++;
    if (int rc = one(i, fb)) return rc;
  }
  HIPCHK(hipStreamSynchronize(s));
  if (worst_margin) *worst_margin = worst;
  return 0;
}

}  // namespace

// A batch captured once and replayed (see include/bluecrunch_gpu.h).
struct bcg_batch_s {
  BatchPlan plan;
  BatchBufs bufs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" {

int bcg_muladd_batch_dev(uint32_t *C, const uint32_t *A, const uint32_t *B, const uint32_t *D, const size_t *c_off,
                         const size_t *cl, const size_t *a_off, const size_t *al, const size_t *b_off, const size_t *bl,
                         const size_t *d_off, const size_t *dl, size_t count, void *stream, double *worst_margin) {
  ENTER();
  hipStream_t s = pick(stream);
  BatchPlan bp;
  if (int rc = batch_plan(bp, C, A, B, D, c_off, cl, a_off, al, b_off, bl, d_off, dl, count)) return rc;
  if (int rc = batch_upload(bp, g_bb, s)) return rc;
  if (int rc = batch_launch(bp, g_bb, s)) return rc;
  return batch_finish(bp, g_bb, s, worst_margin);
}

int bcg_batch_create(bcg_batch *batch, uint32_t *C, const uint32_t *A, const uint32_t *B, const uint32_t *D,
                     const size_t *c_off, const size_t *cl, const size_t *a_off, const size_t *al, const size_t *b_off,
                     const size_t *bl, const size_t *d_off, const size_t *dl, size_t count) {
  ENTER();
  if (!batch) return fail(BCG_E_ARG, "null handle pointer");
  *batch = nullptr;
  bcg_batch_s *bt = new (std::nothrow) bcg_batch_s;
  if (!bt) return fail(BCG_E_NOMEM, "batch handle");
  hipStream_t s = g.stream;
  int rc = batch_plan(bt->plan, C, A, B, D, c_off, cl, a_off, al, b_off, bl, d_off, dl, count);
  if (!rc) rc = batch_upload(bt->plan, bt->bufs, s);
  // one eager run: builds every table the kernels read, checks the launches
  if (!rc) rc = batch_launch(bt->plan, bt->bufs, s);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = fail(BCG_E_HIP, "batch warm-up", e);
  }
  if (!rc) {
    const bool prof = g.prof;
    g.prof = false;  // no timing events inside the graph
    hipError_t e = hipStreamBeginCapture(g.cap, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
      rc = fail(BCG_E_HIP, "begin capture", e);
    } else {
      int lrc = batch_launch(bt->plan, bt->bufs, g.cap);
      hipGraph_t graph = nullptr;
      e = hipStreamEndCapture(g.cap, &graph);
      bt->graph = graph;
      if (lrc) rc = lrc;
      else if (e != hipSuccess || !graph) rc = fail(BCG_E_HIP, "end capture", e);
    }
    g.prof = prof;
  }
  if (!rc) {
    hipError_t e = hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = fail(BCG_E_HIP, "graph instantiate", e);
  }
  if (rc) {
    if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
    if (bt->graph) (void)hipGraphDestroy(bt->graph);
    bt->bufs.release();
    delete bt;
    return rc;
  }
  *batch = bt;
  return 0;
}

int bcg_batch_run(bcg_batch batch, void *stream, double *worst_margin) {
  ENTER();
  if (!batch || !batch->exec) return fail(BCG_E_ARG, "invalid batch handle");
  hipStream_t s = pick(stream);
  HIPCHK(hipGraphLaunch(batch->exec, s));
  return batch_finish(batch->plan, batch->bufs, s, worst_margin);
}

int bcg_batch_destroy(bcg_batch batch) {
  ENTER();
  if (!batch) return 0;
  HIPCHK(hipStreamSynchronize(g.stream));
  if (batch->exec) (void)hipGraphExecDestroy(batch->exec);
  if (batch->graph) (void)hipGraphDestroy(batch->graph);
  batch->bufs.release();
  delete batch;
  return 0;
}

int bcg_mul_batch_dev(uint32_t *C, const uint32_t *A, const uint32_t *B, const size_t *c_off, const size_t *cl,
                      const size_t *a_off, const size_t *al, const size_t *b_off, const size_t *bl, size_t count,
                      void *stream, double *worst_margin) {
  return bcg_muladd_batch_dev(C, A, B, nullptr, c_off, cl, a_off, al, b_off, bl, nullptr, nullptr, count, stream,
                              worst_margin);
}

int bcg_muladd_dev(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL,
                   const uint32_t *D, size_t DL, void *stream, bcg_mul_info *info) {
  ENTER();
  if ((AL && !A) || (BL && !B) || (CL && !C) || (DL && !D)) return fail(BCG_E_ARG, "null pointer");
  return mul_dev_locked(C, CL, A, AL, B, BL, 0, pick(stream), info, D, DL);
}

int bcg_fft_forward_dev(double *z, size_t length, void *stream) {
  ENTER();
  return transform_dev(z, length, false, pick(stream));
}

int bcg_fft_inverse_dev(double *z, size_t length, void *stream) {
  ENTER();
  return transform_dev(z, length, true, pick(stream));
}

int bcg_fft_pointwise_dev(double *a, const double *b, size_t length, void *stream) {
  ENTER();
  if (!length) return 0;
  hipStream_t s = pick(stream);
  LAUNCH(k_pointwise, grid_for(length), 256, s, (cd *)a, (const cd *)b, (unsigned long long)length);
  HIPCHK(hipGetLastError());
  return 0;
}

int bcg_int_to_fft_dev(double *z, size_t length, const uint32_t *A, size_t AL, int bits, size_t *points,
                       void *stream) {
  ENTER();
  if (bits < 1 || bits > 30) return fail(BCG_E_ARG, "int_to_fft: bits per point must be in [1, 30]");
  unsigned long long pts = AL ? digits_of(AL, bits) : 0;
  if (points) *points = (size_t)pts;
  unsigned long long n = length > pts ? length : pts;
  if (!n) return 0;
  hipStream_t s = pick(stream);
  LAUNCH(k_int_to_fft, grid_for(n), 256, s, (cd *)z, (unsigned long long)length, A, (unsigned long long)AL, bits, pts);
  HIPCHK(hipGetLastError());
  return 0;
}

int bcg_fft_to_int_dev(const double *z, size_t length, uint32_t *A, size_t AL, int bits, double *margin,
                       void *stream) {
  ENTER();
  if (bits < 1 || bits > 30) return fail(BCG_E_ARG, "fft_to_int: bits per point must be in [1, 30]");
  if (margin) *margin = 0;
  if (!AL) return 0;
  if (!length) return fail(BCG_E_ARG, "fft_to_int: zero length");
  hipStream_t s = pick(stream);
  unsigned long long pts = digits_of(AL, bits);
  if (int rc = ensure(g.zbuf, pts * sizeof(long long))) return rc;
  if (int rc = zero_margin(s)) return rc;
  LAUNCH(k_round_points, grid_for(pts), 256, s, (const cd *)z, pts, pts, 1.0 / (double)length, (long long *)g.zbuf.p,
         (unsigned int *)g.margin.p);
  HIPCHK(hipGetLastError());
  // r_k are < 2^53; bound the coefficient width by the largest possible value
  if (int rc = run_carry((const long long *)g.zbuf.p, pts, bits, 63, A, AL, s)) return rc;
  double m = 0;
  if (int rc = read_margin(s, &m)) return rc;
  if (margin) *margin = m;
  return 0;
}

int bcg_fft_forward_host(double *z, size_t length) {
  ENTER();
  if (int rc = ensure(g.bufA, length * 16 + 16)) return rc;
  HIPCHK(hipMemcpyAsync(g.bufA.p, z, length * 16, hipMemcpyHostToDevice, g.stream));
  if (int rc = transform_dev((double *)g.bufA.p, length, false, g.stream)) return rc;
  HIPCHK(hipMemcpyAsync(z, g.bufA.p, length * 16, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  return 0;
}

int bcg_fft_inverse_host(double *z, size_t length) {
  ENTER();
  if (int rc = ensure(g.bufA, length * 16 + 16)) return rc;
  HIPCHK(hipMemcpyAsync(g.bufA.p, z, length * 16, hipMemcpyHostToDevice, g.stream));
  if (int rc = transform_dev((double *)g.bufA.p, length, true, g.stream)) return rc;
  HIPCHK(hipMemcpyAsync(z, g.bufA.p, length * 16, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  return 0;
}

int bcg_fft_pointwise_host(double *a, const double *b, size_t length) {
  ENTER();
  if (!length) return 0;
  if (int rc = ensure(g.bufA, length * 16)) return rc;
  if (int rc = ensure(g.bufB, length * 16)) return rc;
  HIPCHK(hipMemcpyAsync(g.bufA.p, a, length * 16, hipMemcpyHostToDevice, g.stream));
  HIPCHK(hipMemcpyAsync(g.bufB.p, b, length * 16, hipMemcpyHostToDevice, g.stream));
  LAUNCH(k_pointwise, grid_for(length), 256, g.stream, (cd *)g.bufA.p, (const cd *)g.bufB.p,
         (unsigned long long)length);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(a, g.bufA.p, length * 16, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  return 0;
}

int bcg_int_to_fft_host(double *z, size_t length, const uint32_t *A, size_t AL, int bits, size_t *points) {
  ENTER();
  if (bits < 1 || bits > 30) return fail(BCG_E_ARG, "int_to_fft: bits per point must be in [1, 30]");
  unsigned long long pts = AL ? digits_of(AL, bits) : 0;
  unsigned long long n = length > pts ? length : pts;
  if (points) *points = (size_t)pts;
  if (!n) return 0;
  if (int rc = ensure(g.bufA, n * 16)) return rc;
  if (int rc = ensure(g.wA, AL * 4 + 4)) return rc;
  if (AL) HIPCHK(hipMemcpyAsync(g.wA.p, A, AL * 4, hipMemcpyHostToDevice, g.stream));
  size_t dummy;
  if (int rc = bcg_int_to_fft_dev((double *)g.bufA.p, length, (const uint32_t *)g.wA.p, AL, bits, &dummy, g.stream))
    return rc;
  HIPCHK(hipMemcpyAsync(z, g.bufA.p, n * 16, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  return 0;
}

int bcg_fft_to_int_host(const double *z, size_t length, uint32_t *A, size_t AL, int bits, double *margin) {
  ENTER();
  if (bits < 1 || bits > 30) return fail(BCG_E_ARG, "fft_to_int: bits per point must be in [1, 30]");
  if (!AL) {
    if (margin) *margin = 0;
    return 0;
  }
  // like fft/convert.c, ceil(32 AL / bits) points are read whatever `length`
  // is (length only sets the 1/length scale)
  unsigned long long pts = digits_of(AL, bits);
  if (int rc = ensure(g.bufA, pts * 16 + 16)) return rc;
  if (int rc = ensure(g.wC, AL * 4 + 4)) return rc;
  HIPCHK(hipMemcpyAsync(g.bufA.p, z, pts * 16, hipMemcpyHostToDevice, g.stream));
  if (int rc = bcg_fft_to_int_dev((const double *)g.bufA.p, length, (uint32_t *)g.wC.p, AL, bits, margin, g.stream))
    return rc;
  HIPCHK(hipMemcpyAsync(A, g.wC.p, AL * 4, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  return 0;
}

int bcg_tft_forward_host(double *z, size_t len, int k) {
  ENTER();
  if (k < 0 || k > 40 || len > (1ull << k)) return fail(BCG_E_ARG, "tft_forward: need len <= 2^k");
  unsigned long long n = 1ull << k;
  if (int rc = ensure(g.bufA, n * 16)) return rc;
  hipStream_t s = g.stream;
  if (len) HIPCHK(hipMemcpyAsync(g.bufA.p, z, len * 16, hipMemcpyHostToDevice, s));
  cd *d = (cd *)g.bufA.p;
  if (n > len) LAUNCH(k_zero_range, grid_for(n - len), 256, s, d, (unsigned long long)len, n);
  if (n >= 2)
    if (int rc = transform_dev((double *)d, n, false, s)) return rc;
  if (n > len) LAUNCH(k_zero_range, grid_for(n - len), 256, s, d, (unsigned long long)len, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(z, d, n * 16, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int bcg_tft_inverse_host(double *z, size_t len, int k) {
  ENTER();
  if (k < 0 || k > 40 || len > (1ull << k)) return fail(BCG_E_ARG, "tft_inverse: need len <= 2^k");
  unsigned long long n = 1ull << k;
  if (int rc = ensure(g.bufA, n * 16)) return rc;
  hipStream_t s = g.stream;
  cd *d = (cd *)g.bufA.p;
  if (len) HIPCHK(hipMemcpyAsync(d, z, len * 16, hipMemcpyHostToDevice, s));
  if (int rc = tft_inverse_rec(d, len, n, d, s)) return rc;
  // unscaled like the reference inverse: 2^k * x, zeros after len
  LAUNCH(k_scale, grid_for(n), 256, s, d, (unsigned long long)len, (double)n);
  if (n > len) LAUNCH(k_zero_range, grid_for(n - len), 256, s, d, (unsigned long long)len, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(z, d, n * 16, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int bcg_plan_query(size_t AL, size_t BL, int *bits, size_t *points, int *passes) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  Plan p;
  if (int rc = make_plan(AL, BL, 0, p)) return rc;
  if (bits) *bits = p.bits;
  if (points) *points = (size_t)p.N;
  if (passes) *passes = p.P;
  return 0;
}


int bcg_plan_candidates(size_t N, int index, int *passes, int *L, int *C) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  if (!valid_length(N) || N < 64) return fail(BCG_E_ARG, "length must be 2^k or 3*2^k, >= 64");
  std::vector<Plan> c = factor_candidates(N);
  if (index < 0) return (int)c.size();
  if (index >= (int)c.size()) return fail(BCG_E_ARG, "candidate index out of range");
  const Plan &p = c[index];
  if (passes) *passes = p.P;
  for (int i = 0; i < 3; i++) {
    if (L) L[i] = i < p.P ? p.L[i] : 0;
    if (C) C[i] = i < p.P ? p.C[i] : 0;
  }
  return (int)c.size();
}

int bcg_plan_digit_bias(int extra_bits) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  if (extra_bits < 0 || extra_bits > 8) return fail(BCG_E_ARG, "digit bias must be in [0, 8]");
  g_digit_bias = extra_bits;
  return 0;
}

int bcg_plan_chunk(size_t bytes, int streams) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  if (streams < 1 || streams > 1 + Engine::NAUX) return fail(BCG_E_ARG, "chunk streams must be in [1, 4]");
  g_chunk_bytes = bytes;
  g_chunk_streams = streams;
  return 0;
}

int bcg_plan_force(size_t N, int index) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  if (index < 0) {
    g_tuned.erase(N);
    return 0;
  }
  if (index >= (int)factor_candidates(N).size()) return fail(BCG_E_ARG, "candidate index out of range");
  g_tuned[N] = index;
  return 0;
}

int bcg_plan_current(size_t N) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  Plan p;
  if (!factor_cached(N, p)) return fail(BCG_E_ARG, "no pass split for this length");
  auto it = g_cands.find(N);
  for (size_t i = 0; i < it->second.size(); i++) {
    const Plan &c = it->second[i];
    if (c.P == p.P && c.L[0] == p.L[0] && c.L[1] == p.L[1] && c.L[2] == p.L[2] && c.C[0] == p.C[0] &&
        c.C[1] == p.C[1])
      return (int)i;
  }
  return fail(BCG_E_ARG, "no pass split for this length");
}

int bcg_autotune(size_t AL, size_t BL, int square, int reps, int *chosen, double *times_ms, int max_out) {
  ENTER();
  if (AL == 0 || BL == 0 || AL + BL < 100) return fail(BCG_E_ARG, "autotune needs an FFT-sized product");
  if (square) BL = AL;
  if (reps < 1) reps = 1;
  Plan p0;
  if (int rc = make_plan(AL, BL, 0, p0)) return rc;
  const unsigned long long N = p0.N;
  std::vector<Plan> cands = factor_candidates(N);
  hipStream_t s = g.stream;
  uint32_t *dA = nullptr, *dB = nullptr, *dC = nullptr;
  HIPCHK(hipMalloc(&dA, AL * 4));
  HIPCHK(hipMalloc(&dB, BL * 4));
  HIPCHK(hipMalloc(&dC, (AL + BL) * 4));
  LAUNCH(k_fill_hash, grid_for(AL), 256, s, dA, (unsigned long long)AL, 1ull);
  LAUNCH(k_fill_hash, grid_for(BL), 256, s, dB, (unsigned long long)BL, 2ull);
  const uint32_t *pB = square ? dA : dB;
  hipEvent_t e0 = get_event(), e1 = get_event();
  int best = -1;
  double best_ms = 0;
  int rc = 0;
  auto saved = g_tuned.find(N) != g_tuned.end() ? g_tuned[N] : -1;
  for (int i = 0; i < (int)cands.size() && !rc; i++) {
    g_tuned[N] = i;
    bcg_mul_info inf;
    rc = mul_dev_locked(dC, AL + BL, dA, AL, pB, BL, 0, s, &inf);  // warm-up (tables, buffers)
    if (rc) break;
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps && !rc; r++) rc = mul_dev_locked(dC, AL + BL, dA, AL, pB, BL, 0, s, &inf);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    double per = ms / reps;
    if (times_ms && i < max_out) times_ms[i] = per;
    if (best < 0 || per < best_ms) {
      best = i;
      best_ms = per;
    }
  }
  if (rc || best < 0) {
    if (saved >= 0) g_tuned[N] = saved; else g_tuned.erase(N);
  } else {
    g_tuned[N] = best;
  }
  g.free_events.push_back(e0);
  g.free_events.push_back(e1);
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  if (rc) return rc;
  if (chosen) *chosen = best;
  return 0;
}

}  // extern "C"

namespace {

// Operand words (square products) whose plan has transform length N, the
// longest such (0 if no square product lands on N).
size_t words_for_class(unsigned long long N) {
  size_t lo = 1, hi = (size_t)N;  // 32 w / b digits fill at most 2N slots
  Plan p;
  if (make_plan(lo, lo, 0, p) || p.N > N) return 0;
  while (hi - lo > 1) {
    const size_t mid = lo + (hi - lo) / 2;
    if (!make_plan(mid, mid, 0, p) && p.N <= N)
      lo = mid;
    else
      hi = mid;
  }
  if (make_plan(lo, lo, 0, p) || p.N != N) return 0;
  return lo;
}

// Milliseconds of one batched call's launches over `count` synthetic products
// of w x w words (plan and upload untimed), averaged over reps.
int time_batch(size_t w, int count, int reps, uint32_t *dA, uint32_t *dB, uint32_t *dC, double *ms) {
  std::vector<size_t> off(count), cl(count, 2 * w), ab_off(count), len(count, w);
  for (int i = 0; i < count; i++) {
    off[i] = (size_t)i * 2 * w;
    ab_off[i] = (size_t)i * w;
  }
  hipStream_t s = g.stream;
  BatchPlan bp;
  if (int rc = batch_plan(bp, dC, dA, dB, nullptr, off.data(), cl.data(), ab_off.data(), len.data(), ab_off.data(),
                          len.data(), nullptr, nullptr, (size_t)count))
    return rc;
  if (int rc = batch_upload(bp, g_bb, s)) return rc;
  if (int rc = batch_launch(bp, g_bb, s)) return rc;  // warm-up
  HIPCHK(hipStreamSynchronize(s));
  hipEvent_t e0 = get_event(), e1 = get_event();
  int rc = 0;
  (void)hipEventRecord(e0, s);
  for (int r = 0; r < reps && !rc; r++) rc = batch_launch(bp, g_bb, s);
  (void)hipEventRecord(e1, s);
  (void)hipEventSynchronize(e1);
  float t = 0;
  (void)hipEventElapsedTime(&t, e0, e1);
  g.free_events.push_back(e0);
  g.free_events.push_back(e1);
  *ms = t / reps;
  return rc;
}

}  // namespace

extern "C" {

int bcg_autotune_classes(size_t lo, size_t hi, int reps, int *tuned) {
  ENTER();
  if (reps < 1) reps = 1;
  if (hi > ((size_t)1 << 27)) return fail(BCG_E_ARG, "length classes above 2^27 points are not tuned");
  std::vector<unsigned long long> classes;
  for (unsigned long long b = 32; 2 * b <= hi; b *= 2)
    for (unsigned long long N : {2 * b, 3 * b})
      if (N >= lo && N <= hi) classes.push_back(N);
  // the largest class sets the scratch: about 2^23 points per timed batch
  size_t maxw = 0;
  std::vector<size_t> words(classes.size());
  std::vector<int> counts(classes.size());
  for (size_t k = 0; k < classes.size(); k++) {
    words[k] = words_for_class(classes[k]);
    const unsigned long long N = classes[k];
    counts[k] = (int)std::max(1ull, std::min(256ull, (8ull << 20) / N));
    maxw = std::max(maxw, words[k] * (size_t)counts[k]);
  }
  uint32_t *dA = nullptr, *dB = nullptr, *dC = nullptr;
  HIPCHK(hipMalloc(&dA, maxw * 4 + 4));
  HIPCHK(hipMalloc(&dB, maxw * 4 + 4));
  HIPCHK(hipMalloc(&dC, maxw * 8 + 8));
  LAUNCH(k_fill_hash, grid_for(maxw), 256, g.stream, dA, (unsigned long long)maxw, 11ull);
  LAUNCH(k_fill_hash, grid_for(maxw), 256, g.stream, dB, (unsigned long long)maxw, 12ull);
  int rc = 0, done = 0;
  for (size_t k = 0; k < classes.size() && !rc; k++) {
    const unsigned long long N = classes[k];
    if (!words[k]) continue;
    const int nc = (int)factor_candidates(N).size();
    if (nc < 2) continue;  // nothing to choose
    // each product's top word is nonzero in the hash fill only at the end of
    // the whole array; the plan depends on lengths alone
    int best = -1;
    double best_ms = 0;
    const auto saved = g_tuned_batch.find(N) != g_tuned_batch.end() ? g_tuned_batch[N] : -1;
    for (int i = 0; i < nc && !rc; i++) {
      g_tuned_batch[N] = i;
      double ms = 0;
      rc = time_batch(words[k], counts[k], reps, dA, dB, dC, &ms);
      if (!rc && (best < 0 || ms < best_ms)) {
        best = i;
        best_ms = ms;
      }
    }
    if (rc || best < 0) {
      if (saved >= 0) g_tuned_batch[N] = saved;
      else g_tuned_batch.erase(N);
    } else {
      g_tuned_batch[N] = best;
      done++;
    }
  }
  (void)hipStreamSynchronize(g.stream);
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  if (rc) return rc;
  if (tuned) *tuned = done;
  return 0;
}

int bcg_plan_force_batch(size_t N, int index) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  if (index < 0) {
    g_tuned_batch.erase(N);
    return 0;
  }
  if (index >= (int)factor_candidates(N).size()) return fail(BCG_E_ARG, "candidate index out of range");
  g_tuned_batch[N] = index;
  return 0;
}

int bcg_plan_current_batch(size_t N) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  Plan p;
  if (!factor_cached(N, p, true)) return fail(BCG_E_ARG, "no pass split for this length");
  auto it = g_cands.find(N);
  for (size_t i = 0; i < it->second.size(); i++) {
    const Plan &c = it->second[i];
    if (c.P == p.P && c.L[0] == p.L[0] && c.L[1] == p.L[1] && c.L[2] == p.L[2] && c.C[0] == p.C[0] &&
        c.C[1] == p.C[1])
      return (int)i;
  }
  return fail(BCG_E_ARG, "no pass split for this length");
}

size_t bcg_e_terms(size_t words) { return e_terms_of(words); }

int bcg_e_bsr_dev(size_t terms, uint32_t **pq, size_t *pl, size_t *ql) {
  ENTER();
  if (!pq || !pl || !ql) return fail(BCG_E_ARG, "null pointer");
  const int rc = bsr_device(terms, pq, pl, ql, g.stream);
  if (!rc) ep_detach(*pq);
  ep_trim(EP_KEEP);
  return rc;
}

int bcg_div_fraction_dev(const uint32_t *P, size_t pl, const uint32_t *Q, size_t ql, size_t nw, uint32_t **F) {
  ENTER();
  if (!P || !Q || !pl || !ql || !nw || !F) return fail(BCG_E_ARG, "bad arguments");
  DInt f;
  int rc = d_divide_fraction(P, pl, Q, ql, nw, f, g.stream);
  if (!rc) rc = d_floor_fix(P, pl, Q, ql, nw, f, g.stream);
  if (!rc && hipStreamSynchronize(g.stream) != hipSuccess) rc = fail(BCG_E_HIP, "division");
  if (!rc) {
    ep_detach(f.p);
    *F = f.p;
  } else {
    dfree(f);
  }
  ep_trim(EP_KEEP);
  return rc;
}

int bcg_e_digits(size_t digits, char *out, size_t cap, double *ms) {
  ENTER();
  if (!out) return fail(BCG_E_ARG, "null output");
  const int rc = e_digits(digits, out, cap, ms, g.stream);
  ep_trim(EP_KEEP);
  return rc;
}

int bcg_e_digits_ex(size_t digits, char *out, size_t cap, double *ms, int flags) {
  ENTER();
  if (!out) return fail(BCG_E_ARG, "null pointer");
  if (flags & ~BCG_E_PER_MULTIPLY) return fail(BCG_E_ARG, "unknown flags");
  const int rc = e_digits(digits, out, cap, ms, g.stream, (flags & BCG_E_PER_MULTIPLY) != 0);
  ep_trim(EP_KEEP);
  return rc;
}

int bcg_radix_dev(const uint32_t *F, size_t nf, size_t groups, uint32_t *digits) {
  ENTER();
  if (!F || !nf || !digits) return fail(BCG_E_ARG, "bad arguments");
  if (!groups) return 0;
  const int rc = radix_convert(F, nf, groups, digits, g.stream);
  ep_trim(EP_KEEP);
  return rc;
}

// Device bigfloats (bigfloat_dev.inc): operands copied so that the target may
// be one of them; the target's old mantissa is released by the operation.
int bcg_bf_set(bcg_bigfloat *x, const uint32_t *words, size_t len, int64_t exp, int sign) {
  ENTER();
  if (!x || (len && !words)) return fail(BCG_E_ARG, "null pointer");
  uint32_t *w = nullptr;
  if (len) {
    hipError_t e = ep_malloc(&w, len * sizeof(uint32_t));
    if (e != hipSuccess) return fail(BCG_E_NOMEM, "bigfloat words", e);
    if (int rc = h2d(w, words, len * sizeof(uint32_t), g.stream)) {
      bf_release(w);
      return rc;
    }
  }
  bf_release(x->coef);
  x->coef = w;
  x->len = len;
  x->exp = len ? exp : 0;
  x->sign = len ? (sign ? 1 : 0) : 1;
  return 0;
}

int bcg_bf_get(const bcg_bigfloat *x, uint32_t *words) {
  ENTER();
  if (!x || (x->len && (!words || !x->coef))) return fail(BCG_E_ARG, "null pointer");
  if (!x->len) return 0;
  return d2h(words, x->coef, x->len * sizeof(uint32_t), g.stream);
}

int bcg_bf_free(bcg_bigfloat *x) {
  ENTER();
  if (!x) return 0;
  bf_release(x->coef);
  bf_make_zero(*x);
  ep_trim(EP_KEEP);
  return 0;
}

}  // extern "C"

namespace {
template <class F>
int bf_call(bcg_bigfloat *t, const bcg_bigfloat *a, const bcg_bigfloat *b, F op) {
  if (!t || !a || (a->len && !a->coef) || (b && b->len && !b->coef)) return fail(BCG_E_ARG, "null pointer");
  bcg_bigfloat T = *t, A = *a, B = b ? *b : bcg_bigfloat();
  const int rc = op(T, A, B);
  *t = T;
  ep_trim(EP_KEEP);
  return rc;
}
}  // namespace

extern "C" {

int bcg_bf_add(bcg_bigfloat *t, const bcg_bigfloat *a, const bcg_bigfloat *b, size_t p) {
  ENTER();
  if (!b) return fail(BCG_E_ARG, "null pointer");
  return bf_call(t, a, b, [&](bcg_bigfloat &T, bcg_bigfloat &A, bcg_bigfloat &B) {
    return bf_add(T, A, B, p, g.stream);
  });
}

int bcg_bf_sub(bcg_bigfloat *t, const bcg_bigfloat *a, const bcg_bigfloat *b, size_t p) {
  ENTER();
  if (!b) return fail(BCG_E_ARG, "null pointer");
  return bf_call(t, a, b, [&](bcg_bigfloat &T, bcg_bigfloat &A, bcg_bigfloat &B) {
    return bf_sub(T, A, B, p, g.stream);
  });
}

int bcg_bf_mul(bcg_bigfloat *t, const bcg_bigfloat *a, const bcg_bigfloat *b, size_t p) {
  ENTER();
  if (!b) return fail(BCG_E_ARG, "null pointer");
  return bf_call(t, a, b, [&](bcg_bigfloat &T, bcg_bigfloat &A, bcg_bigfloat &B) {
    return bf_mul(T, A, B, p, g.stream);
  });
}

int bcg_bf_rcp(bcg_bigfloat *t, const bcg_bigfloat *a, size_t p) {
  ENTER();
  return bf_call(t, a, nullptr, [&](bcg_bigfloat &T, bcg_bigfloat &A, bcg_bigfloat &) {
    return bf_rcp(T, A, p, g.stream);
  });
}

int bcg_bf_div(bcg_bigfloat *t, const bcg_bigfloat *a, const bcg_bigfloat *b, size_t p) {
  ENTER();
  if (!b) return fail(BCG_E_ARG, "null pointer");
  return bf_call(t, a, b, [&](bcg_bigfloat &T, bcg_bigfloat &A, bcg_bigfloat &B) {
    return bf_div(T, A, B, p, g.stream);
  });
}

int bcg_test_fail_alloc(long nth) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  if (nth < 0) return fail(BCG_E_ARG, "nth must be >= 0");
  g_ep_fail_at = nth;
  return 0;
}

int bcg_prof_enable(int on) {
  ENTER();
  if (!on) prof_collect();
  g.prof = on != 0;
  return 0;
}

int bcg_stats_reset(void) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  g_stats = bcg_stats();
  return 0;
}

int bcg_defer_begin(void) {
  ENTER();
  return defer_begin();
}

int bcg_defer_end(double *worst_margin) {
  ENTER();
  return defer_end(worst_margin);
}

int bcg_stats_get(bcg_stats *out) {
  std::lock_guard<std::recursive_mutex> lk(g_lock);
  if (!out) return fail(BCG_E_ARG, "null pointer");
  *out = g_stats;
  return 0;
}

int bcg_prof_select(unsigned kinds) {
  ENTER();
  g.prof_kinds = kinds;
  return 0;
}

int bcg_prof_read(int kind, long *launches, double *ms) {
  ENTER();
  if (kind < 0 || kind >= BCG_PROF_KINDS) return fail(BCG_E_ARG, "bad kernel kind");
  prof_collect();
  if (launches) *launches = g.prof_n[kind];
  if (ms) *ms = g.prof_ms[kind];
  return 0;
}

int bcg_prof_reset(void) {
  ENTER();
  prof_collect();
  for (int i = 0; i < BCG_PROF_KINDS; i++) {
    g.prof_n[i] = 0;
    g.prof_ms[i] = 0;
  }
  return 0;
}

}  // extern "C"
