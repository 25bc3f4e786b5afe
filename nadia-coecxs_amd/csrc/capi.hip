// capi.hip -- implementation of the C ABI in include/nadia_amd.h.
//
// Device data of one reconstruction (nx x ny, real type T):
//   X  : the caller's estimate (a nadia_field), complex, x-major
//   W  : work field(s) of the fused path (2 fields when the algorithm has a
//        P_F P_S term), complex
//   S  : normalised support, real
//   Q  : sqrt(I) with the beam-stop mask in the sign bit, real
//   L  : (Fresnel) white field in the detector plane, Ls at the sample, complex
//   t0..t3, conv : temporaries of the unfused path and of shrinkwrap
// Fused iteration (Planar, Fresnel): column pass (forward column DFT, modulus
// projection, inverse column DFT) then row pass (inverse row DFT, algorithm
// update of X, forward row DFT of the new X); W holds the row DFT of X between
// iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nadia_amd.h"
#include "kernels.h"
#include "plan.h"

using namespace nadia;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(NADIA_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define LAUNCHCHK()                                                                         \
  do {                                                                                      \
    hipError_t e_ = hipGetLastError();                                                      \
    if (e_ != hipSuccess) return fail(NADIA_EHIP, "kernel launch: %s", hipGetErrorString(e_)); \
  } while (0)

// Page-locked staging for uploads from the caller's (pageable) host arrays:
// the array is copied into it on the host, the device copy is queued from it
// and the call returns without waiting for the device.  Regions are handed
// out in ring order and reused only after the copies from them completed.
constexpr size_t kStageBytes = (size_t)256 << 20;

struct Staging {
  char* base = nullptr;
  size_t head = 0;
  struct Use {
    size_t off, len;
    hipEvent_t ev;
  };
  std::deque<Use> uses;            // in-flight copies, oldest first
  std::vector<hipEvent_t> spare;   // completed events for reuse
};

// Device tables of a transform of any length (kern_any.inc): the stage list
// and twiddles of the len-point plan and, for Bluestein, the chirp and the
// kernel transform.  One allocation, kept until the process ends.
struct AnyPlan {
  int n = 0, len = 0, nst = 0;
  int radix[kAnyMaxStages] = {};
  bool blue = false;
  const int* stages = nullptr;
  const void* tw = nullptr;
  const void* chirp = nullptr;
  const void* kern = nullptr;
};

struct DeviceCtx {
  hipStream_t stream = nullptr;
  Staging stage;
  std::mutex stage_mu;
  // this device's address of the file-scope `scratch` array (set once, by
  // the first allocation-free call on the device)
  void* scratch = nullptr;
  std::map<std::pair<int, int>, void*> twiddles;  // (n, precision) -> table
  std::map<std::pair<int, int>, struct AnyPlan> anyplans;  // (n, precision) -> any-length plan
  // side streams (default priority) and events that fork independent groups
  // of a batch off `stream` while a graph is captured
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  // page-locked host slots that the update pass writes each error into, one
  // per reconstruction (get_error() after a stream sync reads its slot
  // instead of copying from the device); allocated once, never freed,
  // slots recycled
  double* herr_pool = nullptr;  // host address
  double* herr_dev = nullptr;   // the same memory as the device addresses it
  std::vector<int> herr_free;
};
constexpr int kHostErrSlots = 8192;

std::mutex g_mu;
std::map<int, DeviceCtx> g_dev;

int current_device() {
  int d = 0;
  (void)hipGetDevice(&d);
  return d;
}

DeviceCtx* dev_ctx(int dev) {
  std::lock_guard<std::mutex> lk(g_mu);
  DeviceCtx& c = g_dev[dev];
  if (!c.stream) {
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) return nullptr;
  }
  return &c;
}

// A page-locked error slot of device context c: its host address (and
// device address in *dev, index in *slot), or null when none is left or the
// pool cannot be allocated (get_error() then copies from the device).
double* herr_acquire(DeviceCtx* c, double** dev, int* slot) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c->herr_pool) {
    void* p = nullptr;
    if (hipHostMalloc(&p, sizeof(double) * kHostErrSlots, hipHostMallocCoherent) != hipSuccess || !p) {
      (void)hipGetLastError();
      return nullptr;
    }
    void* d = nullptr;
    if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess || !d) {
      (void)hipGetLastError();
      (void)hipHostFree(p);
      return nullptr;
    }
    c->herr_pool = (double*)p;
    c->herr_dev = (double*)d;
    for (int i = kHostErrSlots - 1; i >= 0; --i) c->herr_free.push_back(i);
  }
  if (c->herr_free.empty()) return nullptr;
  *slot = c->herr_free.back();
  c->herr_free.pop_back();
  *dev = c->herr_dev + *slot;
  return c->herr_pool + *slot;
}

void herr_release(DeviceCtx* c, int slot) {
  if (slot < 0) return;
  std::lock_guard<std::mutex> lk(g_mu);
  c->herr_free.push_back(slot);
}

// Host copy into the staging ring: large arrays are split over a few
// threads (one thread's memcpy reaches about a third of the host's memory
// bandwidth).
void copy_host(char* dst, const void* src, size_t bytes) {
  constexpr size_t kChunk = (size_t)4 << 20;
  const int nt = bytes >= 4 * kChunk ? 4 : 1;
  if (nt == 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t part = ((bytes / nt) + 255) & ~(size_t)255;
  std::thread th[3];
  for (int t = 1; t < nt; t++) {
    const size_t a = t * part, b = std::min(bytes, (t + 1) * part);
    th[t - 1] = std::thread([=] { std::memcpy(dst + a, (const char*)src + a, b - a); });
  }
  std::memcpy(dst, src, std::min(bytes, part));
  for (int t = 1; t < nt; t++) th[t - 1].join();
}

// Queue a copy of `bytes` of pageable host memory to device memory `dst` on
// the device's stream and return without waiting for the device (through
// the staging ring; uploads larger than half of it copy directly and wait).
int upload(DeviceCtx* c, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  if (bytes > kStageBytes / 2) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
  }
  std::lock_guard<std::mutex> lk(c->stage_mu);
  Staging& st = c->stage;
  if (!st.base) HIPCHK(hipHostMalloc((void**)&st.base, kStageBytes, hipHostMallocDefault));
  const size_t len = (bytes + 255) & ~(size_t)255;
  size_t off = st.head;
  if (off + len > kStageBytes) off = 0;
  auto overlaps = [&](const Staging::Use& u) { return u.off < off + len && off < u.off + u.len; };
  while (std::any_of(st.uses.begin(), st.uses.end(), overlaps)) {
    HIPCHK(hipEventSynchronize(st.uses.front().ev));
    st.spare.push_back(st.uses.front().ev);
    st.uses.pop_front();
  }
  hipEvent_t ev;
  if (!st.spare.empty()) {
    ev = st.spare.back();
    st.spare.pop_back();
  } else {
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  copy_host(st.base + off, src, bytes);
  HIPCHK(hipMemcpyAsync(dst, st.base + off, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipEventRecord(ev, c->stream));
  st.uses.push_back({off, len, ev});
  st.head = off + len;
  return 0;
}

// Twiddle table of the Stockham plan for length n: per stage s, L*(r-1)
// values W_{L r}^{k m}, computed in long double and rounded once.
void* twiddles(DeviceCtx* c, int n, int prec, bool row) {
  const int p = row ? row_points(n) : col_points(n);
  const int key_n = row ? n : -n;
  auto key = std::make_pair(key_n, prec);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = c->twiddles.find(key);
    if (it != c->twiddles.end()) return it->second;
  }
  PlanDesc d = make_plan(n, p);
  if (d.nst <= 0) return nullptr;
  const int total = plan_twiddles(d);
  std::vector<double> re(total), im(total);
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int s = 0; s < d.nst; ++s) {
    const int r = d.r[s], L = d.L[s];
    for (int k = 0; k < L; ++k)
      for (int m = 1; m < r; ++m) {
        const long long t = ((long long)k * m) % ((long long)L * r);
        const long double a = -twopi * (long double)t / (long double)(L * r);
        re[d.tw_off[s] + k * (r - 1) + m - 1] = (double)cosl(a);
        im[d.tw_off[s] + k * (r - 1) + m - 1] = (double)sinl(a);
      }
  }
  void* dptr = nullptr;
  size_t esz = prec == NADIA_FP64 ? 16 : 8;
  if (hipMalloc(&dptr, (size_t)(total > 0 ? total : 1) * esz) != hipSuccess) return nullptr;
  hipError_t e = hipSuccess;
  if (prec == NADIA_FP64) {
    std::vector<double> h(2 * (size_t)total);
    for (int i = 0; i < total; ++i) { h[2 * i] = re[i]; h[2 * i + 1] = im[i]; }
    if (total) e = hipMemcpy(dptr, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
  } else {
    std::vector<float> h(2 * (size_t)total);
    for (int i = 0; i < total; ++i) { h[2 * i] = (float)re[i]; h[2 * i + 1] = (float)im[i]; }
    if (total) e = hipMemcpy(dptr, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {  // an unfilled table is never cached
    (void)hipFree(dptr);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  auto ins = c->twiddles.emplace(key, dptr);
  if (!ins.second) {  // another thread built it meanwhile
    (void)hipFree(dptr);
    return ins.first->second;
  }
  return dptr;
}

// ---- transforms of any length ---------------------------------------------
// Lengths without compiled single-launch passes run kern_any.inc's tile
// kernel per axis.  A length whose prime factors are all <= the precision's
// largest radix (13 in FP32, 7 in FP64) gets a direct mixed-radix plan
// (radix 8 first, then 4, 2, 3, 5, 7, 11, 13); any other length Bluestein's:
// X_k = c_k sum_j (x_j c_j) conj(c_{k-j}), c_j = exp(-i pi j^2 / n), the
// convolution computed with transforms of the smallest length >= 2n - 1
// whose prime factors are <= 7.
constexpr int kAnyMaxLen = 4096;

bool any_factor(int n, int maxr, std::vector<int>& rad) {
  rad.clear();
  int m = n;
  for (int r : {8, 4, 2}) while (m % r == 0) { rad.push_back(r); m /= r; }
  for (int r : {3, 5, 7, 11, 13}) {
    if (r > maxr) break;
    while (m % r == 0) { rad.push_back(r); m /= r; }
  }
  return m == 1 && (int)rad.size() <= kAnyMaxStages;
}

// DFT (sign -1) in long double of a length whose prime factors are <= 7:
// recursive decimation in time on the smallest prime factor p,
// X[k + m n/p] = sum_j W_n^{j (k + m n/p)} Y_j[k] over the p sub-transforms.
void host_dft(std::vector<long double>& re, std::vector<long double>& im) {
  const int n = (int)re.size();
  if (n == 1) return;
  int p = 2;
  while (n % p) ++p;
  const int m = n / p;
  std::vector<std::vector<long double>> sr(p, std::vector<long double>(m)), si(p, std::vector<long double>(m));
  for (int j = 0; j < p; ++j)
    for (int k = 0; k < m; ++k) {
      sr[j][k] = re[k * p + j];
      si[j][k] = im[k * p + j];
    }
  for (int j = 0; j < p; ++j) host_dft(sr[j], si[j]);
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int q = 0; q < n; ++q) {
    const int k = q % m;
    long double ar = 0, ai = 0;
    for (int j = 0; j < p; ++j) {
      const long long t = ((long long)j * q) % n;
      const long double a = -twopi * (long double)t / (long double)n;
      const long double c = cosl(a), sn = sinl(a);
      ar += sr[j][k] * c - si[j][k] * sn;
      ai += sr[j][k] * sn + si[j][k] * c;
    }
    re[q] = ar;
    im[q] = ai;
  }
}

// smallest length >= m whose prime factors are all <= 7
int next_smooth(int m) {
  for (int v = m;; ++v) {
    int t = v;
    for (int r : {2, 3, 5, 7})
      while (t % r == 0) t /= r;
    if (t == 1) return v;
  }
}

const AnyPlan* any_plan(DeviceCtx* c, int n, int prec) {
  const auto key = std::make_pair(n, prec);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = c->anyplans.find(key);
    if (it != c->anyplans.end()) return &it->second;
  }
  if (n < 2 || n > kAnyMaxLen) return nullptr;
  AnyPlan p;
  p.n = n;
  std::vector<int> rad;
  const int maxr = any_max_radix(prec == NADIA_FP64 ? 8 : 4);
  p.blue = !any_factor(n, maxr, rad);
  p.len = n;
  if (p.blue) {
    p.len = next_smooth(2 * n - 1);
    any_factor(p.len, maxr, rad);
  }
  p.nst = (int)rad.size();
  // stage list and twiddles W_{L r}^{k m} of the len-point plan
  std::vector<int> st(3 * kAnyMaxStages, 0);
  std::vector<long double> wr, wi;
  const long double twopi = 6.283185307179586476925286766559005768L, pi = twopi / 2;
  int L = 1;
  for (int s = 0; s < p.nst; ++s) {
    const int r = rad[s];
    p.radix[s] = r;
    st[3 * s] = r;
    st[3 * s + 1] = L;
    st[3 * s + 2] = (int)wr.size();
    for (int k = 0; k < L; ++k)
      for (int m = 1; m < r; ++m) {
        const long long t = ((long long)k * m) % ((long long)L * r);
        const long double a = -twopi * (long double)t / (long double)(L * r);
        wr.push_back(cosl(a));
        wi.push_back(sinl(a));
      }
    L *= r;
  }
  const size_t ntw = wr.size();
  std::vector<long double> cr, ci, kr, ki;
  if (p.blue) {
    cr.resize(n); ci.resize(n);
    for (int j = 0; j < n; ++j) {  // exp(-i pi j^2 / n), j^2 reduced mod 2n exactly
      const long long q = ((long long)j * j) % (2LL * n);
      const long double a = -pi * (long double)q / (long double)n;
      cr[j] = cosl(a);
      ci[j] = sinl(a);
    }
    kr.assign(p.len, 0.0L); ki.assign(p.len, 0.0L);
    for (int j = 0; j < n; ++j) {
      kr[j] = cr[j]; ki[j] = -ci[j];
      if (j > 0) { kr[p.len - j] = cr[j]; ki[p.len - j] = -ci[j]; }
    }
    host_dft(kr, ki);
    for (int j = 0; j < p.len; ++j) { kr[j] /= p.len; ki[j] /= p.len; }
  }
  const size_t esz = prec == NADIA_FP64 ? 16 : 8;
  const size_t o_tw = 256, o_ch = o_tw + ((ntw * esz + 255) & ~(size_t)255),
               o_k = o_ch + ((cr.size() * esz + 255) & ~(size_t)255), total = o_k + kr.size() * esz;
  std::vector<unsigned char> h(total, 0);
  std::memcpy(h.data(), st.data(), st.size() * sizeof(int));
  auto put = [&](size_t off, const std::vector<long double>& re, const std::vector<long double>& im) {
    for (size_t i = 0; i < re.size(); ++i) {
      if (prec == NADIA_FP64) {
        const double v[2] = {(double)re[i], (double)im[i]};
        std::memcpy(h.data() + off + i * esz, v, sizeof(v));
      } else {
        const float v[2] = {(float)re[i], (float)im[i]};
        std::memcpy(h.data() + off + i * esz, v, sizeof(v));
      }
    }
  };
  put(o_tw, wr, wi);
  put(o_ch, cr, ci);
  put(o_k, kr, ki);
  void* d = nullptr;
  if (hipMalloc(&d, total) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), total, hipMemcpyHostToDevice) != hipSuccess) {  // never cache an unfilled table
    (void)hipFree(d);
    return nullptr;
  }
  p.stages = (const int*)d;
  p.tw = (const char*)d + o_tw;
  p.chirp = p.blue ? (const char*)d + o_ch : nullptr;
  p.kern = p.blue ? (const char*)d + o_k : nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  auto ins = c->anyplans.emplace(key, p);
  if (!ins.second) (void)hipFree(d);  // another thread built it meanwhile
  return &ins.first->second;
}

// Launch shape of kern_any.inc's kernel for the sequences of one axis:
// sequences per workgroup, LDS stride, LDS class and workgroup size such
// that every stage's butterflies fit the threads' registers.
struct AnyShape {
  int cb, stride, lds_class, threads;
};

bool any_shape(const AnyPlan& p, int nseq, bool rows, int prec, AnyShape& sh) {
  const int esz = prec == NADIA_FP64 ? 16 : 8;
  const int V = any_vals(esz / 2);
  // rows: about 1k elements per workgroup; columns: 64-byte row segments,
  // down to 32 bytes while that leaves fewer workgroups than CUs (256)
  int cb = rows ? std::max(1, 1024 / p.len) : 64 / esz;
  if (!rows)
    while (cb * esz > 32 && (nseq + cb - 1) / cb < 256) cb /= 2;
  cb = std::max(1, std::min(cb, nseq));
  for (; cb >= 1; cb = cb > 1 ? cb / 2 : 0) {
    const int stride = (rows || cb == 1) ? p.len : p.len + 1;
    const size_t bytes = (size_t)cb * stride * esz;
    int cls = -1;
    for (int k = 0; k < 3 && cls < 0; ++k)
      if (bytes <= (size_t)kAnyLds[k]) cls = k;
    if (cls < 0) continue;
    int t = 64;
    for (int s = 0; s < p.nst; ++s) {
      const int r = p.radix[s], nb = std::max(1, V / r);
      const int need = ((p.len / r) * cb + nb - 1) / nb;
      t = std::max(t, need);
    }
    t = (t + 63) / 64 * 64;
    if (t > 1024) continue;
    sh = AnyShape{cb, stride, cls, t};
    return true;
  }
  return false;
}

// the transforms of length n along one axis of an nx x ny field
template <typename T>
int any_axis(DeviceCtx* c, void* buf, int nx, int ny, int dir, bool rows, hipStream_t st) {
  const int prec = sizeof(T) == 8 ? NADIA_FP64 : NADIA_FP32;
  const int n = rows ? ny : nx;
  if (n == 1) return 0;
  const AnyPlan* p = any_plan(c, n, prec);
  if (!p) return fail(NADIA_EHIP, "transform tables for length %d: allocation failed", n);
  AnyArgs a{};
  a.buf = buf;
  a.rows = rows ? 1 : 0;
  a.nseq = rows ? nx : ny;
  a.es = rows ? 1 : ny;
  a.ss = rows ? ny : 1;
  a.n = n;
  a.len = p->len;
  a.dir = dir;
  a.nst = p->nst;
  a.stages = p->stages;
  a.tw = p->tw;
  a.chirp = p->chirp;
  a.kern = p->kern;
  AnyShape sh{};
  if (!any_shape(*p, a.nseq, rows, prec, sh)) return fail(NADIA_ESIZE, "no launch shape for length %d", n);
  a.cb = sh.cb;
  a.stride = sh.stride;
  launch_any<T>(a, sh.lds_class, sh.threads, st);
  LAUNCHCHK();
  return 0;
}

// Both tables of an nx x ny transform, built before any stream capture (the
// allocation and the synchronous upload are not allowed inside one).
int ensure_twiddles(DeviceCtx* c, int nx, int ny, int prec) {
  if (!twiddles(c, ny, prec, true) || !twiddles(c, nx, prec, false))
    return fail(NADIA_EHIP, "twiddle table allocation failed");
  return 0;
}

enum { A_PSF = 0, A_PFS = 1, A_PS = 2, A_PF = 3, A_PI = 4 };

// Size of the device array behind nadia_field_magnitude_noalloc: the
// magnitudes of a 2048 x 2048 FP64 (or 4096 x 2048 FP32) field.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;

int alloc(void** p, size_t bytes);
int env_int(const char* name, int dflt);

}  // namespace

// Staging of nadia_field_magnitude_noalloc: part of the library's code
// object, so every GPU the library runs on holds its own copy from module
// load to process end and no call allocates it.  In host code this name is
// a host-side stand-in: kernels and copies use the per-device address from
// hipGetSymbolAddress (DeviceCtx::scratch).
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// This device's address of `scratch`, looked up once per device with that
// device current.
int noalloc_scratch(int dev, void** out) {
  DeviceCtx* c = dev_ctx(dev);
  if (!c) return fail(NADIA_EHIP, "stream creation failed");
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (c->scratch) {
      *out = c->scratch;
      return 0;
    }
  }
  int cur = 0;
  HIPCHK(hipGetDevice(&cur));
  if (cur != dev) HIPCHK(hipSetDevice(dev));
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != dev) (void)hipSetDevice(cur);
  if (e != hipSuccess || !p) return fail(NADIA_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  std::lock_guard<std::mutex> lk(g_mu);
  c->scratch = p;
  *out = p;
  return 0;
}

}  // namespace

struct nadia_field {
  int nx, ny, prec, dev;
  size_t esz;  // bytes per complex element
  void* d;
  unsigned long long version;  // bumped by every write
};

struct nadia_recon {
  nadia_field* est;
  int kind, prec, nx, ny, dev, n_best;
  DeviceCtx* ctx;
  size_t esz, rsz;
  long long npix;
  void* arena;  // one allocation holding S, Q, W, conv, rtmp and the small per-job buffers
  void *S, *Q, *beam;
  void* S8;     // the support as bytes (k_normalise; see JobDesc::s8)
  bool s_soft;  // the support was softened (S8 not kept)
  int has_beam, has_intensity;
  void* W;  // 2 fields
  void* t[4];
  void* conv;
  void* rtmp;
  double* partials;  // col-pass / reduction partial sums
  double* errs;      // ring of per-iteration errors
  double* scal;      // device scalars
  int ring_cap;
  double* last_err;  // device slot holding get_error()
  double h_err_cache;
  bool h_err_valid;
  // page-locked slot (device-visible) the update pass also writes each error
  // into; herr_fresh: it holds get_error() once the stream's work is done
  int herr_slot = -1;
  double* herr = nullptr;      // host address
  double* herr_dev = nullptr;  // device address
  bool herr_fresh = false;
  // algorithm state (BaseCDI)
  double beta;
  int algorithm;
  double alg[5];
  int path;
  // fused-path caches
  bool w_valid, w2_valid;
  unsigned long long w_version, w2_version, w2_sversion, s_version;
  // transmission constraint (not owned) and its Fresnel work field
  nadia_constraint* tc;
  void* ttmp;
  // Fresnel
  void *L, *Ls;
  double norm, wl, zfd, zfs, pixel, phase_factor;
  void *phx, *phy;  // per-axis factors of the Fresnel quadratic phase (complex T)
  bool fresnel_ready;
  // white field
  double z12, z23, wf_sx, wf_sy, wf_k, wf_norm;
  bool wf_ready;
  // n best: snapshot buffers (fixed), and on the device the kept errors by
  // rank, the rank -> buffer map, the buffer pointers and the last copy target
  std::vector<void*> best;
  double* dbest_err;
  int* dbest_perm;
  void** dbest_bufs;
  int* dbest_target;
  // tile variants of the fused passes (-1: not chosen yet)
  int col_v, row_v;
  bool tile_fixed;  // set by the caller, not the autotuner
  // timing
  bool timing;
  double pass_ms[2];
  long long pass_n[2];
  std::vector<hipEvent_t> ev_pool;
  // job descriptors of the fused passes (device): [0] this reconstruction,
  // [1] scratch copy used by the autotuner
  JobDesc* djob;
  JobDesc djob_host;  // what djob[0] holds (valid when djob_set)
  bool djob_set;
  double* tune_errs;
};

struct nadia_constraint {
  int dev, nx, ny;
  int nreg;
  std::vector<unsigned char> hmap;  // region number per pixel (host copy)
  std::vector<double> a1, a2, cm;   // per region; cm holds fixed values
  std::vector<int> fixed;
  int charge_flip, flip_sign, unity;
  unsigned char* map;  // device
  double* dpar;        // device: a1[MAXR], a2[MAXR], c_mean[MAXR]
  int* dfixed;         // device
  double* partials;    // device: kEwBlocks * MAXR * 2
  bool dirty, map_dirty;
  nadia_custom_constraint_fn custom;
  void* user;
};

struct PdFrame {
  nadia_recon* r;
  double x, y, alpha;
  void* result;  // the frame's transmission (Fresnel) or estimate (planar), lnx x lny
  void* weight;  // lnx x lny, element type of the frame
};

struct nadia_pd {
  int dev, prec;
  double beta, gamma;
  int parallel, scale, ipc;
  std::vector<PdFrame> frames;
  void* object;
  int nx, ny, x_min, y_min;
  bool weights_set;
  PdFrameDev* dframes;  // device copy of the frame geometry (scale_object)
  int dframes_cap;
  unsigned char* covered;  // scale_object's covered map of the object, per set of weights
  long long covered_cap;
  double* derr;            // the frames' errors gathered on the device, and a pinned host copy
  double* herr;
  int err_cap;
  nadia_batch* batch;   // parallel mode: the frames' iterations run as one batch
};

struct nadia_batch {
  int dev;
  std::vector<nadia_recon*> jobs;
  JobDesc* djobs;  // device descriptors, grouped
  int cap;
  JobDesc* hjobs;  // pinned staging of the descriptors; hev marks the last copy from it
  int hcap;
  hipEvent_t hev;
  bool hev_pending;
  // graph mode: uniform sweeps run as replays of a captured graph of
  // `gsteps` iterations; gsig describes the launches it was captured for
  // (membership generation, groups, jobs and their n-best state: the graph
  // holds each job's raw n-best and error-ring pointers)
  int graph_steps;
  hipGraphExec_t gexec;
  int gsteps;
  std::vector<long long> gsig;
  long long graph_replays;
  unsigned long long gen;  // bumped by every add / remove
};

namespace {

hipStream_t S_(const nadia_recon* r) { return r->ctx->stream; }

template <typename T>
int fft2d_T(DeviceCtx* c, void* buf, int nx, int ny, int dir) {
  const int prec = sizeof(T) == 8 ? NADIA_FP64 : NADIA_FP32;
  if (!size_native(nx) || !size_native(ny)) {  // lengths without compiled passes
    int rc;
    if ((rc = any_axis<T>(c, buf, nx, ny, dir, true, c->stream))) return rc;
    return any_axis<T>(c, buf, nx, ny, dir, false, c->stream);
  }
  void* twy = twiddles(c, ny, prec, true);
  void* twx = twiddles(c, nx, prec, false);
  if (!twy || !twx) return fail(NADIA_EHIP, "twiddle table allocation failed");
  RowArgs ra{};
  ra.mode = ROW_FFT;
  ra.nrows = nx;
  ra.ny = ny;
  ra.tw = twy;
  ra.src = buf;
  ra.dst = buf;
  ra.dir = dir;
  ra.scale = 1.0;
  launch_row<T>(ra, 1, 0, c->stream);
  LAUNCHCHK();
  ColArgs ca{};
  ca.mode = COL_FFT;
  ca.nx = nx;
  ca.ny = ny;
  ca.tw = twx;
  ca.buf = buf;
  ca.dir = dir;
  ca.scale = 1.0;
  static const int fft_cv = env_int("NADIA_FFT_COL_VARIANT", 0);
  launch_col<T>(ca, 1, fft_cv, c->stream);
  LAUNCHCHK();
  return 0;
}

int fft2d(DeviceCtx* c, void* buf, int nx, int ny, int prec, int dir) {
  return prec == NADIA_FP64 ? fft2d_T<double>(c, buf, nx, ny, dir) : fft2d_T<float>(c, buf, nx, ny, dir);
}

#define DISPATCH(prec, call) \
  do {                       \
    if ((prec) == NADIA_FP64) { using T = double; call; } \
    else { using T = float; call; }                        \
  } while (0)

int check_dims(int nx, int ny) {
  if (nx < 1 || ny < 1 || nx > kAnyMaxLen || ny > kAnyMaxLen)
    return fail(NADIA_ESIZE, "unsupported size %dx%d (supported per axis: 1..%d)", nx, ny, kAnyMaxLen);
  return 0;
}

// ---- operators of one reconstruction on an arbitrary field buffer -------

int op_invert(nadia_recon* r, void* buf) {
  DISPATCH(r->prec, k_invert<T>(buf, r->nx, r->ny, 1, S_(r)));
  LAUNCHCHK();
  return 0;
}

int op_mulfac(nadia_recon* r, void* buf, int dir) {
  if (r->kind == NADIA_FRESNEL)
    DISPATCH(r->prec, k_mulfac_fresnel<T>(buf, r->nx, r->ny, r->phase_factor, dir, S_(r)));
  else
    DISPATCH(r->prec, k_mulfac_wf<T>(buf, r->nx, r->ny, r->wf_sx, r->wf_sy, r->z12, r->z23, r->wf_k, r->wf_norm,
                                     dir, S_(r)));
  LAUNCHCHK();
  return 0;
}

int need_setup(nadia_recon* r) {
  if (r->kind == NADIA_FRESNEL && !r->fresnel_ready)
    return fail(NADIA_ESTATE, "FresnelCDI parameters not set (nadia_fresnel_setup)");
  if (r->kind == NADIA_FRESNEL_WF && !r->wf_ready)
    return fail(NADIA_ESTATE, "FresnelCDI_WF parameters not set (nadia_wf_setup)");
  return 0;
}

int op_to_detector(nadia_recon* r, void* buf) {
  int rc;
  if ((rc = need_setup(r))) return rc;
  if (r->kind == NADIA_PLANAR) {
    if ((rc = fft2d(r->ctx, buf, r->nx, r->ny, r->prec, -1))) return rc;
    return op_invert(r, buf);
  }
  if (r->kind == NADIA_FRESNEL) {
    if ((rc = op_invert(r, buf))) return rc;
    if ((rc = fft2d(r->ctx, buf, r->nx, r->ny, r->prec, -1))) return rc;
    return op_mulfac(r, buf, -1);
  }
  if ((rc = fft2d(r->ctx, buf, r->nx, r->ny, r->prec, -1))) return rc;
  if ((rc = op_mulfac(r, buf, -1))) return rc;
  if ((rc = op_invert(r, buf))) return rc;
  return fft2d(r->ctx, buf, r->nx, r->ny, r->prec, -1);
}

int op_from_detector(nadia_recon* r, void* buf) {
  int rc;
  if ((rc = need_setup(r))) return rc;
  if (r->kind == NADIA_PLANAR) {
    if ((rc = op_invert(r, buf))) return rc;
    return fft2d(r->ctx, buf, r->nx, r->ny, r->prec, +1);
  }
  if (r->kind == NADIA_FRESNEL) {
    if ((rc = op_mulfac(r, buf, +1))) return rc;
    if ((rc = fft2d(r->ctx, buf, r->nx, r->ny, r->prec, +1))) return rc;
    return op_invert(r, buf);
  }
  if ((rc = fft2d(r->ctx, buf, r->nx, r->ny, r->prec, +1))) return rc;
  if ((rc = op_invert(r, buf))) return rc;
  if ((rc = op_mulfac(r, buf, +1))) return rc;
  return fft2d(r->ctx, buf, r->nx, r->ny, r->prec, +1);
}

// scale_intensity writing the error into *slot
int op_scale_intensity(nadia_recon* r, void* buf, double* slot) {
  if (!r->has_intensity) return fail(NADIA_ESTATE, "intensity not set");
  int rc;
  if ((rc = need_setup(r))) return rc;
  if (r->kind == NADIA_FRESNEL) DISPATCH(r->prec, k_fresnel_add<T>(buf, r->L, 1.0, r->npix, S_(r)));
  DISPATCH(r->prec, k_scale_intensity<T>(buf, r->Q, r->nx, r->ny, r->partials, kEwBlocks, S_(r)));
  k_final_sum2(r->partials, kEwBlocks, slot, nullptr, S_(r));
  if (r->kind == NADIA_FRESNEL) DISPATCH(r->prec, k_fresnel_add<T>(buf, r->L, -1.0, r->npix, S_(r)));
  LAUNCHCHK();
  r->last_err = slot;
  r->herr_fresh = false;
  r->h_err_valid = false;
  return 0;
}

int op_project(nadia_recon* r, void* buf, double* slot) {
  int rc;
  if ((rc = op_to_detector(r, buf))) return rc;
  if ((rc = op_scale_intensity(r, buf, slot))) return rc;
  return op_from_detector(r, buf);
}

int tc_apply(nadia_constraint* tc, hipStream_t st, void* buf, int prec, const void* ill = nullptr, int nx = 0,
             int ny = 0);
int tc_stats(nadia_constraint* tc, hipStream_t st, const void* buf, int prec, const void* ill, int nx, int ny,
             TcDev* out);
int tc_view(nadia_constraint* tc, hipStream_t st, const void* ill, int nx, int ny, TcDev* out);
constexpr int kRowStatsRegions = 4;  // the update pass sums up to this many regions itself

// BaseCDI::apply_support (NADIA/src/BaseCDI.c++:130-134) and
// FresnelCDI::apply_support (NADIA/src/FresnelCDI.c++:203-215): support,
// then the transmission constraint on the estimate (planar) or on its
// transmission function (Fresnel: get_transmission_function applies it,
// set_transmission_function writes the exit-surface wave back).
int op_constraint(nadia_recon* r, void* buf);

int op_support(nadia_recon* r, void* buf) {
  DISPATCH(r->prec, k_support<T>(buf, r->S, r->npix, S_(r)));
  LAUNCHCHK();
  return op_constraint(r, buf);
}

// the transmission-constraint part of apply_support
int op_constraint(nadia_recon* r, void* buf) {
  if (!r->tc) return 0;
  if (r->kind == NADIA_FRESNEL) {
    // without a host function the kernels read T from the exit-surface wave
    // and write the wave back themselves (the same roundings as the three
    // steps below)
    if (!r->tc->custom) return tc_apply(r->tc, S_(r), buf, r->prec, r->Ls, r->nx, r->ny);
    int rc;
    if (!r->ttmp && (rc = alloc(&r->ttmp, r->npix * r->esz))) return rc;
    DISPATCH(r->prec, k_trans_from_esw<T>(buf, r->Ls, r->ttmp, r->npix, S_(r)));
    if ((rc = tc_apply(r->tc, S_(r), r->ttmp, r->prec))) return rc;
    DISPATCH(r->prec, k_esw_from_trans<T>(r->ttmp, r->Ls, buf, r->nx, r->ny, S_(r)));
    LAUNCHCHK();
    return 0;
  }
  return tc_apply(r->tc, S_(r), buf, r->prec);
}

int copy_field(nadia_recon* r, void* dst, const void* src) {
  HIPCHK(hipMemcpyAsync(dst, src, r->npix * r->esz, hipMemcpyDeviceToDevice, S_(r)));
  return 0;
}

// Unfused iteration following BaseCDI::iterate's data flow
// (NADIA/src/BaseCDI.c++:297-387) one kernel per reference operation.
int iterate_plain(nadia_recon* r, double* slot) {
  int rc;
  void* x = r->est->d;
  if (r->kind == NADIA_FRESNEL_WF) {  // FresnelCDI_WF::iterate, src/FresnelCDI_WF.c++:122-150
    if ((rc = op_from_detector(r, x))) return rc;
    if ((rc = op_support(r, x))) return rc;
    if ((rc = op_to_detector(r, x))) return rc;
    return op_scale_intensity(r, x, slot);
  }
  if (r->algorithm == NADIA_ER) {
    if ((rc = op_project(r, x, slot))) return rc;
    return op_support(r, x);
  }
  for (int n = 0; n < 4; n++)
    if (r->alg[n] != 0 && !r->t[n] && (rc = alloc(&r->t[n], r->npix * r->esz))) return rc;
  void *tPSF = r->t[0], *tPFS = r->t[1], *tPS = r->t[2], *tPF = r->t[3];
  const double* a = r->alg;
  if (a[A_PFS] != 0) {
    if ((rc = copy_field(r, tPFS, x))) return rc;
    if ((rc = op_support(r, tPFS))) return rc;
    if ((rc = op_project(r, tPFS, slot))) return rc;
  }
  if (a[A_PF] != 0) {
    if ((rc = copy_field(r, tPF, x))) return rc;
    if ((rc = op_project(r, tPF, slot))) return rc;
  }
  if (a[A_PS] != 0) {
    if ((rc = copy_field(r, tPS, x))) return rc;
    if ((rc = op_support(r, tPS))) return rc;
  }
  if (a[A_PSF] != 0) {
    if (a[A_PF] != 0) {
      if ((rc = copy_field(r, tPSF, tPF))) return rc;
      if ((rc = op_support(r, tPSF))) return rc;
    } else {
      if ((rc = copy_field(r, tPSF, x))) return rc;
      if ((rc = op_project(r, tPSF, slot))) return rc;
      if ((rc = op_support(r, tPSF))) return rc;
    }
  }
  // coefficient order expected by k_combine: a(PSF), b(PFS), c(PS), d(PF), e(PI)
  double coef[5] = {a[A_PSF], a[A_PFS], a[A_PS], a[A_PF], a[A_PI]};
  DISPATCH(r->prec, k_combine<T>(x, tPFS, tPF, tPS, tPSF, coef, r->npix, S_(r)));
  LAUNCHCHK();
  return 0;
}

bool has_projection(const nadia_recon* r) {
  if (r->kind == NADIA_FRESNEL_WF || r->algorithm == NADIA_ER) return true;
  const double* a = r->alg;
  return a[A_PF] != 0 || a[A_PSF] != 0 || a[A_PFS] != 0;
}

// a transmission constraint needs whole-field reductions between the
// projection and the update, so a constrained job does not share the fused
// launches (ER takes fused_tc_ok's path, the rest the unfused path)
// (lengths without compiled passes always take the unfused path)
bool native_dims(const nadia_recon* r) { return size_native(r->nx) && size_native(r->ny); }
bool fused_ok(const nadia_recon* r) {
  return r->path == NADIA_PATH_FUSED && native_dims(r) && has_projection(r) && !r->tc;
}

// ER with a transmission constraint: the fused passes give support(P_F x),
// the constraint then runs on the new estimate between iterations
bool fused_tc_ok(const nadia_recon* r) {
  return r->path == NADIA_PATH_FUSED && native_dims(r) && has_projection(r) && r->tc && r->algorithm == NADIA_ER &&
         r->kind != NADIA_FRESNEL_WF;
}

struct Ev {
  hipEvent_t a, b;
  int kind;
};

int record_start(nadia_recon* r, std::vector<Ev>& evs, int kind) {
  if (!r->timing) return 0;
  Ev e{};
  HIPCHK(hipEventCreate(&e.a));
  HIPCHK(hipEventCreate(&e.b));
  e.kind = kind;
  HIPCHK(hipEventRecord(e.a, S_(r)));
  evs.push_back(e);
  return 0;
}
int record_stop(nadia_recon* r, std::vector<Ev>& evs) {
  if (!r->timing) return 0;
  HIPCHK(hipEventRecord(evs.back().b, S_(r)));
  return 0;
}
int collect_events(nadia_recon* r, std::vector<Ev>& evs) {
  if (evs.empty()) return 0;
  HIPCHK(hipEventSynchronize(evs.back().b));
  for (auto& e : evs) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
    r->pass_ms[e.kind] += ms;
    r->pass_n[e.kind] += 1;
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  evs.clear();
  return 0;
}

// the row transform that makes W from the estimate (the fused passes' state
// between iterations)
RowArgs w_rebuild_args(const nadia_recon* r, const void* twy) {
  RowArgs ra{};
  ra.mode = ROW_FFT;
  ra.nrows = r->nx;
  ra.ny = r->ny;
  ra.tw = twy;
  ra.scale = 1.0;
  if (r->kind == NADIA_FRESNEL_WF) {
    ra.dir = +1;
    ra.pre_sign = 1;
  } else {
    ra.dir = -1;
  }
  return ra;
}

bool w_stale(const nadia_recon* r) { return !r->w_valid || r->w_version != r->est->version; }

template <typename T>
int fused_prepare(nadia_recon* r, bool two) {
  void* twy = twiddles(r->ctx, r->ny, r->prec, true);
  if (!twy) return fail(NADIA_EHIP, "twiddle table allocation failed");
  const long long n = r->npix;
  const int rv = r->row_v < 0 ? 0 : r->row_v;
  if (w_stale(r)) {
    RowArgs ra = w_rebuild_args(r, twy);
    ra.src = r->est->d;
    ra.dst = r->W;
    launch_row<T>(ra, 1, rv, S_(r));
    LAUNCHCHK();
    r->w_valid = true;
    r->w_version = r->est->version;
  }
  if (two && (!r->w2_valid || r->w2_version != r->est->version || r->w2_sversion != r->s_version)) {
    void* W2 = (char*)r->W + n * r->esz;
    HIPCHK(hipMemcpyAsync(W2, r->est->d, n * r->esz, hipMemcpyDeviceToDevice, S_(r)));
    k_support<T>(W2, r->S, n, S_(r));
    RowArgs ra{};
    ra.mode = ROW_FFT;
    ra.nrows = r->nx;
    ra.ny = r->ny;
    ra.tw = twy;
    ra.src = W2;
    ra.dst = W2;
    ra.dir = -1;
    ra.scale = 1.0;
    launch_row<T>(ra, 1, rv, S_(r));
    LAUNCHCHK();
    r->w2_valid = true;
    r->w2_version = r->est->version;
    r->w2_sversion = r->s_version;
  }
  return 0;
}

bool two_fields(const nadia_recon* r) {
  return r->kind != NADIA_FRESNEL_WF && r->algorithm != NADIA_ER && r->alg[A_PFS] != 0;
}

// Split work fields (PlanarCDI and FresnelCDI with one work field and no
// transmission constraint): the column pass leaves the projection in the
// second field, the first keeps the row transform of x, and the update pass
// stores x only when a sequence of iterations ends.  ER never reads x (17
// instead of 25 bytes per pixel in the update pass); the other algorithms
// rebuild x by one more inverse row transform instead of reading it (25
// instead of 33 bytes), which pays for rows of up to 1024 points (64-job
// batch of 256..1024: +5% FP32, +7% FP64) but not for 2048-point rows, whose
// update pass is already half transform work (2048^2 HIO: -7..-10%).  A
// fixed property of the reconstruction's set-up, so its trajectory does not
// depend on how its iterations are split into calls or batches.
bool split_fields(const nadia_recon* r) {
  return (r->kind == NADIA_PLANAR || r->kind == NADIA_FRESNEL) && !two_fields(r) && !r->tc &&
         (r->algorithm == NADIA_ER || r->ny <= 1024);
}

bool wants_n_best(const nadia_recon* r);

// Host copy of the reconstruction's job descriptor for the fused passes, with
// the estimate X and work fields W given explicitly (the autotuner points
// them at scratch copies).
JobDesc make_jobdesc(const nadia_recon* r, void* X, void* W, double* errs) {
  JobDesc j{};
  j.x = X;
  j.w = W;
  j.s = r->S;
  j.s8 = r->s_soft ? nullptr : (const unsigned char*)r->S8;
  j.q = r->Q;
  j.ill = r->L;
  j.partials = r->partials;
  j.errs = errs;
  // the fused update pass also writes each error into get_error()'s slot
  // (not the tuner's scratch descriptors, whose errors go to tune_errs)
  j.cur_err = errs == r->errs ? r->scal + 8 : nullptr;
  j.host_err = errs == r->errs ? r->herr_dev : nullptr;
  for (int i = 0; i < 5; i++) j.coef[i] = r->alg[i];
  j.er = r->algorithm == NADIA_ER;
  j.store_every = wants_n_best(r) ? 1 : 0;
  if (wants_n_best(r)) {
    j.nb_err = r->dbest_err;
    j.nb_perm = r->dbest_perm;
    j.nb_target = r->dbest_target;
    j.nb_n = r->n_best;
  }
  // get_error() reports the last projection the reference evaluates: P_F x
  // when a PF or PSF term exists, else P_F P_S x
  j.err_field = (r->algorithm == NADIA_ER || r->alg[A_PF] != 0 || r->alg[A_PSF] != 0) ? 0 : 1;
  j.phase_factor = r->phase_factor;
  j.fx = r->phx;
  j.fy = r->phy;
  j.wf_sx = r->wf_sx;
  j.wf_sy = r->wf_sy;
  j.wf_z12 = r->z12;
  j.wf_z23 = r->z23;
  j.wf_k = r->wf_k;
  j.wf_norm = r->wf_norm;
  return j;
}

// Reconstructions whose fused passes can share launches.
struct GroupKey {
  int prec, kind, nx, ny, two, split;
  bool operator<(const GroupKey& o) const {
    if (prec != o.prec) return prec < o.prec;
    if (kind != o.kind) return kind < o.kind;
    if (nx != o.nx) return nx < o.nx;
    if (ny != o.ny) return ny < o.ny;
    if (two != o.two) return two < o.two;
    return split < o.split;
  }
};

GroupKey group_key(const nadia_recon* r) {
  return GroupKey{r->prec, r->kind, r->nx, r->ny, two_fields(r) ? 1 : 0, split_fields(r) ? 1 : 0};
}

// Queue one fused iteration of `njobs` reconstructions sharing key g, their
// descriptors in djobs (device); error ring slot `iter`; column / row tile
// variants cv / rv; on stream st (default: the device's stream).  PlanarCDI and FresnelCDI: column pass (modulus
// projection) + row pass (update).  FresnelCDI_WF: four passes and a final
// error reduction.  evs (may be null) receives per-pass timing events.
// FresnelCDI_WF: the error of an iteration is reduced by the next
// iteration's first column pass; jobs [fin_first, fin_first + fin_count)
// (fin_count < 0: all) run no further iteration in this sequence and get a
// separate reduction after this one.
template <typename T>
int launch_fused(DeviceCtx* c, const GroupKey& g, const JobDesc* djobs, int njobs, int iter, int cv, int rv,
                 nadia_recon* timer, std::vector<Ev>* evs, hipStream_t st = nullptr, const TcDev* tcs = nullptr,
                 double* tc_part = nullptr, int fin_first = 0, int fin_count = -1, int rank = 0) {
  void* twy = twiddles(c, g.ny, g.prec, true);
  void* twx = twiddles(c, g.nx, g.prec, false);
  if (!twx || !twy) return fail(NADIA_EHIP, "twiddle table allocation failed");
  if (!st) st = c->stream;
  ColArgs ca{};
  ca.jobs = djobs;
  ca.nfields = g.two ? 2 : 1;
  ca.split = g.split;
  ca.nx = g.nx;
  ca.ny = g.ny;
  ca.tw = twx;
  ca.sf = 1.0 / std::sqrt((double)g.nx * (double)g.ny);
  ca.xmid = (g.nx - 1) / 2.0;
  ca.ymid = (g.ny - 1) / 2.0;
  RowArgs ra{};
  ra.jobs = djobs;
  ra.iter = iter;
  ra.nrows = g.nx;
  ra.ny = g.ny;
  ra.tw = twy;
  ra.two = g.two;
  ra.split = g.split;
  ra.xs_first = fin_first;
  ra.xs_count = fin_count;
  ra.rank = rank;
  ra.post = ca.sf;
  int rc;
  auto col = [&](int mode) -> int {
    ca.mode = mode;
    if (evs && (rc = record_start(timer, *evs, 0))) return rc;
    launch_col<T>(ca, njobs, cv, st);
    if (evs && (rc = record_stop(timer, *evs))) return rc;
    LAUNCHCHK();
    return 0;
  };
  auto row = [&](int mode) -> int {
    ra.mode = mode;
    if (evs && (rc = record_start(timer, *evs, 1))) return rc;
    launch_row<T>(ra, njobs, rv, st);
    if (evs && (rc = record_stop(timer, *evs))) return rc;
    LAUNCHCHK();
    return 0;
  };
  if (g.kind != NADIA_FRESNEL_WF) {
    ra.fin_n = g.ny / col_group<T>(g.nx);
    if ((rc = col(g.kind == NADIA_FRESNEL ? COL_FRESNEL : COL_PLANAR))) return rc;
    if (tcs) {  // ER with a constraint: the update pass also sums its regions
      ra.tc = *tcs;
      ra.tc_part = tc_part;
      return row(ROW_UPDTC);
    }
    return row(ROW_UPDATE);
  }
  ca.iter = iter;
  ca.fin_n = row_blocks<T>(g.nx, g.ny, rv);
  if ((rc = col(COL_WF1)) || (rc = row(ROW_WF2)) || (rc = col(COL_WF3)) || (rc = row(ROW_WF4))) return rc;
  if (fin_count < 0) fin_count = njobs - fin_first;
  if (fin_count > 0) {
    k_final_jobs(djobs + fin_first, fin_count, ca.fin_n, iter, st);
    LAUNCHCHK();
  }
  return 0;
}

int upload_jobs(hipStream_t st, JobDesc* dst, const JobDesc* src, int n) {
  HIPCHK(hipMemcpyAsync(dst, src, sizeof(JobDesc) * n, hipMemcpyHostToDevice, st));
  // the host array may be reused right after: wait for the copy
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

// ---- autotuner -------------------------------------------------------------
// Picks the column and row tile variants of the fused passes for one
// (precision, kind, nx, ny, field count) by timing real fused iterations on
// scratch copies of the estimate and work fields (the reconstruction's own
// state is not touched).  Choices are cached for the life of the process.
// NADIA_TUNE=0 disables it (default variants); NADIA_COL_VARIANT /
// NADIA_ROW_VARIANT force a choice.
std::map<GroupKey, std::pair<int, int>> g_tuned;

int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  return (v && *v) ? atoi(v) : dflt;
}

template <typename T>
int time_variant(nadia_recon* r, const JobDesc* dj, int cv, int rv, int reps, float* ms) {
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  int rc = 0;
  const GroupKey g = group_key(r);
  HIPCHK(hipEventRecord(a, S_(r)));
  for (int k = 0; k < reps && !rc; ++k)
    rc = launch_fused<T>(r->ctx, g, dj, 1, k, cv, rv, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                         k == reps - 1 ? 1 : 0);
  HIPCHK(hipEventRecord(b, S_(r)));
  HIPCHK(hipEventSynchronize(b));
  HIPCHK(hipEventElapsedTime(ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

constexpr int kTuneReps = 16;

template <typename T>
int autotune(nadia_recon* r) {
  if (r->tile_fixed && r->col_v >= 0 && r->row_v >= 0) return 0;
  const int fc = env_int("NADIA_COL_VARIANT", -1), fr = env_int("NADIA_ROW_VARIANT", -1);
  if (fc >= 0 || fr >= 0 || env_int("NADIA_TUNE", 1) == 0 || (!tuned_size(r->nx) && !tuned_size(r->ny))) {
    r->col_v = (fc >= 0 && fc < kColVariants) ? fc : 0;
    r->row_v = (fr >= 0 && fr < kRowVariants) ? fr : 0;
    return 0;
  }
  const GroupKey key = group_key(r);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      r->col_v = it->second.first;
      r->row_v = it->second.second;
      return 0;
    }
  }
  void *X = nullptr, *W = nullptr;
  int rc;
  if ((rc = alloc(&X, r->npix * r->esz))) return rc;
  if ((rc = alloc(&W, 2 * r->npix * r->esz))) {
    (void)hipFree(X);
    return rc;
  }
  HIPCHK(hipMemcpyAsync(X, r->est->d, r->npix * r->esz, hipMemcpyDeviceToDevice, S_(r)));
  HIPCHK(hipMemcpyAsync(W, r->W, 2 * r->npix * r->esz, hipMemcpyDeviceToDevice, S_(r)));
  const JobDesc hj = make_jobdesc(r, X, W, r->tune_errs);
  if ((rc = upload_jobs(S_(r), r->djob + 1, &hj, 1))) return rc;
  // each candidate: a warm-up, then the best of two timed runs
  auto measure = [&](int cv, int rv, float* best) -> int {
    float ms = 0;
    *best = 1e30f;
    if ((rc = time_variant<T>(r, r->djob + 1, cv, rv, 2, &ms))) return rc;
    for (int k = 0; k < 2; ++k) {
      if ((rc = time_variant<T>(r, r->djob + 1, cv, rv, kTuneReps, &ms))) return rc;
      if (ms < *best) *best = ms;
    }
    return 0;
  };
  int best_c = 0, best_r = 0;
  float best = 1e30f, ms = 0;
  for (int cv = 0; cv < kColVariants && !rc; ++cv) {
    if ((rc = measure(cv, 0, &ms))) break;
    if (ms < best) {
      best = ms;
      best_c = cv;
    }
  }
  best = 1e30f;
  for (int rv = 0; rv < kRowVariants && !rc; ++rv) {
    if ((rc = measure(best_c, rv, &ms))) break;
    if (ms < best) {
      best = ms;
      best_r = rv;
    }
  }
  (void)hipStreamSynchronize(S_(r));
  (void)hipFree(X);
  (void)hipFree(W);
  if (rc) return rc;
  r->col_v = best_c;
  r->row_v = best_r;
  std::lock_guard<std::mutex> lk(g_mu);
  g_tuned[key] = std::make_pair(best_c, best_r);
  return 0;
}

// Make W current and the tile variants chosen; upload the descriptor.
template <typename T>
int fused_ready(nadia_recon* r, bool own_desc = true) {
  int rc;
  if ((rc = fused_prepare<T>(r, two_fields(r)))) return rc;
  if ((rc = autotune<T>(r))) return rc;
  if (!own_desc) return 0;  // a batch launches from its own descriptors
  const JobDesc hj = make_jobdesc(r, r->est->d, r->W, r->errs);
  // unchanged since the last upload (a loop of single iterations): no copy
  if (r->djob_set && std::memcmp(&hj, &r->djob_host, sizeof(JobDesc)) == 0) return 0;
  r->djob_host = hj;
  r->djob_set = true;
  return upload_jobs(S_(r), r->djob, &hj, 1);
}

// After fused iterations: X changed, W is the row transform of the new X.
void fused_done(nadia_recon* r, int n) {
  r->est->version++;
  r->w_version = r->est->version;
  if (two_fields(r)) {
    r->w2_version = r->est->version;
    r->w2_sversion = r->s_version;
  }
  // PlanarCDI / FresnelCDI update passes wrote every error into get_error()'s
  // slot as well (no copy out of the ring needed); other paths leave it in
  // the ring
  const bool direct = fused_ok(r) && r->kind != NADIA_FRESNEL_WF;
  r->last_err = direct ? r->scal + 8 : r->errs + (n - 1);
  r->herr_fresh = direct && r->herr;
  r->h_err_valid = false;
}

bool wants_n_best(const nadia_recon* r);
int update_n_best(nadia_recon* r, const double* err, hipStream_t st);
int n_best_copy(nadia_recon* r, hipStream_t st);

template <typename T>
int iterate_fused(nadia_recon* r, int n) {
  int rc;
  if ((rc = fused_ready<T>(r))) return rc;
  std::vector<Ev> evs;
  const GroupKey g = group_key(r);
  const bool nb = wants_n_best(r);
  for (int k = 0; k < n; ++k) {
    // the update pass ranks the error for the kept best estimates; the
    // snapshot copy follows it
    if ((rc = launch_fused<T>(r->ctx, g, r->djob, 1, k, r->col_v, r->row_v, r, r->timing ? &evs : nullptr,
                              nullptr, nullptr, nullptr, 0, k == n - 1 ? 1 : 0, 1)))
      return rc;
    if (nb && (rc = n_best_copy(r, S_(r)))) return rc;
  }
  fused_done(r, n);
  return collect_events(r, evs);
}

// ER iterations with a transmission constraint (BaseCDI::iterate's ER
// branch, NADIA/src/BaseCDI.c++:303-308, with FresnelCDI::apply_support,
// NADIA/src/FresnelCDI.c++:203-215): the fused column and row passes give
// x = support(P_F x); the constraint then runs on x as in op_support, and
// the row transform W is rebuilt from the constrained x for the next
// column pass.
// One such iteration, its error into ring slot k (after fused_ready): the
// region means, then one row pass applying the constraint and rebuilding W
// (a host constraint function takes the separate steps).
template <typename T>
int fused_tc_step(nadia_recon* r, int k) {
  int rc;
  nadia_constraint* tc = r->tc;
  const bool fres = r->kind == NADIA_FRESNEL;
  const int rv = r->row_v < 0 ? 0 : r->row_v;
  if (!tc->custom && tc->nreg <= kRowStatsRegions) {
    // column pass; update pass storing x and summing the regions; region
    // means; one pass applying the constraint and rebuilding W
    TcDev d{};
    if ((rc = tc_view(tc, S_(r), fres ? r->Ls : nullptr, r->nx, r->ny, &d))) return rc;
    if ((rc = launch_fused<T>(r->ctx, group_key(r), r->djob, 1, k, r->col_v, rv, r, nullptr, nullptr, &d,
                              tc->partials)))
      return rc;
    if (tc->nreg > 0) k_tc_final(d, tc->partials, row_blocks<T>(r->nx, r->ny, rv), S_(r));
    void* twy = twiddles(r->ctx, r->ny, r->prec, true);
    if (!twy) return fail(NADIA_EHIP, "twiddle table allocation failed");
    RowArgs ra = w_rebuild_args(r, twy);
    ra.mode = ROW_TC;
    ra.src = r->est->d;
    ra.dst = r->W;
    ra.tc = d;
    launch_row<T>(ra, 1, rv, S_(r));
    LAUNCHCHK();
    r->w_valid = true;
    r->w_version = r->est->version;
    if (wants_n_best(r) && (rc = update_n_best(r, r->errs + k, S_(r)))) return rc;
    return 0;
  }
  if ((rc = launch_fused<T>(r->ctx, group_key(r), r->djob, 1, k, r->col_v, r->row_v, r, nullptr))) return rc;
  if (r->tc->custom) {
    if ((rc = op_constraint(r, r->est->d))) return rc;
    r->w_valid = false;
    if ((rc = fused_prepare<T>(r, false))) return rc;
  } else {
    const bool fres = r->kind == NADIA_FRESNEL;
    TcDev d{};
    if ((rc = tc_stats(r->tc, S_(r), r->est->d, r->prec, fres ? r->Ls : nullptr, r->nx, r->ny, &d))) return rc;
    void* twy = twiddles(r->ctx, r->ny, r->prec, true);
    if (!twy) return fail(NADIA_EHIP, "twiddle table allocation failed");
    RowArgs ra = w_rebuild_args(r, twy);
    ra.mode = ROW_TC;
    ra.src = r->est->d;
    ra.dst = r->W;
    ra.tc = d;
    launch_row<T>(ra, 1, r->row_v < 0 ? 0 : r->row_v, S_(r));
    LAUNCHCHK();
    r->w_valid = true;
    r->w_version = r->est->version;
  }
  if (wants_n_best(r) && (rc = update_n_best(r, r->errs + k, S_(r)))) return rc;
  return 0;
}

template <typename T>
int iterate_fused_tc(nadia_recon* r, int n) {
  int rc;
  if ((rc = fused_ready<T>(r))) return rc;
  for (int k = 0; k < n; ++k)
    if ((rc = fused_tc_step<T>(r, k))) return rc;
  fused_done(r, n);
  return 0;
}

// ---- fused iterations at sizes without compiled passes ----------------------
// The column and row passes of the fused iteration on kern_any.inc's tiles:
// W = row DFT of X between iterations, a column pass (forward DFT,
// modulus projection, inverse) and a row pass (inverse DFT, update of X,
// forward DFT), one job per launch.  PlanarCDI and FresnelCDI with
// algorithms of one work field; the rest take the unfused path.
bool any_fused_ok(const nadia_recon* r) {
  return r->path == NADIA_PATH_FUSED && !native_dims(r) && has_projection(r) && !r->tc && !two_fields(r) &&
         (r->kind == NADIA_PLANAR || r->kind == NADIA_FRESNEL);
}

// one any-length pass of reconstruction r: mode, error slot, column-pass
// workgroups (for the row pass's error); *nblk = workgroups launched
template <typename T>
int any_pass(nadia_recon* r, int mode, int iter, int fin_n, int* nblk) {
  const int prec = sizeof(T) == 8 ? NADIA_FP64 : NADIA_FP32;
  const bool rows = mode == ANY_ROW_UPDATE;
  const int n = rows ? r->ny : r->nx;
  const AnyPlan* p = any_plan(r->ctx, n, prec);
  if (!p) return fail(NADIA_EHIP, "transform tables for length %d: allocation failed", n);
  AnyArgs a{};
  a.mode = mode;
  a.jd = r->djob;
  a.nx = r->nx;
  a.ny = r->ny;
  a.sf = 1.0 / std::sqrt((double)r->nx * (double)r->ny);
  a.iter = iter;
  a.fin_n = fin_n;
  a.rows = rows ? 1 : 0;
  a.nseq = rows ? r->nx : r->ny;
  a.es = rows ? 1 : r->ny;
  a.ss = rows ? r->ny : 1;
  a.n = n;
  a.len = p->len;
  a.nst = p->nst;
  a.stages = p->stages;
  a.tw = p->tw;
  a.chirp = p->chirp;
  a.kern = p->kern;
  AnyShape sh{};
  if (!any_shape(*p, a.nseq, rows, prec, sh)) return fail(NADIA_ESIZE, "no launch shape for length %d", n);
  a.cb = sh.cb;
  a.stride = sh.stride;
  launch_any<T>(a, sh.lds_class, sh.threads, S_(r));
  LAUNCHCHK();
  if (nblk) *nblk = (a.nseq + a.cb - 1) / a.cb;
  return 0;
}

// W current (rebuilt from X when X changed) and the job descriptor uploaded
template <typename T>
int any_ready(nadia_recon* r) {
  int rc;
  if (w_stale(r)) {
    HIPCHK(hipMemcpyAsync(r->W, r->est->d, r->npix * r->esz, hipMemcpyDeviceToDevice, S_(r)));
    if ((rc = any_axis<T>(r->ctx, r->W, r->nx, r->ny, -1, true, S_(r)))) return rc;
    r->w_valid = true;
    r->w_version = r->est->version;
  }
  const JobDesc hj = make_jobdesc(r, r->est->d, r->W, r->errs);
  if (r->djob_set && std::memcmp(&hj, &r->djob_host, sizeof(JobDesc)) == 0) return 0;
  r->djob_host = hj;
  r->djob_set = true;
  return upload_jobs(S_(r), r->djob, &hj, 1);
}

// one iteration, its error into ring slot k (after any_ready)
template <typename T>
int any_step(nadia_recon* r, int k) {
  int rc, nb = 0;
  if ((rc = any_pass<T>(r, r->kind == NADIA_FRESNEL ? ANY_COL_FRESNEL : ANY_COL_PLANAR, k, 0, &nb))) return rc;
  if ((rc = any_pass<T>(r, ANY_ROW_UPDATE, k, nb, nullptr))) return rc;
  if (wants_n_best(r) && (rc = update_n_best(r, r->errs + k, S_(r)))) return rc;
  return 0;
}

template <typename T>
int iterate_any(nadia_recon* r, int n) {
  int rc;
  if ((rc = any_ready<T>(r))) return rc;
  for (int k = 0; k < n; ++k)
    if ((rc = any_step<T>(r, k))) return rc;
  fused_done(r, n);
  return 0;
}

int fetch_error(nadia_recon* r, double* out) {
  if (r->h_err_valid) {
    *out = r->h_err_cache;
    return 0;
  }
  if (!r->last_err) {
    // current_error is never initialised by the reference before the first
    // projection; report 0
    *out = 0.0;
    return 0;
  }
  if (r->herr_fresh) {
    // the update pass wrote the error into the page-locked slot as well: no
    // copy, only the wait
    HIPCHK(hipStreamSynchronize(S_(r)));
    r->h_err_cache = *(volatile const double*)r->herr;
    r->h_err_valid = true;
    *out = r->h_err_cache;
    return 0;
  }
  HIPCHK(hipMemcpyAsync(&r->h_err_cache, r->last_err, sizeof(double), hipMemcpyDeviceToHost, S_(r)));
  HIPCHK(hipStreamSynchronize(S_(r)));
  r->h_err_valid = true;
  *out = r->h_err_cache;
  return 0;
}

// BaseCDI::update_n_best (NADIA/src/BaseCDI.c++:390-413), queued on the
// device after an iteration: `err` is the device slot holding that
// iteration's get_error() (its ring slot; for a projection-free algorithm the
// unchanged current error).  No host round trip.
bool wants_n_best(const nadia_recon* r) { return r->n_best > 0 && r->kind != NADIA_FRESNEL_WF; }

const double* current_error_slot(const nadia_recon* r, const double* ring_slot) {
  if (has_projection(r)) return ring_slot;
  return r->last_err ? r->last_err : r->scal + 10;  // never projected: 0, as get_error() reports
}

int update_n_best(nadia_recon* r, const double* err, hipStream_t st) {
  if (!wants_n_best(r)) return 0;
  k_nbest(err, r->dbest_err, r->dbest_perm, r->n_best, r->dbest_target, r->est->d, r->dbest_bufs,
          (long long)(r->npix * r->esz), st);
  LAUNCHCHK();
  return 0;
}

// The snapshot copy of update_n_best after a fused update pass launched with
// rank = 1, which ranked the error itself (the job descriptor carries the
// kept errors, the rank -> buffer map and the copy target).
int n_best_copy(nadia_recon* r, hipStream_t st) {
  if (!wants_n_best(r)) return 0;
  k_nbest_copy(r->dbest_target, r->est->d, r->dbest_bufs, (long long)(r->npix * r->esz), st);
  LAUNCHCHK();
  return 0;
}

// BaseCDI::convolve's window g(i, j) = exp(-(i^2 + j^2) / (2 sigma^2)),
// 0 <= i, j <= cut (src/BaseCDI.c++:441-447), rounded to the field's real
// type as the reference stores it
void conv_window(const nadia_recon* r, double sigma, int cut, double* g, float* gf) {
  for (int i = 0; i <= cut; i++)
    for (int j = 0; j <= cut; j++) {
      const double v = exp(-1 * (i * i + j * j) / (2.0 * sigma * sigma));
      g[i * (cut + 1) + j] = r->prec == NADIA_FP64 ? v : (double)(float)v;
      gf[i * (cut + 1) + j] = (float)v;
    }
}

// set_support on a device real array already in r->rtmp (src/BaseCDI.c++:95-108)
int support_from_rtmp(nadia_recon* r, int soften) {
  DISPATCH(r->prec, k_real_max<T>(r->rtmp, r->npix, r->partials, kEwBlocks, S_(r)));
  k_final_max(r->partials, kEwBlocks, r->scal + 1, S_(r));
  DISPATCH(r->prec, k_normalise<T>(r->rtmp, r->S, r->scal + 1, r->npix, S_(r), (unsigned char*)r->S8));
  LAUNCHCHK();
  r->s_soft = soften != 0;
  if (soften) {
    // convolve(support, 3, 5): 11x11 window, sigma 3, left unnormalised
    ConvArgs ca{};
    conv_window(r, 3.0, 5, ca.g, ca.gf);
    ca.src = r->S;
    ca.src_complex = 0;
    ca.dst = r->conv;
    ca.nx = r->nx;
    ca.ny = r->ny;
    ca.partials = r->partials;
    int nb = 0;
    DISPATCH(r->prec, k_conv<T>(ca, 5, &nb, S_(r)));
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(r->S, r->conv, r->npix * r->rsz, hipMemcpyDeviceToDevice, S_(r)));
  }
  r->s_version++;
  return 0;
}

int alloc(void** p, size_t bytes) {
  HIPCHK(hipMalloc(p, bytes ? bytes : 16));
  return 0;
}

}  // namespace

// =========================== exported C ABI ===============================

extern "C" {

const char* nadia_last_error(void) { return g_err.c_str(); }
int nadia_version(void) { return 100; }

int nadia_device_count(int* count) {
  HIPCHK(hipGetDeviceCount(count));
  return 0;
}

int nadia_set_device(int device) {
  HIPCHK(hipSetDevice(device));
  return 0;
}

int nadia_size_supported(int n) { return n >= 1 && n <= kAnyMaxLen ? 1 : 0; }

int nadia_size_native(int n) { return size_native(n) ? 1 : 0; }

int nadia_synchronize(void) {
  DeviceCtx* c = dev_ctx(current_device());
  if (!c) return fail(NADIA_EHIP, "stream creation failed");
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// ---- fields ---------------------------------------------------------------

int nadia_field_create(int nx, int ny, int precision, nadia_field** out) {
  if (!out) return fail(NADIA_EINVAL, "null output pointer");
  if (precision != NADIA_FP32 && precision != NADIA_FP64) return fail(NADIA_EINVAL, "bad precision %d", precision);
  int rc;
  if ((rc = check_dims(nx, ny))) return rc;
  nadia_field* f = new nadia_field{};
  f->nx = nx;
  f->ny = ny;
  f->prec = precision;
  f->dev = current_device();
  f->esz = precision == NADIA_FP64 ? 16 : 8;
  DeviceCtx* c = dev_ctx(f->dev);
  if (!c) {
    delete f;
    return fail(NADIA_EHIP, "stream creation failed");
  }
  if (hipMalloc(&f->d, (size_t)nx * ny * f->esz) != hipSuccess) {
    delete f;
    return fail(NADIA_EHIP, "hipMalloc of %dx%d field failed", nx, ny);
  }
  HIPCHK(hipMemsetAsync(f->d, 0, (size_t)nx * ny * f->esz, c->stream));
  *out = f;
  return 0;
}

int nadia_field_destroy(nadia_field* f) {
  if (!f) return 0;
  DeviceCtx* c = dev_ctx(f->dev);
  if (c) (void)hipStreamSynchronize(c->stream);
  (void)hipFree(f->d);
  delete f;
  return 0;
}

int nadia_field_dims(const nadia_field* f, int* nx, int* ny, int* precision) {
  if (!f) return fail(NADIA_EINVAL, "null field");
  if (nx) *nx = f->nx;
  if (ny) *ny = f->ny;
  if (precision) *precision = f->prec;
  return 0;
}

int nadia_field_upload(nadia_field* f, const void* host) {
  if (!f || !host) return fail(NADIA_EINVAL, "null argument");
  int rc;
  if ((rc = upload(dev_ctx(f->dev), f->d, host, (size_t)f->nx * f->ny * f->esz))) return rc;
  f->version++;
  return 0;
}

int nadia_field_download(const nadia_field* f, void* host) {
  if (!f || !host) return fail(NADIA_EINVAL, "null argument");
  DeviceCtx* c = dev_ctx(f->dev);
  HIPCHK(hipMemcpyAsync(host, f->d, (size_t)f->nx * f->ny * f->esz, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int nadia_field_copy(nadia_field* dst, const nadia_field* src) {
  if (!dst || !src) return fail(NADIA_EINVAL, "null argument");
  if (dst->nx != src->nx || dst->ny != src->ny || dst->prec != src->prec)
    return fail(NADIA_EINVAL, "copy between fields of different dimensions or precision");
  DeviceCtx* c = dev_ctx(dst->dev);
  HIPCHK(hipMemcpyAsync(dst->d, src->d, (size_t)dst->nx * dst->ny * dst->esz, hipMemcpyDeviceToDevice, c->stream));
  dst->version++;
  return 0;
}

int nadia_field_fft(nadia_field* f, int direction) {
  if (!f || (direction != -1 && direction != 1)) return fail(NADIA_EINVAL, "bad argument");
  int rc = fft2d(dev_ctx(f->dev), f->d, f->nx, f->ny, f->prec, direction);
  f->version++;
  return rc;
}

int nadia_field_invert(nadia_field* f, int scale) {
  if (!f) return fail(NADIA_EINVAL, "null field");
  DeviceCtx* c = dev_ctx(f->dev);
  DISPATCH(f->prec, k_invert<T>(f->d, f->nx, f->ny, scale, c->stream));
  LAUNCHCHK();
  f->version++;
  return 0;
}

int nadia_field_scale(nadia_field* f, double factor) {
  if (!f) return fail(NADIA_EINVAL, "null field");
  DeviceCtx* c = dev_ctx(f->dev);
  DISPATCH(f->prec, k_scale<T>(f->d, factor, (long long)f->nx * f->ny, c->stream));
  LAUNCHCHK();
  f->version++;
  return 0;
}

int nadia_field_add(nadia_field* f, const nadia_field* other, double scale) {
  if (!f || !other) return fail(NADIA_EINVAL, "null field");
  if (f->nx != other->nx || f->ny != other->ny || f->prec != other->prec)
    return fail(NADIA_EINVAL, "add: dimensions of the input do not match");
  DeviceCtx* c = dev_ctx(f->dev);
  DISPATCH(f->prec, k_fresnel_add<T>(f->d, other->d, scale, (long long)f->nx * f->ny, c->stream));
  LAUNCHCHK();
  f->version++;
  return 0;
}

int nadia_field_norm(const nadia_field* f, double* out) {
  if (!f || !out) return fail(NADIA_EINVAL, "null argument");
  DeviceCtx* c = dev_ctx(f->dev);
  double* part = nullptr;
  HIPCHK(hipMalloc(&part, (kEwBlocks + 1) * sizeof(double)));
  DISPATCH(f->prec, k_norm2<T>(f->d, (long long)f->nx * f->ny, part, kEwBlocks, c->stream));
  k_final_sum1(part, kEwBlocks, part + kEwBlocks, c->stream);
  double s = 0;
  HIPCHK(hipMemcpyAsync(&s, part + kEwBlocks, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  (void)hipFree(part);
  *out = std::sqrt(s);
  return 0;
}

int nadia_field_magnitude(const nadia_field* f, void* host_real) {
  if (!f || !host_real) return fail(NADIA_EINVAL, "null argument");
  DeviceCtx* c = dev_ctx(f->dev);
  void* tmp = nullptr;
  size_t rsz = f->esz / 2;
  HIPCHK(hipMalloc(&tmp, (size_t)f->nx * f->ny * rsz));
  DISPATCH(f->prec, k_magnitude<T>(f->d, tmp, (long long)f->nx * f->ny, c->stream));
  HIPCHK(hipMemcpyAsync(host_real, tmp, (size_t)f->nx * f->ny * rsz, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  (void)hipFree(tmp);
  return 0;
}

// nadia_field_magnitude staged through the library's fixed device array
// instead of an allocation: no hipMalloc / hipFree.  Calls that overlap on
// one GPU share the array (race); every element it stages is written before
// it is read, so nothing needs clearing.
int nadia_field_magnitude_noalloc(const nadia_field* f, void* host_real) {
  if (!f || !host_real) return fail(NADIA_EINVAL, "null argument");
  const long long n = (long long)f->nx * f->ny;
  const size_t bytes = (size_t)n * (f->esz / 2);
  if (bytes > kNoAllocScratchBytes)
    return fail(NADIA_ESIZE,
                "nadia_field_magnitude_noalloc: a %dx%d field needs %zu bytes of staging, more than the %zu it holds "
                "(use nadia_field_magnitude)",
                f->nx, f->ny, bytes, kNoAllocScratchBytes);
  void* p = nullptr;
  int rc;
  if ((rc = noalloc_scratch(f->dev, &p))) return rc;
  DeviceCtx* c = dev_ctx(f->dev);
  DISPATCH(f->prec, k_magnitude<T>(f->d, p, n, c->stream));
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(host_real, p, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// ---- reconstructions --------------------------------------------------------

int nadia_recon_set_algorithm(nadia_recon* r, int alg);

int nadia_recon_create(nadia_field* est, int kind, int n_best, nadia_recon** out) {
  if (!est || !out) return fail(NADIA_EINVAL, "null argument");
  if (kind < NADIA_PLANAR || kind > NADIA_FRESNEL_WF) return fail(NADIA_EINVAL, "bad kind %d", kind);
  if (n_best < 0) return fail(NADIA_EINVAL, "negative n_best");
  if ((est->nx & 1) || (est->ny & 1)) return fail(NADIA_ESIZE, "odd dimensions are not supported");
  nadia_recon* r = new nadia_recon{};
  r->est = est;
  r->kind = kind;
  r->prec = est->prec;
  r->nx = est->nx;
  r->ny = est->ny;
  r->dev = est->dev;
  r->n_best = n_best;
  r->ctx = dev_ctx(r->dev);
  if (!r->ctx) {
    delete r;
    return fail(NADIA_EHIP, "stream creation failed");
  }
  r->esz = est->esz;
  r->rsz = est->esz / 2;
  r->npix = (long long)r->nx * r->ny;
  r->ring_cap = 4096;
  r->path = NADIA_PATH_FUSED;
  r->beta = 0.9;
  r->col_v = r->row_v = -1;
  r->tile_fixed = false;
  const size_t cb = r->npix * r->esz, rb = r->npix * r->rsz;
  int rc = 0;
  // the fixed buffers share one allocation (one hipMalloc per
  // reconstruction), each on a 256-byte boundary
  {
    void** dst[] = {&r->S, &r->Q, &r->W, &r->conv, &r->rtmp, (void**)&r->partials, (void**)&r->errs,
                    (void**)&r->scal, (void**)&r->djob, (void**)&r->tune_errs, &r->S8};
    const size_t len[] = {rb, rb, 2 * cb, rb, rb, sizeof(double) * 2 * (size_t)(kEwBlocks + 2 * 4096 + 16),
                          sizeof(double) * (r->ring_cap + 8), sizeof(double) * 16, sizeof(JobDesc) * 2,
                          sizeof(double) * 64, (size_t)r->npix};
    size_t total = 0;
    for (size_t b : len) total += (b + 255) & ~(size_t)255;
    if ((rc = alloc(&r->arena, total))) {
      nadia_recon_destroy(r);  // frees what was allocated so far
      return rc;
    }
    char* p = (char*)r->arena;
    for (int i = 0; i < 11; i++) {
      *dst[i] = p;
      p += (len[i] + 255) & ~(size_t)255;
    }
  }
  (void)hipMemsetAsync(r->S, 0, rb, S_(r));
  (void)hipMemsetAsync(r->S8, 0, r->npix, S_(r));
  r->s_soft = false;
  (void)hipMemsetAsync(r->Q, 0, rb, S_(r));
  (void)hipMemsetAsync(r->scal, 0, sizeof(double) * 16, S_(r));
  for (int i = 0; i < n_best; i++) {
    void* p = nullptr;
    if (hipMalloc(&p, cb) != hipSuccess) {
      nadia_recon_destroy(r);  // frees what was allocated so far
      return fail(NADIA_EHIP, "hipMalloc failed for n_best snapshot");
    }
    (void)hipMemsetAsync(p, 0, cb, S_(r));
    r->best.push_back(p);
  }
  if (n_best > 0) {
    if ((rc = alloc((void**)&r->dbest_err, sizeof(double) * n_best)) ||
        (rc = alloc((void**)&r->dbest_perm, sizeof(int) * n_best)) ||
        (rc = alloc((void**)&r->dbest_bufs, sizeof(void*) * n_best)) || (rc = alloc((void**)&r->dbest_target, sizeof(int)))) {
      nadia_recon_destroy(r);  // frees what was allocated so far
      return rc;
    }
    // BaseCDI constructor: every best error starts at 1 (src/BaseCDI.c++:50-56)
    std::vector<double> e(n_best, 1.0);
    std::vector<int> perm(n_best);
    for (int i = 0; i < n_best; i++) perm[i] = i;
    HIPCHK(hipMemcpyAsync(r->dbest_err, e.data(), sizeof(double) * n_best, hipMemcpyHostToDevice, S_(r)));
    HIPCHK(hipMemcpyAsync(r->dbest_perm, perm.data(), sizeof(int) * n_best, hipMemcpyHostToDevice, S_(r)));
    HIPCHK(hipMemcpyAsync(r->dbest_bufs, r->best.data(), sizeof(void*) * n_best, hipMemcpyHostToDevice, S_(r)));
    // the host arrays above end with this block: the copies must be done
    HIPCHK(hipStreamSynchronize(S_(r)));
  }
  // BaseCDI::BaseCDI: algorithm = CUSTOM; set_algorithm(ER) (src/BaseCDI.c++:63-64)
  r->algorithm = NADIA_CUSTOM;
  nadia_recon_set_algorithm(r, NADIA_ER);
  // last, so that no failure above leaves a slot taken
  r->herr = herr_acquire(r->ctx, &r->herr_dev, &r->herr_slot);
  *out = r;
  return 0;
}

}  // extern "C"

namespace {
// Device bytes owned by a reconstruction of npix pixels (not counting its
// estimate field): S, Q, conv, rtmp (real), W (2 complex fields) and the small
// per-job buffers; beam stop, unfused temporaries, Fresnel fields and n-best
// snapshots when present.
size_t base_bytes(long long npix, size_t esz) {
  const size_t cb = npix * esz, rb = npix * (esz / 2);
  // (+ the support bytes, and the 256-byte alignment of the eleven buffers
  // of the shared allocation)
  return 4 * rb + 2 * cb + (size_t)npix + sizeof(double) * 2 * (size_t)(kEwBlocks + 2 * 4096 + 16) +
         sizeof(double) * (4096 + 8) + sizeof(double) * 16 + sizeof(JobDesc) * 2 + sizeof(double) * 64 + 11 * 256;
}
}  // namespace

extern "C" {

int nadia_recon_device_bytes(const nadia_recon* r, size_t* out) {
  if (!r || !out) return fail(NADIA_EINVAL, "null argument");
  const size_t cb = r->npix * r->esz, rb = r->npix * r->rsz;
  size_t b = base_bytes(r->npix, r->esz);
  if (r->beam) b += rb;
  for (void* t : r->t)
    if (t) b += cb;
  if (r->L) b += cb;
  if (r->Ls) b += cb;
  b += r->best.size() * cb;
  *out = b;
  return 0;
}

int nadia_recon_bytes_estimate(int nx, int ny, int precision, int kind, int n_best, size_t* out) {
  if (!out || nx <= 0 || ny <= 0 || n_best < 0) return fail(NADIA_EINVAL, "bad argument");
  if (precision != NADIA_FP32 && precision != NADIA_FP64) return fail(NADIA_EINVAL, "bad precision");
  const long long npix = (long long)nx * ny;
  const size_t esz = precision == NADIA_FP64 ? 16 : 8;
  const size_t cb = npix * esz;
  size_t b = cb + base_bytes(npix, esz) + npix * (esz / 2) + (size_t)n_best * cb;
  if (kind == NADIA_FRESNEL || kind == NADIA_FRESNEL_WF) b += 2 * cb;
  *out = b;
  return 0;
}

int nadia_recon_destroy(nadia_recon* r) {
  if (!r) return 0;
  (void)hipStreamSynchronize(S_(r));
  void* ptrs[] = {r->arena, r->beam, r->t[0], r->t[1], r->t[2], r->t[3], r->L, r->Ls, r->ttmp,
                  r->dbest_err, r->dbest_perm, r->dbest_bufs, r->dbest_target, r->phx, r->phy};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (void* p : r->best) (void)hipFree(p);
  herr_release(r->ctx, r->herr_slot);
  delete r;
  return 0;
}

int nadia_recon_set_path(nadia_recon* r, int path) {
  if (!r || (path != NADIA_PATH_FUSED && path != NADIA_PATH_PLAIN)) return fail(NADIA_EINVAL, "bad argument");
  r->path = path;
  r->w_valid = r->w2_valid = false;
  return 0;
}

// The setters copy the caller's array (set_* copy their inputs,
// src/BaseCDI.c++:95-124) into the staging ring and queue the rest on the
// device: they return without waiting for it.
int nadia_recon_set_support(nadia_recon* r, const void* host, int soften) {
  if (!r || !host) return fail(NADIA_EINVAL, "null argument");
  int rc;
  if ((rc = upload(r->ctx, r->rtmp, host, r->npix * r->rsz))) return rc;
  return support_from_rtmp(r, soften);
}

int nadia_recon_set_intensity(nadia_recon* r, const void* host) {
  if (!r || !host) return fail(NADIA_EINVAL, "null argument");
  int rc;
  if ((rc = upload(r->ctx, r->rtmp, host, r->npix * r->rsz))) return rc;
  DISPATCH(r->prec, k_make_q<T>(r->rtmp, r->beam, r->has_beam, r->Q, r->npix, S_(r)));
  LAUNCHCHK();
  r->has_intensity = 1;
  return 0;
}

int nadia_recon_set_beam_stop(nadia_recon* r, const void* host) {
  if (!r || !host) return fail(NADIA_EINVAL, "null argument");
  int rc;
  if (!r->beam && (rc = alloc(&r->beam, r->npix * r->rsz))) return rc;
  if ((rc = upload(r->ctx, r->beam, host, r->npix * r->rsz))) return rc;
  r->has_beam = 1;
  // re-mask sqrt(I) already on the device
  DISPATCH(r->prec, k_make_q<T>(nullptr, r->beam, 1, r->Q, r->npix, S_(r)));
  LAUNCHCHK();
  return 0;
}

int nadia_recon_set_custom_algorithm(nadia_recon* r, const double m[10]) {
  if (!r || !m) return fail(NADIA_EINVAL, "null argument");
  if (r->kind == NADIA_FRESNEL_WF) return 0;  // hidden no-op, FresnelCDI_WF.h:176-181
  // src/BaseCDI.c++:245-249
  r->alg[A_PSF] = m[0] + m[1] + m[2] + m[3];
  r->alg[A_PFS] = -m[0] + m[4] + m[5] + m[6];
  r->alg[A_PS] = -m[2] - m[5] - m[7] + m[9];
  r->alg[A_PF] = -m[1] - m[4] + m[7] + m[8];
  r->alg[A_PI] = -m[3] - m[6] - m[8] - m[9];
  // reallocate_temp_complex_memory (src/BaseCDI.c++:257-275): temporaries of
  // terms that vanished are freed here; the unfused path allocates the ones it
  // needs on first use (the fused path never needs them)
  for (int n = 0; n < 4; n++) {
    if (r->alg[n] == 0 && r->t[n]) {
      (void)hipStreamSynchronize(S_(r));
      (void)hipFree(r->t[n]);
      r->t[n] = nullptr;
    }
  }
  r->algorithm = NADIA_CUSTOM;
  r->w2_valid = false;
  return 0;
}

int nadia_recon_set_algorithm(nadia_recon* r, int alg) {
  if (!r) return fail(NADIA_EINVAL, "null recon");
  // FresnelCDI_WF hides set_algorithm (FresnelCDI_WF.h:147-150); only the
  // constructor's BaseCDI call takes effect
  if (r->kind == NADIA_FRESNEL_WF && r->algorithm != NADIA_CUSTOM) return 0;
  if (r->algorithm == alg) return 0;  // src/BaseCDI.c++:194-195
  const double b = r->beta;
  double m[10] = {0};
  switch (alg) {
    case NADIA_ER: break;
    case NADIA_BIO: m[1] = b; break;
    case NADIA_BOO: m[1] = b; m[8] = 1; break;
    case NADIA_HIO: m[1] = b; m[2] = 1; break;
    case NADIA_DM: m[0] = b; m[2] = 1; m[4] = 1; break;
    case NADIA_SF: m[1] = 1; m[3] = 1; break;
    case NADIA_ASR: m[1] = 1; m[2] = 1; break;
    case NADIA_HPR: m[1] = b; m[2] = 1; break;
    case NADIA_RAAR: m[1] = 1; m[2] = b; m[8] = 1 - b; break;
    default:
      fprintf(stdout, "Algorithm unknown\n");
      r->algorithm = alg;
      return 0;
  }
  int saved_kind = r->kind;
  r->kind = NADIA_PLANAR;  // let set_custom_algorithm run for every kind here
  int rc = nadia_recon_set_custom_algorithm(r, m);
  r->kind = saved_kind;
  if (rc) return rc;
  r->algorithm = alg;
  return 0;
}

int nadia_recon_set_relaxation_parameter(nadia_recon* r, double beta) {
  if (!r) return fail(NADIA_EINVAL, "null recon");
  if (r->kind == NADIA_FRESNEL_WF) return 0;  // FresnelCDI_WF.h:168-172
  r->beta = beta;
  return nadia_recon_set_algorithm(r, r->algorithm);  // include/BaseCDI.h:264-267
}

int nadia_recon_get_coefficients(const nadia_recon* r, double out5[5], int* algorithm) {
  if (!r) return fail(NADIA_EINVAL, "null recon");
  if (out5)
    for (int i = 0; i < 5; i++) out5[i] = r->alg[i];
  if (algorithm) *algorithm = r->algorithm;
  return 0;
}

int nadia_recon_iterate(nadia_recon* r, int n, double* errors) {
  if (!r) return fail(NADIA_EINVAL, "null recon");
  if (n <= 0) return 0;
  if (!r->has_intensity) return fail(NADIA_ESTATE, "intensity not set");
  int rc;
  if ((rc = need_setup(r))) return rc;
  const bool nb = wants_n_best(r);
  int done = 0;
  while (done < n) {
    int chunk = n - done;
    if (chunk > r->ring_cap) chunk = r->ring_cap;
    if (fused_ok(r)) {
      DISPATCH(r->prec, rc = iterate_fused<T>(r, chunk));
      if (rc) return rc;
    } else if (fused_tc_ok(r)) {
      DISPATCH(r->prec, rc = iterate_fused_tc<T>(r, chunk));
      if (rc) return rc;
    } else if (any_fused_ok(r)) {
      DISPATCH(r->prec, rc = iterate_any<T>(r, chunk));
      if (rc) return rc;
    } else {
      for (int k = 0; k < chunk; ++k) {
        if ((rc = iterate_plain(r, r->errs + k))) return rc;
        if (nb && (rc = update_n_best(r, current_error_slot(r, r->errs + k), S_(r)))) return rc;
      }
      r->est->version++;
    }
    if (errors) {
      if (chunk == 1 && r->herr_fresh) {
        // one iteration through the fused passes: its error is in the
        // page-locked slot once the stream's work is done
        HIPCHK(hipStreamSynchronize(S_(r)));
        errors[done] = r->h_err_cache = *(volatile const double*)r->herr;
        r->h_err_valid = true;
      } else if (has_projection(r)) {
        HIPCHK(hipMemcpyAsync(errors + done, r->errs, chunk * sizeof(double), hipMemcpyDeviceToHost, S_(r)));
      } else {
        // a projection-free custom algorithm leaves get_error() unchanged
        double v = 0;
        if ((rc = fetch_error(r, &v))) return rc;
        for (int k = 0; k < chunk; ++k) errors[done + k] = v;
      }
    }
    done += chunk;
    if (done < n || errors) HIPCHK(hipStreamSynchronize(S_(r)));
  }
  // keep get_error() valid after the ring is reused by the next call
  if (r->last_err && r->last_err != r->scal + 8) {
    HIPCHK(hipMemcpyAsync(r->scal + 8, r->last_err, sizeof(double), hipMemcpyDeviceToDevice, S_(r)));
    r->last_err = r->scal + 8;
    r->herr_fresh = false;
  }
  return 0;
}

int nadia_recon_get_error(nadia_recon* r, double* out) {
  if (!r || !out) return fail(NADIA_EINVAL, "null argument");
  return fetch_error(r, out);
}

int nadia_recon_apply_shrinkwrap(nadia_recon* r, double gauss_width, double threshold) {
  if (!r) return fail(NADIA_EINVAL, "null recon");
  // BaseCDI::apply_shrinkwrap (src/BaseCDI.c++:416-429): convolve |x| with
  // the 9x9 window, threshold at thr * max, set_support.  Everything is
  // queued on the device (the window travels in the kernel arguments): the
  // call returns without waiting.
  ConvArgs ca{};
  conv_window(r, gauss_width, 4, ca.g, ca.gf);
  ca.src = r->est->d;
  ca.src_complex = 1;
  ca.dst = r->conv;
  ca.nx = r->nx;
  ca.ny = r->ny;
  ca.partials = r->partials;
  int nb = 0;
  DISPATCH(r->prec, k_conv<T>(ca, 4, &nb, S_(r)));
  k_final_max(r->partials, nb, r->scal, S_(r));
  if (threshold <= 1.0 && r->ny <= r->nx) {
    // Every pixel is thresholded (apply_threshold's j < nx covers each row)
    // and the largest value stays 1 (it is not below threshold * max when
    // threshold <= 1; with a NaN max every comparison fails and all are 1),
    // so set_support's maximum is 1 and its division the identity: the
    // threshold writes the support and its byte copy directly.
    DISPATCH(r->prec, k_threshold<T>(r->conv, r->S, r->nx, r->ny, r->scal, threshold, S_(r),
                                     (unsigned char*)r->S8));
    LAUNCHCHK();
    r->s_soft = false;
    r->s_version++;
    return 0;
  }
  DISPATCH(r->prec, k_threshold<T>(r->conv, r->rtmp, r->nx, r->ny, r->scal, threshold, S_(r)));
  LAUNCHCHK();
  return support_from_rtmp(r, 0);
}

int nadia_recon_get_support(const nadia_recon* r, void* host) {
  if (!r || !host) return fail(NADIA_EINVAL, "null argument");
  HIPCHK(hipMemcpyAsync(host, r->S, r->npix * r->rsz, hipMemcpyDeviceToHost, r->ctx->stream));
  HIPCHK(hipStreamSynchronize(r->ctx->stream));
  return 0;
}

int nadia_recon_initialise_estimate(nadia_recon* r, int seed) {
  if (!r) return fail(NADIA_EINVAL, "null recon");
  int rc;
  const long long n = r->npix;
  std::vector<double> s(n), q(n);
  // download support and sqrt(I) in the field's real type
  {
    std::vector<char> bs(n * r->rsz), bq(n * r->rsz);
    HIPCHK(hipMemcpyAsync(bs.data(), r->S, n * r->rsz, hipMemcpyDeviceToHost, S_(r)));
    HIPCHK(hipMemcpyAsync(bq.data(), r->Q, n * r->rsz, hipMemcpyDeviceToHost, S_(r)));
    HIPCHK(hipStreamSynchronize(S_(r)));
    for (long long p = 0; p < n; p++) {
      if (r->prec == NADIA_FP64) {
        s[p] = ((double*)bs.data())[p];
        q[p] = std::fabs(((double*)bq.data())[p]);
      } else {
        s[p] = ((float*)bs.data())[p];
        q[p] = std::fabs(((float*)bq.data())[p]);
      }
    }
  }
  std::vector<double> x(2 * n, 0.0);
  if (r->kind == NADIA_PLANAR) {
    // PlanarCDI::initialise_estimate (src/PlanarCDI.c++:39-62)
    srand(seed);
    double qmax = n ? q[0] : 0;
    for (long long p = 0; p < n; p++)
      if (q[p] > qmax) qmax = q[p];
    if (r->prec == NADIA_FP32) qmax = (float)qmax;
    const int max_value = (int)qmax;
    for (long long p = 0; p < n; p++) {
      if (!s[p]) continue;
      // max_value*rand() is an int product in the reference; it wraps
      const int pr = (int)((unsigned)max_value * (unsigned)rand());
      const double re = s[p] * (pr / (double)RAND_MAX);
      const int pi = (int)((unsigned)max_value * (unsigned)rand());
      const double im = s[p] * (pi / (double)RAND_MAX);
      x[2 * p] = re;
      x[2 * p + 1] = im;
    }
  } else if (r->kind == NADIA_FRESNEL) {
    // FresnelCDI::initialise_estimate (src/FresnelCDI.c++:152-201)
    if ((rc = need_setup(r))) return rc;
    srand(seed);
    std::vector<double> tr(2 * n, 0.0);
    for (long long p = 0; p < n; p++) tr[2 * p] = 1 - 0.2 * rand() / ((double)RAND_MAX);
    void* tbuf = r->t[0] ? r->t[0] : nullptr;
    bool own = false;
    if (!tbuf) {
      if ((rc = alloc(&tbuf, n * r->esz))) return rc;
      own = true;
    }
    if (r->prec == NADIA_FP64) HIPCHK(hipMemcpyAsync(tbuf, tr.data(), n * 16, hipMemcpyHostToDevice, S_(r)));
    else {
      std::vector<float> f(tr.begin(), tr.end());
      HIPCHK(hipMemcpyAsync(tbuf, f.data(), n * 8, hipMemcpyHostToDevice, S_(r)));
      HIPCHK(hipStreamSynchronize(S_(r)));
    }
    DISPATCH(r->prec, k_esw_from_trans<T>(tbuf, r->Ls, r->est->d, r->nx, r->ny, S_(r)));
    if ((rc = op_support(r, r->est->d))) return rc;
    HIPCHK(hipStreamSynchronize(S_(r)));
    if (own) (void)hipFree(tbuf);
    r->est->version++;
    return 0;
  } else {
    // FresnelCDI_WF::initialise_estimate (src/FresnelCDI_WF.c++:99-117)
    srand(seed);
    for (long long p = 0; p < n; p++)
      if (s[p]) x[2 * p] = q[p];
  }
  if (r->prec == NADIA_FP64) HIPCHK(hipMemcpyAsync(r->est->d, x.data(), n * 16, hipMemcpyHostToDevice, S_(r)));
  else {
    std::vector<float> f(x.begin(), x.end());
    HIPCHK(hipMemcpyAsync(r->est->d, f.data(), n * 8, hipMemcpyHostToDevice, S_(r)));
    HIPCHK(hipStreamSynchronize(S_(r)));
  }
  HIPCHK(hipStreamSynchronize(S_(r)));
  r->est->version++;
  return 0;
}

static int same_shape(nadia_recon* r, nadia_field* c) {
  if (!r || !c) return fail(NADIA_EINVAL, "null argument");
  if (c->nx != r->nx || c->ny != r->ny || c->prec != r->prec)
    return fail(NADIA_EINVAL, "field does not match the reconstruction's size or precision");
  return 0;
}

int nadia_recon_apply_support(nadia_recon* r, nadia_field* c) {
  int rc;
  if ((rc = same_shape(r, c))) return rc;
  c->version++;
  return op_support(r, c->d);
}

int nadia_recon_project_intensity(nadia_recon* r, nadia_field* c) {
  int rc;
  if ((rc = same_shape(r, c))) return rc;
  c->version++;
  return op_project(r, c->d, r->scal + 9);
}

int nadia_recon_scale_intensity(nadia_recon* r, nadia_field* c) {
  int rc;
  if ((rc = same_shape(r, c))) return rc;
  c->version++;
  return op_scale_intensity(r, c->d, r->scal + 9);
}

int nadia_recon_propagate_to_detector(nadia_recon* r, nadia_field* c) {
  int rc;
  if ((rc = same_shape(r, c))) return rc;
  c->version++;
  return op_to_detector(r, c->d);
}

int nadia_recon_propagate_from_detector(nadia_recon* r, nadia_field* c) {
  int rc;
  if ((rc = same_shape(r, c))) return rc;
  c->version++;
  return op_from_detector(r, c->d);
}

int nadia_recon_get_best_result(nadia_recon* r, int index, double* error, void* host) {
  if (!r) return fail(NADIA_EINVAL, "null recon");
  if (index < 0 || index >= r->n_best) {
    fprintf(stdout, "index out of bounds in BaseCDI::get_best_result\n");
    return fail(NADIA_EINVAL, "index out of bounds in get_best_result");
  }
  double e = 0;
  int buf = 0;
  HIPCHK(hipMemcpyAsync(&e, r->dbest_err + index, sizeof(double), hipMemcpyDeviceToHost, S_(r)));
  HIPCHK(hipMemcpyAsync(&buf, r->dbest_perm + index, sizeof(int), hipMemcpyDeviceToHost, S_(r)));
  HIPCHK(hipStreamSynchronize(S_(r)));
  if (error) *error = e;
  if (host) {
    HIPCHK(hipMemcpyAsync(host, r->best[buf], r->npix * r->esz, hipMemcpyDeviceToHost, S_(r)));
    HIPCHK(hipStreamSynchronize(S_(r)));
  }
  return 0;
}

// reset_best (NADIA/include/BaseCDI.h:476-479): every kept error back to 1
int nadia_recon_reset_best(nadia_recon* r) {
  if (!r) return fail(NADIA_EINVAL, "null recon");
  if (r->n_best <= 0) return 0;
  std::vector<double> e(r->n_best, 1.0);
  HIPCHK(hipMemcpyAsync(r->dbest_err, e.data(), sizeof(double) * r->n_best, hipMemcpyHostToDevice, S_(r)));
  HIPCHK(hipStreamSynchronize(S_(r)));
  return 0;
}

int nadia_recon_get_intensity_autocorrelation(nadia_recon* r, void* host_real) {
  if (!r || !host_real) return fail(NADIA_EINVAL, "null argument");
  void* tmp = nullptr;
  int rc;
  if ((rc = alloc(&tmp, r->npix * r->esz))) return rc;
  DISPATCH(r->prec, k_autocorr_prep<T>(r->Q, tmp, r->npix, S_(r)));
  if ((rc = fft2d(r->ctx, tmp, r->nx, r->ny, r->prec, +1))) return rc;
  if ((rc = op_invert(r, tmp))) return rc;
  DISPATCH(r->prec, k_magnitude<T>(tmp, r->conv, r->npix, S_(r)));
  HIPCHK(hipMemcpyAsync(host_real, r->conv, r->npix * r->rsz, hipMemcpyDeviceToHost, S_(r)));
  HIPCHK(hipStreamSynchronize(S_(r)));
  (void)hipFree(tmp);
  return 0;
}

// ---- FresnelCDI ---------------------------------------------------------

int nadia_fresnel_set_experimental_parameters(nadia_recon* r, double wl, double zfd, double zfs, double pixel) {
  if (!r || r->kind != NADIA_FRESNEL) return fail(NADIA_EINVAL, "not a FresnelCDI reconstruction");
  if (!r->L) return fail(NADIA_ESTATE, "white field not set (nadia_fresnel_setup)");
  // src/FresnelCDI.c++:57-95
  r->wl = wl;
  r->pixel = pixel;
  r->zfd = zfd;
  r->zfs = zfs;
  const double zsd = zfd - zfs;
  r->phase_factor = pixel * pixel * M_PI / (wl) * ((1 / zfd) - (1 / zsd));
  // ready only once the factor tables and the illumination at the sample
  // exist (the fused column pass reads the tables through raw pointers)
  r->fresnel_ready = false;
  int rc;
  // the quadratic phase about ((nx-1)/2, (ny-1)/2) is separable:
  // exp(i f (dx^2 + dy^2)) = exp(i f dx^2) exp(i f dy^2); the fused column
  // pass multiplies the two per-axis factors instead of evaluating cos/sin
  // per pixel
  if (!r->phx && (rc = alloc(&r->phx, (size_t)r->nx * r->esz))) return rc;
  if (!r->phy && (rc = alloc(&r->phy, (size_t)r->ny * r->esz))) return rc;
  for (int axis = 0; axis < 2; axis++) {
    const int n = axis ? r->ny : r->nx;
    std::vector<double> h(2 * (size_t)n);
    for (int i = 0; i < n; i++) {
      const double dd = (n - 1) / 2.0 - i;
      const double ph = r->phase_factor * (dd * dd);
      h[2 * i] = std::cos(ph);
      h[2 * i + 1] = std::sin(ph);
    }
    void* dst = axis ? r->phy : r->phx;
    if (r->prec == NADIA_FP64) {
      HIPCHK(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, S_(r)));
      HIPCHK(hipStreamSynchronize(S_(r)));
    } else {
      std::vector<float> f(h.begin(), h.end());
      HIPCHK(hipMemcpyAsync(dst, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, S_(r)));
      HIPCHK(hipStreamSynchronize(S_(r)));
    }
  }
  if (!r->Ls && (rc = alloc(&r->Ls, r->npix * r->esz))) return rc;
  HIPCHK(hipMemcpyAsync(r->Ls, r->L, r->npix * r->esz, hipMemcpyDeviceToDevice, S_(r)));
  r->fresnel_ready = true;  // op_from_detector checks it
  if ((rc = op_from_detector(r, r->Ls)) == 0) {
    hipError_t e = hipStreamSynchronize(S_(r));
    if (e != hipSuccess) rc = fail(NADIA_EHIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
  }
  if (rc) r->fresnel_ready = false;
  return rc;
}

int nadia_fresnel_setup(nadia_recon* r, const void* white_field, double wl, double zfd, double zfs, double pixel,
                        double normalisation) {
  if (!r || r->kind != NADIA_FRESNEL) return fail(NADIA_EINVAL, "not a FresnelCDI reconstruction");
  if (!white_field) return fail(NADIA_EINVAL, "null white field");
  int rc;
  if (!r->L && (rc = alloc(&r->L, r->npix * r->esz))) return rc;
  // src/FresnelCDI.c++:20-44: illumination = white_field * norm
  r->norm = normalisation;
  HIPCHK(hipMemcpyAsync(r->L, white_field, r->npix * r->esz, hipMemcpyHostToDevice, S_(r)));
  DISPATCH(r->prec, k_scale<T>(r->L, normalisation, r->npix, S_(r)));
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(S_(r)));
  return nadia_fresnel_set_experimental_parameters(r, wl, zfd, zfs, pixel);
}

int nadia_fresnel_set_norm(nadia_recon* r, double nn) {
  if (!r || r->kind != NADIA_FRESNEL || !r->L) return fail(NADIA_ESTATE, "FresnelCDI not set up");
  // src/FresnelCDI.c++:139-148
  const double old = r->norm;
  r->norm = nn;
  DISPATCH(r->prec, k_scale<T>(r->L, nn / old, r->npix, S_(r)));
  if (r->Ls) DISPATCH(r->prec, k_scale<T>(r->Ls, nn / old, r->npix, S_(r)));
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(S_(r)));
  return 0;
}

int nadia_fresnel_set_normalisation(nadia_recon* r, double nn) {
  if (!r || r->kind != NADIA_FRESNEL) return fail(NADIA_EINVAL, "not a FresnelCDI reconstruction");
  r->norm = nn;  // include/FresnelCDI.h:154-156 stores the value only
  return 0;
}

int nadia_fresnel_auto_set_norm(nadia_recon* r) {
  if (!r || r->kind != NADIA_FRESNEL || !r->L) return fail(NADIA_ESTATE, "FresnelCDI not set up");
  // src/FresnelCDI.c++:129-137: sum sqrt(I) / sum |white field|, sums in the
  // field's real type like Real_2D::get_sum
  const long long n = r->npix;
  std::vector<char> bl(n * r->esz), bq(n * r->rsz);
  HIPCHK(hipMemcpyAsync(bl.data(), r->L, n * r->esz, hipMemcpyDeviceToHost, S_(r)));
  HIPCHK(hipMemcpyAsync(bq.data(), r->Q, n * r->rsz, hipMemcpyDeviceToHost, S_(r)));
  HIPCHK(hipStreamSynchronize(S_(r)));
  double wf, in;
  if (r->prec == NADIA_FP64) {
    double a = 0, b = 0;
    const double* L = (const double*)bl.data();
    const double* q = (const double*)bq.data();
    for (long long p = 0; p < n; p++) {
      a += std::sqrt(L[2 * p] * L[2 * p] + L[2 * p + 1] * L[2 * p + 1]);
      b += std::fabs(q[p]);
    }
    wf = a;
    in = b;
  } else {
    float a = 0, b = 0;
    const float* L = (const float*)bl.data();
    const float* q = (const float*)bq.data();
    for (long long p = 0; p < n; p++) {
      a += std::sqrt(L[2 * p] * L[2 * p] + L[2 * p + 1] * L[2 * p + 1]);
      b += std::fabs(q[p]);
    }
    wf = a;
    in = b;
  }
  return nadia_fresnel_set_norm(r, in / wf);
}

int nadia_fresnel_get_illumination_at_sample(nadia_recon* r, void* host) {
  if (!r || r->kind != NADIA_FRESNEL || !r->Ls) return fail(NADIA_ESTATE, "FresnelCDI not set up");
  HIPCHK(hipMemcpyAsync(host, r->Ls, r->npix * r->esz, hipMemcpyDeviceToHost, S_(r)));
  HIPCHK(hipStreamSynchronize(S_(r)));
  return 0;
}

int nadia_fresnel_get_transmission_function(nadia_recon* r, const nadia_field* esw, void* host) {
  if (!r || r->kind != NADIA_FRESNEL || !r->Ls) return fail(NADIA_ESTATE, "FresnelCDI not set up");
  if (!host) return fail(NADIA_EINVAL, "null output");
  const nadia_field* e = esw ? esw : r->est;
  void* tmp = nullptr;
  int rc;
  if ((rc = alloc(&tmp, r->npix * r->esz))) return rc;
  DISPATCH(r->prec, k_trans_from_esw<T>(e->d, r->Ls, tmp, r->npix, S_(r)));
  // the reference applies the transmission constraint here
  // (NADIA/src/FresnelCDI.c++:349-352)
  if (r->tc && (rc = tc_apply(r->tc, S_(r), tmp, r->prec))) {
    (void)hipFree(tmp);
    return rc;
  }
  HIPCHK(hipMemcpyAsync(host, tmp, r->npix * r->esz, hipMemcpyDeviceToHost, S_(r)));
  HIPCHK(hipStreamSynchronize(S_(r)));
  (void)hipFree(tmp);
  return 0;
}

int nadia_fresnel_set_transmission_function(nadia_recon* r, const void* host, nadia_field* esw) {
  if (!r || r->kind != NADIA_FRESNEL || !r->Ls) return fail(NADIA_ESTATE, "FresnelCDI not set up");
  if (!host) return fail(NADIA_EINVAL, "null input");
  nadia_field* e = esw ? esw : r->est;
  void* tmp = nullptr;
  int rc;
  if ((rc = alloc(&tmp, r->npix * r->esz))) return rc;
  HIPCHK(hipMemcpyAsync(tmp, host, r->npix * r->esz, hipMemcpyHostToDevice, S_(r)));
  DISPATCH(r->prec, k_esw_from_trans<T>(tmp, r->Ls, e->d, r->nx, r->ny, S_(r)));
  HIPCHK(hipStreamSynchronize(S_(r)));
  (void)hipFree(tmp);
  e->version++;
  return 0;
}

// ---- FresnelCDI_WF ----------------------------------------------------

int nadia_wf_setup(nadia_recon* r, double wl, double z12, double z23, double pixel) {
  if (!r || r->kind != NADIA_FRESNEL_WF) return fail(NADIA_EINVAL, "not a FresnelCDI_WF reconstruction");
  // src/FresnelCDI_WF.c++:36-58
  r->wl = wl;
  r->z12 = z12;
  r->z23 = z23;
  r->pixel = pixel;
  r->wf_sx = wl * z23 / (pixel * r->nx);
  r->wf_sy = wl * z23 / (pixel * r->ny);
  r->wf_k = M_PI / wl;
  r->wf_norm = 1 / (std::sqrt((double)(r->nx * r->ny)));
  r->wf_ready = false;
  // phi = (pi/lambda)(rho^2/z12 + rho^2/z23), rho^2 = ax^2 + ay^2, is a sum
  // of per-axis terms: the fused column passes multiply two factor tables
  // exp(i phi_x) exp(i phi_y) instead of evaluating cos / sin per pixel
  int rc;
  if (!r->phx && (rc = alloc(&r->phx, (size_t)r->nx * r->esz))) return rc;
  if (!r->phy && (rc = alloc(&r->phy, (size_t)r->ny * r->esz))) return rc;
  for (int axis = 0; axis < 2; axis++) {
    const int n = axis ? r->ny : r->nx;
    const double sc = axis ? r->wf_sy : r->wf_sx;
    std::vector<double> h(2 * (size_t)n);
    for (int i = 0; i < n; i++) {
      const double a = sc * ((n - 1) / 2.0 - i);
      const double ph = (a * a / z12 + a * a / z23) * r->wf_k;
      h[2 * i] = std::cos(ph);
      h[2 * i + 1] = std::sin(ph);
    }
    void* dst = axis ? r->phy : r->phx;
    if (r->prec == NADIA_FP64) {
      HIPCHK(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, S_(r)));
      HIPCHK(hipStreamSynchronize(S_(r)));
    } else {
      std::vector<float> f(h.begin(), h.end());
      HIPCHK(hipMemcpyAsync(dst, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, S_(r)));
      HIPCHK(hipStreamSynchronize(S_(r)));
    }
  }
  r->wf_ready = true;
  return 0;
}

int nadia_wf_set_support(nadia_recon* r, double z_diameter, double size) {
  if (!r || r->kind != NADIA_FRESNEL_WF || !r->wf_ready) return fail(NADIA_ESTATE, "FresnelCDI_WF not set up");
  // src/FresnelCDI_WF.c++:177-196: integer centre (nx-1)/2
  const double centre = (double)((r->nx - 1) / 2);
  const double dz = (r->z12 / r->z23) * r->pixel;
  r->s_soft = true;  // convolved below: the byte copy is not kept
  DISPATCH(r->prec, k_wf_disk<T>(r->S, r->nx, r->ny, centre, dz, size * z_diameter / 2.0, S_(r)));
  LAUNCHCHK();
  // convolve(support, 3, 5), left unnormalised
  HIPCHK(hipMemcpyAsync(r->rtmp, r->S, r->npix * r->rsz, hipMemcpyDeviceToDevice, S_(r)));
  ConvArgs ca{};
  conv_window(r, 3.0, 5, ca.g, ca.gf);
  ca.src = r->rtmp;
  ca.src_complex = 0;
  ca.dst = r->S;
  ca.nx = r->nx;
  ca.ny = r->ny;
  ca.partials = r->partials;
  int nb = 0;
  DISPATCH(r->prec, k_conv<T>(ca, 5, &nb, S_(r)));
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(S_(r)));
  r->s_version++;
  return 0;
}

// ---- measurement ---------------------------------------------------------

// ---- batches ------------------------------------------------------------

// ---- TransmissionConstraint -------------------------------------------

}  // extern "C"

namespace {
constexpr int kMaxRegions = 255;  // region numbers are stored in one byte

int tc_upload(nadia_constraint* tc, hipStream_t st) {
  const size_t npix = (size_t)tc->nx * tc->ny;
  if (tc->map_dirty && tc->nreg > 0) {
    HIPCHK(hipMemcpyAsync(tc->map, tc->hmap.data(), npix, hipMemcpyHostToDevice, st));
  }
  if (tc->dirty) {
    std::vector<double> par(3 * kMaxRegions, 0.0);
    std::vector<int> fx(kMaxRegions, 0);
    for (int k = 0; k < tc->nreg; ++k) {
      par[k] = tc->a1[k];
      par[kMaxRegions + k] = tc->a2[k];
      par[2 * kMaxRegions + k] = tc->cm[k];
      fx[k] = tc->fixed[k];
    }
    HIPCHK(hipMemcpyAsync(tc->dpar, par.data(), par.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(tc->dfixed, fx.data(), fx.size() * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));  // par / fx live on this stack frame
  } else if (tc->map_dirty) {
    HIPCHK(hipStreamSynchronize(st));
  }
  tc->dirty = tc->map_dirty = false;
  return 0;
}

TcDev tc_dev(const nadia_constraint* tc) {
  TcDev d{};
  d.map = tc->map;
  d.nregions = tc->nreg;
  d.alpha1 = tc->dpar;
  d.alpha2 = tc->dpar + kMaxRegions;
  d.c_mean = tc->dpar + 2 * kMaxRegions;
  d.fixed = tc->dfixed;
  d.charge_flip = tc->charge_flip;
  d.flip_sign = tc->flip_sign;
  d.unity = tc->unity;
  return d;
}

// The device view of a constraint (uploaded first), acting on T of the exit
// wave when ill is set.
int tc_view(nadia_constraint* tc, hipStream_t st, const void* ill, int nx, int ny, TcDev* out) {
  int rc;
  if ((rc = tc_upload(tc, st))) return rc;
  TcDev d = tc_dev(tc);
  if (ill) {
    if (tc->custom) return fail(NADIA_EINVAL, "a host constraint function needs the transmission function array");
    d.ill = ill;
    d.nx = nx;
    d.ny = ny;
  }
  *out = d;
  return 0;
}

// The region means of apply_constraint (TransmissionConstraint.c++:81-120)
// over buf, or (ill set) over the transmission function of the exit-surface
// wave buf holds; *out receives the device view for the apply step.
int tc_stats(nadia_constraint* tc, hipStream_t st, const void* buf, int prec, const void* ill, int nx, int ny,
             TcDev* out) {
  int rc;
  if ((rc = tc_upload(tc, st))) return rc;
  const long long n = (long long)tc->nx * tc->ny;
  TcDev d = tc_dev(tc);
  if (ill) {
    if (tc->custom) return fail(NADIA_EINVAL, "a host constraint function needs the transmission function array");
    d.ill = ill;
    d.nx = nx;
    d.ny = ny;
  }
  if (tc->nreg > 0) {
    DISPATCH(prec, k_tc_stats<T>(buf, d, n, tc->partials, st));
    k_tc_final(d, tc->partials, kEwBlocks, st);
    LAUNCHCHK();
  }
  *out = d;
  return 0;
}

// TransmissionConstraint::apply_constraint (NADIA/src/TransmissionConstraint.c++:81-172)
// on buf, or (ill set, no host function) on the transmission function of the
// exit-surface wave buf holds, written back as the wave
int tc_apply(nadia_constraint* tc, hipStream_t st, void* buf, int prec, const void* ill, int nx, int ny) {
  int rc;
  const long long n = (long long)tc->nx * tc->ny;
  TcDev d{};
  if ((rc = tc_stats(tc, st, buf, prec, ill, nx, ny, &d))) return rc;
  // with ill the pass also does the T -> ESW round trip the reference does
  // even when nothing changes T
  if (tc->nreg > 0 || tc->charge_flip || tc->unity || ill) {
    DISPATCH(prec, k_tc_apply<T>(buf, d, n, st));
    LAUNCHCHK();
  }
  if (tc->custom) {
    // the custom constraint is a host function: hand it the array
    const size_t bytes = (size_t)n * (prec == NADIA_FP64 ? 16 : 8);
    std::vector<char> h(bytes);
    HIPCHK(hipMemcpyAsync(h.data(), buf, bytes, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    tc->custom(h.data(), tc->nx, tc->ny, prec, tc->user);
    HIPCHK(hipMemcpyAsync(buf, h.data(), bytes, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return 0;
}
}  // namespace

extern "C" {

int nadia_constraint_create(int nx, int ny, nadia_constraint** out) {
  if (!out || nx <= 0 || ny <= 0) return fail(NADIA_EINVAL, "bad argument");
  nadia_constraint* tc = new nadia_constraint{};
  tc->dev = current_device();
  tc->nx = nx;
  tc->ny = ny;
  // TransmissionConstraint::TransmissionConstraint: unity and charge flipping on
  tc->unity = 1;
  tc->charge_flip = 1;
  tc->flip_sign = 1;
  tc->dirty = true;
  int rc;
  if ((rc = alloc((void**)&tc->map, (size_t)nx * ny)) || (rc = alloc((void**)&tc->dpar, 3 * kMaxRegions * sizeof(double))) ||
      (rc = alloc((void**)&tc->dfixed, kMaxRegions * sizeof(int))) ||
      (rc = alloc((void**)&tc->partials, (size_t)kEwBlocks * kMaxRegions * 2 * sizeof(double)))) {
    nadia_constraint_destroy(tc);
    return rc;
  }
  *out = tc;
  return 0;
}

int nadia_constraint_destroy(nadia_constraint* tc) {
  if (!tc) return 0;
  DeviceCtx* c = dev_ctx(tc->dev);
  if (c) (void)hipStreamSynchronize(c->stream);
  void* ptrs[] = {tc->map, tc->dpar, tc->dfixed, tc->partials};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete tc;
  return 0;
}

int nadia_constraint_set_charge_flipping(nadia_constraint* tc, int enable, int flip_sign) {
  if (!tc) return fail(NADIA_EINVAL, "null constraint");
  tc->charge_flip = enable ? 1 : 0;
  tc->flip_sign = flip_sign > 0 ? 1 : -1;
  return 0;
}

int nadia_constraint_set_enforce_unity(nadia_constraint* tc, int enable) {
  if (!tc) return fail(NADIA_EINVAL, "null constraint");
  tc->unity = enable ? 1 : 0;
  return 0;
}

int nadia_constraint_add_region(nadia_constraint* tc, const double* region, double alpha1, double alpha2,
                                int* index) {
  if (!tc || !region) return fail(NADIA_EINVAL, "null argument");
  if (tc->nreg >= kMaxRegions) return fail(NADIA_EINVAL, "at most %d complex-constraint regions", kMaxRegions);
  const size_t npix = (size_t)tc->nx * tc->ny;
  if (tc->hmap.empty()) tc->hmap.assign(npix, 0);
  const int k = tc->nreg++;
  bool overlap = false;
  for (size_t p = 0; p < npix; ++p)
    if (region[p] > 0) {
      if (tc->hmap[p] > 0) overlap = true;  // the earlier region keeps the pixel
      else tc->hmap[p] = (unsigned char)(k + 1);
    }
  if (overlap)
    fprintf(stdout, "Overlap between complex constraint region detected. The last constraint to be set will be "
                    "ignored in the overlapping region.\n");
  tc->a1.push_back(alpha1);
  tc->a2.push_back(alpha2);
  tc->cm.push_back(0.0);
  tc->fixed.push_back(0);
  tc->dirty = tc->map_dirty = true;
  if (index) *index = k;
  return 0;
}

int nadia_constraint_set_alphas(nadia_constraint* tc, int index, double alpha1, double alpha2) {
  if (!tc || index < 0 || index >= tc->nreg) return fail(NADIA_EINVAL, "bad region index");
  tc->a1[index] = alpha1;
  tc->a2[index] = alpha2;
  tc->dirty = true;
  return 0;
}

int nadia_constraint_set_fixed_c(nadia_constraint* tc, int index, double c_mean) {
  if (!tc || index < 0 || index >= tc->nreg) return fail(NADIA_EINVAL, "bad region index");
  tc->fixed[index] = 1;
  tc->cm[index] = c_mean;
  tc->dirty = true;
  return 0;
}

int nadia_constraint_get_c_mean(nadia_constraint* tc, int index, double* c_mean) {
  if (!tc || !c_mean || index < 0 || index >= tc->nreg) return fail(NADIA_EINVAL, "bad region index");
  DeviceCtx* c = dev_ctx(tc->dev);
  if (!c) return fail(NADIA_EHIP, "stream creation failed");
  if (tc->dirty) {
    *c_mean = tc->cm[index];
    return 0;
  }
  HIPCHK(hipMemcpyAsync(c_mean, tc->dpar + 2 * kMaxRegions + index, sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int nadia_constraint_clear_regions(nadia_constraint* tc) {
  if (!tc) return fail(NADIA_EINVAL, "null constraint");
  tc->nreg = 0;
  tc->hmap.clear();
  tc->a1.clear();
  tc->a2.clear();
  tc->cm.clear();
  tc->fixed.clear();
  tc->dirty = true;
  return 0;
}

int nadia_constraint_set_custom(nadia_constraint* tc, nadia_custom_constraint_fn fn, void* user) {
  if (!tc) return fail(NADIA_EINVAL, "null constraint");
  tc->custom = fn;
  tc->user = user;
  return 0;
}

int nadia_constraint_apply(nadia_constraint* tc, nadia_field* f) {
  if (!tc || !f) return fail(NADIA_EINVAL, "null argument");
  if (f->nx != tc->nx || f->ny != tc->ny) return fail(NADIA_ESIZE, "field and constraint sizes differ");
  DeviceCtx* c = dev_ctx(f->dev);
  if (!c) return fail(NADIA_EHIP, "stream creation failed");
  int rc = tc_apply(tc, c->stream, f->d, f->prec);
  f->version++;
  return rc;
}

int nadia_recon_set_complex_constraint(nadia_recon* r, nadia_constraint* tc) {
  if (!r) return fail(NADIA_EINVAL, "null recon");
  if (tc && (tc->nx != r->nx || tc->ny != r->ny)) return fail(NADIA_ESIZE, "constraint and reconstruction sizes differ");
  if (tc && tc->dev != r->dev) return fail(NADIA_EINVAL, "constraint lives on another device");
  r->tc = tc;
  r->w_valid = r->w2_valid = false;
  return 0;
}

// ---- PhaseDiverseCDI (NADIA/src/PhaseDiverseCDI.c++) ---------------------

}  // extern "C"

namespace {

int pd_gx(const nadia_pd* d, int x, double off) { return (int)(x - off - d->x_min); }
int pd_gy(const nadia_pd* d, int y, double off) { return (int)(y - off - d->y_min); }

// reallocate_object_memory (PhaseDiverseCDI.c++:72-82): a fresh zero array
int pd_realloc(nadia_pd* d, int nx, int ny) {
  DeviceCtx* c = dev_ctx(d->dev);
  if (d->object) {
    HIPCHK(hipStreamSynchronize(c->stream));
    (void)hipFree(d->object);
    d->object = nullptr;
  }
  const size_t b = (size_t)nx * ny * (d->prec == NADIA_FP64 ? 16 : 8);
  int rc;
  if ((rc = alloc(&d->object, b))) return rc;
  HIPCHK(hipMemsetAsync(d->object, 0, b, c->stream));
  d->nx = nx;
  d->ny = ny;
  return 0;
}

template <typename T>
std::vector<double> download_real(nadia_recon* r, const void* dev) {
  std::vector<T> h(r->npix);
  (void)hipMemcpyAsync(h.data(), dev, r->npix * sizeof(T), hipMemcpyDeviceToHost, S_(r));
  (void)hipStreamSynchronize(S_(r));
  return std::vector<double>(h.begin(), h.end());
}

// |illumination at sample| in the element type, as Complex_2D::get_2d(MAG)
template <typename T>
std::vector<double> illum_mag(nadia_recon* r) {
  std::vector<T> h(2 * r->npix);
  (void)hipMemcpyAsync(h.data(), r->Ls, r->npix * 2 * sizeof(T), hipMemcpyDeviceToHost, S_(r));
  (void)hipStreamSynchronize(S_(r));
  std::vector<double> m(r->npix);
  for (long long p = 0; p < r->npix; p++) {
    const T a = h[2 * p], b = h[2 * p + 1];
    m[p] = (double)(T)std::sqrt(a * a + b * b);
  }
  return m;
}

// set_up_weights (PhaseDiverseCDI.c++:721-848), on the host in double with
// the results stored in the frames' element type
template <typename T>
int pd_weights_T(nadia_pd* d) {
  const int nf = (int)d->frames.size();
  std::vector<std::vector<double>> w(nf);
  for (int n = 0; n < nf; n++) {
    nadia_recon* r = d->frames[n].r;
    const std::vector<double> sup = download_real<T>(r, r->S);
    std::vector<double> mag;
    double mx = 0;
    const bool fres = r->kind == NADIA_FRESNEL;
    if (fres) {
      mag = illum_mag<T>(r);
      mx = mag[0];
      for (double v : mag) mx = v > mx ? v : mx;
    }
    w[n].resize(r->npix);
    for (long long p = 0; p < r->npix; p++) {
      double value = fres ? mag[p] / mx : 1;
      value = d->frames[n].alpha * std::pow(value, d->gamma);
      if (sup[p] <= 0.0) value = 0;
      if (!d->parallel) value *= d->beta;
      w[n][p] = (double)(T)value;
    }
  }
  if (d->parallel) {
    std::vector<double> norm((size_t)d->nx * d->ny, 0.0);
    for (int n = 0; n < nf; n++) {
      const nadia_recon* r = d->frames[n].r;
      for (int i_ = 0; i_ < r->nx; i_++)
        for (int j_ = 0; j_ < r->ny; j_++) {
          const int i = pd_gx(d, i_, d->frames[n].x), j = pd_gy(d, j_, d->frames[n].y);
          if (i >= 0 && j >= 0 && i < d->nx && j < d->ny) {
            double& g = norm[(size_t)i * d->ny + j];
            const double v = w[n][(size_t)i_ * r->ny + j_];
            g = (double)(T)(n == 0 ? v : v + g);
          }
        }
    }
    for (int n = 0; n < nf; n++) {
      const nadia_recon* r = d->frames[n].r;
      for (int i_ = 0; i_ < r->nx; i_++)
        for (int j_ = 0; j_ < r->ny; j_++) {
          const int i = pd_gx(d, i_, d->frames[n].x), j = pd_gy(d, j_, d->frames[n].y);
          if (i >= 0 && j >= 0 && i < d->nx && j < d->ny) {
            double& v = w[n][(size_t)i_ * r->ny + j_];
            const double nm = norm[(size_t)i * d->ny + j];
            v = nm <= 0 ? 0 : (double)(T)(d->beta * v / nm);
          }
        }
    }
  }
  DeviceCtx* c = dev_ctx(d->dev);
  for (int n = 0; n < nf; n++) {
    std::vector<T> h(w[n].begin(), w[n].end());
    HIPCHK(hipMemcpyAsync(d->frames[n].weight, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

PdArgs pd_args(nadia_pd* d, int n);

int pd_weights(nadia_pd* d) {
  if (d->weights_set) return 0;
  d->weights_set = true;
  int rc;
  DISPATCH(d->prec, rc = pd_weights_T<T>(d));
  if (rc) return rc;
  // frame geometry for scale_object
  const int nf = (int)d->frames.size();
  DeviceCtx* c = dev_ctx(d->dev);
  if (d->dframes_cap < nf) {
    if (d->dframes) {
      HIPCHK(hipStreamSynchronize(c->stream));
      (void)hipFree(d->dframes);
      d->dframes = nullptr;
    }
    if ((rc = alloc((void**)&d->dframes, sizeof(PdFrameDev) * (nf > 0 ? nf : 1)))) return rc;
    d->dframes_cap = nf;
  }
  std::vector<PdFrameDev> h(nf);
  for (int n = 0; n < nf; n++) h[n] = PdFrameDev{d->frames[n].weight, d->frames[n].r->nx, d->frames[n].r->ny, d->frames[n].x, d->frames[n].y};
  if (nf) {
    HIPCHK(hipMemcpyAsync(d->dframes, h.data(), sizeof(PdFrameDev) * nf, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  // scale_object's covered map: fixed while the weights and positions are
  const long long np = (long long)d->nx * d->ny;
  if (d->covered_cap < np) {
    if (d->covered) {
      HIPCHK(hipStreamSynchronize(c->stream));
      (void)hipFree(d->covered);
      d->covered = nullptr;
    }
    if ((rc = alloc((void**)&d->covered, (size_t)np))) return rc;
    d->covered_cap = np;
  }
  const PdArgs a = pd_args(d, -1);
  DISPATCH(d->prec, k_pd_cover<T>(a, d->dframes, nf, d->covered, c->stream));
  LAUNCHCHK();
  return 0;
}

PdArgs pd_args(nadia_pd* d, int n) {
  PdArgs a{};
  a.object = d->object;
  a.nx = d->nx;
  a.ny = d->ny;
  a.scale = d->scale;
  a.x_min = d->x_min;
  a.y_min = d->y_min;
  a.parallel = d->parallel;
  if (n >= 0) {
    const PdFrame& f = d->frames[n];
    a.f = PdFrameDev{f.weight, f.r->nx, f.r->ny, f.x, f.y};
    a.small = f.result;
  }
  return a;
}

// get_result / set_result (PhaseDiverseCDI.c++:317-335)
int pd_get_result(nadia_pd* d, int n) {
  nadia_recon* r = d->frames[n].r;
  void* res = d->frames[n].result;
  if (r->kind == NADIA_FRESNEL) {
    DISPATCH(r->prec, k_trans_from_esw<T>(r->est->d, r->Ls, res, r->npix, S_(r)));
    LAUNCHCHK();
    if (r->tc) return tc_apply(r->tc, S_(r), res, r->prec);
    return 0;
  }
  HIPCHK(hipMemcpyAsync(res, r->est->d, r->npix * r->esz, hipMemcpyDeviceToDevice, S_(r)));
  return 0;
}

int pd_set_result(nadia_pd* d, int n) {
  nadia_recon* r = d->frames[n].r;
  const void* res = d->frames[n].result;
  if (r->kind == NADIA_FRESNEL) {
    DISPATCH(r->prec, k_esw_from_trans<T>(res, r->Ls, r->est->d, r->nx, r->ny, S_(r)));
    LAUNCHCHK();
  } else {
    HIPCHK(hipMemcpyAsync(r->est->d, res, r->npix * r->esz, hipMemcpyDeviceToDevice, S_(r)));
  }
  r->est->version++;
  return 0;
}

// get_result then add_to_object (PhaseDiverseCDI.c++:317-335, 851-1016); at
// scale 1 without a frame constraint both run as one pass
int pd_add(nadia_pd* d, int n) {
  int rc;
  nadia_recon* r = d->frames[n].r;
  if (d->scale == 1 && !r->tc) {
    if ((rc = pd_weights(d))) return rc;
    const PdArgs a = pd_args(d, n);
    const void* ill = r->kind == NADIA_FRESNEL ? r->Ls : nullptr;
    DISPATCH(d->prec, k_pd_get_add<T>(a, r->est->d, ill, dev_ctx(d->dev)->stream));
    LAUNCHCHK();
    return 0;
  }
  if ((rc = pd_get_result(d, n)) || (rc = pd_weights(d))) return rc;
  const PdArgs a = pd_args(d, n);
  DISPATCH(d->prec, k_pd_add<T>(a, dev_ctx(d->dev)->stream));
  LAUNCHCHK();
  return 0;
}

// get_result then add_to_object of every frame, in frame order.  At scale 1
// with integral positions, no frame constraint and every frame inside the
// object, one pass over the object takes kPdSetMax frames at a time (each
// object pixel adds its frames in order, as the frame-by-frame launches do);
// otherwise frame by frame.
int pd_add_all(nadia_pd* d) {
  int rc;
  const int nf = (int)d->frames.size();
  bool one_pass = d->scale == 1;
  for (const PdFrame& f : d->frames) {
    const int lnx = f.r->nx, lny = f.r->ny;
    if (f.r->tc || f.x != std::floor(f.x) || f.y != std::floor(f.y) || pd_gx(d, 0, f.x) < 0 ||
        pd_gy(d, 0, f.y) < 0 || pd_gx(d, lnx - 1, f.x) >= d->nx || pd_gy(d, lny - 1, f.y) >= d->ny)
      one_pass = false;
  }
  if (!one_pass) {
    for (int n = 0; n < nf; n++)
      if ((rc = pd_add(d, n))) return rc;
    return 0;
  }
  if ((rc = pd_weights(d))) return rc;
  const PdArgs a = pd_args(d, -1);
  for (int k0 = 0; k0 < nf; k0 += kPdSetMax) {
    PdSetList l{};
    l.n = std::min(kPdSetMax, nf - k0);
    for (int k = 0; k < l.n; k++) {
      const PdFrame& f = d->frames[k0 + k];
      l.f[k] = PdFrameDev{f.weight, f.r->nx, f.r->ny, f.x, f.y};
      l.small[k] = f.result;
      l.est[k] = f.r->est->d;
      l.ill[k] = f.r->kind == NADIA_FRESNEL ? f.r->Ls : nullptr;
    }
    DISPATCH(d->prec, k_pd_get_add_all<T>(a, l, dev_ctx(d->dev)->stream));
    LAUNCHCHK();
  }
  return 0;
}

// update_from_object then set_result (PhaseDiverseCDI.c++:1066-1130,
// 317-335) of frames first..first+count-1, kPdSetMax frames per launch
int pd_update_range(nadia_pd* d, int first, int count) {
  int rc;
  if ((rc = pd_weights(d))) return rc;
  const PdArgs a = pd_args(d, -1);
  for (int k0 = first; k0 < first + count; k0 += kPdSetMax) {
    PdSetList l{};
    l.n = std::min(kPdSetMax, first + count - k0);
    for (int k = 0; k < l.n; k++) {
      const PdFrame& f = d->frames[k0 + k];
      l.f[k] = PdFrameDev{f.weight, f.r->nx, f.r->ny, f.x, f.y};
      l.small[k] = f.result;
      l.est[k] = f.r->est->d;
      l.ill[k] = f.r->kind == NADIA_FRESNEL ? f.r->Ls : nullptr;
      f.r->est->version++;
    }
    DISPATCH(d->prec, k_pd_update_set<T>(a, l, dev_ctx(d->dev)->stream));
    LAUNCHCHK();
  }
  return 0;
}

int pd_update(nadia_pd* d, int n) { return pd_update_range(d, n, 1); }

// scale_object (PhaseDiverseCDI.c++:270-313)
int pd_scale(nadia_pd* d, double factor) {
  int rc;
  if ((rc = pd_weights(d))) return rc;
  const PdArgs a = pd_args(d, -1);
  DISPATCH(d->prec, k_pd_scale<T>(a, d->dframes, (int)d->frames.size(), d->covered, factor, dev_ctx(d->dev)->stream));
  LAUNCHCHK();
  return 0;
}

// ---- position refinement (PhaseDiverseCDI.c++:337-719, utils.c:1151-1435) --

// Device buffers of one adjust_positions / align call, freed on scope exit.
struct DevScratch {
  std::vector<void*> p;
  DeviceCtx* c;
  explicit DevScratch(DeviceCtx* ctx) : c(ctx) {}
  int get(void** out, size_t bytes) {
    int rc;
    if ((rc = alloc(out, bytes ? bytes : 1))) return rc;
    p.push_back(*out);
    HIPCHK(hipMemsetAsync(*out, 0, bytes ? bytes : 1, c->stream));
    return 0;
  }
  ~DevScratch() {
    if (p.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (void* q : p) (void)hipFree(q);
  }
};

// align (utils.c:1151-1343) on device arrays of element type T: the images
// and weights zero-padded to twice the larger size, their correlations by
// 2-D DFTs (complex transforms of the real arrays), then the best ratio.
// *found = 0 leaves *ox, *oy unchanged.
template <typename T>
int align_T(DeviceCtx* c, const void* img1, int nx1, int ny1, const void* img2, int nx2, int ny2, int* ox, int* oy,
            int* found, int min_x, int max_x, int min_y, int max_y, const void* w1, const void* w2,
            double overlap_fraction) {
  const int prec = sizeof(T) == 8 ? NADIA_FP64 : NADIA_FP32;
  const int nx = 2 * (nx1 > nx2 ? nx1 : nx2), ny = 2 * (ny1 > ny2 ? ny1 : ny2);
  int rc;
  if ((rc = check_dims(nx, ny))) return rc;
  if (min_x == max_x || min_y == max_y) {
    min_x = (int)(1 - nx / 2.0);
    max_x = (int)(nx / 2.0);
    min_y = (int)(1 - ny / 2.0);
    max_y = (int)(ny / 2.0);
  }
  const size_t np = (size_t)nx * ny, cb = np * sizeof(T) * 2;
  DevScratch sc(c);
  void *i1, *i2, *c1, *c2, *cnt, *part, *best;
  if ((rc = sc.get(&i1, cb)) || (rc = sc.get(&i2, cb)) || (rc = sc.get(&c1, cb)) || (rc = sc.get(&c2, cb)) ||
      (rc = sc.get(&cnt, sizeof(unsigned))) || (rc = sc.get(&part, sizeof(XcBest) * kEwBlocks)) ||
      (rc = sc.get(&best, sizeof(XcBest))))
    return rc;
  XcPlace pl{img1, w1, img2, w2, nx1, ny1, nx2, ny2, i1, i2, c1, c2, nx, ny, (unsigned*)cnt};
  k_xc_place<T>(pl, c->stream);
  LAUNCHCHK();
  for (void* b : {i1, i2, c1, c2})
    if ((rc = fft2d(c, b, nx, ny, prec, -1))) return rc;
  k_xc_mulconj<T>(i1, i2, (long long)np, c->stream);
  k_xc_mulconj<T>(c1, c2, (long long)np, c->stream);
  LAUNCHCHK();
  if ((rc = fft2d(c, i1, nx, ny, prec, +1)) || (rc = fft2d(c, c1, nx, ny, prec, +1))) return rc;
  unsigned count = 0;
  HIPCHK(hipMemcpyAsync(&count, cnt, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  XcScan xs{i1, c1, nx, ny, (double)count * nx * ny, overlap_fraction, min_x, max_x, min_y, max_y};
  k_xc_scan<T>(xs, (XcBest*)part, c->stream);
  k_xc_final((const XcBest*)part, kEwBlocks, (XcBest*)best, c->stream);
  LAUNCHCHK();
  XcBest hb{};
  HIPCHK(hipMemcpyAsync(&hb, best, sizeof(XcBest), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *found = hb.index >= 0;
  if (*found) {
    const int i = (int)(hb.index / ny), j = (int)(hb.index % ny);
    *ox = (i >= nx / 2.0) ? -i + nx : -i;
    *oy = (j >= ny / 2.0) ? -j + ny : -j;
  }
  return 0;
}

int pd_iterate_frame(nadia_pd* d, int n) { return nadia_recon_iterate(d->frames[n].r, 1, nullptr); }

// check_position (PhaseDiverseCDI.c++:582-719): the 3 x 3 positions around
// the current one, each tried with one iteration of the frame from the
// object merged at beta 0.5; the position of least error wins, the step is
// halved when the centre wins; more than 10 moves give up and restore.
int pd_check_position(nadia_pd* d, int n, double step, int min_x, int max_x, int min_y, int max_y, int tries,
                      int* gave_up, int* status) {
  PdFrame& f = d->frames[n];
  const double x = f.x, y = f.y;
  if (step < 1.0 / d->scale) {
    *status = 1;
    return 0;
  }
  if (tries > 10 || x < x + min_x || x > x + max_x || y < y + min_y || y > y + max_y) {
    if (gave_up) *gave_up = 1;
    *status = 0;
    return 0;
  }
  DeviceCtx* c = dev_ctx(d->dev);
  const size_t esz = d->prec == NADIA_FP64 ? 16 : 8;
  const size_t ob = (size_t)d->nx * d->ny * esz, fb = (size_t)f.r->npix * f.r->esz;
  DevScratch sc(c);
  void *obj_copy, *single_copy;
  int rc;
  if ((rc = sc.get(&obj_copy, ob)) || (rc = sc.get(&single_copy, fb))) return rc;
  HIPCHK(hipMemcpyAsync(obj_copy, d->object, ob, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(single_copy, f.result, fb, hipMemcpyDeviceToDevice, c->stream));
  double best_x = x, best_y = y, best_error = 100;
  d->beta = 0.5;
  d->weights_set = false;
  for (int i = -1; i < 2; i++)
    for (int j = -1; j < 2; j++) {
      const double nxp = x + i * step, nyp = y + j * step;
      f.x = nxp;
      f.y = nyp;
      double e = 0;
      if ((rc = pd_add(d, n)) || (rc = pd_update(d, n)) || (rc = pd_iterate_frame(d, n)) ||
          (rc = fetch_error(f.r, &e)))
        return rc;
      if (e < best_error) {
        best_error = e;
        best_x = nxp;
        best_y = nyp;
      }
      HIPCHK(hipMemcpyAsync(d->object, obj_copy, ob, hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(f.result, single_copy, fb, hipMemcpyDeviceToDevice, c->stream));
      if ((rc = pd_set_result(d, n))) return rc;
    }
  f.x = best_x;
  f.y = best_y;
  if (best_x == x && best_y == y) step = step / 2.0;
  int st = 0;
  if ((rc = pd_check_position(d, n, step, min_x, max_x, min_y, max_y, tries + 1, gave_up, &st))) return rc;
  if (!st) {
    f.x = x;
    f.y = y;
    *status = 0;
    return 0;
  }
  d->beta = 1.0;
  d->weights_set = false;
  *status = 1;
  return 0;
}

// adjust_positions (PhaseDiverseCDI.c++:337-579), on a temporary global
// object of ones; the real object is left as it was.
template <typename T>
int pd_adjust_T(nadia_pd* d, int type, bool forward, int min_x, int max_x, int min_y, int max_y, double step,
                int* gave_up) {
  const int nf = (int)d->frames.size();
  DeviceCtx* c = dev_ctx(d->dev);
  const int lnx = d->frames[0].r->nx, lny = d->frames[0].r->ny, sc = d->scale;
  const int mx = lnx * sc, my = lny * sc;
  const int parallel_c = d->parallel, nx_c = d->nx, ny_c = d->ny, xmin_c = d->x_min, ymin_c = d->y_min;
  const double beta_c = d->beta;
  void* const obj_c = d->object;
  DevScratch scr(c);
  void* temp_object;
  int rc;
  if ((rc = scr.get(&temp_object, (size_t)d->nx * d->ny * sizeof(T) * 2))) return rc;
  auto restore = [&]() {
    d->object = obj_c;
    d->nx = nx_c;
    d->ny = ny_c;
    d->x_min = xmin_c;
    d->y_min = ymin_c;
    d->parallel = parallel_c;
    d->beta = beta_c;
    d->weights_set = false;
  };
  d->parallel = 0;
  d->object = temp_object;
  {
    const PdArgs a = pd_args(d, -1);
    k_pd_scale<T>(a, nullptr, 0, nullptr, 0.0, c->stream);  // every pixel 1 + 0i
  }
  d->beta = 1.0;
  d->weights_set = false;
  const int first = forward ? 0 : nf - 1;
  if ((rc = pd_iterate_frame(d, first)) || (rc = pd_get_result(d, first)) || (rc = pd_add(d, first))) {
    restore();
    return rc;
  }
  for (int n = forward ? 1 : nf - 2; n != (forward ? nf : -1); n += forward ? 1 : -1) {
    PdFrame& f = d->frames[n];
    const double bx = f.x, by = f.y;
    if ((rc = pd_iterate_frame(d, n)) || (rc = pd_get_result(d, n))) break;
    XcMaps m{d->object, f.result, nullptr, nullptr, d->nx, d->ny, lnx, lny, mx, my, sc, d->x_min, d->y_min, bx, by};
    DevScratch it(c);
    if (type == 0) {
      void *tothers, *tsingle;
      if ((rc = it.get(&tothers, (size_t)mx * my * sizeof(T))) || (rc = it.get(&tsingle, (size_t)lnx * lny * sizeof(T))))
        break;
      m.others = tothers;
      m.tsingle = tsingle;
      k_xc_maps<T>(m, c->stream);
      const void *img = tsingle, *wgt = f.r->S;
      if (sc != 1) {
        void *ibig, *wbig;
        if ((rc = it.get(&ibig, (size_t)mx * my * sizeof(T))) || (rc = it.get(&wbig, (size_t)mx * my * sizeof(T))))
          break;
        k_interp<T>(tsingle, lnx, lny, ibig, mx, my, c->stream);
        k_interp<T>(f.r->S, lnx, lny, wbig, mx, my, c->stream);
        img = ibig;
        wgt = wbig;
      }
      LAUNCHCHK();
      int new_x = 0, new_y = 0, found = 0;
      if ((rc = align_T<T>(c, img, mx, my, tothers, mx, my, &new_x, &new_y, &found, min_x, max_x, min_y, max_y, wgt,
                           tothers, 0.2)))
        break;
      f.x = bx + new_x / (double)sc;
      f.y = by + new_y / (double)sc;
    } else {
      void* temp;
      if ((rc = it.get(&temp, (size_t)mx * my * sizeof(T) * 2))) break;
      m.others = temp;
      k_pd_subgrid<T>(m, c->stream);
      LAUNCHCHK();
      d->object = temp;
      d->nx = mx;
      d->ny = my;
      d->x_min = (int)(-f.x);
      d->y_min = (int)(-f.y);
      int st = 0;
      rc = pd_check_position(d, n, step, min_x, max_x, min_y, max_y, 0, gave_up ? &gave_up[n] : nullptr, &st);
      d->object = temp_object;
      d->nx = nx_c;
      d->ny = ny_c;
      d->x_min = xmin_c;
      d->y_min = ymin_c;
      if (rc) break;
    }
    d->weights_set = false;
    if ((rc = pd_add(d, n))) break;
  }
  restore();
  return rc;
}

}  // namespace

extern "C" {

int nadia_pd_create(double beta, double gamma, int parallel, int granularity, nadia_pd** out) {
  if (!out || granularity < 1) return fail(NADIA_EINVAL, "bad argument");
  nadia_pd* d = new nadia_pd{};
  d->dev = current_device();
  d->prec = -1;
  d->beta = beta;
  d->gamma = gamma;
  d->parallel = parallel ? 1 : 0;
  d->scale = granularity;
  d->ipc = 1;
  *out = d;
  return 0;
}

int nadia_pd_destroy(nadia_pd* d) {
  if (!d) return 0;
  DeviceCtx* c = dev_ctx(d->dev);
  if (c) (void)hipStreamSynchronize(c->stream);
  for (PdFrame& f : d->frames) {
    if (f.result) (void)hipFree(f.result);
    if (f.weight) (void)hipFree(f.weight);
  }
  if (d->object) (void)hipFree(d->object);
  if (d->dframes) (void)hipFree(d->dframes);
  if (d->covered) (void)hipFree(d->covered);
  if (d->derr) (void)hipFree(d->derr);
  if (d->herr) (void)hipHostFree(d->herr);
  if (d->batch) nadia_batch_destroy(d->batch);
  delete d;
  return 0;
}

// add_new_position (PhaseDiverseCDI.c++:118-178)
int nadia_pd_add_position(nadia_pd* d, nadia_recon* r, double x, double y, double alpha) {
  if (!d || !r) return fail(NADIA_EINVAL, "null argument");
  if (r->kind == NADIA_FRESNEL_WF) return fail(NADIA_EINVAL, "frames are PlanarCDI or FresnelCDI objects");
  if (r->dev != d->dev) return fail(NADIA_EINVAL, "frame lives on another device");
  if (d->prec >= 0 && r->prec != d->prec) return fail(NADIA_EINVAL, "all frames must have one precision");
  if (r->kind == NADIA_FRESNEL && !r->Ls) return fail(NADIA_ESTATE, "FresnelCDI frame not set up");
  d->prec = r->prec;
  d->weights_set = false;
  const int lnx = r->nx, lny = r->ny;
  PdFrame f{r, x, y, alpha, nullptr, nullptr};
  int rc;
  if ((rc = alloc(&f.result, r->npix * r->esz)) || (rc = alloc(&f.weight, r->npix * r->rsz))) {
    if (f.result) (void)hipFree(f.result);
    return rc;
  }
  DeviceCtx* c = dev_ctx(d->dev);
  HIPCHK(hipMemsetAsync(f.result, 0, r->npix * r->esz, c->stream));
  d->frames.push_back(f);
  if (!d->object) {
    d->x_min = (int)(-x);
    d->y_min = (int)(-y);
    if ((rc = pd_realloc(d, lnx * d->scale, lny * d->scale))) return rc;
  }
  int extra_x = 0, extra_y = 0;
  const int gxmin = pd_gx(d, 0, x) * d->scale, gymin = pd_gy(d, 0, y) * d->scale;
  const int gxmax = pd_gx(d, lnx, x) * d->scale, gymax = pd_gy(d, lny, y) * d->scale;
  if (gxmin < 0) {
    d->x_min = (int)(-x);
    extra_x += -gxmin;
  }
  if (gymin < 0) {
    d->y_min = (int)(-y);
    extra_y += -gymin;
  }
  if (gxmax > d->nx) extra_x += gxmax - d->nx;
  if (gymax > d->ny) extra_y += gymax - d->ny;
  if (extra_x || extra_y) return pd_realloc(d, d->nx + extra_x, d->ny + extra_y);
  return 0;
}

int nadia_pd_remove_positions(nadia_pd* d) {
  if (!d) return fail(NADIA_EINVAL, "null argument");
  DeviceCtx* c = dev_ctx(d->dev);
  HIPCHK(hipStreamSynchronize(c->stream));
  for (PdFrame& f : d->frames) {
    (void)hipFree(f.result);
    (void)hipFree(f.weight);
  }
  d->frames.clear();
  d->weights_set = false;
  if (d->batch) {
    nadia_batch_destroy(d->batch);
    d->batch = nullptr;
  }
  return 0;
}

int nadia_pd_set_parameters(nadia_pd* d, int iterations_per_cycle, double beta, double gamma) {
  if (!d || iterations_per_cycle < 0) return fail(NADIA_EINVAL, "bad argument");
  d->ipc = iterations_per_cycle;
  if (beta != d->beta || gamma != d->gamma) d->weights_set = false;
  d->beta = beta;
  d->gamma = gamma;
  return 0;
}

int nadia_pd_set_probe_scaling(nadia_pd* d, int n, double alpha) {
  if (!d) return fail(NADIA_EINVAL, "null argument");
  if (n < 0 || n >= (int)d->frames.size()) {
    fprintf(stdout, "In PhaseDiverseCDI::set_probe_scaling, the probe number given is too large. Are you trying "
                    "to set the value of alpha before calling add_new_position?\n");
    return 0;
  }
  d->frames[n].alpha = alpha;
  d->weights_set = false;
  return 0;
}

// initialise_estimate (PhaseDiverseCDI.c++:182-199)
int nadia_pd_initialise_estimate(nadia_pd* d) {
  if (!d || !d->object) return fail(NADIA_ESTATE, "no frames");
  const PdArgs a = pd_args(d, -1);
  // object = 1 + 0i: scale_object with no frame covering any pixel
  DISPATCH(d->prec, k_pd_scale<T>(a, nullptr, 0, nullptr, 0.0, dev_ctx(d->dev)->stream));
  LAUNCHCHK();
  int rc;
  for (int n = 0; n < (int)d->frames.size(); n++)
    if ((rc = pd_add(d, n))) return rc;
  return 0;
}

// iterate (PhaseDiverseCDI.c++:202-267).  In parallel mode the frames'
// iterations are independent of each other and run as one batch.
int nadia_pd_iterate(nadia_pd* d, double* frame_errors) {
  if (!d || !d->object) return fail(NADIA_ESTATE, "no frames");
  const int nf = (int)d->frames.size();
  int rc;
  if (d->parallel) {
    if ((rc = pd_update_range(d, 0, nf))) return rc;
    if (d->ipc > 0) {
      if (!d->batch) {
        if ((rc = nadia_batch_create(&d->batch))) return rc;
        for (PdFrame& f : d->frames)
          if ((rc = nadia_batch_add(d->batch, f.r))) return rc;
      }
      if ((rc = nadia_batch_iterate(d->batch, d->ipc, nullptr))) return rc;
    }
    if ((rc = pd_scale(d, 1 - d->beta))) return rc;
    if ((rc = pd_add_all(d))) return rc;
  } else {
    for (int n = 0; n < nf; n++) {
      if ((rc = pd_update(d, n))) return rc;
      if (d->ipc > 0 && (rc = nadia_recon_iterate(d->frames[n].r, d->ipc, nullptr))) return rc;
      if ((rc = pd_add(d, n))) return rc;
    }
  }
  if (frame_errors) {
    // every frame's get_error(): gathered on the device, one copy into
    // pinned memory, one synchronisation
    DeviceCtx* c = dev_ctx(d->dev);
    if (d->err_cap < nf) {
      HIPCHK(hipStreamSynchronize(c->stream));
      if (d->derr) (void)hipFree(d->derr);
      if (d->herr) (void)hipHostFree(d->herr);
      d->derr = nullptr;
      d->herr = nullptr;
      d->err_cap = 0;
      if ((rc = alloc((void**)&d->derr, sizeof(double) * nf))) return rc;
      HIPCHK(hipHostMalloc((void**)&d->herr, sizeof(double) * nf, hipHostMallocDefault));
      d->err_cap = nf;
    }
    CopyList l{};
    std::vector<int> which;
    for (int n = 0; n < nf; n++) {
      nadia_recon* r = d->frames[n].r;
      if (r->h_err_valid || !r->last_err) continue;
      l.src[l.n] = r->last_err;
      l.dst[l.n] = d->derr + which.size();
      which.push_back(n);
      if (++l.n == kCopyMax) {
        k_copy_doubles(l, c->stream);
        l.n = 0;
      }
    }
    k_copy_doubles(l, c->stream);
    LAUNCHCHK();
    if (!which.empty())
      HIPCHK(hipMemcpyAsync(d->herr, d->derr, sizeof(double) * which.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t q = 0; q < which.size(); q++) {
      nadia_recon* r = d->frames[which[q]].r;
      r->h_err_cache = d->herr[q];
      r->h_err_valid = true;
    }
    for (int n = 0; n < nf; n++)
      if ((rc = fetch_error(d->frames[n].r, &frame_errors[n]))) return rc;
  }
  return 0;
}

// adjust_positions (PhaseDiverseCDI.c++:337-579)
int nadia_pd_adjust_positions(nadia_pd* d, int type, int forward, int x_min, int x_max, int y_min, int y_max,
                              double step_size, int* gave_up) {
  if (!d || !d->object || d->frames.empty()) return fail(NADIA_ESTATE, "no frames");
  if (type != 0 && type != 1)
    return fail(NADIA_EINVAL, "The alignment type given to PhaseDiverseCDI::adjust_positions is not a known type");
  for (const PdFrame& f : d->frames)
    if (f.r->nx != d->frames[0].r->nx || f.r->ny != d->frames[0].r->ny)
      return fail(NADIA_ESIZE, "adjust_positions needs frames of one size");
  if (gave_up)
    for (size_t n = 0; n < d->frames.size(); n++) gave_up[n] = 0;
  int rc;
  DISPATCH(d->prec, rc = pd_adjust_T<T>(d, type, forward != 0, x_min, x_max, y_min, y_max, step_size, gave_up));
  return rc;
}

int nadia_pd_get_position(const nadia_pd* d, int n, double* x, double* y) {
  if (!d || !x || !y) return fail(NADIA_EINVAL, "null argument");
  if (n < 0 || n >= (int)d->frames.size()) return fail(NADIA_EINVAL, "no frame %d", n);
  *x = d->frames[n].x;
  *y = d->frames[n].y;
  return 0;
}

// align (utils.c:1151-1343) on host arrays (x-major, element type of
// `precision`); w1 / w2 may be null
int nadia_align(int precision, const void* img1, int nx1, int ny1, const void* img2, int nx2, int ny2, int min_x,
                int max_x, int min_y, int max_y, const void* w1, const void* w2, double overlap_fraction,
                int* offset_x, int* offset_y, int* found) {
  if (!img1 || !img2 || !offset_x || !offset_y || !found || nx1 <= 0 || ny1 <= 0 || nx2 <= 0 || ny2 <= 0)
    return fail(NADIA_EINVAL, "bad argument");
  if (precision != NADIA_FP32 && precision != NADIA_FP64) return fail(NADIA_EINVAL, "unknown precision");
  DeviceCtx* c = dev_ctx(current_device());
  const size_t rs = precision == NADIA_FP64 ? 8 : 4;
  DevScratch sc(c);
  void *a, *b, *wa = nullptr, *wb = nullptr;
  int rc;
  if ((rc = sc.get(&a, (size_t)nx1 * ny1 * rs)) || (rc = sc.get(&b, (size_t)nx2 * ny2 * rs))) return rc;
  HIPCHK(hipMemcpyAsync(a, img1, (size_t)nx1 * ny1 * rs, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(b, img2, (size_t)nx2 * ny2 * rs, hipMemcpyHostToDevice, c->stream));
  if (w1) {
    if ((rc = sc.get(&wa, (size_t)nx1 * ny1 * rs))) return rc;
    HIPCHK(hipMemcpyAsync(wa, w1, (size_t)nx1 * ny1 * rs, hipMemcpyHostToDevice, c->stream));
  }
  if (w2) {
    if ((rc = sc.get(&wb, (size_t)nx2 * ny2 * rs))) return rc;
    HIPCHK(hipMemcpyAsync(wb, w2, (size_t)nx2 * ny2 * rs, hipMemcpyHostToDevice, c->stream));
  }
  DISPATCH(precision, rc = align_T<T>(c, a, nx1, ny1, b, nx2, ny2, offset_x, offset_y, found, min_x, max_x, min_y,
                                      max_y, wa, wb, overlap_fraction));
  return rc;
}

int nadia_pd_get_size(const nadia_pd* d, int* nx, int* ny) {
  if (!d || !nx || !ny) return fail(NADIA_EINVAL, "null argument");
  *nx = d->nx;
  *ny = d->ny;
  return 0;
}

int nadia_pd_get_transmission(nadia_pd* d, void* host) {
  if (!d || !host || !d->object) return fail(NADIA_EINVAL, "no object");
  DeviceCtx* c = dev_ctx(d->dev);
  HIPCHK(hipMemcpyAsync(host, d->object, (size_t)d->nx * d->ny * (d->prec == NADIA_FP64 ? 16 : 8),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// set_transmission (PhaseDiverseCDI.c++:94-113)
int nadia_pd_set_transmission(nadia_pd* d, int nx, int ny, const void* host) {
  if (!d || !host) return fail(NADIA_EINVAL, "null argument");
  if (d->prec < 0) return fail(NADIA_ESTATE, "add a frame first (it sets the precision)");
  int rc;
  if (!d->object && (rc = pd_realloc(d, nx, ny))) return rc;
  if (nx != d->nx || ny != d->ny)
    return fail(NADIA_ESIZE, "the transmission function does not have the dimensions of the object");
  DeviceCtx* c = dev_ctx(d->dev);
  HIPCHK(hipMemcpyAsync(d->object, host, (size_t)nx * ny * (d->prec == NADIA_FP64 ? 16 : 8), hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int nadia_batch_create(nadia_batch** out) {
  if (!out) return fail(NADIA_EINVAL, "null output pointer");
  nadia_batch* b = new nadia_batch{};
  b->dev = current_device();
  *out = b;
  return 0;
}

int nadia_batch_destroy(nadia_batch* b) {
  if (!b) return 0;
  DeviceCtx* c = dev_ctx(b->dev);
  if (c) (void)hipStreamSynchronize(c->stream);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->djobs) (void)hipFree(b->djobs);
  if (b->hjobs) (void)hipHostFree(b->hjobs);
  if (b->hev) (void)hipEventDestroy(b->hev);
  delete b;
  return 0;
}

int nadia_batch_add(nadia_batch* b, nadia_recon* r) {
  if (!b || !r) return fail(NADIA_EINVAL, "null argument");
  if (r->dev != b->dev) return fail(NADIA_EINVAL, "reconstruction lives on another device than the batch");
  for (nadia_recon* j : b->jobs)
    if (j == r) return fail(NADIA_EINVAL, "reconstruction already in the batch");
  b->jobs.push_back(r);
  b->gen++;  // a captured graph never replays for another membership
  return 0;
}

int nadia_batch_remove(nadia_batch* b, nadia_recon* r) {
  if (!b || !r) return fail(NADIA_EINVAL, "null argument");
  for (size_t i = 0; i < b->jobs.size(); i++)
    if (b->jobs[i] == r) {
      b->jobs.erase(b->jobs.begin() + i);
      b->gen++;
      return 0;
    }
  return fail(NADIA_EINVAL, "reconstruction not in the batch");
}

int nadia_batch_size(const nadia_batch* b, int* n) {
  if (!b || !n) return fail(NADIA_EINVAL, "null argument");
  *n = (int)b->jobs.size();
  return 0;
}

int nadia_batch_set_graph(nadia_batch* b, int steps) {
  if (!b) return fail(NADIA_EINVAL, "null batch");
  if (steps < 0 || steps > 4096) return fail(NADIA_EINVAL, "graph length %d outside 0..4096", steps);
  b->graph_steps = steps;
  return 0;
}

int nadia_batch_graph_replays(const nadia_batch* b, long long* n) {
  if (!b || !n) return fail(NADIA_EINVAL, "null argument");
  *n = b->graph_replays;
  return 0;
}

}  // extern "C"

namespace {

// One launch group of a batch chunk: jobs of one GroupKey, their descriptors
// at d (device), sorted by the iterations they still run in this chunk
// (descending), so iteration k launches the prefix of jobs with rem > k.
struct BatchLaunch {
  GroupKey g;
  JobDesc* d;
  int cv, rv;
  std::vector<int> rem;
  std::vector<nadia_recon*> jobs;  // in descriptor order
  bool rebuild_w = false;          // every job's W is rebuilt in one launch first
  bool record = false;             // time the column variants on this call's first iterations
  // after iteration k of the first a jobs: the snapshot copies of their
  // kept best estimates (the shared update pass, launched with rank = 1,
  // ranked each job's error), on the group's stream
  int n_best_after(int a, hipStream_t st) const {
    int rc;
    for (int q = 0; q < a; ++q)
      if (wants_n_best(jobs[q]) && (rc = n_best_copy(jobs[q], st))) return rc;
    return 0;
  }
  int active(int k) const {
    int a = 0;
    while (a < (int)rem.size() && rem[a] > k) a++;
    return a;
  }
};

// ---- column tiles of shared launches ----------------------------------------
// The tuner above times one job alone.  A launch shared by several jobs can
// run faster on other column tiles than the job alone (2048^2 FP32: two jobs
// on half-width tiles, three or more on full-width ones; groups of 1024^2 on
// the register-limited ones), so the first eager iterations of a launch group
// of a new (key, job count) run the column variants in turn — every variant
// gives the same estimates and errors — in blocks of kGroupBlock iterations
// timed by events, in the order 0 1 2 2 1 0 (a clock still rising favours no
// variant), and a later call, once those events have completed, keeps the
// variant with the fastest block.  Nothing waits for the device: until then
// the group runs the full-width tiles.  Only a batch that is one launch
// group times and uses them: groups that run side by side disturb each
// other's timings (the 64-job batch of 256..1024 lost 1% in FP32 and 5% in
// FP64 with choices timed there), so they keep the full-width tiles.
struct GroupTune {
  int cv = -1;  // the chosen variant (-1: not yet)
  std::vector<hipEvent_t> ev;  // start/stop pairs around the timed blocks
  std::vector<int> trial_cv;   // the variant of each block
};
std::map<std::pair<GroupKey, int>, GroupTune> g_group_tune;
constexpr int kGroupBlock = 3;  // iterations per timed block
constexpr int kGroupTimed = 2 * kColVariants * kGroupBlock;

// the variant of timed block q (0 1 2 2 1 0)
int group_trial_variant(int q) { return q < kColVariants ? q : 2 * kColVariants - 1 - q; }

// The column variant a shared launch of `njobs` jobs of key g uses on this
// call, and whether this call records the variants' trial iterations.
int group_variant(const GroupKey& g, int njobs, bool* record) {
  *record = false;
  if (njobs < 2 || !tuned_size(g.nx) || env_int("NADIA_TUNE", 1) == 0) return 0;
  std::lock_guard<std::mutex> lk(g_mu);
  GroupTune& t = g_group_tune[std::make_pair(g, njobs)];
  if (t.cv >= 0) return t.cv;
  if (t.ev.empty()) {
    *record = true;
    return 0;
  }
  for (hipEvent_t e : t.ev)
    if (hipEventQuery(e) != hipSuccess) return 0;  // still running (or failed: keep the default)
  std::vector<float> best(kColVariants, 1e30f);
  for (size_t i = 0; i < t.trial_cv.size(); ++i) {
    float ms = 1e30f;
    if (hipEventElapsedTime(&ms, t.ev[2 * i], t.ev[2 * i + 1]) == hipSuccess && ms < best[t.trial_cv[i]])
      best[t.trial_cv[i]] = ms;
  }
  int cv = 0;
  for (int v = 1; v < kColVariants; ++v)
    if (best[v] < best[cv]) cv = v;
  for (hipEvent_t e : t.ev) (void)hipEventDestroy(e);
  t.ev.clear();
  t.trial_cv.clear();
  t.cv = cv;
  return cv;
}

int ensure_side(DeviceCtx* c, int n) {
  if (!c->fork) HIPCHK(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
  while ((int)c->side.size() < n) {
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    c->side.push_back(st);
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    c->join.push_back(e);
  }
  return 0;
}

// Capture `steps` iterations of every launch group into a graph: each group
// is an independent chain of passes (groups touch disjoint jobs), forked off
// the device stream so the graph holds them as parallel branches.
int capture_batch_graph(nadia_batch* b, DeviceCtx* c, const std::vector<BatchLaunch>& launches, int steps) {
  std::vector<long long> sig{steps, (long long)(intptr_t)b->djobs, (long long)b->gen};
  for (const BatchLaunch& L : launches) {
    long long v[] = {L.g.prec, L.g.kind, L.g.nx, L.g.ny, L.g.two, L.g.split, (long long)(L.d - b->djobs),
                     (long long)L.rem.size(), L.cv, L.rv};
    sig.insert(sig.end(), v, v + 10);
    for (const nadia_recon* r : L.jobs) {
      sig.push_back((long long)(intptr_t)r);
      sig.push_back(wants_n_best(r) ? r->n_best : 0);
    }
  }
  if (b->gexec && sig == b->gsig) return 0;
  if (b->gexec) {
    HIPCHK(hipStreamSynchronize(c->stream));
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  const int nl = (int)launches.size();
  int rc;
  if ((rc = ensure_side(c, nl - 1))) return rc;
  hipStream_t st = c->stream;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  rc = 0;
  if (nl > 1 && hipEventRecord(c->fork, st) != hipSuccess) rc = fail(NADIA_EHIP, "event record in capture failed");
  for (int i = 0; i < nl && !rc; ++i) {
    const BatchLaunch& L = launches[i];
    hipStream_t s2 = i == 0 ? st : c->side[i - 1];
    if (i > 0 && hipStreamWaitEvent(s2, c->fork, 0) != hipSuccess) {
      rc = fail(NADIA_EHIP, "stream wait in capture failed");
      break;
    }
    for (int k = 0; k < steps && !rc; ++k) {
      DISPATCH(L.g.prec, rc = launch_fused<T>(c, L.g, L.d, (int)L.rem.size(), k, L.cv, L.rv, nullptr, nullptr, s2,
                                              nullptr, nullptr, 0, k == steps - 1 ? -1 : 0, 1));
      if (!rc) rc = L.n_best_after((int)L.rem.size(), s2);
    }
    if (!rc && i > 0 && hipEventRecord(c->join[i - 1], s2) != hipSuccess)
      rc = fail(NADIA_EHIP, "event record in capture failed");
  }
  for (int i = 1; i < nl && !rc; ++i)
    if (hipStreamWaitEvent(st, c->join[i - 1], 0) != hipSuccess) rc = fail(NADIA_EHIP, "stream join in capture failed");
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(st, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(NADIA_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    b->gexec = nullptr;
    return fail(NADIA_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  b->gsig = sig;
  return 0;
}

// counts[j]: iterations of job j (in the order added).  errors (only when
// every count is equal) receives errors[k*size + j].
int batch_run(nadia_batch* b, const std::vector<int>& counts, double* errors) {
  const int nj = (int)b->jobs.size();
  int nmax = 0;
  bool uniform = true;
  for (int j = 0; j < nj; ++j) {
    if (counts[j] < 0) return fail(NADIA_EINVAL, "negative iteration count for job %d", j);
    if (counts[j] > nmax) nmax = counts[j];
    if (counts[j] != counts[0]) uniform = false;
  }
  if (nmax <= 0) return 0;
  if (!uniform) errors = nullptr;
  DeviceCtx* c = dev_ctx(b->dev);
  if (!c) return fail(NADIA_EHIP, "stream creation failed");
  int rc;
  for (int j = 0; j < nj; ++j) {
    nadia_recon* r = b->jobs[j];
    if (counts[j] == 0) continue;
    if (!r->has_intensity) return fail(NADIA_ESTATE, "a reconstruction in the batch has no intensity");
    if ((rc = need_setup(r))) return rc;
  }
  if (b->cap < nj) {
    // grow geometrically: freeing device memory waits for the device, so a
    // batch that gains a few jobs per call (continuous batching) must not
    // reallocate on every call
    const int cap = std::max(nj, 2 * b->cap);
    if (b->gexec) {  // the graph addresses the old descriptor array
      HIPCHK(hipStreamSynchronize(c->stream));
      (void)hipGraphExecDestroy(b->gexec);
      b->gexec = nullptr;
    }
    if (b->djobs) (void)hipFree(b->djobs);
    b->djobs = nullptr;
    if ((rc = alloc((void**)&b->djobs, sizeof(JobDesc) * cap))) return rc;
    b->cap = cap;
  }
  std::vector<double> tmp;
  // copy the ring slots 0..m-1 of every job to errors rows base..base+m-1
  auto emit = [&](int base, int m) -> int {
    if (!errors) return 0;
    tmp.resize(m);
    for (int j = 0; j < nj; ++j) {
      nadia_recon* r = b->jobs[j];
      if (has_projection(r)) {
        HIPCHK(hipMemcpyAsync(tmp.data(), r->errs, m * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
      } else {
        double v = 0;
        if ((rc = fetch_error(r, &v))) return rc;
        for (int k = 0; k < m; ++k) tmp[k] = v;
      }
      for (int k = 0; k < m; ++k) errors[(size_t)(base + k) * nj + j] = tmp[k];
    }
    return 0;
  };
  int done = 0;
  while (done < nmax) {
    int chunk = nmax - done;
    if (chunk > 4096) chunk = 4096;
    std::vector<int> rem(nj);
    for (int j = 0; j < nj; ++j) rem[j] = std::max(0, std::min(chunk, counts[j] - done));
    // group the fused-path jobs; the rest run the unfused path one by one
    std::map<GroupKey, std::vector<int>> groups;
    std::vector<int> plain;
    for (int j = 0; j < nj; ++j) {
      if (rem[j] == 0) continue;
      if (fused_ok(b->jobs[j])) groups[group_key(b->jobs[j])].push_back(j);
      else plain.push_back(j);
    }
    std::vector<BatchLaunch> launches;
    std::vector<JobDesc> host;
    host.reserve(nj);
    for (auto& kv : groups) {
      std::stable_sort(kv.second.begin(), kv.second.end(), [&](int x, int y) { return rem[x] > rem[y]; });
      // every job's estimate changed since its last iteration (e.g. frames
      // updated from a PhaseDiverseCDI object): one W rebuild for the group
      // after the descriptors are up, instead of one launch per job
      bool all_stale = kv.second.size() > 1;
      for (int j : kv.second) all_stale = all_stale && w_stale(b->jobs[j]);
      for (int j : kv.second) {
        nadia_recon* r = b->jobs[j];
        if (all_stale) {
          r->w_valid = true;
          r->w_version = r->est->version;
        }
        DISPATCH(r->prec, rc = fused_ready<T>(r, false));
        if (rc) return rc;
      }
      nadia_recon* r0 = b->jobs[kv.second[0]];
      // both twiddle tables exist before a graph capture can start
      if ((rc = ensure_twiddles(c, kv.first.nx, kv.first.ny, kv.first.prec))) return rc;
      // the autotuner times one job alone; a launch shared by several jobs
      // takes the column tiles timed for its key and job count (full-width
      // tiles until then), unless the caller fixed the variants.  Every
      // variant gives the same estimates and errors.
      bool record = false;
      const int group_cv = (kv.second.size() > 1 && !r0->tile_fixed)
                               ? (groups.size() == 1 ? group_variant(kv.first, (int)kv.second.size(), &record) : 0)
                               : r0->col_v;
      BatchLaunch L{kv.first, b->djobs + host.size(), group_cv, r0->row_v, {}, {}};
      L.rebuild_w = all_stale;
      L.record = record;
      for (int j : kv.second) {
        host.push_back(make_jobdesc(b->jobs[j], b->jobs[j]->est->d, b->jobs[j]->W, b->jobs[j]->errs));
        L.rem.push_back(rem[j]);
        L.jobs.push_back(b->jobs[j]);
      }
      launches.push_back(std::move(L));
    }
    if (!host.empty()) {
      // through pinned staging: the copy is asynchronous, and the staging is
      // reused only after the previous copy from it has completed
      if (b->hcap < (int)host.size()) {
        const int cap = std::max((int)host.size(), 2 * b->hcap);
        if (b->hev_pending) HIPCHK(hipEventSynchronize(b->hev));
        if (b->hjobs) (void)hipHostFree(b->hjobs);
        b->hjobs = nullptr;
        b->hcap = 0;
        HIPCHK(hipHostMalloc((void**)&b->hjobs, sizeof(JobDesc) * cap, hipHostMallocDefault));
        b->hcap = cap;
      } else if (b->hev_pending) {
        HIPCHK(hipEventSynchronize(b->hev));
      }
      if (!b->hev) HIPCHK(hipEventCreateWithFlags(&b->hev, hipEventDisableTiming));
      std::memcpy(b->hjobs, host.data(), sizeof(JobDesc) * host.size());
      HIPCHK(hipMemcpyAsync(b->djobs, b->hjobs, sizeof(JobDesc) * host.size(), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipEventRecord(b->hev, c->stream));
      b->hev_pending = true;
    }
    for (const BatchLaunch& L : launches) {
      if (!L.rebuild_w) continue;
      nadia_recon* r0 = L.jobs[0];
      void* twy = twiddles(c, r0->ny, r0->prec, true);
      if (!twy) return fail(NADIA_EHIP, "twiddle table allocation failed");
      RowArgs ra = w_rebuild_args(r0, twy);
      ra.jobs = L.d;
      DISPATCH(r0->prec, launch_row<T>(ra, (int)L.jobs.size(), L.rv, c->stream));
      LAUNCHCHK();
    }
    // graph replays for the uniform, fused-only part; the rest eagerly
    const int G = b->graph_steps;
    const bool graph = G > 0 && uniform && plain.empty() && !launches.empty() && chunk >= G;
    int eager_from = 0, last_slot = chunk - 1;
    if (graph) {
      if ((rc = capture_batch_graph(b, c, launches, G))) return rc;
      const int reps = chunk / G;
      for (int q = 0; q < reps; ++q) {
        HIPCHK(hipGraphLaunch(b->gexec, c->stream));
        b->graph_replays++;
        if ((rc = emit(done + q * G, G))) return rc;
      }
      eager_from = reps * G;
      last_slot = eager_from < chunk ? chunk - eager_from - 1 : G - 1;
    }
    // eager iterations: each launch group runs its chain of passes on its own
    // stream (groups share no memory), forked off the device stream and
    // joined back before the errors are read; unfused jobs stay on the
    // device stream
    const int nl = (int)launches.size();
    const bool fork = nl > 1 && eager_from < chunk;
    if (fork) {
      if ((rc = ensure_side(c, nl - 1))) return rc;
      HIPCHK(hipEventRecord(c->fork, c->stream));
      for (int i = 1; i < nl; ++i) HIPCHK(hipStreamWaitEvent(c->side[i - 1], c->fork, 0));
    }
    for (int i = 0; i < nl; ++i) {
      const BatchLaunch& L = launches[i];
      hipStream_t s2 = (fork && i > 0) ? c->side[i - 1] : c->stream;
      // timing the column variants: the first kGroupTimed iterations, when
      // every job of the group runs all of them, in blocks
      const int njl = (int)L.jobs.size();
      const bool rec = L.record && chunk - eager_from >= kGroupTimed && L.active(eager_from + kGroupTimed - 1) == njl;
      std::vector<hipEvent_t> tev;
      std::vector<int> tcv;
      for (int k = eager_from; k < chunk; ++k) {
        const int a = L.active(k);
        if (a == 0) break;  // jobs are sorted by remaining count: none later either
        // the jobs whose last iteration this is: [active(k + 1), a)
        const int a_next = k + 1 < chunk ? L.active(k + 1) : 0;
        const int q = k - eager_from;
        const bool timed = rec && q < kGroupTimed;
        const int cv = timed ? group_trial_variant(q / kGroupBlock) : L.cv;
        if (timed && q % kGroupBlock == 0) {
          hipEvent_t e0, e1;
          HIPCHK(hipEventCreate(&e0));
          HIPCHK(hipEventCreate(&e1));
          tev.push_back(e0);
          tev.push_back(e1);
          tcv.push_back(cv);
          HIPCHK(hipEventRecord(e0, s2));
        }
        DISPATCH(L.g.prec, rc = launch_fused<T>(c, L.g, L.d, a, k - eager_from, cv, L.rv, nullptr, nullptr, s2,
                                                nullptr, nullptr, a_next, a - a_next, 1));
        if (rc || (rc = L.n_best_after(a, s2))) return rc;
        if (timed && q % kGroupBlock == kGroupBlock - 1) HIPCHK(hipEventRecord(tev.back(), s2));
      }
      if (rec) {
        std::lock_guard<std::mutex> lk(g_mu);
        GroupTune& t = g_group_tune[std::make_pair(L.g, njl)];
        if (t.cv < 0 && t.ev.empty()) {
          t.ev = std::move(tev);
          t.trial_cv = std::move(tcv);
        } else {
          for (hipEvent_t e : tev) (void)hipEventDestroy(e);
        }
      }
    }
    // jobs outside the shared launches: ER with a transmission constraint
    // takes its own fused iterations (as when it runs alone), the rest the
    // unfused path
    for (int j : plain) {
      if (fused_tc_ok(b->jobs[j])) DISPATCH(b->jobs[j]->prec, rc = fused_ready<T>(b->jobs[j]));
      else if (any_fused_ok(b->jobs[j])) DISPATCH(b->jobs[j]->prec, rc = any_ready<T>(b->jobs[j]));
      if (rc) return rc;
    }
    for (int k = eager_from; k < chunk; ++k)
      for (int j : plain) {
        nadia_recon* r = b->jobs[j];
        if (rem[j] <= k) continue;
        if (fused_tc_ok(r)) {
          DISPATCH(r->prec, rc = fused_tc_step<T>(r, k - eager_from));
          if (rc) return rc;
          continue;
        }
        if (any_fused_ok(r)) {
          DISPATCH(r->prec, rc = any_step<T>(r, k - eager_from));
          if (rc) return rc;
          continue;
        }
        if ((rc = iterate_plain(r, r->errs + (k - eager_from)))) return rc;
        if (wants_n_best(r) && (rc = update_n_best(r, current_error_slot(r, r->errs + (k - eager_from)), c->stream)))
          return rc;
      }
    if (fork) {
      for (int i = 1; i < nl; ++i) {
        HIPCHK(hipEventRecord(c->join[i - 1], c->side[i - 1]));
        HIPCHK(hipStreamWaitEvent(c->stream, c->join[i - 1], 0));
      }
    }
    if (eager_from < chunk && (rc = emit(done + eager_from, chunk - eager_from))) return rc;
    for (int j = 0; j < nj; ++j) {
      if (rem[j] == 0) continue;
      nadia_recon* r = b->jobs[j];
      if (fused_ok(r)) fused_done(r, graph ? last_slot + 1 : rem[j]);
      else if (fused_tc_ok(r) || any_fused_ok(r)) fused_done(r, rem[j]);
      else r->est->version++;
    }
    done += chunk;
    if (done < nmax) HIPCHK(hipStreamSynchronize(c->stream));
  }
  // keep get_error() valid after the rings are reused: one launch per
  // kCopyMax jobs
  CopyList l{};
  for (int j = 0; j < nj; ++j) {
    nadia_recon* r = b->jobs[j];
    if (counts[j] > 0 && r->last_err && r->last_err != r->scal + 8) {
      l.src[l.n] = r->last_err;
      l.dst[l.n] = r->scal + 8;
      r->last_err = r->scal + 8;
      r->herr_fresh = false;
      if (++l.n == kCopyMax) {
        k_copy_doubles(l, c->stream);
        l.n = 0;
      }
    }
  }
  k_copy_doubles(l, c->stream);
  LAUNCHCHK();
  return 0;
}

}  // namespace

extern "C" {

int nadia_batch_iterate(nadia_batch* b, int n, double* errors) {
  if (!b) return fail(NADIA_EINVAL, "null batch");
  if (n <= 0 || b->jobs.empty()) return 0;
  return batch_run(b, std::vector<int>(b->jobs.size(), n), errors);
}

int nadia_batch_iterate_each(nadia_batch* b, const int* counts) {
  if (!b || (!counts && !b->jobs.empty())) return fail(NADIA_EINVAL, "null argument");
  if (b->jobs.empty()) return 0;
  return batch_run(b, std::vector<int>(counts, counts + b->jobs.size()), nullptr);
}

// ---- asynchronous use ------------------------------------------------------

}  // extern "C"

struct nadia_event {
  int dev;
  hipEvent_t ev;
  bool recorded;
};

extern "C" {

int nadia_event_create(nadia_event** out) {
  if (!out) return fail(NADIA_EINVAL, "null output pointer");
  nadia_event* e = new nadia_event{};
  e->dev = current_device();
  if (!dev_ctx(e->dev)) {
    delete e;
    return fail(NADIA_EHIP, "stream creation failed");
  }
  if (hipEventCreateWithFlags(&e->ev, hipEventDisableTiming) != hipSuccess) {
    delete e;
    return fail(NADIA_EHIP, "hipEventCreateWithFlags failed");
  }
  *out = e;
  return 0;
}

int nadia_event_destroy(nadia_event* e) {
  if (!e) return 0;
  (void)hipEventDestroy(e->ev);
  delete e;
  return 0;
}

int nadia_event_record(nadia_event* e) {
  if (!e) return fail(NADIA_EINVAL, "null event");
  HIPCHK(hipEventRecord(e->ev, dev_ctx(e->dev)->stream));
  e->recorded = true;
  return 0;
}

int nadia_event_query(nadia_event* e, int* done) {
  if (!e || !done) return fail(NADIA_EINVAL, "null argument");
  if (!e->recorded) {
    *done = 1;
    return 0;
  }
  const hipError_t q = hipEventQuery(e->ev);
  if (q == hipSuccess) *done = 1;
  else if (q == hipErrorNotReady) *done = 0;
  else return fail(NADIA_EHIP, "hipEventQuery: %s", hipGetErrorString(q));
  return 0;
}

int nadia_event_synchronize(nadia_event* e) {
  if (!e) return fail(NADIA_EINVAL, "null event");
  if (e->recorded) HIPCHK(hipEventSynchronize(e->ev));
  return 0;
}

int nadia_host_alloc(size_t bytes, void** out) {
  if (!out) return fail(NADIA_EINVAL, "null output pointer");
  HIPCHK(hipHostMalloc(out, bytes ? bytes : 8, hipHostMallocDefault));
  return 0;
}

int nadia_host_free(void* p) {
  if (p) HIPCHK(hipHostFree(p));
  return 0;
}

int nadia_recon_get_error_async(nadia_recon* r, double* host) {
  if (!r || !host) return fail(NADIA_EINVAL, "null argument");
  if (r->h_err_valid) {
    *host = r->h_err_cache;
    return 0;
  }
  if (!r->last_err) {  // never projected: get_error() reports 0
    *host = 0.0;
    return 0;
  }
  HIPCHK(hipMemcpyAsync(host, r->last_err, sizeof(double), hipMemcpyDeviceToHost, S_(r)));
  return 0;
}

int nadia_tile_variant_count(int* col_variants, int* row_variants) {
  if (!col_variants || !row_variants) return fail(NADIA_EINVAL, "null argument");
  *col_variants = kColVariants;
  *row_variants = kRowVariants;
  return 0;
}

int nadia_recon_set_tile_variants(nadia_recon* r, int col_variant, int row_variant) {
  if (!r) return fail(NADIA_EINVAL, "null recon");
  if (col_variant < 0 || row_variant < 0) {  // back to the autotuner
    r->tile_fixed = false;
    r->col_v = r->row_v = -1;
    return 0;
  }
  if (col_variant >= kColVariants || row_variant >= kRowVariants)
    return fail(NADIA_EINVAL, "tile variant out of range (column < %d, row < %d)", kColVariants, kRowVariants);
  r->col_v = col_variant;
  r->row_v = row_variant;
  r->tile_fixed = true;
  return 0;
}

int nadia_recon_get_tile_variants(const nadia_recon* r, int* col_variant, int* row_variant) {
  if (!r) return fail(NADIA_EINVAL, "null recon");
  if (col_variant) *col_variant = r->col_v;
  if (row_variant) *row_variant = r->row_v;
  return 0;
}

int nadia_recon_set_timing(nadia_recon* r, int on) {
  if (!r) return fail(NADIA_EINVAL, "null recon");
  r->timing = on != 0;
  r->pass_ms[0] = r->pass_ms[1] = 0;
  r->pass_n[0] = r->pass_n[1] = 0;
  return 0;
}

int nadia_recon_pass_times(nadia_recon* r, double ms[2], long long count[2]) {
  if (!r) return fail(NADIA_EINVAL, "null recon");
  for (int i = 0; i < 2; i++) {
    if (ms) ms[i] = r->pass_ms[i];
    if (count) count[i] = r->pass_n[i];
  }
  return 0;
}

}  // extern "C"
