// fft3d.hip -- 3D real FFT plan, kernel dispatch by line length, and the
// mm_fft_* entry points that stand in for f_trans / inv_trans (ModFunc.f:15-40).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "fft3d.h"
#include "fft_pass.h"
#include "mm_common.h"

namespace mm {

bool fft_length_ok(int n) {
  switch (n) {
    case 8: case 12: case 16: case 24: case 32: case 48: case 64: case 96:
    case 128: case 192: case 256: case 384: case 512:
      return true;
  }
  return false;
}

static bool xlen_ok(int l) {
  switch (l) {
    case 4: case 6: case 8: case 12: case 16: case 24: case 32: case 48: case 64:
    case 96: case 128: case 192: case 256:
      return true;
  }
  return false;
}

void make_twiddles(int n, double* out) {
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int e = 0; e <= n / 2; ++e) {
    double c, s;
    long q4 = 4L * e;
    if (q4 % n == 0) {
      static const double cc[4] = {1.0, 0.0, -1.0, 0.0}, ss[4] = {0.0, 1.0, 0.0, -1.0};
      int q = (int)(q4 / n);
      c = cc[q];
      s = ss[q];
    } else {
      long double a = twopi * (long double)e / (long double)n;
      c = (double)cosl(a);
      s = (double)sinl(a);
    }
    out[2 * e] = c;
    out[2 * e + 1] = -s;
  }
  for (int e = n / 2 + 1; e < n; ++e) {
    out[2 * e] = out[2 * (n - e)];
    out[2 * e + 1] = -out[2 * (n - e) + 1];
  }
}

static int upload_tw(int n, int prec, void** dst) {
  std::vector<double> d(2 * (size_t)n);
  make_twiddles(n, d.data());
  size_t bytes;
  std::vector<float> f;
  const void* src;
  if (prec == MM_FP64) {
    bytes = d.size() * sizeof(double);
    src = d.data();
  } else {
    f.resize(d.size());
    for (size_t i = 0; i < d.size(); ++i) f[i] = (float)d[i];
    bytes = f.size() * sizeof(float);
    src = f.data();
  }
  MM_HIP(hipMalloc(dst, bytes));
  MM_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
  return 0;
}

int Fft3::init(int nx_, int ny_, int nz_, int prec_) {
  if (!fft_length_ok(nx_) || !fft_length_ok(ny_) || !fft_length_ok(nz_)) return MM_ERR_ARG;
  if (!xlen_ok(nx_ / 2)) return MM_ERR_ARG;
  if (prec_ != MM_FP64 && prec_ != MM_MIXED) return MM_ERR_ARG;
  nx = nx_; ny = ny_; nz = nz_; H = nx / 2 + 1; prec = prec_;
  MM_TRY(upload_tw(nx / 2, prec, &twLx));
  MM_TRY(upload_tw(nx, prec, &twNx));
  MM_TRY(upload_tw(ny, prec, &twNy));
  MM_TRY(upload_tw(nz, prec, &twNz));
  return 0;
}

void Fft3::destroy() {
  void* p[4] = {twLx, twNx, twNy, twNz};
  for (void* q : p)
    if (q) (void)hipFree(q);
  twLx = twNx = twNy = twNz = nullptr;
}

// ---------------------------------------------------------------- probe
void Probe::begin(int k, hipStream_t s) {
  if (k != kind) return;
  if ((int)ev.size() < 2 * (used + 1)) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    ev.push_back(a);
    ev.push_back(b);
  }
  (void)hipEventRecord(ev[2 * used], s);
}
void Probe::end(int k, hipStream_t s, double b) {
  if (k != kind || (int)ev.size() < 2 * (used + 1)) return;
  (void)hipEventRecord(ev[2 * used + 1], s);
  used += 1;
  bytes += b;
  launches += 1;
}
int Probe::collect() {
  for (int i = 0; i < used; ++i) {
    float t = 0.f;
    MM_HIP(hipEventSynchronize(ev[2 * i + 1]));
    MM_HIP(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]));
    ms += (double)t;
  }
  used = 0;
  return 0;
}
void Probe::reset() { used = 0; bytes = 0.0; ms = 0.0; launches = 0; }
void Probe::destroy() {
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  ev.clear();
  reset();
}

// ------------------------------------------------------------- dispatch
#define MM_XLENS(X) X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256)
#define MM_NLENS(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

template <typename T>
static Batch<T> make_batch(const void* const* in, void* const* out, int nf, const int* gate) {
  Batch<T> b;
  for (int i = 0; i < kMaxBatch; ++i) { b.in[i] = nullptr; b.out[i] = nullptr; }
  for (int i = 0; i < nf; ++i) { b.in[i] = in[i]; b.out[i] = out[i]; }
  b.gate = gate;
  return b;
}

// plan variant v of length L: 0 = Plan, 1 = PlanLean, 2 = PlanLean4
template <int L, int V> struct PV { using type = Plan<L>; };
template <int L> struct PV<L, 1> { using type = PlanLean<L>; };
template <int L> struct PV<L, 2> { using type = PlanLean4<L>; };

template <typename T, class P>
static void go_x_r2c(const Batch<T>& b, int nf, const void* twL, const void* twN, int ny, int nz,
                     T scale, hipStream_t s) {
  using TL = TileLM<T, P>;
  dim3 grid((ny + TL::B - 1) / TL::B, nz, nf);
  hipLaunchKernelGGL((k_x_r2c<T, P>), grid, dim3(TL::THREADS), 0, s, b, (const C2<T>*)twL,
                     (const C2<T>*)twN, ny, scale);
}
template <typename T, class P>
static void go_x_c2r(const Batch<T>& b, int nf, const void* twL, const void* twN, int ny, int nz,
                     hipStream_t s) {
  using TL = TileLM<T, P>;
  dim3 grid((ny + TL::B - 1) / TL::B, nz, nf);
  hipLaunchKernelGGL((k_x_c2r<T, P>), grid, dim3(TL::THREADS), 0, s, b, (const C2<T>*)twL,
                     (const C2<T>*)twN, ny);
}
template <typename T, class P>
static void go_y(const Batch<T>& b, int nf, const void* tw, int nlines, int dir, hipStream_t s) {
  using TL = TileLM<T, P>;
  dim3 grid((nlines + TL::B - 1) / TL::B, 1, nf);
  if (dir < 0)
    hipLaunchKernelGGL((k_y_pass<T, P, -1>), grid, dim3(TL::THREADS), 0, s, b, (const C2<T>*)tw, nlines);
  else
    hipLaunchKernelGGL((k_y_pass<T, P, 1>), grid, dim3(TL::THREADS), 0, s, b, (const C2<T>*)tw, nlines);
}
template <typename T, class P>
static void go_z(const Batch<T>& b, int nf, const void* tw, int ncols, int dir, hipStream_t s) {
  using TE = TileEM<T, P>;
  dim3 grid((ncols + TE::B - 1) / TE::B, 1, nf);
  if (dir < 0)
    hipLaunchKernelGGL((k_z_pass<T, P, -1>), grid, dim3(TE::THREADS), 0, s, b, (const C2<T>*)tw, ncols);
  else
    hipLaunchKernelGGL((k_z_pass<T, P, 1>), grid, dim3(TE::THREADS), 0, s, b, (const C2<T>*)tw, ncols);
}

// variant dispatch: lengths below 64 have a single plan
#define MM_VARIANTS(GO, L_, var, ...)                                              \
  do {                                                                            \
    if constexpr (L_ >= 64) {                                                     \
      if ((var) == 1) { GO<T, typename PV<L_, 1>::type>(__VA_ARGS__); break; }    \
      if ((var) == 2) { GO<T, typename PV<L_, 2>::type>(__VA_ARGS__); break; }    \
    }                                                                             \
    GO<T, typename PV<L_, 0>::type>(__VA_ARGS__);                                 \
  } while (0)

template <typename T>
int Fft3::x_r2c(const T* const* in, void* const* out, int nf, hipStream_t s, T scale) const {
  if (nf < 1 || nf > kMaxBatch) return MM_ERR_ARG;
  Batch<T> b = make_batch<T>((const void* const*)in, out, nf, gate);
  if (probe) probe->begin(MM_PROBE_X_R2C, s);
  switch (nx / 2) {
#define X(L_) case L_: MM_VARIANTS(go_x_r2c, L_, var[0], b, nf, twLx, twNx, ny, nz, scale, s); break;
    MM_XLENS(X)
#undef X
    default: return MM_ERR_ARG;
  }
  if (probe) probe->end(MM_PROBE_X_R2C, s, (double)nf * (nreal() * sizeof(T) + ncplx() * 2 * sizeof(T)));
  MM_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int Fft3::x_c2r(void* const* in, T* const* out, int nf, hipStream_t s) const {
  if (nf < 1 || nf > kMaxBatch) return MM_ERR_ARG;
  Batch<T> b = make_batch<T>((const void* const*)in, (void* const*)out, nf, gate);
  if (probe) probe->begin(MM_PROBE_X_C2R, s);
  switch (nx / 2) {
#define X(L_) case L_: MM_VARIANTS(go_x_c2r, L_, var[1], b, nf, twLx, twNx, ny, nz, s); break;
    MM_XLENS(X)
#undef X
    default: return MM_ERR_ARG;
  }
  if (probe) probe->end(MM_PROBE_X_C2R, s, (double)nf * (nreal() * sizeof(T) + ncplx() * 2 * sizeof(T)));
  MM_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int Fft3::y_pass(void* const* in, void* const* out, int nf, int dir, hipStream_t s) const {
  if (nf < 1 || nf > kMaxBatch) return MM_ERR_ARG;
  Batch<T> b = make_batch<T>((const void* const*)in, out, nf, gate);
  int nlines = nz * H;
  if (probe) probe->begin(MM_PROBE_Y, s);
  switch (ny) {
#define X(L_) case L_: MM_VARIANTS(go_y, L_, var[2], b, nf, twNy, nlines, dir, s); break;
    MM_NLENS(X)
#undef X
    default: return MM_ERR_ARG;
  }
  if (probe) probe->end(MM_PROBE_Y, s, (double)nf * 2.0 * ncplx() * 2 * sizeof(T));
  MM_HIP(hipGetLastError());
  return 0;
}

template <typename T, class P>
static void go_yt(const Batch<T>& b, int nf, const void* tw, int nz, int H, int dir, int mode, hipStream_t s) {
  using TL = TileLM<T, P>;
  dim3 grid((nz * H + TL::B - 1) / TL::B, 1, nf);
  const C2<T>* t = (const C2<T>*)tw;
  if (dir < 0 && mode == 1) hipLaunchKernelGGL((k_y_pass_t<T, P, -1, 1>), grid, dim3(TL::THREADS), 0, s, b, t, nz, H);
  else if (dir < 0) hipLaunchKernelGGL((k_y_pass_t<T, P, -1, 2>), grid, dim3(TL::THREADS), 0, s, b, t, nz, H);
  else if (mode == 1) hipLaunchKernelGGL((k_y_pass_t<T, P, 1, 1>), grid, dim3(TL::THREADS), 0, s, b, t, nz, H);
  else hipLaunchKernelGGL((k_y_pass_t<T, P, 1, 2>), grid, dim3(TL::THREADS), 0, s, b, t, nz, H);
}

template <typename T>
int Fft3::y_pass_t(void* const* in, void* const* out, int nf, int dir, int mode, hipStream_t s) const {
  if (nf < 1 || nf > kMaxBatch || (mode != 1 && mode != 2) || ny % kTW) return MM_ERR_ARG;
  Batch<T> b = make_batch<T>((const void* const*)in, out, nf, gate);
  if (probe) probe->begin(MM_PROBE_Y, s);
  switch (ny) {
#define X(L_) case L_: MM_VARIANTS(go_yt, L_, var[4], b, nf, twNy, nz, H, dir, mode, s); break;
    MM_NLENS(X)
#undef X
    default: return MM_ERR_ARG;
  }
  if (probe) probe->end(MM_PROBE_Y, s, (double)nf * 2.0 * ncplx() * 2 * sizeof(T));
  MM_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int Fft3::perm_sz(const void* in, void* out, int nf, int mode, hipStream_t s) const {
  if (nf < 1 || (mode != 1 && mode != 2) || ny % kTW) return MM_ERR_ARG;
  const size_t n = ncplx() * nf;
  const int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
  if (mode == 1)
    hipLaunchKernelGGL((k_perm_sz<T, 1>), dim3(grid), dim3(256), 0, s, (const C2<T>*)in, (C2<T>*)out, nf, ny, nz, H);
  else
    hipLaunchKernelGGL((k_perm_sz<T, 2>), dim3(grid), dim3(256), 0, s, (const C2<T>*)in, (C2<T>*)out, nf, ny, nz, H);
  MM_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int Fft3::z_pass(void* const* in, void* const* out, int nf, int dir, hipStream_t s) const {
  if (nf < 1 || nf > kMaxBatch) return MM_ERR_ARG;
  Batch<T> b = make_batch<T>((const void* const*)in, out, nf, gate);
  int ncols = H * ny;
  if (probe) probe->begin(MM_PROBE_Z, s);
  switch (nz) {
#define X(L_) case L_: MM_VARIANTS(go_z, L_, var[3], b, nf, twNz, ncols, dir, s); break;
    MM_NLENS(X)
#undef X
    default: return MM_ERR_ARG;
  }
  if (probe) probe->end(MM_PROBE_Z, s, (double)nf * 2.0 * ncplx() * 2 * sizeof(T));
  MM_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int Fft3::forward(const T* const* in, void* const* out, int nf, hipStream_t s, T scale) const {
  MM_TRY(x_r2c<T>(in, out, nf, s, scale));
  MM_TRY(y_pass<T>(out, out, nf, -1, s));
  MM_TRY(z_pass<T>(out, out, nf, -1, s));
  return 0;
}

template <typename T>
int Fft3::inverse(void* const* spec, T* const* out, int nf, hipStream_t s) const {
  MM_TRY(z_pass<T>(spec, spec, nf, +1, s));
  MM_TRY(y_pass<T>(spec, spec, nf, +1, s));
  MM_TRY(x_c2r<T>(spec, out, nf, s));
  return 0;
}

template int Fft3::forward<double>(const double* const*, void* const*, int, hipStream_t, double) const;
template int Fft3::forward<float>(const float* const*, void* const*, int, hipStream_t, float) const;
template int Fft3::inverse<double>(void* const*, double* const*, int, hipStream_t) const;
template int Fft3::inverse<float>(void* const*, float* const*, int, hipStream_t) const;
template int Fft3::x_r2c<double>(const double* const*, void* const*, int, hipStream_t, double) const;
template int Fft3::x_r2c<float>(const float* const*, void* const*, int, hipStream_t, float) const;
template int Fft3::x_c2r<double>(void* const*, double* const*, int, hipStream_t) const;
template int Fft3::x_c2r<float>(void* const*, float* const*, int, hipStream_t) const;
template int Fft3::y_pass<double>(void* const*, void* const*, int, int, hipStream_t) const;
template int Fft3::y_pass<float>(void* const*, void* const*, int, int, hipStream_t) const;
template int Fft3::z_pass<double>(void* const*, void* const*, int, int, hipStream_t) const;
template int Fft3::y_pass_t<double>(void* const*, void* const*, int, int, int, hipStream_t) const;
template int Fft3::y_pass_t<float>(void* const*, void* const*, int, int, int, hipStream_t) const;
template int Fft3::perm_sz<double>(const void*, void*, int, int, hipStream_t) const;
template int Fft3::perm_sz<float>(const void*, void*, int, int, hipStream_t) const;
template int Fft3::z_pass<float>(void* const*, void* const*, int, int, hipStream_t) const;

// ---------------------------------------------------------------- autotuner
// Times every plan variant of every pass on scratch fields of the plan's own
// size (4 fields per launch, the typical batch) and keeps the fastest.
template <typename T>
static int autotune_t(Fft3& f, hipStream_t s) {
  const int nf = 4;
  void* re = nullptr;
  void* sp = nullptr;
  if (hipMalloc(&re, nf * f.nreal() * sizeof(T)) != hipSuccess) return MM_ERR_NOMEM;
  if (hipMalloc(&sp, 2 * nf * f.ncplx() * 2 * sizeof(T)) != hipSuccess) { (void)hipFree(re); return MM_ERR_NOMEM; }
  MM_HIP(hipMemsetAsync(re, 0, nf * f.nreal() * sizeof(T), s));
  MM_HIP(hipMemsetAsync(sp, 0, 2 * nf * f.ncplx() * 2 * sizeof(T), s));
  const T* rin[4];
  T* rout[4];
  void* sps[4];
  void* sps2[4];
  for (int i = 0; i < nf; ++i) {
    rin[i] = (const T*)re + i * f.nreal();
    rout[i] = (T*)re + i * f.nreal();
    sps[i] = (C2<T>*)sp + i * f.ncplx();
    sps2[i] = (C2<T>*)sp + (nf + i) * f.ncplx();
  }
  hipEvent_t a, b;
  MM_HIP(hipEventCreate(&a));
  MM_HIP(hipEventCreate(&b));
  Probe* keep = f.probe;
  f.probe = nullptr;
  const int lens[5] = {f.nx / 2, f.nx / 2, f.ny, f.nz, f.ny};
  int rc = 0;
  for (int k = 0; k < 5 && !rc; ++k) {
    if (lens[k] < 64 || (k == 4 && f.ny % kTW)) { f.var[k] = 0; continue; }
    float best = 1e30f;
    int bestv = 0;
    for (int v = 0; v < 3 && !rc; ++v) {
      f.var[k] = v;
      float tmin = 1e30f;
      for (int rep = 0; rep < 3 && !rc; ++rep) {
        MM_HIP(hipEventRecord(a, s));
        if (k == 0) rc = f.x_r2c<T>(rin, sps, nf, s, (T)1);
        else if (k == 1) rc = f.x_c2r<T>(sps, rout, nf, s);
        else if (k == 2) rc = f.y_pass<T>(sps, sps, nf, -1, s);
        else if (k == 3) rc = f.z_pass<T>(sps, sps, nf, -1, s);
        else {
          // the tiled y pass of the fused path: forward S -> Z and inverse Z -> S
          rc = f.y_pass_t<T>(sps, sps2, nf, -1, 1, s);
          if (!rc) rc = f.y_pass_t<T>(sps2, sps, nf, +1, 2, s);
        }
        MM_HIP(hipEventRecord(b, s));
        MM_HIP(hipEventSynchronize(b));
        float t = 0.f;
        MM_HIP(hipEventElapsedTime(&t, a, b));
        if (rep > 0 && t < tmin) tmin = t;
      }
      f.tune_ms[k][v] = tmin;
      if (tmin < best) { best = tmin; bestv = v; }
    }
    f.var[k] = bestv;
  }
  f.probe = keep;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(re);
  (void)hipFree(sp);
  return rc;
}

// The plans chosen for a grid size and precision are kept for the process:
// every later context of the same shape on the same device reuses them, so
// contexts of one process (ensemble members and single runs alike) round the
// same way, and the timing runs once per shape.
namespace {
struct Tuned {
  int var[5];
  float ms[5][3];
};
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int>, Tuned> g_tuned;
}  // namespace

int Fft3::autotune(hipStream_t s) {
  int dev = 0;
  MM_HIP(hipGetDevice(&dev));
  const auto key = std::make_tuple(nx, ny, nz, prec, dev);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tuned.find(key);
  if (it != g_tuned.end()) {
    std::memcpy(var, it->second.var, sizeof(var));
    std::memcpy(tune_ms, it->second.ms, sizeof(tune_ms));
    return 0;
  }
  const int rc = prec == MM_FP64 ? autotune_t<double>(*this, s) : autotune_t<float>(*this, s);
  if (rc) return rc;
  Tuned t;
  std::memcpy(t.var, var, sizeof(var));
  std::memcpy(t.ms, tune_ms, sizeof(tune_ms));
  g_tuned[key] = t;
  return 0;
}

}  // namespace mm

// ---------------------------------------------------------------- C ABI
struct mm_fft {
  mm::Fft3 f;
};

extern "C" int mm_fft_create(int32_t nx, int32_t ny, int32_t nz, int32_t precision, mm_fft** out) {
  if (!out) return MM_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MM_ERR_NODEV;
  mm_fft* h = new mm_fft();
  int rc = h->f.init(nx, ny, nz, precision);
  const char* at = std::getenv("MM_AUTOTUNE");
  const char* fv = std::getenv("MM_FFT_VARIANT");  // 0, 1, 2: every pass on that plan, no timing
  if (!rc && fv && fv[0] >= '0' && fv[0] <= '2') {
    for (int k = 0; k < 5; ++k) h->f.var[k] = fv[0] - '0';
  } else if (!rc && !(at && at[0] == '0')) {
    rc = h->f.autotune(nullptr);
  }
  if (rc) {
    h->f.destroy();
    delete h;
    return rc;
  }
  *out = h;
  return 0;
}

extern "C" void mm_fft_destroy(mm_fft* h) {
  if (!h) return;
  h->f.destroy();
  delete h;
}

extern "C" size_t mm_fft_spectrum_elems(const mm_fft* h) { return h ? h->f.ncplx() : 0; }

extern "C" int mm_fft_forward(mm_fft* h, const void* real_dev, void* spec_dev, void* stream) {
  if (!h || !real_dev || !spec_dev) return MM_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const double n3 = (double)(float)((long)h->f.nx * h->f.ny * h->f.nz);
  void* out[1] = {spec_dev};
  if (h->f.prec == MM_FP64) {
    const double* in[1] = {(const double*)real_dev};
    return h->f.forward<double>(in, out, 1, s, 1.0 / n3);
  }
  const float* in[1] = {(const float*)real_dev};
  return h->f.forward<float>(in, out, 1, s, (float)(1.0 / n3));
}

extern "C" int mm_fft_inverse(mm_fft* h, void* spec_dev, void* real_dev, void* stream) {
  if (!h || !real_dev || !spec_dev) return MM_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  void* sp[1] = {spec_dev};
  if (h->f.prec == MM_FP64) {
    double* out[1] = {(double*)real_dev};
    return h->f.inverse<double>(sp, out, 1, s);
  }
  float* out[1] = {(float*)real_dev};
  return h->f.inverse<float>(sp, out, 1, s);
}

// ------------------------------------- forward transform that allocates nothing
// mm_fft_forward_noalloc runs the passes of mm_fft_forward with the default
// plans; the twiddle tables, which mm_fft_create allocates per plan, come from
// `scratch`, a file-scope device array that holds the tables of every
// supported length in both precisions (about 62 KiB).  The host never names
// `scratch` in a kernel argument or a copy: at the first call on a device,
// with that device current, hipGetSymbolAddress gives that device's address
// of it, the tables are copied there once, and the address is kept per
// device.  After that the array is only read, so no call clears or writes
// it, and every call on a device reads the same tables.
namespace {
constexpr int kNLen[13] = {8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512};
constexpr int kXLen[13] = {4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256};
constexpr size_t sum_len(const int (&a)[13]) {
  size_t s = 0;
  for (int i = 0; i < 13; ++i) s += (size_t)a[i];
  return s;
}
// complex entries over all tables; FP64 tables first, then the FP32 ones
constexpr size_t kTwEntries = sum_len(kNLen) + sum_len(kXLen);
constexpr size_t kScratchBytes = kTwEntries * 16 + kTwEntries * 8;
constexpr size_t kScratchDoubles = 8192;  // 64 KiB
static_assert(kScratchBytes <= kScratchDoubles * sizeof(double), "twiddle tables exceed scratch");
constexpr int kMaxDevices = 64;

// byte offset of the table of length n (xlen: the x-pass half length table)
size_t tw_offset(int prec, bool xlen, int n) {
  size_t off = prec == MM_FP64 ? 0 : kTwEntries * 16;
  const size_t esz = prec == MM_FP64 ? 16 : 8;
  for (int i = 0; i < 13; ++i) {
    if (!xlen && kNLen[i] == n) return off;
    off += (size_t)kNLen[i] * esz;
  }
  for (int i = 0; i < 13; ++i) {
    if (xlen && kXLen[i] == n) return off;
    off += (size_t)kXLen[i] * esz;
  }
  return (size_t)-1;
}

std::mutex g_na_mu;
void* g_na_ptr[kMaxDevices] = {};
double g_na_host[kScratchDoubles];  // host image of the tables, filled under g_na_mu
}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

// the device address of `scratch` on the current device, with the tables in it
static int noalloc_scratch(char** base) {
  int dev = 0;
  MM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return MM_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    MM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    double t[2 * 512];  // one table at a time
    char* h = (char*)g_na_host;
    for (int prec = 0; prec < 2; ++prec)
      for (int x = 0; x < 2; ++x)
        for (int i = 0; i < 13; ++i) {
          const int n = x ? kXLen[i] : kNLen[i];
          mm::make_twiddles(n, t);
          char* dst = h + tw_offset(prec, x != 0, n);
          if (prec == MM_FP64) {
            std::memcpy(dst, t, 2 * (size_t)n * sizeof(double));
          } else {
            for (int e = 0; e < 2 * n; ++e) ((float*)dst)[e] = (float)t[e];
          }
        }
    MM_HIP(hipMemcpy(p, g_na_host, kScratchBytes, hipMemcpyHostToDevice));
    g_na_ptr[dev] = p;
  }
  *base = (char*)g_na_ptr[dev];
  return 0;
}

extern "C" int mm_fft_forward_noalloc(int32_t nx, int32_t ny, int32_t nz, int32_t precision,
                                      const void* real_dev, void* spec_dev, void* stream) {
  if (!real_dev || !spec_dev) return MM_ERR_ARG;
  if (!mm::fft_length_ok(nx) || !mm::fft_length_ok(ny) || !mm::fft_length_ok(nz) || !mm::xlen_ok(nx / 2))
    return MM_ERR_ARG;
  if (precision != MM_FP64 && precision != MM_MIXED) return MM_ERR_ARG;
  char* base = nullptr;
  MM_TRY(noalloc_scratch(&base));
  mm::Fft3 f;  // a plan view over the tables in scratch; nothing to destroy
  f.nx = nx; f.ny = ny; f.nz = nz; f.H = nx / 2 + 1; f.prec = precision;
  f.twLx = base + tw_offset(precision, true, nx / 2);
  f.twNx = base + tw_offset(precision, false, nx);
  f.twNy = base + tw_offset(precision, false, ny);
  f.twNz = base + tw_offset(precision, false, nz);
  hipStream_t s = (hipStream_t)stream;
  const double n3 = (double)(float)((long)nx * ny * nz);
  void* out[1] = {spec_dev};
  if (precision == MM_FP64) {
    const double* in[1] = {(const double*)real_dev};
    return f.forward<double>(in, out, 1, s, 1.0 / n3);
  }
  const float* in[1] = {(const float*)real_dev};
  return f.forward<float>(in, out, 1, s, (float)(1.0 / n3));
}
