// fft3d.h -- host-side 3D real FFT plan over the pass kernels of fft_pass.h.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

namespace mm {

// Supported line lengths: y/z axes take N, the x pass takes Nx/2.
bool fft_length_ok(int n);  // 8..512, 2^a 3^b (a >= 2 for x/2 to stay even-free of 2-length)

// Live timing of one kernel class with HIP events recorded on the stream the
// kernel is launched on.  Classes: the MM_PROBE_* ids of meumapps_ss.h.
struct Probe {
  int kind = -1;
  std::vector<hipEvent_t> ev;  // begin/end pairs
  int used = 0;
  double bytes = 0.0, ms = 0.0;
  long long launches = 0;
  void begin(int k, hipStream_t s);
  void end(int k, hipStream_t s, double b);
  int collect();  // after the stream is idle: accumulate elapsed time
  void reset();
  void destroy();
};

struct Fft3 {
  int nx = 0, ny = 0, nz = 0, H = 0, prec = 0;
  Probe* probe = nullptr;
  // plan variant per pass (x_r2c, x_c2r, y, z, tiled y): 0 default, 1 lean,
  // 2 lean-4; chosen by autotune() from the kernel times in tune_ms
  int var[5] = {0, 0, 0, 0, 0};
  float tune_ms[5][3] = {};
  // when non-null every pass launched through this plan is gated on *gate
  // (set while recording the graph-mode step)
  const int* gate = nullptr;
  int autotune(hipStream_t s);
  void* twLx = nullptr;  // exp(-2 pi i e/(Nx/2)), e < Nx/2
  void* twNx = nullptr;  // exp(-2 pi i e/Nx),     e < Nx
  void* twNy = nullptr;
  void* twNz = nullptr;
  size_t nreal() const { return (size_t)nx * ny * nz; }
  size_t ncplx() const { return (size_t)H * ny * nz; }
  int init(int nx, int ny, int nz, int prec);  // 0 or MM_ERR_*
  void destroy();
  // real fields -> spectra, scaled by `scale` (f_trans uses 1/(Nx Ny Nz))
  template <typename T>
  int forward(const T* const* in, void* const* out, int nf, hipStream_t s, T scale) const;
  // spectra -> real fields; the spectra are overwritten (used as scratch)
  template <typename T>
  int inverse(void* const* spec, T* const* out, int nf, hipStream_t s) const;
  // the three passes separately, for fused pipelines
  template <typename T>
  int x_r2c(const T* const* in, void* const* out, int nf, hipStream_t s, T scale) const;
  template <typename T>
  int x_c2r(void* const* in, T* const* out, int nf, hipStream_t s) const;
  template <typename T>
  int y_pass(void* const* in, void* const* out, int nf, int dir, hipStream_t s) const;
  // y pass with the layout change of the fused path: mode 1 S -> Z, mode 2 Z -> S (out of place)
  template <typename T>
  int y_pass_t(void* const* in, void* const* out, int nf, int dir, int mode, hipStream_t s) const;
  // re-order nf spectra between S and Z (mode 1 S -> Z, mode 2 Z -> S), out of place
  template <typename T>
  int perm_sz(const void* in, void* out, int nf, int mode, hipStream_t s) const;
  template <typename T>
  int z_pass(void* const* in, void* const* out, int nf, int dir, hipStream_t s) const;
};

// forward twiddle table exp(-2 pi i e / n), e in [0, n), interleaved doubles,
// with the symmetries w[n-e] = conj w[e] and exact values at multiples of pi/2
void make_twiddles(int n, double* out);

}  // namespace mm
