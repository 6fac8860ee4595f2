// zfused.hip -- launchers of the fused z-pass kernels (zfused.h), dispatched
// on the z length.
#include "fft3d.h"
#include "mm_common.h"
#include "zfused.h"

namespace mm {

template <typename T, int L, class Op>
static int go_zf(const Fft3& f, const Op& op, int ngroups, hipStream_t s, int* nblk) {
  using TL = TileZF<T, typename ZPlan<L, Op::PLAN>::type, Op::MAXT>;
  const int ncols = f.H * f.ny;
  dim3 grid((ncols + TL::B - 1) / TL::B, ngroups);
  hipLaunchKernelGGL((k_z_fused<T, L, Op>), grid, dim3(TL::THREADS), 0, s, op, op.P, (const C2<T>*)f.twNz, ncols);
  if (nblk) *nblk = (int)grid.x;
  return 0;
}

template <typename T, class Op>
int launch_zfused(const Fft3& f, const Op& op, int ngroups, hipStream_t s, int probe_kind, double bytes,
                  int* nblk) {
  if (f.probe) f.probe->begin(probe_kind, s);
  switch (f.nz) {
#define X(L_) case L_: go_zf<T, L_, Op>(f, op, ngroups, s, nblk); break;
    X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)
#undef X
    default: return MM_ERR_ARG;
  }
  if (f.probe) f.probe->end(probe_kind, s, bytes);
  MM_HIP(hipGetLastError());
  return 0;
}

// variant v of the elastic op: (plan kind, workgroup cap, min waves per SIMD)
template <typename T>
int launch_zf_elastic(const Fft3& f, const OpElastic<T>& op, int v, hipStream_t s, int probe_kind, double bytes,
                      int* nblk) {
  // 192 and 384 have no 6-point plan within five stages; with 12 points per
  // thread the two-waves-per-SIMD shapes spill hundreds of VGPRs, the
  // 512-thread shape at one wave per SIMD the fewest
  if (f.nz == 192 || f.nz == 384) v = 1;
  switch (v) {
    case 1: return launch_zfused<T>(f, OpElasticV<T, 1, 512, 1>(op), 1, s, probe_kind, bytes, nblk);
    case 2: return launch_zfused<T>(f, OpElasticV<T, 1, 256, 2, true>(op), 1, s, probe_kind, bytes, nblk);
    case 3: return launch_zfused<T>(f, OpElasticV<T, 0, 256, 1, true>(op), 1, s, probe_kind, bytes, nblk);
    default: return launch_zfused<T>(f, op, 1, s, probe_kind, bytes, nblk);
  }
}
template int launch_zf_elastic<double>(const Fft3&, const OpElastic<double>&, int, hipStream_t, int, double, int*);
template int launch_zf_elastic<float>(const Fft3&, const OpElastic<float>&, int, hipStream_t, int, double, int*);

// the y pass of the Cahn-Hilliard update with the x and y flux terms folded in
template <typename T>
int launch_y_ch(const Fft3& f, const DevParams* P, const void* in, void* out, int ne, int e0, hipStream_t s) {
  const int H = f.H, nz = f.nz;
  if (f.probe) f.probe->begin(MM_PROBE_Y, s);
  switch (f.ny) {
#define X(L_)                                                                                             \
  case L_: {                                                                                              \
    using TL = TileLM<T, PlanLean<L_>>;                                                                   \
    dim3 grid((nz * H + TL::B - 1) / TL::B, 1, ne);                                                       \
    hipLaunchKernelGGL((k_y_pass_ch<T, PlanLean<L_>>), grid, dim3(TL::THREADS), 0, s, P, (const C2<T>*)in, \
                       (C2<T>*)out, f.ncplx(), (const C2<T>*)f.twNy, nz, H, e0, f.gate);                  \
    break;                                                                                                \
  }
    X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)
#undef X
    default: return MM_ERR_ARG;
  }
  if (f.probe) f.probe->end(MM_PROBE_Y, s, (double)ne * 6.0 * f.ncplx() * 2 * sizeof(T));
  MM_HIP(hipGetLastError());
  return 0;
}
template int launch_y_ch<double>(const Fft3&, const DevParams*, const void*, void*, int, int, hipStream_t);
template int launch_y_ch<float>(const Fft3&, const DevParams*, const void*, void*, int, int, hipStream_t);

// the inverse y pass of the gradient block after OpGrad3 (nv variants)
template <typename T>
int launch_y_grad(const Fft3& f, const DevParams* P, const void* in, void* out, int nv, hipStream_t s) {
  const int H = f.H, nz = f.nz;
  if (f.probe) f.probe->begin(MM_PROBE_Y, s);
  switch (f.ny) {
#define X(L_)                                                                                             \
  case L_: {                                                                                              \
    using TL = TileLM<T, PlanLean<L_>>;                                                                   \
    dim3 grid((nz * H + TL::B - 1) / TL::B, 1, nv);                                                       \
    hipLaunchKernelGGL((k_y_pass_grad<T, PlanLean<L_>>), grid, dim3(TL::THREADS), 0, s, P,                \
                       (const C2<T>*)in, (C2<T>*)out, f.ncplx(), (const C2<T>*)f.twNy, nz, H, f.gate);    \
    break;                                                                                                \
  }
    X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)
#undef X
    default: return MM_ERR_ARG;
  }
  if (f.probe) f.probe->end(MM_PROBE_Y, s, (double)nv * 7.0 * f.ncplx() * 2 * sizeof(T));
  MM_HIP(hipGetLastError());
  return 0;
}
template int launch_y_grad<double>(const Fft3&, const DevParams*, const void*, void*, int, hipStream_t);
template int launch_y_grad<float>(const Fft3&, const DevParams*, const void*, void*, int, hipStream_t);

#define INST(T_, OP_) \
  template int launch_zfused<T_, OP_<T_>>(const Fft3&, const OP_<T_>&, int, hipStream_t, int, double, int*);
INST(double, OpGrad)
INST(float, OpGrad)
INST(double, OpElastic)
INST(float, OpElastic)
INST(double, OpCH)
INST(float, OpCH)
INST(double, OpCH2)
INST(float, OpCH2)
INST(double, OpGrad3)
INST(float, OpGrad3)
INST(double, OpAC)
INST(float, OpAC)
#undef INST

}  // namespace mm
