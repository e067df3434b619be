// Spectral kurtosis sums on the device: Mode::process accumulates, for every recorded band and
// channel, s1 += |X|^2 and s2 += |X|^4 over the FFTs of its block range (mpifxcorr/src/
// mode.cpp:1237-1253, magnitude then two squares).  Here one lane per (band, channel) walks the
// FFTs of the spectra chunk in order and adds its partial sums to the band's accumulators; the
// chunks of a subint are launched in order on one stream, so no two lanes share a sum.
#include "fxg_internal.h"
#include <hip/hip_fp16.h>

namespace {

__device__ __forceinline__ float2 spec_value(const KurtArgs& a, const KurtBand& b, int c, int lt) {
  if (a.mode == FXG_X_FP32) return a.spec32[b.off + (long long)lt * b.C + c];
  const uint32_t u = a.spec16[b.off + ((long long)(c >> 3) * a.chunkcap + lt) * 8 + (c & 7)];
  float re, im;
  if (a.mode == FXG_X_FP16) {
    const __half2 h = *reinterpret_cast<const __half2*>(&u);
    const float2 f = __half22float2(h);
    re = f.x;
    im = f.y;
  } else {
    re = __uint_as_float((u & 0xFFFFu) << 16);
    im = __uint_as_float(u & 0xFFFF0000u);
  }
  const float sc = a.xscale[b.scale_idx];
  return make_float2(re / sc, im / sc);
}

__global__ void __launch_bounds__(256) k_kurt(KurtArgs a) {
  const KurtBand b = a.tab[blockIdx.x];
  const int c = blockIdx.y * 256 + threadIdx.x;
  if (c >= b.C) return;
  float s1 = 0.f, s2 = 0.f;
  for (int lt = 0; lt < a.nffts; ++lt) {
    const float2 x = spec_value(a, b, c, lt);
    const float m = sqrtf(x.x * x.x + x.y * x.y);     // vectorMagnitude_cf32
    const float p = m * m;                             // first square
    s1 += p;
    s2 += p * p;                                       // second square
  }
  float2* acc = a.acc + b.acc_off + c;
  float2 v = *acc;
  v.x += s1;
  v.y += s2;
  *acc = v;
}

}  // namespace

int fxg_launch_kurt(const KurtArgs& a, int nband, int maxC, hipStream_t st) {
  if (nband <= 0) return 0;
  hipLaunchKernelGGL(k_kurt, dim3(nband, (maxC + 255) / 256), dim3(256), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
