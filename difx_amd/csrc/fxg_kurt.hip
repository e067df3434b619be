// Spectral kurtosis sums on the device: Mode::process accumulates, for every recorded band and
// channel, s1 += |X|^2 and s2 += |X|^4 over the FFTs of its block range (mpifxcorr/src/
// mode.cpp:1237-1253, magnitude then two squares).  Here the sums of each spectra chunk are
// added to the band's accumulators by one lane per channel; the chunks of a subint are launched
// in order on one stream, so no two lanes share a sum.
#include "fxg_internal.h"
#include <hip/hip_fp16.h>

namespace {

// the stored value of channel c of FFT lt: FP16/BF16 spectra carry the row's scale, which the
// sums take out once at the end (|X/sc|^2 = |X|^2 / sc^2)
__device__ __forceinline__ float2 spec_raw(const KurtArgs& a, const KurtBand& b, int c, int lt) {
  if (a.mode == FXG_X_FP32) return a.spec32[b.off + (long long)lt * b.C + c];
  const uint32_t u = a.spec16[b.off + ((((long long)(c >> b.twlog) * a.chunkcap + lt)) << b.twlog) +
                              (c & ((1 << b.twlog) - 1))];
  if (a.mode == FXG_X_FP16) {
    const __half2 h = *reinterpret_cast<const __half2*>(&u);
    return __half22float2(h);
  }
  return make_float2(__uint_as_float((u & 0xFFFFu) << 16), __uint_as_float(u & 0xFFFF0000u));
}

// One workgroup per (band, tile of channels): its lanes cover FFTS_PER_PASS FFTs x CH channels
// of the spectra layout (FP16/BF16: the 8 or 16 channels of a tile, so a pass reads 32 FFTs x
// 32 bytes or 16 FFTs x 64 bytes of contiguous runs; FP32: 64 channels of 8 bytes, 4 FFTs per
// pass), each lane summing every FFTS_PER_PASS-th FFT, and the partial sums are added over the
// FFTs in LDS.
template <int CH>
__global__ void __launch_bounds__(256) k_kurt(KurtArgs a) {
  constexpr int FPP = 256 / CH;                          // FFTs per pass
  const KurtBand b = a.tab[blockIdx.x];
  const int cl = threadIdx.x % CH, fl = threadIdx.x / CH;
  const int c = blockIdx.y * CH + cl;
  __shared__ float2 part[256];
  float s1 = 0.f, s2 = 0.f;
  if (c < b.C) {
#pragma unroll 8
    for (int lt = fl; lt < a.nffts; lt += FPP) {
      const float2 x = spec_raw(a, b, c, lt);
      const float p = x.x * x.x + x.y * x.y;           // |X|^2 (the reference squares |X|)
      s1 += p;
      s2 += p * p;                                     // |X|^4
    }
  }
  part[threadIdx.x] = make_float2(s1, s2);
  __syncthreads();
  if (threadIdx.x < CH && c < b.C) {
    float2 t = part[threadIdx.x];
    for (int f = 1; f < FPP; ++f) {
      const float2 q = part[f * CH + threadIdx.x];
      t.x += q.x;
      t.y += q.y;
    }
    if (a.mode != FXG_X_FP32) {
      const float sc = a.xscale[b.scale_idx];
      const float i2 = 1.0f / (sc * sc);
      t.x *= i2;
      t.y *= i2 * i2;
    }
    float2* acc = a.acc + b.acc_off + c;
    float2 v = *acc;
    v.x += t.x;
    v.y += t.y;
    *acc = v;
  }
}

// FP16/BF16 spectra whose bands are all in 16-channel tiles: one workgroup per (band, tile),
// each lane reading 4 channels of an FFT with one 16-byte load (4 lanes per FFT, 64 FFTs per
// pass), so a wave streams 16 FFTs x 64 bytes of one contiguous run per load instruction.
__global__ void __launch_bounds__(256) k_kurt16x4(KurtArgs a) {
  const KurtBand b = a.tab[blockIdx.x];
  const int q = threadIdx.x & 3, fl = threadIdx.x >> 2;
  const int c0 = blockIdx.y * 16 + 4 * q;
  __shared__ float2 part[64][17];                        // [FFT lane][channel], padded
  float s1[4] = {0.f, 0.f, 0.f, 0.f}, s2[4] = {0.f, 0.f, 0.f, 0.f};
  if (c0 < b.C) {                                        // C is a multiple of 16
    const uint4* __restrict__ row = reinterpret_cast<const uint4*>(
        a.spec16 + b.off + (((long long)blockIdx.y * a.chunkcap) << 4) + 4 * q);
#pragma unroll 4
    for (int lt = fl; lt < a.nffts; lt += 64) {
      const uint4 u = row[(long long)lt * 4];
      const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        float2 x;
        if (a.mode == FXG_X_FP16) {
          x = __half22float2(*reinterpret_cast<const __half2*>(&w[j]));
        } else {
          x = make_float2(__uint_as_float((w[j] & 0xFFFFu) << 16), __uint_as_float(w[j] & 0xFFFF0000u));
        }
        const float p = x.x * x.x + x.y * x.y;           // |X|^2 (the reference squares |X|)
        s1[j] += p;
        s2[j] += p * p;                                  // |X|^4
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) part[fl][4 * q + j] = make_float2(s1[j], s2[j]);
  __syncthreads();
  const int c = blockIdx.y * 16 + threadIdx.x;
  if (threadIdx.x < 16 && c < b.C) {
    float2 t = part[0][threadIdx.x];
    for (int f = 1; f < 64; ++f) {
      const float2 v = part[f][threadIdx.x];
      t.x += v.x;
      t.y += v.y;
    }
    const float sc = a.xscale[b.scale_idx];
    const float i2 = 1.0f / (sc * sc);
    float2* acc = a.acc + b.acc_off + c;
    float2 v = *acc;
    v.x += t.x * i2;
    v.y += t.y * (i2 * i2);
    *acc = v;
  }
}

}  // namespace

int fxg_launch_kurt(const KurtArgs& a, int nband, int maxC, int mintwlog, hipStream_t st) {
  if (nband <= 0) return 0;
  if (a.mode == FXG_X_FP32)
    hipLaunchKernelGGL(k_kurt<64>, dim3(nband, (maxC + 63) / 64), dim3(256), 0, st, a);
  else if (mintwlog >= 4)                                // every band in 16-channel tiles
    hipLaunchKernelGGL(k_kurt16x4, dim3(nband, (maxC + 15) / 16), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(k_kurt<8>, dim3(nband, (maxC + 7) / 8), dim3(256), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
