// Amplitude scaling of a slot's (or an integration's) results, Visibility::writedata
// (mpifxcorr/src/visibility.cpp:353-646), on the device: one workgroup per segment computes
// the segment's scale from the weights stored in the results, in the reference's float and
// double order, and multiplies the segment's channels.  A first kernel forms the band means
// of the parallel autocorrelations that the Tsys calibration divides by.
#include "fxg_internal.h"

namespace {

// autocorrcalibs[ds][band].re: the mean over the band's channels (vectorMean_cf32,
// visibility.cpp:466-476), real part
__global__ void __launch_bounds__(256) k_norm_cal(NormArgs a) {
  const NormCal c = a.cal[blockIdx.x];
  __shared__ float part[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < c.nout; i += 256) s += a.res[2LL * (c.off + i)];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.calmean[blockIdx.x] = part[0] / (float)c.nout;
}

__device__ float seg_scale(const NormArgs& a, const NormSeg& g) {
  if (g.kind == 2) {                                             // phase-cal tones, visibility.cpp:650-675
    float acw = g.wf >= 0 ? (float)(a.res[g.wf] / a.fftsperint) : 0.0f;
    if (acw == 0.0f && g.pcz) acw = 1.0f;
    if (!(acw > 0.0f)) return 0.0f;                              // left unscaled
    return (float)(1.0 / ((double)acw * a.msi * (double)g.chanfactor));
  }
  const float w = (float)(a.res[g.wf] / a.fftsperint);          // baselineweights / autocorrweights
  if (g.kind == 0) {
    const double t1 = a.tsys[g.ds1], t2 = a.tsys[g.ds2];
    if (t1 > 0.0 && t2 > 0.0) {                                  // visibility.cpp:510-515
      const float aw1 = g.aw1 >= 0 ? (float)(a.res[g.aw1] / a.fftsperint) : 0.0f;
      const float aw2 = g.aw2 >= 0 ? (float)(a.res[g.aw2] / a.fftsperint) : 0.0f;
      const float divisor = a.dpct[g.ds1] * a.dpct[g.ds2] * a.calmean[g.cal1] * a.calmean[g.cal2] / (aw1 * aw2);
      if (divisor > 0.0f && w > 0) return (float)(sqrt(t1 * t2 / (double)divisor) / (double)w);
      return 0.0f;
    }
    if (!(w > 0.0f)) return 0.0f;                                // visibility.cpp:520-541
    float sc = (float)(1.0 / ((double)w * a.msi * (double)g.chanfactor));
    sc *= g.zs1;
    sc *= g.zs2;
    return sc;
  }
  const double t = a.tsys[g.ds1];
  if (t > 0.0) {                                                 // visibility.cpp:597-606
    const float divisor = sqrtf(a.calmean[g.cal1] * a.calmean[g.cal2]);
    return divisor > 0.0f ? (float)(t / (double)divisor) : 0.0f;
  }
  if (!(w > 0.0f)) return 0.0f;                                  // visibility.cpp:607-622
  float sc = (float)(1.0 / ((double)w * a.msi * (double)g.chanfactor));
  sc *= g.zs1;
  return sc;
}

__global__ void __launch_bounds__(256) k_norm_apply(NormArgs a) {
  const NormSeg g = a.seg[blockIdx.x];
  __shared__ float s_scale;
  if (threadIdx.x == 0) {
    float sc = seg_scale(a, g);
    if (g.decf >= 0) {          // uv-shifted phase centre: divided by its shift decorrelation
      const float dec = (float)((double)a.res[g.decf] / (1000000000.0 * a.inttime_s));
      sc = sc / dec;
    }
    s_scale = sc;
  }
  __syncthreads();
  const float sc = s_scale;
  if (!(sc > 0.0f)) return;                                      // only a positive scale is applied
  float2* v = reinterpret_cast<float2*>(a.res) + g.off;
  for (int i = threadIdx.x; i < g.nout; i += 256) {
    float2 x = v[i];
    x.x *= sc;
    x.y *= sc;
    v[i] = x;
  }
}

}  // namespace

int fxg_launch_normalise(const NormArgs& a, int nseg, int ncal, hipStream_t st) {
  if (ncal > 0) hipLaunchKernelGGL(k_norm_cal, dim3(ncal), dim3(256), 0, st, a);
  if (nseg > 0) hipLaunchKernelGGL(k_norm_apply, dim3(nseg), dim3(256), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
