// Phase-cal extraction on the device.  Mode::process hands each recorded band's unpacked real
// samples of every valid FFT to its PCal extractor, numbered from the scan start
// (mpifxcorr/src/mode.cpp:758-772); every extractor PCal::getNew picks (pcal.cpp:191-274)
// computes, for tone k at baseband frequency offset + k interval, sum x[m] exp(-2 pi i f m / fs)
// over those samples.  Here the samples are folded by their sample number modulo the tone
// period M (M = fs / gcd(fs, offset, interval)), in LDS per workgroup and then into the band's
// bins; the tones follow from the bins by an M-point transform when they are read.
#include "fxg_fe_common.h"

namespace {

// grid (band entries, groups of FFTS_PER_WG FFTs of the chunk)
constexpr int FFTS_PER_WG = 16;

__global__ void __launch_bounds__(256) k_pcal(PcalArgs a) {
  extern __shared__ float bins[];
  const PcalBand pb = a.tab[blockIdx.x];
  const DevDs& ds = a.ds[pb.ds];
  const DevDsRun& run = a.run[pb.ds];
  const int M = pb.M;
  for (int j = threadIdx.x; j < M; j += 256) bins[j] = 0.f;
  __syncthreads();
  // sample number of the buffer start since the scan start (mode.cpp:760)
  const long long s0 = (long long)(((double)run.datasec * 1e9 + (double)run.datans) / (ds.sampletime_us * 1e3) + 0.5);
  const uint8_t* fv = a.fvalid + ds.fvalid_off;
  const int lt0 = blockIdx.y * FFTS_PER_WG;
  const int lt1 = min(lt0 + FFTS_PER_WG, a.nffts);
  for (int lt = lt0; lt < lt1; ++lt) {
    const FftDelay fd = a.params[(long long)pb.ds * a.chunkcap + lt];
    if (!fd.ok || !run.valid) continue;               // uniform over the workgroup
    const int base = (int)((s0 + fd.n0) % M);       // bin of the FFT's first sample
    for (int n = threadIdx.x; n < ds.L; n += 256) {
      const int g = fd.n0 + n;                        // sample of the buffer (n0 >= 0)
      const int fr = g / ds.spf;
      const int within = g - fr * ds.spf;
      float x = 0.f;
      if (fr < run.nframes && fv[fr]) {
        const long long fbase = (long long)fr * ds.framebytes + ds.headerbytes;
        if (ds.nbits == 2) {
          const long long q = (long long)within * ds.nrecbands + pb.band;
          const uint8_t byte = run.data[fbase + (q >> 2)];
          x = lut2(ds.format, (byte >> (2 * (q & 3))) & 3u);
        } else {
          x = lut8(run.data[fbase + (long long)within * ds.nrecbands + pb.band]);
        }
      }
      if (x != 0.f) atomicAdd(&bins[(base + n) % M], x);
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < M; j += 256) {
    const float v = bins[j];
    if (v != 0.f) atomicAdd(&a.bins[pb.bins_off + j], v);
  }
}

}  // namespace

int fxg_launch_pcal(const PcalArgs& a, int nband, int maxM, hipStream_t st) {
  if (nband <= 0 || a.nffts <= 0) return 0;
  const size_t lds = sizeof(float) * (size_t)maxM;
  static size_t raised = 0;
  if (raise_lds(k_pcal, lds, &raised)) return -1;
  hipLaunchKernelGGL(k_pcal, dim3(nband, (a.nffts + FFTS_PER_WG - 1) / FFTS_PER_WG), dim3(256), lds, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
