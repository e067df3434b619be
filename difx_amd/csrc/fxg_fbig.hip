// F-engine for FFT lengths beyond one workgroup's LDS: L = R * 2048 with R = 8 or 16
// (nchan 8192 / 16384 real, 16384 / 32768 complex).  What one FFT computes is what the fused
// F-engine computes (Mode::process, mode.cpp:634-1374: decode, rotator, L-point FFT, channel
// selection, fractional-sample correction); only the FFT is organised differently, as a
// four-step transform with n = R*m + r and k = k1 + 2048*k2:
//   1. for each r: the 2048-point FFT over m of the rotated samples x[R*m + r] (the LDS
//      Stockham FFT of the fused engine, 256 threads x 8 points), giving Y_r[k1];
//   2. Z_r[k1] = Y_r[k1] * exp(-2 pi i r k1 / L), staged in a per-workgroup scratch row
//      (L * 8 bytes, resident in L2 / MALL);
//   3. for each k1: X[k1 + 2048*k2] = sum_r Z_r[k1] exp(-2 pi i r k2 / R), an R-point DFT in
//      registers.
// Each thread ends with channels k1 = t + 256*s, consecutive across a wavefront, so the stores
// stay as coalesced as the fused engine's.

#include "fxg_fe_common.h"

// One FFT length per launch: workgroup (x, y) processes FFTs x, x + gridDim.x, ... of datastream
// dslist[y]; its scratch row is scratch[(y * gridDim.x + x) * L].
// Waves per SIMD the register allocation is held to: 3 (at most 168 VGPRs, a few spilled)
// measured faster than 2 (up to 256, no spills) on 16384- and 32768-point FFTs.
#ifndef FXG_FB_MINWAVES
#define FXG_FB_MINWAVES 3
#endif

template <int L, bool CPLX, int ORDER>
__global__ void __launch_bounds__(BNT, FXG_FB_MINWAVES) k_fbig(FengineArgs args, float2* __restrict__ scratch,
                                                  const float2* __restrict__ twL) {
  extern __shared__ float2 fx_smem[];
  const int d = args.dslist[blockIdx.y];
  float2* scr = scratch + ((long long)blockIdx.y * gridDim.x + blockIdx.x) * L;
  fb_range<L, CPLX, ORDER, BNT>(args, args.twiddle, twL, d, blockIdx.x, args.nffts_pad, gridDim.x, fx_smem, scr, true);
}

namespace {

template <int L, bool CPLX, int ORDER>
int launch_big(const FengineArgs& a, dim3 grid, float2* scratch, const float2* twL, hipStream_t st) {
  static size_t raised = 0;
  const size_t lds = fb_lds_bytes<BNT>(a.phcap);
  if (raise_lds(k_fbig<L, CPLX, ORDER>, lds, &raised)) return -1;
  hipLaunchKernelGGL((k_fbig<L, CPLX, ORDER>), grid, dim3(BNT), lds, st, a, scratch, twL);
  return 0;
}

template <int L, bool CPLX>
int launch_big_o(const FengineArgs& a, dim3 grid, float2* scratch, const float2* twL, hipStream_t st) {
  switch (a.rotorder) {
    case 1: return launch_big<L, CPLX, 1>(a, grid, scratch, twL, st);
    case 2: return launch_big<L, CPLX, 2>(a, grid, scratch, twL, st);
    case 0: return CPLX ? -2 : launch_big<L, false, 0>(a, grid, scratch, twL, st);
  }
  return -2;
}

}  // namespace

// FFT lengths 16384 and 32768: the per-FFT parameters, then the four-step F-engine with
// `wgx` workgroups per datastream (the scratch holds wgx * nds rows of L cf32).
int fxg_launch_fbig(const FengineArgs& a, int L, int iscomplex, int nds, int wgx, float2* scratch,
                    const float2* twL, hipStream_t st) {
  if (fxg_launch_fftparams(a, nds, st)) return -1;
  const dim3 grid(wgx, nds);
  int rc = -2;
  if (L == 16384) rc = iscomplex ? launch_big_o<16384, true>(a, grid, scratch, twL, st)
                                 : launch_big_o<16384, false>(a, grid, scratch, twL, st);
  else if (L == 32768) rc = iscomplex ? launch_big_o<32768, true>(a, grid, scratch, twL, st)
                                      : launch_big_o<32768, false>(a, grid, scratch, twL, st);
  if (rc) return rc;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

size_t fxg_fbig_lds(int ntb, int phcap) {
  return ntb == 512 ? fb_lds_bytes<512>(phcap) : fb_lds_bytes<256>(phcap);
}
