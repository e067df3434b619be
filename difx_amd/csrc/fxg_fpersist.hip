// The F-engine of every FFT length of a configuration in one persistent launch (see
// fe_range in fxg_fe_common.h for what one FFT computes).
#include "fxg_fe_common.h"

namespace {

// The persistent launch calls each length's body out of line, so that every body gets its own
// register allocation (inlined side by side they spill).
template <int L, bool CPLX, int ORDER, int NTB>
__device__ __attribute__((noinline)) void fe_range_call(const FengineArgs& args, const float2* twiddle, int d,
                                                        int u0, int u1, float2* smem, bool load_tw) {
  fe_range<L, CPLX, ORDER, NTB>(args, twiddle, d, u0, u1, smem, load_tw);
}

template <int L, bool CPLX, int ORDER, int NTB>
__device__ __attribute__((noinline)) void fb_range_call(const FengineArgs& args, const float2* twL, int d,
                                                        int lt0, int lt1, float2* smem, float2* scr, bool load_tw) {
  fb_range<L, CPLX, ORDER, NTB>(args, args.tw2048, twL, d, lt0, lt1, 1, smem, scr, load_tw);
}

}  // namespace

// Every FFT length of a configuration in one persistent launch: workgroups take work items
// (FFT length, datastream, part of the chunk) from a device counter until none are left; the
// twiddles are reloaded when the length changes.  Items are ordered longest FFT first.  Items
// of length 16384 and 32768 run the four-step body with this workgroup's scratch rows.
template <bool CPLX, int ORDER, int NTB>
__global__ void __launch_bounds__(NTB, 1) k_fpersist(FengineArgs args, const FItem* __restrict__ items,
                                                       int nitems, int* counter) {
  extern __shared__ float2 fx_smem[];
  __shared__ int s_item;
  int curL = 0;
  const int nunits = args.nffts_pad / 8;
  while (true) {
    __syncthreads();                              // s_item consumed by every thread
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = s_item;
    if (it >= nitems) break;                      // every wave of the workgroup leaves here
    const FItem w = items[it];
    const int u0 = (int)((long long)w.part * nunits / w.nparts);
    const int u1 = (int)((long long)(w.part + 1) * nunits / w.nparts);
    const bool load = w.L != curL;
    curL = w.L;
    const float2* tw = args.tw_by_log[w.log2L - 8];
    switch (w.L) {
      case 256: fe_range_call<256, CPLX, ORDER, NTB>(args, tw, w.d, u0, u1, fx_smem, load); break;
      case 512: fe_range_call<512, CPLX, ORDER, NTB>(args, tw, w.d, u0, u1, fx_smem, load); break;
      case 1024: fe_range_call<1024, CPLX, ORDER, NTB>(args, tw, w.d, u0, u1, fx_smem, load); break;
      case 2048: fe_range_call<2048, CPLX, ORDER, NTB>(args, tw, w.d, u0, u1, fx_smem, load); break;
      case 4096: fe_range_call<4096, CPLX, ORDER, NTB>(args, tw, w.d, u0, u1, fx_smem, load); break;
      case 8192: fe_range_call<8192, CPLX, ORDER, NTB>(args, tw, w.d, u0, u1, fx_smem, load); break;
      case 16384:
      case 32768: {
        float2* scr = args.bigscratch + (long long)blockIdx.x * (NTB / 256) * args.bigL;
        if (w.L == 16384) fb_range_call<16384, CPLX, ORDER, NTB>(args, tw, w.d, u0 * 8, u1 * 8, fx_smem, scr, load);
        else fb_range_call<32768, CPLX, ORDER, NTB>(args, tw, w.d, u0 * 8, u1 * 8, fx_smem, scr, load);
        break;
      }
      default: break;
    }
  }
}

template <bool CPLX, int ORDER, int NTB>
static int launch_p(const FengineArgs& a, const FItem* items, int nitems, int* counter, int nwg, size_t lds,
                    hipStream_t st) {
  static size_t raised = 0;
  if (raise_lds(k_fpersist<CPLX, ORDER, NTB>, lds, &raised)) return -1;
  hipLaunchKernelGGL((k_fpersist<CPLX, ORDER, NTB>), dim3(nwg), dim3(NTB), lds, st, a, items, nitems, counter);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fxg_launch_fpersist(const FengineArgs& a, const FItem* items, int nitems, int* counter, int nwg, int ntb,
                        int iscomplex, size_t lds, hipStream_t st) {
  if (fxg_launch_zero(counter, sizeof(int), st)) return -1;
#define FXG_P(C_, O_, N_) return launch_p<C_, O_, N_>(a, items, nitems, counter, nwg, lds, st)
  if (ntb == 256) {
    if (iscomplex) { if (a.rotorder == 1) FXG_P(true, 1, 256); if (a.rotorder == 2) FXG_P(true, 2, 256); return -2; }
    if (a.rotorder == 0) FXG_P(false, 0, 256);
    if (a.rotorder == 1) FXG_P(false, 1, 256);
    if (a.rotorder == 2) FXG_P(false, 2, 256);
  } else if (ntb == 512) {
    if (iscomplex) { if (a.rotorder == 1) FXG_P(true, 1, 512); if (a.rotorder == 2) FXG_P(true, 2, 512); return -2; }
    if (a.rotorder == 0) FXG_P(false, 0, 512);
    if (a.rotorder == 1) FXG_P(false, 1, 512);
    if (a.rotorder == 2) FXG_P(false, 2, 512);
  }
#undef FXG_P
  return -2;
}
