// The F-engine of every FFT length of a configuration in one persistent launch (see
// fe_range in fxg_fe_common.h for what one FFT computes).
#include "fxg_fe_common.h"

namespace {

// The work items of one FFT length: workgroups take items (datastream, part of the chunk) from
// the length's own counter until none are left, then move on to the next length.  Each length
// has its own loop, so the compiler keeps each body's loop-invariant values to that loop (one
// loop around every body holds them all at once and spills).  The twiddles are loaded with
// the length's first item.
template <int L, bool CPLX, int ORDER, int NTB>
__device__ __forceinline__ void p_class(const FengineArgs& args, const FXG_CONST FItem* it0, int n, int* ctr,
                                        int nunits, const float2* tw, float2* smem, int& s_item) {
  bool load = true;
  while (true) {
    __syncthreads();                              // s_item consumed by every thread
    if (threadIdx.x == 0) s_item = atomicAdd(ctr, 1);
    __syncthreads();
    const int k = s_item;
    if (k >= n) break;                            // every wave of the workgroup leaves here
    const int wd = it0[k].d, part = it0[k].part, nparts = it0[k].nparts;
    const int u0 = (int)((long long)part * nunits / nparts);
    const int u1 = (int)((long long)(part + 1) * nunits / nparts);
    if constexpr (L <= FXG_MAX_L) {
      fe_range<L, CPLX, ORDER, NTB>(args, tw, wd, u0, u1, smem, load);
    } else {
      float2* scr = args.bigscratch + (long long)blockIdx.x * (NTB / 256) * args.bigL;
      fb_range<L, CPLX, ORDER, NTB>(args, args.tw2048, tw, wd, u0 * 8, u1 * 8, 1, smem, scr, load);
    }
    load = false;
  }
}

}  // namespace

// Every FFT length of a configuration in one persistent launch.  `cls` lists the lengths
// longest first with their ranges in `items`; counters[c] is length c's item counter (zeroed
// before the launch).  Items of length 16384 and 32768 run the four-step body with this
// workgroup's scratch rows.
template <bool CPLX, int ORDER, int NTB>
__global__ void __launch_bounds__(NTB, FXG_P_MINWAVES) k_fpersist(FengineArgs args, const FItem* __restrict__ items,
                                                                  const IClass* __restrict__ cls, int ncls, int* counters) {
  extern __shared__ float2 fx_smem[];
  __shared__ int s_item;
  const int nunits = args.nffts_pad / 8;
  const FXG_CONST IClass* ccls = (const FXG_CONST IClass*)cls;
  const FXG_CONST FItem* citems = (const FXG_CONST FItem*)items;
  for (int c = 0; c < ncls; ++c) {
    const int L = ccls[c].L, b = ccls[c].begin, n = ccls[c].count;
    int* ctr = counters + c;
    // (the twiddle table is picked with a constant index: a run-time index into the kernel
    // arguments would copy them to scratch memory)
#define FXG_PC(L_, LOG_) case L_: p_class<L_, CPLX, ORDER, NTB>(args, citems + b, n, ctr, nunits, args.tw_by_log[LOG_ - 8], fx_smem, s_item); break
    switch (L) {
      FXG_PC(256, 8);
      FXG_PC(512, 9);
      FXG_PC(1024, 10);
      FXG_PC(2048, 11);
      FXG_PC(4096, 12);
      FXG_PC(8192, 13);
      FXG_PC(16384, 14);
      FXG_PC(32768, 15);
      default: break;
    }
#undef FXG_PC
  }
}

template <bool CPLX, int ORDER, int NTB>
static int launch_p(const FengineArgs& a, const FItem* items, const IClass* cls, int ncls, int* counters, int nwg,
                    size_t lds, hipStream_t st) {
  static size_t raised = 0;
  if (raise_lds(k_fpersist<CPLX, ORDER, NTB>, lds, &raised)) return -1;
  hipLaunchKernelGGL((k_fpersist<CPLX, ORDER, NTB>), dim3(nwg), dim3(NTB), lds, st, a, items, cls, ncls, counters);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fxg_launch_fpersist(const FengineArgs& a, const FItem* items, const IClass* cls, int ncls, int* counters,
                        int nwg, int ntb, int iscomplex, size_t lds, hipStream_t st) {
  if (ncls > FXG_MAX_ICLASS) return -2;
  if (fxg_launch_zero(counters, sizeof(int) * ncls, st)) return -1;
#define FXG_P(C_, O_, N_) return launch_p<C_, O_, N_>(a, items, cls, ncls, counters, nwg, lds, st)
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
