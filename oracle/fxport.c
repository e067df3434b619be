/*
 * fxport.c -- CPU restatement of mpifxcorr's per-subint FX path in the reference's FP32
 * arithmetic (test infrastructure and timed CPU baseline; never part of the product).
 *
 * Per FFT block it follows Mode::process (mpifxcorr/src/mode.cpp:634-1374) for fringe
 * rotation orders 0, 1 and 2 with the GENERIC architecture primitives
 * (architecture.h.in:399-1063):
 *   - Mk5Mode::unpack through the 2-bit / 8-bit LUTs (mk5mode.cpp:94-174, format_vdif.c:66-145,
 *     mark5_format_mark5b.c:49-86) with frame blanking;
 *   - the FP64 per-FFT delay chain (mode.cpp:671-675, 743-757, 776-825);
 *   - the strided FP32 rotator: FP64 phases at S sub points and NS step points, wrapped,
 *     FP32 sin/cos, complexrotator[s*S + j] = sub[j] * step[s] (order 1, mode.cpp:883-955),
 *     or with the sub phasors multiplied by the quadratic term after every step (order 2,
 *     mode.cpp:956-1026); order 0 has no pre-F rotator;
 *   - the strided FP32 fractional-sample phasors (mode.cpp:1030-1073), carrying the post-F
 *     fringe phase for order 0 (mode.cpp:1040-1049);
 *   - an unnormalised forward FFT (FFTW3 API via MKL when available, mode.cpp:1174-1183;
 *     a real-to-complex FFT with the LSB conj-flip for order 0, mode.cpp:1139-1156),
 *     USB / real-LSB channel selection (mode.cpp:1185-1217), frac multiply (:1257-1261);
 * and per buffered batch the cross-multiply-accumulate of Core::processdata
 * (core.cpp:884-979, :963) and the autocorrelation products (mode.cpp:1272, 1333-1339), all
 * as sequential FP32 vectorAddProduct over FFTs.  Worker threads own contiguous FFT-block
 * ranges of the subint, as Core's process threads do (core.cpp:2161-2166), and their
 * accumulators are summed in thread order at the end.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define TWO_PI 6.283185307179586476925286766559   /* architecture.h.in:430 */
#define HIMAG 3.3359f                               /* mark5_stream.h:65 */
#define FLAGS_PER_INT 30                            /* mpifxcorr.h:29 */
#define TINY 0.000000001                            /* mode.cpp:27 */

typedef struct { float re, im; } cf32;

typedef struct port_freq {
  double lofreq, lo_offset_hz, fraclooffset, clockoff, halfdelta;
  int lsb, looff, isfraclooffset, nbands, band[4];
  double bhd[4];                  /* per band: +halfdelta (R, fracsamprotatorA) or -halfdelta */
  float pcre[4], pcim[4];         /* per band: pol phase-offset phasor (1, 0) unless rotated */
} port_freq;

typedef struct port_ds {
  int fmt, nbits, cplx, framebytes, headerbytes, spf, nrec, nrecfreqs, L, C, S, NS, NFS;
  int bpb_num, bpb_den, fractionallo, valid;
  int dsb;                        /* complex double sideband (mode.cpp:41-49) */
  double sampletime_us;
  double interp[3];
  int datasec, datans, datalen, nframes;
  const uint8_t* data;
  const int32_t* vflags;
  const float* chanfreq;          /* sub[S], usb, lsb, dsb steps [NFS] each (mode.cpp:303-324) */
  port_freq rf[32];
} port_ds;

typedef struct port_prod {
  int ds1, band1, choff1, ds2, band2, choff2, nchan;
  long long acc_off;
} port_prod;

/* ---------------- FFT: FFTW3 API (MKL) when loadable, else a radix-2 restatement ---------- */
typedef void* (*plan_fn)(int, float*, float*, int, unsigned);
typedef void* (*rplan_fn)(int, float*, float*, unsigned);
typedef void (*exec_fn)(void*);
typedef void (*destroy_fn)(void*);
static plan_fn p_plan = 0;
static rplan_fn p_rplan = 0;
static exec_fn p_exec = 0;
static destroy_fn p_destroy = 0;
static int fft_backend = -1;      /* 0 = own radix-2, 1 = FFTW3 API */
static pthread_mutex_t plan_mu = PTHREAD_MUTEX_INITIALIZER;

static void load_fftw(void) {
  if (fft_backend >= 0) return;
  const char* names[] = {"libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so.1",
                         "/opt/conda/lib/libmkl_rt.so.1", "libfftw3f.so.3", 0};
  for (int i = 0; names[i]; ++i) {
    void* h = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
    if (!h) continue;
    p_plan = (plan_fn)dlsym(h, "fftwf_plan_dft_1d");
    p_exec = (exec_fn)dlsym(h, "fftwf_execute");
    p_destroy = (destroy_fn)dlsym(h, "fftwf_destroy_plan");
    p_rplan = (rplan_fn)dlsym(h, "fftwf_plan_dft_r2c_1d");
    if (p_plan && p_exec && p_destroy) { fft_backend = 1; return; }
  }
  fft_backend = 0;
}

int fxport_fft_backend(void) { load_fftw(); return fft_backend; }

typedef struct {
  int L;
  void* plan;
  void* rplan;  /* real-to-complex plan (order 0) */
  float* rin;
  cf32* in;
  cf32* out;
  cf32* tw;     /* radix-2 twiddles */
  int* rev;
} fftctx;

static void fft_init(fftctx* f, int L) {
  f->L = L;
  f->in = (cf32*)aligned_alloc(64, sizeof(cf32) * L);
  f->out = (cf32*)aligned_alloc(64, sizeof(cf32) * L);
  f->rin = (float*)aligned_alloc(64, sizeof(float) * L);
  f->plan = 0;
  f->rplan = 0;
  f->tw = 0;
  f->rev = 0;
  if (fft_backend == 1) {
    pthread_mutex_lock(&plan_mu);
    f->plan = p_plan(L, (float*)f->in, (float*)f->out, -1 /* FFTW_FORWARD */, 1u << 6 /* FFTW_ESTIMATE */);
    if (p_rplan) f->rplan = p_rplan(L, f->rin, (float*)f->out, 1u << 6);
    pthread_mutex_unlock(&plan_mu);
  }
  if (!f->plan) {
    f->tw = (cf32*)malloc(sizeof(cf32) * L / 2);
    f->rev = (int*)malloc(sizeof(int) * L);
    for (int m = 0; m < L / 2; ++m) {
      double a = -TWO_PI * m / L;
      f->tw[m].re = (float)cos(a);
      f->tw[m].im = (float)sin(a);
    }
    int lg = 0;
    while ((1 << lg) < L) ++lg;
    for (int i = 0; i < L; ++i) {
      int r = 0;
      for (int b = 0; b < lg; ++b) r |= ((i >> b) & 1) << (lg - 1 - b);
      f->rev[i] = r;
    }
  }
}

static void fft_free(fftctx* f) {
  pthread_mutex_lock(&plan_mu);
  if (f->plan) p_destroy(f->plan);
  if (f->rplan) p_destroy(f->rplan);
  pthread_mutex_unlock(&plan_mu);
  free(f->rin);
  free(f->in); free(f->out); free(f->tw); free(f->rev);
}

static void fft_run(fftctx* f) {   /* out = FFT(in), unnormalised */
  if (f->plan) { p_exec(f->plan); return; }
  const int L = f->L;
  for (int i = 0; i < L; ++i) f->out[f->rev[i]] = f->in[i];
  for (int len = 2; len <= L; len <<= 1) {
    const int h = len >> 1, st = L / len;
    for (int b = 0; b < L; b += len)
      for (int j = 0; j < h; ++j) {
        const cf32 w = f->tw[j * st];
        const cf32 x = f->out[b + j], y = f->out[b + j + h];
        const cf32 t = {y.re * w.re - y.im * w.im, y.re * w.im + y.im * w.re};
        f->out[b + j].re = x.re + t.re; f->out[b + j].im = x.im + t.im;
        f->out[b + j + h].re = x.re - t.re; f->out[b + j + h].im = x.im - t.im;
      }
  }
}

/* out[0..L/2] = FFT of the real rin (L/2 + 1 bins, CCS order) */
static void fft_run_real(fftctx* f) {
  if (f->rplan) { p_exec(f->rplan); return; }
  for (int n = 0; n < f->L; ++n) { f->in[n].re = f->rin[n]; f->in[n].im = 0.0f; }
  fft_run(f);
}

/* ---------------------------------------------------------------- unpack helpers ---- */
static float lut2(int fmt, unsigned field) {
  static const float vdif[4] = {-HIMAG, -1.0f, 1.0f, HIMAG};
  static const float m5b[4] = {-HIMAG, 1.0f, -1.0f, HIMAG};
  return fmt == 1 ? m5b[field & 3] : vdif[field & 3];
}
static float lut8(unsigned b) { return (float)(((double)b - 128.0) / 3.3); }

/* Frame flags: 0 blanked, 1 every sample decodes, 2 samples [zone[2 fr], zone[2 fr + 1]) of
 * the frame decode (a Mark5B payload with fill pattern at one end, blanker_mark5.c:112-172). */
static int sample_ok(const port_ds* d, const uint8_t* fv, const int* zone, long long fr, long long within) {
  if (fr < 0 || fr >= d->nframes || !fv[fr]) return 0;
  return fv[fr] == 1 || (within >= zone[2 * fr] && within < zone[2 * fr + 1]);
}

static uint64_t word64(const uint8_t* p, int k) { uint64_t w; memcpy(&w, p + 8 * (long long)k, 8); return w; }
static uint32_t word32(const uint8_t* p, int k) { uint32_t w; memcpy(&w, p + 4 * (long long)k, 4); return w; }

/* blanker_mark5's halving search over payload words [0, n) (blanker_mark5.c:43-111): probes
 * start at n/2 and move by n/4, n/8, ..., 1, down after a probe whose word is (fill == want)
 * and up otherwise; the lowest such probe, n - 1 if none, in bytes. */
static int zone_edge(const uint8_t* pay, int n, int want) {
  int probe = n / 2, hit = n - 1;
  for (int step = n / 4; step > 0; step /= 2) {
    const int m = (word64(pay, probe) == 0x1122334411223344ULL) == want;
    if (m) { hit = probe; probe -= step; } else { probe += step; }
  }
  return 8 * hit;
}

/* Zone 0 of a Mark5B payload as the time samples [lo, hi) whose first byte
 * (n * bits per sample) / 8 lies in the blanker's byte range (mark5_format_mark5b.c:241-243). */
static void mark5_zone(const port_ds* d, const uint8_t* pay, int* lo, int* hi) {
  const int nword = (d->framebytes - d->headerbytes) / 8;
  const int e = nword < 4096 ? nword : 4096;
  const uint64_t fill = 0x1122334411223344ULL;
  const int sok = word64(pay, 0) != fill, eok = word64(pay, e - 1) != fill;
  long long sv = 0, ev = 1LL << 30;
  if (!sok && !eok) { sv = 1LL << 30; ev = 0; }
  else if (sok && !eok) {
    ev = zone_edge(pay, e, 1);
    if (word32(pay, (int)(ev / 4) - 1) == 0x11223344u) ev -= 4;   /* straggler half word */
  } else if (!sok) {
    sv = zone_edge(pay, e, 0);
    if (word32(pay, (int)(sv / 4) - 2) == 0x11223344u) sv += 4;   /* as tested at blanker_mark5.c:104 */
  }
  const long long bps = (long long)d->nbits * (d->cplx ? 2 : 1) * d->nrec;
  long long a = (8 * sv + bps - 1) / bps, b = (8 * ev + bps - 1) / bps;
  *lo = (int)(a < d->spf ? a : d->spf);
  *hi = (int)(b < d->spf ? b : d->spf);
}

/* ------------------------------------------------------------------ worker ---------- */
typedef struct {
  int nds, B, offsec, offns, nbuffered, order;
  double scan_start;
  const port_ds* ds;
  const uint8_t** fvalid;
  const int** fzone;
  int nprod;
  const port_prod* prods;
  long long acclen;
  float* weights;
  int b0, b1;
  cf32* acc;      /* thread accumulator */
} job;

static cf32 cmulf(cf32 a, cf32 b) {
  cf32 r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}

static void* worker(void* arg) {
  job* J = (job*)arg;
  const int nds = J->nds, NB = J->nbuffered;
  /* per datastream buffers */
  cf32*** spec = (cf32***)calloc(nds, sizeof(cf32**));
  float* dweight = (float*)calloc((size_t)nds * NB, sizeof(float));
  fftctx* fc = (fftctx*)calloc(nds, sizeof(fftctx));
  int maxL = 0, maxrec = 0;
  for (int d = 0; d < nds; ++d) {
    const port_ds* D = &J->ds[d];
    spec[d] = (cf32**)calloc((size_t)D->nrec * NB, sizeof(cf32*));
    for (int k = 0; k < D->nrec * NB; ++k) spec[d][k] = (cf32*)calloc(D->C, sizeof(cf32));
    fft_init(&fc[d], D->L);
    if (D->L > maxL) maxL = D->L;
    if (D->nrec > maxrec) maxrec = D->nrec;
  }
  cf32* unp = (cf32*)malloc(sizeof(cf32) * (size_t)maxL * maxrec);
  cf32* rot = (cf32*)malloc(sizeof(cf32) * maxL);
  cf32* frac = (cf32*)malloc(sizeof(cf32) * maxL);
  cf32* tab = (cf32*)malloc(sizeof(cf32) * 2 * maxL);
  cf32* quad = (cf32*)malloc(sizeof(cf32) * maxL);
  const int order = J->order;

  for (int i0 = J->b0; i0 < J->b1; i0 += NB) {
    const int nsub = (J->b1 - i0) < NB ? (J->b1 - i0) : NB;
    for (int d = 0; d < nds; ++d) {
      const port_ds* D = &J->ds[d];
      const uint8_t* fv = J->fvalid[d];
      const int* zn = J->fzone[d];
      const int L = D->L, C = D->C, S = D->S, NS = D->NS, NFS = D->NFS, nrec = D->nrec;
      const double dt = D->sampletime_us;
      for (int sl = 0; sl < nsub; ++sl) {
        const int i = i0 + sl;
        float w = 0.0f;
        for (int b = 0; b < nrec; ++b) memset(spec[d][b * NB + sl], 0, sizeof(cf32) * C);
        dweight[d * NB + sl] = 0.0f;
        if (!D->valid || D->datalen <= 1) continue;
        if (((D->vflags[i / FLAGS_PER_INT] >> (i % FLAGS_PER_INT)) & 1) == 0) continue;
        const double* q = D->interp;
        const double fcn = i + 0.5;
        const double avgdelay = q[0] * fcn * fcn + q[1] * fcn + q[2];
        const double fftstart = i * L * dt;
        const double starttime = (J->offsec - D->datasec) * 1000000.0 +
                                 ((long long)J->offns - (long long)D->datans) / 1000.0 + fftstart - avgdelay;
        int n0 = (int)(starttime / dt + 0.5);
        if (n0 < -1 || (((long long)(n0 + L)) * D->bpb_num) / D->bpb_den > D->datalen) continue;
        if (n0 == -1) n0 = 0;
        /* unpack and count valid samples */
        long long good = 0;
        long long fr = (long long)n0 / D->spf;
        long long within = (long long)n0 - fr * D->spf - 1;
        for (int n = 0; n < L; ++n) {
          if (++within == D->spf) { within = 0; ++fr; }
          const int ok = sample_ok(D, fv, zn, fr, within);
          good += ok;
          const uint8_t* pay = D->data + fr * D->framebytes + D->headerbytes;
          for (int b = 0; b < nrec; ++b) {
            cf32 x = {0.0f, 0.0f};
            if (ok) {
              if (D->cplx) {
                const uint8_t* p = pay + (within * nrec + b) * 2;
                x.re = lut8(p[0]); x.im = lut8(p[1]);
              } else if (D->nbits == 2) {
                const long long qf = within * nrec + b;
                x.re = lut2(D->fmt, (pay[qf >> 2] >> (2 * (qf & 3))) & 3u);
              } else {
                x.re = lut8(pay[within * nrec + b]);
              }
            }
            unp[(size_t)b * L + n] = x;
          }
        }
        w = (float)good / (float)L;
        if (!(w > 0.0f)) continue;
        dweight[d * NB + sl] = w;
        const float fse = (float)(starttime - n0 * dt);
        double a = 0.0, bb = 0.0, cc = 0.0;
        int intdelay = 0;
        if (order == 0) {                                         /* mode.cpp:778-780 */
          intdelay = (int)avgdelay;
        } else if (order == 1) {                                  /* mode.cpp:781-799 */
          const double d0 = q[0] * i * (double)i + q[1] * i + q[2];
          const double d1 = q[0] * (i + 0.5) * (i + 0.5) + q[1] * (i + 0.5) + q[2];
          const double d2 = q[0] * (i + 1.0) * (i + 1.0) + q[1] * (i + 1.0) + q[2];
          a = d2 - d0;
          bb = d0 + (d1 - (a * 0.5 + d0)) / 3.0;
          intdelay = (int)bb;
          bb -= intdelay;
        } else {                                                  /* mode.cpp:800-825 */
          a = q[0];
          bb = q[1] + i * q[0] * 2.0;
          cc = q[2] + i * q[1] + (double)((long long)i * i) * q[0];
          intdelay = (int)cc;
          cc -= intdelay;
        }
        const double walltime = J->scan_start + J->offsec + J->offns / 1.0e9 + fftstart / 1.0e6;
        const int intwall = (int)walltime;
        const double fracwall = walltime - intwall;
        for (int f = 0; f < D->nrecfreqs; ++f) {
          const port_freq* R = &D->rf[f];
          if (R->nbands == 0) continue;
          /* rotator: sub (tab[0..S)) and step (tab[S..S+NS)) phasors, quad[0..S) for order 2 */
          const double sx1 = (double)S / (double)L;
          for (int j = 0; order != 0 && j < S + NS; ++j) {
            double ph;
            if (j < S) {
              const double xoff = (double)j / (double)L;
              const double xval = order == 1 ? xoff * a + bb : xoff * (bb + a * sx1) + cc;
              ph = xval * R->lofreq;
              if (order == 2) {
                const double qph = (xoff * (2 * a * sx1)) * R->lofreq;
                const float qarg = (float)(-TWO_PI * (qph - (int)qph));
                quad[j].re = cosf(qarg);
                quad[j].im = sinf(qarg);
              }
              if (D->fractionallo) ph += (R->lofreq - (int)R->lofreq) * (double)intdelay;
              if (R->looff) {
                ph += (j * dt / 1e6) * (-R->lo_offset_hz);
                ph += -R->lo_offset_hz * fracwall;
                if (R->isfraclooffset) ph += -R->fraclooffset * intwall;
              }
            } else {
              const int s = j - S;
              const double xoff = (double)(s * S) / (double)L;
              ph = order == 1 ? xoff * a * R->lofreq : (xoff * bb + (xoff * xoff) * a) * R->lofreq;
              if (R->looff) ph += (s * S * dt / 1e6) * (-R->lo_offset_hz);
            }
            const float arg = (float)(-TWO_PI * (ph - (int)ph));
            tab[j].re = cosf(arg);
            tab[j].im = sinf(arg);
          }
          if (order == 1) {
            for (int j = 0; j < S; ++j) rot[j] = tab[j];
            for (int s = 1; s < NS; ++s)
              for (int j = 0; j < S; ++j) rot[s * S + j] = cmulf(tab[j], tab[S + s]);
          } else if (order == 2) {                    /* mode.cpp:1017-1024 */
            for (int s = 0; s < NS; ++s) {
              for (int j = 0; j < S; ++j) rot[s * S + j] = cmulf(tab[j], tab[S + s]);
              for (int j = 0; j < S; ++j) tab[j] = cmulf(quad[j], tab[j]);
            }
          }
          /* fractional-sample phasors: built for the first band, and again for a band whose
             clock-delta term differs (fracsamprotatorB of the non-R pol, mode.cpp:1076-1118) */
          const float* sub = D->chanfreq;
          const float* step = D->dsb ? D->chanfreq + S + 2 * NFS
                                     : R->lsb ? D->chanfreq + S + NFS : D->chanfreq + S;   /* mode.cpp:838-857 */
          float prf = 0.0f;
          if (order == 0) {                           /* post-F fringe phase, mode.cpp:1040-1049 */
            double pr = (avgdelay - intdelay) * R->lofreq;
            if (D->fractionallo) pr += intdelay * (R->lofreq - (int)R->lofreq);
            pr -= walltime * R->lo_offset_hz;
            prf = (float)(-TWO_PI * (pr - (int)pr));
          }
          for (int kb = 0; kb < R->nbands; ++kb) {
            const int b = R->band[kb];
            if (kb == 0 || R->bhd[kb] != R->bhd[kb - 1]) {
              const float e = (float)((double)fse - R->clockoff + R->bhd[kb]);
              for (int j = 0; j < S; ++j) {
                float ar = sub[j] * e;
                if (order == 0) ar += prf;
                tab[j].re = cosf(ar); tab[j].im = sinf(ar);
              }
              for (int s = 0; s < NFS; ++s) { const float ar = step[s] * e; tab[S + s].re = cosf(ar); tab[S + s].im = sinf(ar); }
              for (int j = 0; j < S; ++j) frac[j] = tab[j];
              for (int s = 1; s < NFS; ++s)
                for (int j = 0; j < S; ++j) frac[s * S + j] = cmulf(tab[j], tab[S + s]);
              for (int j = 0; j < S; ++j) frac[j] = cmulf(frac[j], tab[S]);
            }
            /* the pol phase offset of the second pol (phasecorrA, mode.cpp:1318-1330) */
            const cf32 pc = {R->pcre[kb], R->pcim[kb]};
            const int rot_pc = !(pc.re == 1.0f && pc.im == 0.0f);
            cf32* in = fc[d].in;
            const cf32* u = unp + (size_t)b * L;
            cf32* F = spec[d][b * NB + sl];
            if (order == 0) {                         /* real FFT (mode.cpp:1139-1156) */
              for (int n = 0; n < L; ++n) fc[d].rin[n] = u[n].re;
              fft_run_real(&fc[d]);
              const cf32* X = fc[d].out;
              if (R->lsb) {                           /* conj-flip of bins 1..C */
                for (int c = 0; c < C; ++c) { F[c].re = X[C - c].re; F[c].im = -X[C - c].im; }
              } else {
                for (int c = 0; c < C; ++c) F[c] = X[c];
              }
              for (int c = 0; c < C; ++c) F[c] = cmulf(F[c], frac[c]);
              if (rot_pc) for (int c = 0; c < C; ++c) F[c] = cmulf(F[c], pc);
              continue;
            }
            if (D->cplx) {
              for (int n = 0; n < L; ++n) in[n] = cmulf(rot[n], u[n]);
            } else {
              for (int n = 0; n < L; ++n) { in[n].re = u[n].re * rot[n].re; in[n].im = u[n].re * rot[n].im; }
            }
            fft_run(&fc[d]);
            const cf32* X = fc[d].out;
            /* channel 0 = lowest sky frequency (mode.cpp:1186-1222) */
            if (!D->cplx) {
              const cf32* Xs = X + (R->lsb ? C : 0);
              for (int c = 0; c < C; ++c) F[c] = Xs[c];
            } else if (D->dsb && R->lsb) {
              for (int c = 0; c <= C / 2; ++c) { F[c].re = X[C / 2 - c].re; F[c].im = -X[C / 2 - c].im; }
              for (int c = C / 2 + 1; c < C; ++c) { F[c].re = X[3 * C / 2 - c].re; F[c].im = -X[3 * C / 2 - c].im; }
            } else if (D->dsb) {
              for (int c = 0; c < C; ++c) F[c] = X[(c + C / 2) % C];
            } else if (R->lsb) {
              F[0] = X[0];
              for (int c = 1; c < C; ++c) { F[c].re = X[C - c].re; F[c].im = -X[C - c].im; }
            } else {
              for (int c = 0; c < C; ++c) F[c] = X[c];
            }
            for (int c = 0; c < C; ++c) F[c] = cmulf(F[c], frac[c]);
            if (rot_pc) for (int c = 0; c < C; ++c) F[c] = cmulf(F[c], pc);
          }
        }
      }
    }
    /* XMAC over the buffered batch (core.cpp:884-979) and autocorrelation products */
    for (int p = 0; p < J->nprod; ++p) {
      const port_prod* P = &J->prods[p];
      const int NB1 = NB;
      cf32* acc = J->acc + P->acc_off;
      const int isauto = (P->ds1 == P->ds2) && (P->band1 == P->band2) && (P->choff1 == P->choff2);
      for (int sl = 0; sl < nsub; ++sl) {
        const cf32* x = spec[P->ds1][P->band1 * NB1 + sl] + P->choff1;
        const cf32* y = spec[P->ds2][P->band2 * NB1 + sl] + P->choff2;
        for (int c = 0; c < P->nchan; ++c) {
          const float yr = y[c].re, yi = -y[c].im;          /* conjfftoutputs */
          acc[c].re += x[c].re * yr - x[c].im * yi;
          acc[c].im += x[c].re * yi + x[c].im * yr;
        }
        if (isauto)
          for (int c = 0; c < P->nchan; ++c) acc[c].im = 0.0f;
      }
    }
    for (int d = 0; d < nds; ++d)
      for (int sl = 0; sl < nsub; ++sl) J->weights[(long long)d * J->B + i0 + sl] = dweight[d * NB + sl];
  }
  for (int d = 0; d < nds; ++d) {
    for (int k = 0; k < J->ds[d].nrec * NB; ++k) free(spec[d][k]);
    free(spec[d]);
    fft_free(&fc[d]);
  }
  free(spec); free(dweight); free(fc); free(unp); free(rot); free(frac); free(tab); free(quad);
  return 0;
}

/* Process one subint: products accumulate into acc (cf32, acclen), per-FFT data weights into
 * weights[nds][B].  Returns the FFT backend used (1 = FFTW3 API, 0 = radix-2). */
int fxport_run(int nds, const port_ds* ds, int B, int offsec, int offns, double scan_start,
               int nprod, const port_prod* prods, long long acclen, float* acc, float* weights,
               int nthreads, int nbuffered, int order) {
  load_fftw();
  if (nthreads < 1) nthreads = 1;
  if (nthreads > B) nthreads = B;
  if (nbuffered < 1) nbuffered = 1;
  /* frame flags per datastream: VDIF word 2 / invalid bit and fill pattern at either end;
   * Mark5B invalid flag and the fill-pattern zone of the payload */
  uint8_t** fvalid = (uint8_t**)calloc(nds, sizeof(uint8_t*));
  int** fzone = (int**)calloc(nds, sizeof(int*));
  for (int d = 0; d < nds; ++d) {
    const port_ds* D = &ds[d];
    fvalid[d] = (uint8_t*)calloc(D->nframes + 1, 1);
    fzone[d] = (int*)calloc(2 * (size_t)D->nframes + 2, sizeof(int));
    for (int f = 0; f < D->nframes; ++f) {
      const uint8_t* fr = D->data + (long long)f * D->framebytes;
      const uint8_t* pay = fr + D->headerbytes;
      uint32_t w0, w2;
      memcpy(&w0, fr, 4);
      memcpy(&w2, fr + 8, 4);
      int ok;
      if (D->fmt == 0) {
        ok = w2 != 0u && ((w0 >> 31) & 1u) == 0u;
        if (word64(pay, 0) == 0x1122334411223344ULL ||
            word64(fr + D->framebytes - 8, 0) == 0x1122334411223344ULL) ok = 0;
      } else {
        ok = (fr[5] & 0x80) == 0;
        if (ok) {
          int lo, hi;
          mark5_zone(D, pay, &lo, &hi);
          ok = hi <= lo ? 0 : (lo == 0 && hi == D->spf) ? 1 : 2;
          fzone[d][2 * f] = lo;
          fzone[d][2 * f + 1] = hi;
        }
      }
      fvalid[d][f] = (uint8_t)ok;
    }
  }
  pthread_t* th = (pthread_t*)calloc(nthreads, sizeof(pthread_t));
  job* jobs = (job*)calloc(nthreads, sizeof(job));
  for (int t = 0; t < nthreads; ++t) {
    job* J = &jobs[t];
    J->nds = nds; J->B = B; J->offsec = offsec; J->offns = offns; J->scan_start = scan_start;
    J->nbuffered = nbuffered; J->order = order; J->ds = ds; J->fvalid = (const uint8_t**)fvalid;
    J->fzone = (const int**)fzone;
    J->nprod = nprod; J->prods = prods; J->acclen = acclen; J->weights = weights;
    J->b0 = (int)((long long)B * t / nthreads);
    J->b1 = (int)((long long)B * (t + 1) / nthreads);
    J->acc = (cf32*)calloc(acclen > 0 ? acclen : 1, sizeof(cf32));
    pthread_create(&th[t], 0, worker, J);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], 0);
  cf32* out = (cf32*)acc;
  for (int t = 0; t < nthreads; ++t) {           /* thread order: deterministic sum */
    for (long long k = 0; k < acclen; ++k) {
      out[k].re += jobs[t].acc[k].re;
      out[k].im += jobs[t].acc[k].im;
    }
    free(jobs[t].acc);
  }
  for (int d = 0; d < nds; ++d) { free(fvalid[d]); free(fzone[d]); }
  free(fvalid); free(fzone); free(th); free(jobs);
  return fft_backend;
}
