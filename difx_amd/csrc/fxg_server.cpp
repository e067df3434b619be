// Serving layer of libfxg: many correlation jobs on one device (the fxg_server_* calls of
// include/fxg.h).
//
// The reference runs one job per mpifxcorr process: FxManager hands subints round-robin to
// Core processes (fxmanager.cpp:331-336), each Core fills and processes a ring of slots
// (core.cpp:205-330), and Visibility::addData sums the subint results of an integration
// (visibility.cpp:192-205).  Here one scheduler thread does that for any number of jobs:
//   - admission: jobs wait in arrival order until their device footprint (the exact bytes
//     fxg_create would hold, from fxg_footprint, plus the job's result, integration and input
//     slots) fits in the budget beside the active jobs;
//   - continuous batching: every active job owns a context and a stream, and the scheduler
//     keeps up to inflight_per_job of its subints enqueued, so the kernels of different jobs
//     overlap on the device while each job's subints stay in order;
//   - device residency: a job's delay-model tables, per-subint delay state and integration
//     accumulator stay in device memory from admission until its last subint; only the
//     integrated results are copied back, once.
// A job's results do not depend on what else runs: its kernels read and write only its own
// buffers, in the same order as when it runs alone.
#include "fxg_internal.h"
#include "../../include/fxg.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

// Deep copy of a job's configuration: the caller's arrays need not outlive fxg_server_add_job.
struct CfgCopy {
  std::vector<fxg_freq> freqs;
  std::vector<fxg_datastream> dss;
  std::vector<fxg_baseline> bls;
  std::vector<std::vector<double>> models;
  fxg_config cfg;

  void take(const fxg_config& c) {
    cfg = c;
    freqs.assign(c.freqs, c.freqs + c.nfreqs);
    dss.assign(c.datastreams, c.datastreams + c.ndatastreams);
    if (c.nbaselines > 0) bls.assign(c.baselines, c.baselines + c.nbaselines);
    models.resize(dss.size());
    for (size_t d = 0; d < dss.size(); ++d) {
      const fxg_datastream& s = dss[d];
      if (s.model_delay_us && s.model_nsamples > 0)
        models[d].assign(s.model_delay_us, s.model_delay_us + (size_t)s.model_nsamples * FXG_MAX_POLY);
      dss[d].model_delay_us = models[d].empty() ? nullptr : models[d].data();
    }
    cfg.freqs = freqs.data();
    cfg.datastreams = dss.data();
    cfg.baselines = bls.empty() ? nullptr : bls.data();
  }
};

struct Request {
  int32_t offsets[3];
  std::vector<const uint8_t*> data;
  std::vector<int32_t> nbytes;
  std::vector<std::vector<int32_t>> control;
  double t_push = 0.0;
};

struct Slot {
  bool busy = false;
  double t_push = 0.0;
  hipEvent_t done = nullptr;
  float* d_res = nullptr;                 // this subint's results (coreresultlength cf32)
  std::vector<uint8_t*> d_in;             // input copies (host-data jobs only)
  std::vector<long long> in_cap;          // their capacities in bytes
};

enum JobState { JOB_PENDING = 0, JOB_ACTIVE = 1, JOB_FINISHED = 2, JOB_FAILED = 3 };

struct Job {
  int id = 0;
  CfgCopy cc;
  int nsub = 0;
  bool on_device = true;
  int pushed = 0, done = 0;
  JobState state = JOB_PENDING;
  long long footprint = 0, crl = 0, slot_bytes = 0, samples_per_subint = 0;
  fxg_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  float* d_integ = nullptr;               // integration accumulator (coreresultlength cf32)
  std::vector<Slot> slots;
  std::deque<Request> queue;
  std::vector<double> lat_ms;
  std::vector<float> result;
  std::string err;
};

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  size_t k = (size_t)std::ceil(q * (double)v.size());
  if (k < 1) k = 1;
  if (k > v.size()) k = v.size();
  return v[k - 1];
}

}  // namespace

struct fxg_server {
  std::mutex mu;
  std::condition_variable wake;           // work for the scheduler
  std::condition_variable changed;        // a job finished or failed
  std::thread sched;
  bool stop = false;
  int device = 0;
  long long budget = 0;
  int max_jobs = 64, inflight = 2;
  Clock::time_point t0;
  std::vector<std::unique_ptr<Job>> jobs;
  std::deque<int> pending;
  std::vector<int> active;
  std::vector<int> retired;               // finished jobs whose device memory is not freed yet
  long long admitted_bytes = 0, peak_bytes = 0;
  int jobs_admitted = 0, jobs_finished = 0, max_active = 0, subints_done = 0;
  long long samples_done = 0;
  double first_launch = -1.0, last_done = 0.0;
  std::vector<double> lat_all;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void release_job(Job* j) {
  for (auto& s : j->slots) {
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.d_res) (void)hipFree(s.d_res);
    for (auto p : s.d_in) if (p) (void)hipFree(p);
  }
  j->slots.clear();
  if (j->d_integ) (void)hipFree(j->d_integ);
  j->d_integ = nullptr;
  if (j->ctx) (void)fxg_destroy(j->ctx);
  j->ctx = nullptr;
  if (j->stream) (void)hipStreamDestroy(j->stream);
  j->stream = nullptr;
}

// Device-side set-up of an admitted job (called without the server lock).
int admit_job(fxg_server* S, Job* j) {
  if (hipStreamCreateWithFlags(&j->stream, hipStreamNonBlocking) != hipSuccess) {
    j->err = "stream create";
    return FXG_E_HIP;
  }
  int rc = fxg_create(&j->cc.cfg, &j->ctx);
  if (rc != FXG_OK) {
    j->err = "context create failed";
    return rc;
  }
  int64_t samples = 0;
  (void)fxg_input_samples(j->ctx, &samples);
  j->samples_per_subint = samples;
  const size_t rbytes = sizeof(float) * 2 * (size_t)std::max<long long>(j->crl, 1);
  if (hipMalloc((void**)&j->d_integ, rbytes) != hipSuccess || hipMemset(j->d_integ, 0, rbytes) != hipSuccess) {
    j->err = "integration buffer";
    return FXG_E_NOMEM;
  }
  j->slots.resize(S->inflight);
  for (auto& s : j->slots) {
    if (hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) return FXG_E_HIP;
    if (hipMalloc((void**)&s.d_res, rbytes) != hipSuccess) { j->err = "result buffer"; return FXG_E_NOMEM; }
    if (!j->on_device) {
      const int nds = j->cc.cfg.ndatastreams;
      s.d_in.assign(nds, nullptr);
      s.in_cap.assign(nds, 0);
      for (int d = 0; d < nds; ++d) {
        int64_t db = 0;
        (void)fxg_data_bytes(j->ctx, d, &db);
        const long long cap = db + 4LL * j->cc.dss[d].framebytes;
        if (hipMalloc((void**)&s.d_in[d], (size_t)cap) != hipSuccess) { j->err = "input slot"; return FXG_E_NOMEM; }
        s.in_cap[d] = cap;
      }
    }
  }
  return FXG_OK;
}

// Enqueue one subint of job j into slot s (server lock held; launches only).
int launch(fxg_server* S, Job* j, Slot& s, Request& r) {
  const int nds = j->cc.cfg.ndatastreams;
  std::vector<const uint8_t*> ptr(nds);
  std::vector<const int32_t*> ctl(nds);
  for (int d = 0; d < nds; ++d) {
    ctl[d] = r.control[d].data();
    if (j->on_device) {
      ptr[d] = r.data[d];
    } else {
      if (r.nbytes[d] < 0 || r.nbytes[d] > s.in_cap[d]) {   // the copy must fit the input slot
        j->err = "subint data larger than the job's input slot";
        return FXG_E_INVALID;
      }
      if (r.nbytes[d] > 0 &&
          hipMemcpyAsync(s.d_in[d], r.data[d], r.nbytes[d], hipMemcpyHostToDevice, j->stream) != hipSuccess)
        return FXG_E_HIP;
      ptr[d] = s.d_in[d];
    }
  }
  int rc = fxg_process_subint_device(j->ctx, r.offsets, ptr.data(), r.nbytes.data(), ctl.data(), s.d_res, j->stream);
  if (rc != FXG_OK) {
    j->err = fxg_last_error(j->ctx);
    return rc;
  }
  if (fxg_launch_add(j->d_integ, s.d_res, 2 * j->crl, j->stream)) return FXG_E_HIP;
  if (hipEventRecord(s.done, j->stream) != hipSuccess) return FXG_E_HIP;
  s.busy = true;
  s.t_push = r.t_push;
  if (S->first_launch < 0) S->first_launch = S->now();
  return FXG_OK;
}

void scheduler(fxg_server* S) {
  (void)hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  while (!S->stop) {
    bool progressed = false;
    // admission in arrival order
    while (!S->pending.empty()) {
      Job* j = S->jobs[S->pending.front()].get();
      if (j->footprint > S->budget) {          // can never fit
        S->pending.pop_front();
        j->state = JOB_FAILED;
        j->err = "job footprint exceeds the server's device-memory budget";
        S->changed.notify_all();
        continue;
      }
      if ((int)S->active.size() < S->max_jobs && S->admitted_bytes + j->footprint > S->budget &&
          !S->retired.empty()) {
        // free retired jobs only under memory pressure: hipFree may wait for the whole device
        std::vector<int> rel;
        rel.swap(S->retired);
        lk.unlock();
        for (int id : rel) release_job(S->jobs[id].get());
        lk.lock();
        for (int id : rel) S->admitted_bytes -= S->jobs[id]->footprint;
        progressed = true;
      }
      if ((int)S->active.size() >= S->max_jobs || S->admitted_bytes + j->footprint > S->budget) break;
      S->pending.pop_front();
      S->admitted_bytes += j->footprint;
      S->peak_bytes = std::max(S->peak_bytes, S->admitted_bytes);
      lk.unlock();
      int rc = admit_job(S, j);
      if (rc != FXG_OK) release_job(j);
      lk.lock();
      if (rc != FXG_OK) {
        j->state = JOB_FAILED;
        S->admitted_bytes -= j->footprint;
        S->changed.notify_all();
        continue;
      }
      j->state = JOB_ACTIVE;
      S->active.push_back(j->id);
      S->jobs_admitted++;
      S->max_active = std::max(S->max_active, (int)S->active.size());
      progressed = true;
    }
    // completions, then launches
    std::vector<int> finished;
    for (int id : S->active) {
      Job* j = S->jobs[id].get();
      for (auto& s : j->slots) {
        if (!s.busy) continue;
        const hipError_t q = hipEventQuery(s.done);
        if (q == hipErrorNotReady) continue;
        s.busy = false;
        progressed = true;
        if (q != hipSuccess) {
          j->err = hipGetErrorString(q);
          j->state = JOB_FAILED;
          continue;
        }
        const double t = S->now();
        const double ms = 1000.0 * (t - s.t_push);
        j->lat_ms.push_back(ms);
        S->lat_all.push_back(ms);
        j->done++;
        S->subints_done++;
        S->samples_done += j->samples_per_subint;
        S->last_done = t;
      }
      if (j->state == JOB_ACTIVE) {
        for (auto& s : j->slots) {
          if (s.busy || j->queue.empty()) continue;
          Request r = std::move(j->queue.front());
          j->queue.pop_front();
          if (launch(S, j, s, r) != FXG_OK) {
            if (j->err.empty()) j->err = "launch failed";
            j->state = JOB_FAILED;
            break;
          }
          progressed = true;
        }
      }
      bool idle = true;
      for (auto& s : j->slots) idle = idle && !s.busy;
      if ((j->done == j->nsub || j->state == JOB_FAILED) && idle) finished.push_back(id);
    }
    for (int id : finished) {
      Job* j = S->jobs[id].get();
      S->active.erase(std::find(S->active.begin(), S->active.end(), id));
      lk.unlock();
      if (j->state != JOB_FAILED) {
        j->result.assign((size_t)2 * j->crl, 0.0f);
        if (hipMemcpyAsync(j->result.data(), j->d_integ, sizeof(float) * j->result.size(), hipMemcpyDeviceToHost,
                           j->stream) != hipSuccess ||
            hipStreamSynchronize(j->stream) != hipSuccess)
          j->state = JOB_FAILED;
      }
      lk.lock();
      if (j->state != JOB_FAILED) {
        j->state = JOB_FINISHED;
        S->jobs_finished++;
      }
      S->retired.push_back(id);               // memory stays counted until released
      S->changed.notify_all();
      progressed = true;
    }
    if (!progressed) {
      // poll in-flight events every 50 us; pushes and new jobs wake the scheduler at once
      S->wake.wait_for(lk, std::chrono::microseconds(50));
    }
  }
}

}  // namespace

extern "C" {

int fxg_server_create(const fxg_server_config* cfg, fxg_server** out) {
  if (!out) return FXG_E_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return FXG_E_NODEVICE;
  fxg_server* S = new fxg_server();
  (void)hipGetDevice(&S->device);
  S->t0 = Clock::now();
  S->max_jobs = cfg && cfg->max_jobs > 0 ? cfg->max_jobs : 64;
  S->inflight = cfg && cfg->inflight_per_job > 0 ? cfg->inflight_per_job : 2;
  if (cfg && cfg->hbm_budget_bytes > 0) {
    S->budget = cfg->hbm_budget_bytes;
  } else {
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) { delete S; return FXG_E_HIP; }
    S->budget = (long long)(0.85 * (double)freeb);
  }
  S->sched = std::thread(scheduler, S);
  *out = S;
  return FXG_OK;
}

int fxg_server_destroy(fxg_server* S) {
  if (!S) return FXG_E_INVALID;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->wake.notify_all();
  S->sched.join();
  for (auto& j : S->jobs) {
    if (j->stream) (void)hipStreamSynchronize(j->stream);
    release_job(j.get());
  }
  delete S;
  return FXG_OK;
}

int fxg_server_add_job(fxg_server* S, const fxg_config* cfg, int32_t nsubints, int32_t data_on_device,
                       int32_t* job) {
  if (!S || !cfg || !job || nsubints < 1) return FXG_E_INVALID;
  std::unique_ptr<Job> j(new Job());
  j->cc.take(*cfg);
  j->nsub = nsubints;
  j->on_device = data_on_device != 0;
  int64_t dev = 0, crl = 0, slot = 0;
  int rc = fxg_footprint(&j->cc.cfg, &dev, &crl, &slot);
  if (rc != FXG_OK) return rc;
  j->crl = crl;
  j->slot_bytes = slot;
  const long long rbytes = 8LL * std::max<long long>(crl, 1);
  j->footprint = dev + rbytes * (1 + S->inflight) + (j->on_device ? 0 : (long long)S->inflight * slot);
  std::lock_guard<std::mutex> lk(S->mu);
  j->id = (int)S->jobs.size();
  *job = j->id;
  S->pending.push_back(j->id);
  S->jobs.push_back(std::move(j));
  S->wake.notify_all();
  return FXG_OK;
}

int fxg_server_push_subint(fxg_server* S, int32_t job, const int32_t offsets[3], const uint8_t* const* data,
                           const int32_t* nbytes, const int32_t* const* control) {
  if (!S || !offsets || !data || !nbytes || !control) return FXG_E_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  if (job < 0 || job >= (int)S->jobs.size()) return FXG_E_INVALID;
  Job* j = S->jobs[job].get();
  if (j->state == JOB_FAILED || j->state == JOB_FINISHED || j->pushed >= j->nsub) return FXG_E_STATE;
  const int nds = j->cc.cfg.ndatastreams;
  // control words: {scan, sec, ns} then ceil(blockspersend / 30) flag words
  const fxg_freq& f0 = j->cc.freqs[j->cc.dss[0].recfreq_index[0]];
  const long long B = (long long)std::llround((double)j->cc.cfg.subint_ns / (1000.0 * f0.nchan / f0.bw_mhz));
  const long long nctl = 3 + (B + 29) / 30;
  Request r;
  std::memcpy(r.offsets, offsets, sizeof(r.offsets));
  r.data.assign(data, data + nds);
  r.nbytes.assign(nbytes, nbytes + nds);
  r.control.resize(nds);
  for (int d = 0; d < nds; ++d) {
    if (!control[d]) return FXG_E_INVALID;
    r.control[d].assign(control[d], control[d] + nctl);
  }
  r.t_push = S->now();
  j->queue.push_back(std::move(r));
  j->pushed++;
  S->wake.notify_all();
  return FXG_OK;
}

int fxg_server_wait_job(fxg_server* S, int32_t job, float* integrated, double* p50_ms, double* p99_ms) {
  if (!S) return FXG_E_INVALID;
  std::unique_lock<std::mutex> lk(S->mu);
  if (job < 0 || job >= (int)S->jobs.size()) return FXG_E_INVALID;
  Job* j = S->jobs[job].get();
  if (j->pushed < j->nsub && j->state != JOB_FAILED) {
    // all subints must have been pushed before waiting, or the wait could never end
    return FXG_E_STATE;
  }
  S->changed.wait(lk, [&] { return j->state == JOB_FINISHED || j->state == JOB_FAILED; });
  if (j->state == JOB_FAILED) return FXG_E_STATE;
  if (integrated) std::memcpy(integrated, j->result.data(), sizeof(float) * j->result.size());
  if (p50_ms) *p50_ms = percentile(j->lat_ms, 0.50);
  if (p99_ms) *p99_ms = percentile(j->lat_ms, 0.99);
  return FXG_OK;
}

int fxg_server_job_error(fxg_server* S, int32_t job, char* buf, int32_t len) {
  if (!S || !buf || len < 1) return FXG_E_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  if (job < 0 || job >= (int)S->jobs.size()) return FXG_E_INVALID;
  std::strncpy(buf, S->jobs[job]->err.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return FXG_OK;
}

int fxg_server_get_stats(fxg_server* S, fxg_server_stats* out) {
  if (!S || !out) return FXG_E_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  std::memset(out, 0, sizeof(*out));
  out->jobs_added = (int32_t)S->jobs.size();
  out->jobs_admitted = S->jobs_admitted;
  out->jobs_finished = S->jobs_finished;
  out->jobs_active = (int32_t)S->active.size();
  out->max_active = S->max_active;
  out->subints_done = S->subints_done;
  out->samples_done = S->samples_done;
  out->admitted_bytes = S->admitted_bytes;
  out->peak_admitted_bytes = S->peak_bytes;
  out->first_launch_s = S->first_launch < 0 ? 0.0 : S->first_launch;
  out->last_done_s = S->last_done;
  out->latency_p50_ms = percentile(S->lat_all, 0.50);
  out->latency_p99_ms = percentile(S->lat_all, 0.99);
  out->latency_max_ms = S->lat_all.empty() ? 0.0 : *std::max_element(S->lat_all.begin(), S->lat_all.end());
  return FXG_OK;
}

}  // extern "C"
