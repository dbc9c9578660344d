// Serving loop on one device: the scheduler of gx_serve_core.hpp driving
// whole-volume plans built inside the server's own memory.
//
// The server holds one device allocation, its budget. An admitted request
// gets two ranges of it: `work` (the plan's buffers, then the copy of a host
// input) for the time it runs, and `keep` (the volume) until the client
// releases it. Each slot is a stream with two events; a request's whole
// path (plan tables, input copy, correction, filter, backprojection) is
// enqueued on its slot's stream, so requests overlap on the device and never
// allocate or synchronise the device.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "guanaco_hip.h"
#include "gx_kernels.hpp"
#include "gx_serve_core.hpp"

namespace gx {

int plan_create_in(const gx_volume_desc *d, void *ws, size_t bytes,
                   hipStream_t s, gx_volume_plan **out);
int fail_status(int code, const char *msg);

namespace {

constexpr size_t kAlign = 256;
size_t round_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// A request as the executor sees it: its own copy of the descriptor.
struct Request {
  gx_volume_desc desc{};
  std::vector<float> angles, weights, centres;
  std::vector<double> defocus;
  const float *input = nullptr;
  bool input_on_host = false;
  size_t input_bytes = 0, plan_bytes = 0;
  gx_volume_plan *plan = nullptr;
  float device_ms = 0;
};

size_t input_elements(const gx_volume_desc *d) {
  const size_t n = d->num_angles * d->num_rows * d->num_pixels;
  return d->input_layout == GX_INPUT_PROJECTIONS ? n : n * d->num_defocus;
}

class HipExec {
 public:
  HipExec(int device, int nslots, char *base)
      : device_(device), base_(base), streams_(nslots), ev0_(nslots),
        ev1_(nslots) {}
  hipError_t init() {
    for (size_t i = 0; i < streams_.size(); ++i) {
      hipError_t e = hipStreamCreateWithFlags(&streams_[i],
                                              hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreate(&ev0_[i]);
      if (e == hipSuccess) e = hipEventCreate(&ev1_[i]);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  ~HipExec() {
    (void)hipSetDevice(device_);
    for (size_t i = 0; i < streams_.size(); ++i) {
      if (streams_[i]) (void)hipStreamSynchronize(streams_[i]);
      if (ev0_[i]) (void)hipEventDestroy(ev0_[i]);
      if (ev1_[i]) (void)hipEventDestroy(ev1_[i]);
      if (streams_[i]) (void)hipStreamDestroy(streams_[i]);
    }
  }
  int slots() const { return (int)streams_.size(); }

  bool start(serve::Job &j, int slot) {
    Request &r = *static_cast<Request *>(j.payload.get());
    if (hipSetDevice(device_) != hipSuccess) {
      j.error = "cannot select the server's device";
      return false;
    }
    hipStream_t s = streams_[slot];
    char *work = base_ + j.work_off;
    float *vol = reinterpret_cast<float *>(base_ + j.keep_off);
    if (hipEventRecord(ev0_[slot], s) != hipSuccess) {
      j.error = "event record failed";
      return false;
    }
    if (plan_create_in(&r.desc, work, r.plan_bytes, s, &r.plan) != GX_OK) {
      j.error = gx_last_error();
      return false;
    }
    const float *in = r.input;
    if (r.input_on_host) {
      float *dst = reinterpret_cast<float *>(work + r.plan_bytes);
      hipError_t e = hipMemcpyAsync(dst, r.input, r.input_bytes,
                                    hipMemcpyHostToDevice, s);
      if (e != hipSuccess) {
        j.error = std::string("input copy: ") + hipGetErrorString(e);
        return false;
      }
      in = dst;
    }
    if (gx_volume_run(r.plan, in, vol, s) != GX_OK) {
      j.error = gx_last_error();
      return false;
    }
    if (hipEventRecord(ev1_[slot], s) != hipSuccess) {
      j.error = "event record failed";
      return false;
    }
    return true;
  }

  int poll(serve::Job &j, int slot) {
    const hipError_t e = hipEventQuery(ev1_[slot]);
    if (e == hipErrorNotReady) return 0;
    Request &r = *static_cast<Request *>(j.payload.get());
    if (e != hipSuccess) {
      j.error = std::string("device: ") + hipGetErrorString(e);
      return -1;
    }
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev0_[slot], ev1_[slot]) == hipSuccess)
      r.device_ms = ms;
    return 1;
  }

  void finish(serve::Job &j) {
    Request &r = *static_cast<Request *>(j.payload.get());
    if (r.plan) gx_volume_plan_destroy(r.plan);  // buffers are borrowed
    r.plan = nullptr;
  }
  void drop(serve::Job &) {}

  float *volume(const serve::Job &j) const {
    return reinterpret_cast<float *>(base_ + j.keep_off);
  }

 private:
  int device_;
  char *base_;
  std::vector<hipStream_t> streams_;
  std::vector<hipEvent_t> ev0_, ev1_;
};

}  // namespace
}  // namespace gx

using namespace gx;

struct gx_server {
  int device = 0;
  char *base = nullptr;
  size_t budget = 0;
  std::unique_ptr<HipExec> exec;
  std::unique_ptr<serve::Scheduler<HipExec>> sched;
  ~gx_server() {
    sched.reset();  // stops the thread after the running requests
    exec.reset();
    if (base) {
      (void)hipSetDevice(device);
      (void)hipFree(base);
    }
  }
};

namespace {

int request_sizes(const gx_volume_desc *d, int input_on_host, size_t *plan,
                  size_t *work, size_t *keep) {
  int rc = gx_volume_plan_bytes(d, plan);
  if (rc) return rc;
  *plan = round_up(*plan);
  *work = *plan +
          (input_on_host ? round_up(input_elements(d) * sizeof(float)) : 0);
  *keep = round_up(d->num_rows * d->depth * d->width * sizeof(float));
  return GX_OK;
}

void fill_info(const gx_server *s, const serve::Job &j, gx_request_info *o) {
  std::memset(o, 0, sizeof(*o));
  o->state = (int)j.state;
  o->submit_s = j.submit_s;
  o->start_s = j.start_s;
  o->end_s = j.end_s;
  o->updates = j.updates;
  o->footprint_bytes = j.work_bytes + j.keep_bytes;
  if (j.payload)
    o->device_ms = static_cast<const Request *>(j.payload.get())->device_ms;
  if (j.state == serve::State::kDone && j.has_keep)
    o->volume = s->exec->volume(j);
  std::snprintf(o->error, sizeof(o->error), "%s", j.error.c_str());
}

}  // namespace

extern "C" int gx_server_create(const gx_server_cfg *cfg, gx_server **out) {
  if (cfg == nullptr || out == nullptr)
    return fail_status(GX_EINVAL, "gx_server_create: NULL argument");
  *out = nullptr;
  if (cfg->gpu_index >= 0 && hipSetDevice(cfg->gpu_index) != hipSuccess)
    return fail_status(GX_EDEVICE, "cannot select the requested device");
  auto srv = std::make_unique<gx_server>();
  if (hipGetDevice(&srv->device) != hipSuccess)
    return fail_status(GX_EDEVICE, "no device");
  hipError_t terr;
  twiddles(&terr);  // the shared transform table, built once per device
  if (terr != hipSuccess) return fail_status(GX_EDEVICE, "twiddle table");
  size_t budget = cfg->budget_bytes;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess)
      return fail_status(GX_EDEVICE, "hipMemGetInfo failed");
    budget = fr / 10 * 8;
  }
  budget = budget / kAlign * kAlign;
  if (budget == 0) return fail_status(GX_EINVAL, "zero server budget");
  if (hipMalloc(&srv->base, budget) != hipSuccess) {
    srv->base = nullptr;
    return fail_status(GX_ENOMEM, "cannot allocate the server budget");
  }
  srv->budget = budget;
  const int n = cfg->max_inflight > 0 ? cfg->max_inflight : 4;
  srv->exec = std::make_unique<HipExec>(srv->device, n, srv->base);
  if (srv->exec->init() != hipSuccess)
    return fail_status(GX_EDEVICE, "cannot create the server's streams");
  srv->sched = std::make_unique<serve::Scheduler<HipExec>>(
      srv->exec.get(), budget, cfg->backfill != 0);
  srv->sched->start_thread();
  *out = srv.release();
  return GX_OK;
}

extern "C" int gx_server_destroy(gx_server *s) {
  delete s;
  return GX_OK;
}

extern "C" int gx_server_request_bytes(const gx_volume_desc *desc,
                                       int input_on_host, size_t *bytes) {
  if (desc == nullptr || bytes == nullptr)
    return fail_status(GX_EINVAL, "gx_server_request_bytes: NULL argument");
  size_t plan, work, keep;
  int rc = request_sizes(desc, input_on_host, &plan, &work, &keep);
  if (rc) return rc;
  *bytes = work + keep;
  return GX_OK;
}

extern "C" int gx_server_submit(gx_server *s, const gx_volume_desc *desc,
                                const float *input, int input_on_host,
                                uint64_t *id) {
  if (s == nullptr || desc == nullptr || input == nullptr || id == nullptr)
    return fail_status(GX_EINVAL, "gx_server_submit: NULL argument");
  auto r = std::make_shared<Request>();
  size_t work, keep;
  int rc = request_sizes(desc, input_on_host, &r->plan_bytes, &work, &keep);
  if (rc) return rc;
  r->desc = *desc;
  r->angles.assign(desc->angles, desc->angles + desc->num_angles);
  r->desc.angles = r->angles.data();
  if (desc->weights) {
    r->weights.assign(desc->weights, desc->weights + desc->num_angles);
    r->desc.weights = r->weights.data();
  }
  if (desc->defocus) {
    r->defocus.assign(desc->defocus, desc->defocus + desc->num_defocus);
    r->desc.defocus = r->defocus.data();
  }
  if (desc->centres) {
    r->centres.assign(desc->centres, desc->centres + desc->num_rows);
    r->desc.centres = r->centres.data();
  }
  r->desc.gpu_index = -1;
  r->input = input;
  r->input_on_host = input_on_host != 0;
  r->input_bytes = input_elements(desc) * sizeof(float);
  const double updates = (double)desc->width * (double)desc->num_rows *
                         (double)desc->depth * (double)desc->num_angles;
  std::string err;
  if (s->sched->submit(work, keep, updates, r, id, &err))
    return fail_status(GX_ENOMEM, err.c_str());
  return GX_OK;
}

extern "C" int gx_server_wait(gx_server *s, uint64_t id, double timeout_s,
                              gx_request_info *info) {
  if (s == nullptr) return fail_status(GX_EINVAL, "gx_server_wait: NULL");
  auto j = s->sched->wait(id, timeout_s > 0 ? timeout_s : -1.0);
  if (!j) return fail_status(GX_EINVAL, "unknown request id");
  if (info) fill_info(s, *j, info);
  return GX_OK;
}

extern "C" int gx_server_info(gx_server *s, uint64_t id,
                              gx_request_info *info) {
  if (s == nullptr || info == nullptr)
    return fail_status(GX_EINVAL, "gx_server_info: NULL argument");
  auto j = s->sched->info(id);
  if (!j) return fail_status(GX_EINVAL, "unknown request id");
  fill_info(s, *j, info);
  return GX_OK;
}

extern "C" int gx_server_cancel(gx_server *s, uint64_t id) {
  if (s == nullptr) return fail_status(GX_EINVAL, "gx_server_cancel: NULL");
  if (!s->sched->cancel(id))
    return fail_status(GX_EINVAL, "unknown request id");
  return GX_OK;
}

extern "C" int gx_server_release(gx_server *s, uint64_t id) {
  if (s == nullptr) return fail_status(GX_EINVAL, "gx_server_release: NULL");
  if (!s->sched->release(id))
    return fail_status(GX_EINVAL,
                       "unknown request id, or the request has not finished");
  return GX_OK;
}

extern "C" int gx_server_get_stats(gx_server *s, gx_server_stats *out) {
  if (s == nullptr || out == nullptr)
    return fail_status(GX_EINVAL, "gx_server_get_stats: NULL argument");
  const serve::Stats st = s->sched->stats();
  out->submitted = st.submitted;
  out->completed = st.completed;
  out->failed = st.failed;
  out->cancelled = st.cancelled;
  out->queued = st.queued;
  out->running = st.running;
  out->p50_ms = st.p50_ms;
  out->p99_ms = st.p99_ms;
  out->mean_ms = st.mean_ms;
  out->max_ms = st.max_ms;
  out->updates_per_s = st.updates_per_s;
  out->span_s = st.span_s;
  out->budget_bytes = st.budget_bytes;
  out->used_bytes = st.used_bytes;
  out->peak_bytes = st.peak_bytes;
  return GX_OK;
}
