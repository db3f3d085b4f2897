// Served decode windows: the request loop behind vhsdecode_hip_serve_*.
//
// One worker thread owns every device call.  It admits queued requests in arrival
// order into slots (a HIP stream, device buffers sized for the largest window, and
// one demod context per decode format, so filter tables stay resident), issues the
// whole chain of an admitted window asynchronously on the slot's stream, and polls
// the slots' completion events (a window's field sequence may take several calls: after
// each, the worker reads where it stands and issues the next from there, as the whole run
// does).  Slots are created on demand while the HBM budget
// allows, and reused.  Windows in different slots overlap on the device; a window's
// kernels and inputs do not depend on the other windows in flight, so its result is
// the same alone and in the mix.
//
// A window decodes as the reference's preview windows do
// (src/VHSDecode.Preview/DecodePreviewSegmentProvider.cs:300-360): a fresh field sequence
// from 0.12 s before the requested sample, whose fields then pass the frame rule of
// PreviewFrameAssembler.Accept (src/VHSDecode.Preview/PreviewFrameAssembler.cs:55-95):
// a second field completes the frame of the first field before it, and the frame is kept
// when the later of the two fields' start samples plus one field duration reaches the
// requested sample.  The kept fields, in order, are the window's result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "host_common.hpp"
#include "vhsdecode_hip.h"

namespace {

using Clock = std::chrono::steady_clock;
using vhsd_host::set_error;

constexpr int kBlock = 32768;
constexpr int kRounds = 4;               // sequence rounds per call
constexpr double kPrerollSeconds = 0.12;  // DecodePreviewSegmentProvider.PrerollSeconds

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

bool is_pinned(const void* p) {
  hipPointerAttribute_t attr{};
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();  // plain pageable memory reports an error; clear it
    return false;
  }
  return attr.type == hipMemoryTypeHost;
}

// Owned copy of one decode format.
struct Format {
  std::vector<double> rf_video, video, vlp05, env_sos, chroma_sos, hp, amp_sos, boost_sos, notch;
  std::vector<float> sinc;
  vhsdecode_hip_filters_v1 filters{};
  vhsdecode_hip_subdeemph_v1 sd{};
  int cut = 0, cut_end = 0, stride = 0;
  vhsdecode_hip_field_params_v1 field{};
  vhsdecode_hip_vhs_sync_v1 sync{};
  int64_t preroll = 0;     // samples decoded before the requested first sample
  int64_t half_frame = 0;  // samples of one field (PreviewFrameAssembler._halfFrameSamples)
  int preroll_fields = 0;  // fields the preroll spans, rounded up
  size_t per_field = 0;    // output samples per field
};

struct Source {
  int format;
  const int16_t* samples;
  int64_t length;
  bool loop;
};

struct Request {
  vhsdecode_hip_window_v1 w{};
  Clock::time_point t_submit, t_admit;
  int32_t status = 0;
  int32_t field_count = 0;
  double queued_ms = 0.0, latency_ms = 0.0;
  bool done = false;
  bool pinned_out = false;
};

struct Slot {
  int64_t max_blocks = 0;  // blocks of the largest window at the smallest stride known when created
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  int16_t* rf = nullptr;
  float* planes = nullptr;
  unsigned* flags = nullptr;
  uint16_t* fields = nullptr;  // the sequence's fields, by record
  uint16_t* kept = nullptr;    // the window's fields
  size_t fields_cap = 0;       // uint16 entries of each
  vhsdecode_hip_vhs_field_v1* records = nullptr;
  int32_t* meta = nullptr;     // record count, next begin (int64 at 8), kept count (16), kept records (32)
  int32_t* meta_host = nullptr;  // pinned
  vhsdecode_hip_vhs_field_v1* records_host = nullptr;  // pinned: the last call's records and stop record
  int32_t* total_dev = nullptr;  // records of the window's sequence so far (device)
  uint16_t* stage = nullptr;     // pinned staging for pageable outputs
  size_t stage_cap = 0;
  std::map<int, vhsdecode_hip_demod_context*> ctx;
  std::shared_ptr<Request> req;
  // the window in flight: its field sequence so far
  vhsdecode_hip_demod_context* cur_ctx = nullptr;
  int phase = 0;             // 1: a sequence call in flight, 2: the selection and downloads
  int32_t budget = 0;        // records the window's sequence may take
  int32_t total = 0;         // records decoded so far
  int64_t begin = 0;         // where the last call started (span coordinates)
  int64_t span = 0;          // span samples
  int64_t origin = 0;        // RF sample of span sample 0
  int stalls = 0;
};

}  // namespace

struct vhsdecode_hip_serve {
  vhsdecode_hip_serve_config_v1 cfg{};
  int min_stride = kBlock;  // smallest block stride of the registered formats
  size_t used_bytes = 0;    // device bytes of the slots and their contexts
  size_t slot_bytes() const;
  size_t context_bytes() const;
  int records_cap() const { return cfg.max_fields + 16; }
  size_t meta_bytes() const { return 32 + 4 * (size_t)records_cap(); }
  std::mutex mu;
  std::condition_variable cv;       // worker wake-ups
  std::condition_variable done_cv;  // waiters
  std::vector<std::unique_ptr<Format>> formats;
  std::vector<Source> sources;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<uint64_t, std::shared_ptr<Request>> tickets;
  std::vector<std::unique_ptr<Slot>> slots;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread worker;

  ~vhsdecode_hip_serve();
  void run();
  int32_t make_slot(Slot** out);
  void release_slot(Slot* s);
  int32_t launch(Slot* s, const std::shared_ptr<Request>& r);
  int32_t issue_call(Slot* s, int64_t begin, bool reset);
  int32_t issue_tail(Slot* s);
  int32_t advance(Slot* s, bool* finished);
  void finish(Slot* s, int32_t status);
};

namespace {

#define SV_HIP(call)                                                                                      \
  do {                                                                                                    \
    hipError_t e_ = (call);                                                                               \
    if (e_ != hipSuccess) return set_error(VHSDECODE_HIP_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// PreviewFrameAssembler.Accept over the window's sequence records: field indices of the
// kept frames, in order (one thread; a window holds a few hundred records at most).
__global__ void set_count_kernel(int32_t* p, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void window_select_kernel(const vhsdecode_hip_vhs_field_v1* rec, const int32_t* count, long long origin,
                                     long long target, long long half_frame, int max_out, int32_t* kept_count,
                                     int32_t* kept) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int n = *count;
  int out = 0, first = -1;
  long long first_start = 0;
  for (int k = 0; k < n && out + 2 <= max_out; k++) {
    if (rec[k].status != 0) continue;  // a failed field renders nothing
    const long long start = origin + rec[k].window_start;
    if (rec[k].is_first_field) {
      first = k;
      first_start = start;
      continue;
    }
    if (first < 0) continue;
    const long long mid = (first_start > start ? first_start : start) + half_frame;
    const int a = first;
    first = -1;
    if (mid < target) continue;
    kept[out++] = a;
    kept[out++] = k;
  }
  *kept_count = out;
}

// The kept fields, gathered in order: one workgroup row per kept field.
__global__ void window_gather_kernel(const uint16_t* fields, size_t per_field, const int32_t* kept_count,
                                     const int32_t* kept, uint16_t* out) {
  const int f = blockIdx.y;
  if (f >= *kept_count) return;
  const uint16_t* src = fields + (size_t)kept[f] * per_field;
  uint16_t* dst = out + (size_t)f * per_field;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < per_field; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

// Device bytes of one slot: RF, the four planes, the sequence's and the window's fields
// and the records.
size_t vhsdecode_hip_serve::slot_bytes() const {
  const int64_t blocks = (cfg.max_window_samples - kBlock) / min_stride + 1;
  const int64_t span = blocks * min_stride;
  return (size_t)cfg.max_window_samples * 2 + (size_t)span * 16 + (size_t)records_cap() * 1135 * 313 * 2 * 2 +
         (size_t)records_cap() * sizeof(vhsdecode_hip_vhs_field_v1) + meta_bytes();
}

// Device bytes of one demod context at the slot's size: per-block spectrum stash and
// repair scratch, per-sample pulse edges, filter and sinc tables, and the field
// sequence's per-field lists (sync edges and pulses of each read window, line tables).
size_t vhsdecode_hip_serve::context_bytes() const {
  const int64_t blocks = (cfg.max_window_samples - kBlock) / min_stride + 1;
  const int64_t span = blocks * min_stride;
  return (size_t)blocks * (8 * (2 * 16384 + 64) + 8 * 32768 + 4 * 65536) + (size_t)span * 10 + ((size_t)8 << 20) +
         (size_t)records_cap() * (5 * 65536 + 4 * 4 * 4096 + 24 * 1024 + 32 * 8 * 330);
}

int32_t vhsdecode_hip_serve::make_slot(Slot** out) {
  auto s = std::make_unique<Slot>();
  s->max_blocks = (cfg.max_window_samples - kBlock) / min_stride + 1;
  const int64_t span = s->max_blocks * min_stride;
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&s->done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc((void**)&s->rf, sizeof(int16_t) * (size_t)cfg.max_window_samples);
  if (e == hipSuccess) e = hipMalloc((void**)&s->planes, sizeof(float) * 4 * (size_t)span);
  if (e == hipSuccess) e = hipMalloc((void**)&s->flags, sizeof(unsigned) * (size_t)s->max_blocks);
  if (e == hipSuccess) e = hipMalloc((void**)&s->meta, meta_bytes());
  if (e == hipSuccess) e = hipHostMalloc((void**)&s->meta_host, meta_bytes(), hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&s->records_host, sizeof(vhsdecode_hip_vhs_field_v1) * (size_t)records_cap(),
                      hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc((void**)&s->total_dev, 256);
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->records, sizeof(vhsdecode_hip_vhs_field_v1) * (size_t)records_cap());
  if (e != hipSuccess) {
    release_slot(s.get());
    return set_error(VHSDECODE_HIP_E_OUT_OF_MEMORY, std::string("slot allocation failed: ") + hipGetErrorString(e));
  }
  used_bytes += slot_bytes();
  *out = s.get();
  slots.push_back(std::move(s));
  return VHSDECODE_HIP_OK;
}

void vhsdecode_hip_serve::release_slot(Slot* s) {
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto& kv : s->ctx) vhsdecode_hip_demod_destroy(kv.second);
  s->ctx.clear();
  if (s->rf) (void)hipFree(s->rf);
  if (s->planes) (void)hipFree(s->planes);
  if (s->flags) (void)hipFree(s->flags);
  if (s->fields) (void)hipFree(s->fields);
  if (s->kept) (void)hipFree(s->kept);
  if (s->records) (void)hipFree(s->records);
  if (s->meta) (void)hipFree(s->meta);
  if (s->meta_host) (void)hipHostFree(s->meta_host);
  if (s->records_host) (void)hipHostFree(s->records_host);
  if (s->total_dev) (void)hipFree(s->total_dev);
  if (s->stage) (void)hipHostFree(s->stage);
  if (s->done) (void)hipEventDestroy(s->done);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  *s = Slot{};
}

int32_t vhsdecode_hip_serve::launch(Slot* s, const std::shared_ptr<Request>& r) {
  const vhsdecode_hip_window_v1& w = r->w;
  const Source& src = sources[w.stream_id];
  Format& f = *formats[src.format];
  auto it = s->ctx.find(src.format);
  vhsdecode_hip_demod_context* ctx = nullptr;
  if (it == s->ctx.end()) {
    // keep within the budget: drop this idle slot's contexts of other formats first
    while (!s->ctx.empty() && used_bytes + context_bytes() > cfg.hbm_budget_bytes) {
      vhsdecode_hip_demod_destroy(s->ctx.begin()->second);
      s->ctx.erase(s->ctx.begin());
      used_bytes -= context_bytes();
    }
    vhsdecode_hip_demod_config_v1 dc{};
    dc.struct_size = sizeof(dc);
    dc.device_id = cfg.device_id;
    dc.block_cut = f.cut;
    dc.block_cut_end = f.cut_end;
    dc.max_blocks = (int32_t)s->max_blocks;
    dc.filters = &f.filters;
    int32_t rc = vhsdecode_hip_demod_create(&dc, &ctx);
    if (rc) return rc;
    if (!f.sinc.empty() && (rc = vhsdecode_hip_set_sinc_table(ctx, f.sinc.data(), (int64_t)f.sinc.size())) != 0) {
      vhsdecode_hip_demod_destroy(ctx);
      return rc;
    }
    s->ctx[src.format] = ctx;
    used_bytes += context_bytes();
  } else {
    ctx = it->second;
  }
  const size_t need = (size_t)records_cap() * f.per_field;
  if (need > s->fields_cap) {
    if (s->fields) (void)hipFree(s->fields);
    if (s->kept) (void)hipFree(s->kept);
    s->fields = s->kept = nullptr;
    s->fields_cap = 0;
    SV_HIP(hipMalloc((void**)&s->fields, sizeof(uint16_t) * need));
    SV_HIP(hipMalloc((void**)&s->kept, sizeof(uint16_t) * need));
    s->fields_cap = need;
  }
  // RF from the preroll start: one copy, or one per wrap of a looped recording
  const int64_t start = std::max<int64_t>(0, w.first_sample - f.preroll) & ~(int64_t)1;
  const int64_t count = w.first_sample - start + w.sample_count;
  int64_t pos = start, done = 0;
  while (done < count) {
    const int64_t p = src.loop ? pos % src.length : pos;
    const int64_t n = std::min<int64_t>(count - done, src.length - p);
    SV_HIP(hipMemcpyAsync(s->rf + done, src.samples + p, sizeof(int16_t) * (size_t)n, hipMemcpyHostToDevice,
                          s->stream));
    done += n;
    pos += n;
  }
  const int stride = f.stride;
  const int32_t nb = (int32_t)((count - kBlock) / stride + 1);
  if (nb > s->max_blocks)
    return set_error(VHSDECODE_HIP_E_CAPACITY, "window needs more blocks than the slot was sized for");
  const int64_t n = (int64_t)nb * stride;
  vhsdecode_hip_span_v1 sp{};
  sp.struct_size = sizeof(sp);
  sp.video = s->planes;
  sp.video_lowpass = s->planes + n;
  sp.envelope = s->planes + 2 * n;
  sp.chroma = s->planes + 3 * n;
  sp.block_flags = s->flags;
  int32_t rc = vhsdecode_hip_demod_span_device(ctx, s->rf, 0, nb, &sp, s->stream);
  if (rc) return rc;
  // the sequence: DecodePreviewSegmentProvider's field budget (output fields, preroll and 8)
  s->cur_ctx = ctx;
  s->budget = std::min(records_cap(), w.max_fields + f.preroll_fields + 8);
  s->total = 0;
  s->span = n;
  s->origin = start + f.cut;  // span sample i is RF sample start + cut + i
  s->stalls = 0;
  return issue_call(s, 0, true);
}

// One call of the window's field sequence from `begin`, its records appended after the
// ones decoded so far; then the records and where the sequence stands to the host.
int32_t vhsdecode_hip_serve::issue_call(Slot* s, int64_t begin, bool reset) {
  const Format& f = *formats[sources[s->req->w.stream_id].format];
  const float* vlp = s->planes + s->span;
  const int32_t room = s->budget - s->total;
  unsigned char* meta = reinterpret_cast<unsigned char*>(s->meta);
  s->begin = begin;
  // the sequence's records and fields stay in one place, this call's after the earlier ones
  int32_t rc = vhsd_decode_vhs_fields_from(s->cur_ctx, vlp, s->planes, s->span, begin, reset ? 1 : 0, 0, &f.sync,
                                           &f.field, s->budget, kRounds, s->fields, s->records, s->meta,
                                           reinterpret_cast<int64_t*>(meta + 8), s->stream, s->total);
  if (rc) return rc;
  SV_HIP(hipMemcpyAsync(s->meta_host, s->meta, 16, hipMemcpyDeviceToHost, s->stream));
  SV_HIP(hipMemcpyAsync(s->records_host, s->records + s->total, sizeof(vhsdecode_hip_vhs_field_v1) * (size_t)room,
                        hipMemcpyDeviceToHost, s->stream));
  SV_HIP(hipEventRecord(s->done, s->stream));
  s->phase = 1;
  return VHSDECODE_HIP_OK;
}

// The frame rule over the whole sequence, the kept fields and their download.
int32_t vhsdecode_hip_serve::issue_tail(Slot* s) {
  const std::shared_ptr<Request>& r = s->req;
  const vhsdecode_hip_window_v1& w = r->w;
  const Format& f = *formats[sources[w.stream_id].format];
  unsigned char* meta = reinterpret_cast<unsigned char*>(s->meta);
  hipLaunchKernelGGL(set_count_kernel, dim3(1), dim3(64), 0, s->stream, s->total_dev, s->total);
  hipLaunchKernelGGL(window_select_kernel, dim3(1), dim3(64), 0, s->stream, s->records, s->total_dev,
                     (long long)s->origin, (long long)w.first_sample, (long long)f.half_frame, w.max_fields,
                     reinterpret_cast<int32_t*>(meta + 16), reinterpret_cast<int32_t*>(meta + 32));
  hipLaunchKernelGGL(window_gather_kernel, dim3(64, w.max_fields), dim3(256), 0, s->stream, s->fields, f.per_field,
                     reinterpret_cast<const int32_t*>(meta + 16), reinterpret_cast<const int32_t*>(meta + 32),
                     s->kept);
  SV_HIP(hipGetLastError());
  SV_HIP(hipMemcpyAsync(s->meta_host, s->meta, 32, hipMemcpyDeviceToHost, s->stream));
  const size_t out_n = (size_t)w.max_fields * f.per_field;
  uint16_t* dst = w.fields_out;
  if (!r->pinned_out) {
    if (out_n > s->stage_cap) {
      if (s->stage) (void)hipHostFree(s->stage);
      s->stage = nullptr;
      s->stage_cap = 0;
      SV_HIP(hipHostMalloc((void**)&s->stage, sizeof(uint16_t) * out_n, hipHostMallocDefault));
      s->stage_cap = out_n;
    }
    dst = s->stage;
  }
  SV_HIP(hipMemcpyAsync(dst, s->kept, sizeof(uint16_t) * out_n, hipMemcpyDeviceToHost, s->stream));
  SV_HIP(hipEventRecord(s->done, s->stream));
  s->phase = 2;
  return VHSDECODE_HIP_OK;
}

// A slot whose last issued work has completed: after a sequence call, the next call (the
// sequence stopped at a window the call did not decode: mispredicted, or not read in its
// rounds) or the tail; after the tail, the window is done.
int32_t vhsdecode_hip_serve::advance(Slot* s, bool* finished) {
  *finished = false;
  if (s->phase == 2) {
    *finished = true;
    return VHSDECODE_HIP_OK;
  }
  const int32_t room = s->budget - s->total;
  const int32_t count = std::max(0, std::min(s->meta_host[0] - s->total, room));
  int64_t next = 0;
  std::memcpy(&next, reinterpret_cast<const unsigned char*>(s->meta_host) + 8, sizeof(next));
  s->total += count;
  bool more = s->total < s->budget;
  if (more) {
    const vhsdecode_hip_vhs_field_v1& stop = s->records_host[count];
    if (stop.status == 5) more = false;  // a window holds more pulses than the lists: decode no further
    if (stop.status == 6 && stop.window_start + stop.window_length > s->span) more = false;  // left the span
  }
  s->stalls = (count == 0 && next == s->begin) ? s->stalls + 1 : 0;
  if (s->stalls > 2) more = false;
  if (more) return issue_call(s, next, false);
  return issue_tail(s);
}

void vhsdecode_hip_serve::finish(Slot* s, int32_t status) {
  const std::shared_ptr<Request> r = s->req;
  s->req.reset();
  s->phase = 0;
  s->cur_ctx = nullptr;
  r->status = status;
  if (status == 0) {
    r->field_count = s->meta_host[4];
    if (!r->pinned_out) {
      const Format& f = *formats[sources[r->w.stream_id].format];
      std::memcpy(r->w.fields_out, s->stage, sizeof(uint16_t) * (size_t)r->field_count * f.per_field);
    }
  }
  r->latency_ms = ms_between(r->t_submit, Clock::now());
  r->done = true;
}

void vhsdecode_hip_serve::run() {
  (void)hipSetDevice(cfg.device_id);
  std::unique_lock<std::mutex> lock(mu);
  for (;;) {
    bool finished = false;
    for (auto& s : slots) {
      if (!s->req) continue;
      const hipError_t q = hipEventQuery(s->done);
      if (q == hipErrorNotReady) {
        (void)hipGetLastError();  // not an error: keep it out of the launchers' error checks
        continue;
      }
      if (q != hipSuccess) {
        finish(s.get(), VHSDECODE_HIP_E_HIP);
        finished = true;
        continue;
      }
      bool done_window = false;
      const int32_t rc = advance(s.get(), &done_window);
      if (rc) {
        (void)hipStreamSynchronize(s->stream);
        finish(s.get(), rc);
        finished = true;
      } else if (done_window) {
        finish(s.get(), 0);
        finished = true;
      }
    }
    if (finished) done_cv.notify_all();
    while (!stop && !queue.empty()) {
      const int want = formats.empty() ? -1 : sources[queue.front()->w.stream_id].format;
      Slot* slot = nullptr;
      Slot* any_free = nullptr;
      for (auto& s : slots) {
        if (s->req) continue;
        if (!any_free) any_free = s.get();
        if (s->ctx.count(want)) {
          slot = s.get();
          break;
        }
      }
      if (!slot && (int)slots.size() < cfg.max_slots &&
          used_bytes + slot_bytes() + context_bytes() <= cfg.hbm_budget_bytes) {
        if (make_slot(&slot) != 0) slot = nullptr;
      }
      if (!slot) slot = any_free;
      if (!slot) break;
      std::shared_ptr<Request> r = queue.front();
      queue.pop_front();
      r->t_admit = Clock::now();
      r->queued_ms = ms_between(r->t_submit, r->t_admit);
      slot->req = r;
      const int32_t rc = launch(slot, r);
      if (rc) {
        (void)hipStreamSynchronize(slot->stream);
        finish(slot, rc);
        done_cv.notify_all();
      }
    }
    bool busy = false;
    for (auto& s : slots) busy |= (bool)s->req;
    if (stop && !busy) break;
    if (busy)
      cv.wait_for(lock, std::chrono::microseconds(100));
    else
      cv.wait(lock, [&] { return stop || !queue.empty(); });
  }
  for (auto& r : queue) {
    r->status = VHSDECODE_HIP_E_INVALID_ARGUMENT;
    r->done = true;
  }
  queue.clear();
  done_cv.notify_all();
}

vhsdecode_hip_serve::~vhsdecode_hip_serve() {
  {
    std::lock_guard<std::mutex> lock(mu);
    stop = true;
  }
  cv.notify_all();
  if (worker.joinable()) worker.join();
  (void)hipSetDevice(cfg.device_id);
  for (auto& s : slots) release_slot(s.get());
}

extern "C" {

int32_t vhsdecode_hip_serve_create(const vhsdecode_hip_serve_config_v1* cfg, vhsdecode_hip_serve** out) {
  try {
    if (!cfg || !out) return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "config and server must not be NULL");
    *out = nullptr;
    if (cfg->struct_size != sizeof(vhsdecode_hip_serve_config_v1))
      return set_error(VHSDECODE_HIP_E_STRUCT_SIZE, "serve_config struct_size mismatch");
    if (cfg->max_slots <= 0 || cfg->max_window_samples < kBlock || cfg->max_fields <= 0 ||
        cfg->max_window_samples > INT32_MAX)
      return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid serve limits");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
      return set_error(VHSDECODE_HIP_E_NO_DEVICE, "no HIP device is available");
    if (cfg->device_id < 0 || cfg->device_id >= count)
      return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "device_id out of range");
    auto* s = new (std::nothrow) vhsdecode_hip_serve();
    if (!s) return set_error(VHSDECODE_HIP_E_OUT_OF_MEMORY, "server allocation failed");
    s->cfg = *cfg;
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return set_error(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_serve_add_format(vhsdecode_hip_serve* s, const vhsdecode_hip_serve_format_v1* fm,
                                       int32_t* format_id) {
  try {
    if (!s || !fm || !format_id || !fm->filters || !fm->field || !fm->sync)
      return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid format");
    if (fm->struct_size != sizeof(vhsdecode_hip_serve_format_v1) ||
        fm->filters->struct_size != sizeof(vhsdecode_hip_filters_v1) ||
        fm->field->struct_size != sizeof(vhsdecode_hip_field_params_v1) ||
        fm->sync->struct_size != sizeof(vhsdecode_hip_vhs_sync_v1))
      return set_error(VHSDECODE_HIP_E_STRUCT_SIZE, "format struct_size mismatch");
    const vhsdecode_hip_filters_v1& src = *fm->filters;
    if (src.laserdisc)
      return set_error(VHSDECODE_HIP_E_UNSUPPORTED, "served windows decode tape formats; LaserDisc is block and span only");
    if (src.block_length != kBlock || !src.rf_video || !src.video || !src.video_lowpass05 || !src.envelope_sos)
      return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "format filters are incomplete");
    const int stride = kBlock - fm->block_cut - fm->block_cut_end;
    if (stride < kBlock / 2) return set_error(VHSDECODE_HIP_E_UNSUPPORTED, "block cuts leave less than half a block");
    auto f = std::make_unique<Format>();
    const size_t spec = 2 * (size_t)kBlock;
    f->rf_video.assign(src.rf_video, src.rf_video + spec);
    f->video.assign(src.video, src.video + spec);
    f->vlp05.assign(src.video_lowpass05, src.video_lowpass05 + spec);
    f->env_sos.assign(src.envelope_sos, src.envelope_sos + 6 * (size_t)src.envelope_sections);
    if (src.chroma_sos && src.chroma_sections > 0)
      f->chroma_sos.assign(src.chroma_sos, src.chroma_sos + 6 * (size_t)src.chroma_sections);
    f->filters = src;
    f->filters.rf_video = f->rf_video.data();
    f->filters.video = f->video.data();
    f->filters.video_lowpass05 = f->vlp05.data();
    f->filters.envelope_sos = f->env_sos.data();
    f->filters.chroma_sos = f->chroma_sos.empty() ? nullptr : f->chroma_sos.data();
    f->filters.chroma_sections = f->chroma_sos.empty() ? 0 : src.chroma_sections;
    f->filters.sub_deemphasis = nullptr;
    if (src.sub_deemphasis) {
      const vhsdecode_hip_subdeemph_v1& sd = *src.sub_deemphasis;
      if (sd.struct_size != sizeof(vhsdecode_hip_subdeemph_v1) || !sd.high_pass || !sd.amplitude_sos)
        return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid sub_deemphasis");
      f->hp.assign(sd.high_pass, sd.high_pass + spec);
      f->amp_sos.assign(sd.amplitude_sos, sd.amplitude_sos + 6);
      f->sd = sd;
      f->sd.high_pass = f->hp.data();
      f->sd.amplitude_sos = f->amp_sos.data();
      f->filters.sub_deemphasis = &f->sd;
    }
    f->filters.rf_boost_sos = nullptr;
    if (src.rf_boost_sos) {
      f->boost_sos.assign(src.rf_boost_sos, src.rf_boost_sos + 6);
      f->filters.rf_boost_sos = f->boost_sos.data();
    }
    f->filters.video_notch = nullptr;
    if (src.video_notch) {
      f->notch.assign(src.video_notch, src.video_notch + 6);
      f->filters.video_notch = f->notch.data();
    }
    if (fm->sinc_table && fm->sinc_count > 0) f->sinc.assign(fm->sinc_table, fm->sinc_table + fm->sinc_count);
    f->cut = fm->block_cut;
    f->cut_end = fm->block_cut_end;
    f->stride = stride;
    f->field = *fm->field;
    f->sync = *fm->sync;
    f->per_field = (size_t)f->field.field_lines * f->field.output_line_length;
    const double fs = f->field.sample_rate_mhz * 1e6, field_rate = f->sync.frames_per_second * 2.0;
    if (!(fs > 0.0) || !(field_rate > 0.0))
      return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "format sample rate and frame rate must be positive");
    f->preroll = (int64_t)std::llround(kPrerollSeconds * fs);
    f->half_frame = std::max<int64_t>(1, (int64_t)std::llround(fs / field_rate));
    f->preroll_fields = (int)std::ceil(kPrerollSeconds * field_rate);
    std::lock_guard<std::mutex> lock(s->mu);
    const int min_stride = std::min(s->min_stride, stride);
    if (s->slots.empty()) {
      const int keep = s->min_stride;
      s->min_stride = min_stride;
      if (s->slot_bytes() + s->context_bytes() > s->cfg.hbm_budget_bytes) {
        s->min_stride = keep;
        return set_error(VHSDECODE_HIP_E_CAPACITY, "the HBM budget does not hold one slot of the largest window");
      }
    } else if (min_stride < s->min_stride) {
      return set_error(VHSDECODE_HIP_E_UNSUPPORTED, "formats with a smaller block stride must be added before serving");
    }
    *format_id = (int32_t)s->formats.size();
    s->formats.push_back(std::move(f));
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return set_error(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_serve_add_stream(vhsdecode_hip_serve* s, int32_t format_id, const int16_t* samples,
                                       int64_t length, int32_t loop, int32_t* stream_id) {
  if (!s || !samples || !stream_id || length < kBlock)
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid stream");
  std::lock_guard<std::mutex> lock(s->mu);
  if (format_id < 0 || format_id >= (int32_t)s->formats.size())
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "unknown format");
  *stream_id = (int32_t)s->sources.size();
  s->sources.push_back(Source{format_id, samples, length, loop != 0});
  return VHSDECODE_HIP_OK;
}

int32_t vhsdecode_hip_serve_submit(vhsdecode_hip_serve* s, const vhsdecode_hip_window_v1* w, uint64_t* ticket) {
  try {
    if (!s || !w || !ticket || !w->fields_out) return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid window");
    if (w->struct_size != sizeof(vhsdecode_hip_window_v1))
      return set_error(VHSDECODE_HIP_E_STRUCT_SIZE, "window struct_size mismatch");
    auto r = std::make_shared<Request>();
    r->w = *w;
    r->t_submit = Clock::now();
    r->pinned_out = is_pinned(w->fields_out);
    std::lock_guard<std::mutex> lock(s->mu);
    if (w->stream_id < 0 || w->stream_id >= (int32_t)s->sources.size())
      return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "unknown stream");
    const Source& src = s->sources[w->stream_id];
    const int64_t preroll = std::min<int64_t>(w->first_sample, s->formats[src.format]->preroll);
    if (w->sample_count < kBlock || w->sample_count + preroll + 1 > s->cfg.max_window_samples ||
        w->first_sample < 0 || (w->first_sample & 1) || w->max_fields <= 0 || w->max_fields > s->cfg.max_fields ||
        (!src.loop && w->first_sample + w->sample_count > src.length))
      return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "window (with its preroll) outside the stream or the server's limits");
    *ticket = s->next_ticket++;
    s->tickets[*ticket] = r;
    s->queue.push_back(r);
    s->cv.notify_all();
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return set_error(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_serve_wait(vhsdecode_hip_serve* s, uint64_t ticket, int32_t timeout_ms,
                                 vhsdecode_hip_window_result_v1* res) {
  if (!s || !res) return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "server and result are required");
  if (res->struct_size != sizeof(vhsdecode_hip_window_result_v1))
    return set_error(VHSDECODE_HIP_E_STRUCT_SIZE, "window_result struct_size mismatch");
  std::unique_lock<std::mutex> lock(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "unknown ticket");
  const std::shared_ptr<Request> r = it->second;
  auto ready = [&] { return r->done; };
  if (timeout_ms < 0)
    s->done_cv.wait(lock, ready);
  else if (!s->done_cv.wait_for(lock, std::chrono::milliseconds(timeout_ms), ready))
    return set_error(VHSDECODE_HIP_E_TIMEOUT, "window still in flight");
  s->tickets.erase(ticket);
  res->status = r->status;
  res->field_count = r->field_count;
  res->queued_ms = r->queued_ms;
  res->latency_ms = r->latency_ms;
  return VHSDECODE_HIP_OK;
}

void vhsdecode_hip_serve_destroy(vhsdecode_hip_serve* s) { delete s; }

}  // extern "C"
