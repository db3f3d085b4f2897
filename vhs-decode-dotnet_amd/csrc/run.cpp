// Whole-run decode behind vhsdecode_hip_run: the counterpart of vhsdecode_cuda_fast_run
// (src/VHSDecode.CudaFast.Native/src/vhsdecode_cuda_fast.cpp:238-370) for the vhs field
// sequence of TbcFieldSequenceDecodeEngine (Decode/TbcFieldSequenceDecodeEngine.cs:161-171,
// 804-846).  The RF is read through the caller's callback one span at a time into pinned
// host memory; each span runs the device chain (vhsdecode_hip_demod_span_device,
// vhsdecode_hip_decode_vhs_fields_device and, when asked,
// vhsdecode_hip_detect_dropouts_device) on one stream, and the decoded fields are
// appended in sequence order to <base>.tbc (and their colour-under chroma, when the config
// carries chroma parameters, to <base>_chroma.tbc, through
// vhsdecode_hip_decode_vhs_fields_chroma_device); <base>.tbc.json is written at the end.
//
// Span placement: span k starts at RF sample k * advance, a multiple of lcm(block length,
// block stride) (491520 samples for 1024/1024 cuts).  Plane sample i of a span starting at
// RF sample p is global plane sample p + i, so the blocks two spans share are the same
// blocks and a field's read window (aligned down to a block length, then to a stride) is
// the same window in every span that holds it.  One field sequence runs through all the
// spans: a span's calls decode fields until the next read window leaves the span, and
// the next span (overlapping by the longest read window and a grid step, so it holds that
// window whole) continues from there with the carried level, parity and sync state.
// Because the places are fixed, a prefetch thread reads span k + 1 through the callback
// while the device decodes span k (the callbacks run on that thread for every span but
// the first).  A failed field (the reference's recovery kinds) writes nothing and the
// sequence continues at its recovery offset.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <future>
#include <numeric>
#include <string>
#include <vector>

#include "host_common.hpp"
#include "vhsdecode_hip.h"

namespace {

using vhsd_host::set_error;
constexpr int kBlock = 32768;
constexpr int kRounds = 2;  // field sequence rounds per call (the span loop calls again after a mispredicted window)

#define RUN_HIP(expr)                                                                         \
  do {                                                                                        \
    const hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return set_error(VHSDECODE_HIP_E_HIP, std::string(#expr) + ": " +   \
                                                                    hipGetErrorString(e_));   \
  } while (0)

// Owners that release on every exit path.
struct DeviceBuf {
  void* p = nullptr;
  ~DeviceBuf() {
    if (p) (void)hipFree(p);
  }
};
struct PinnedBuf {
  void* p = nullptr;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
};
struct OwnedFile {
  std::FILE* f = nullptr;
  ~OwnedFile() {
    if (f) std::fclose(f);
  }
};
struct OwnedContext {
  vhsdecode_hip_demod_context* c = nullptr;
  ~OwnedContext() { vhsdecode_hip_demod_destroy(c); }
};
struct OwnedStream {
  hipStream_t s = nullptr;
  ~OwnedStream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

struct FieldRecord {
  bool first = false;
  uint64_t file_loc = 0;  // RF sample where the field's read window starts (TbcDecodedField.StartSample)
  std::vector<int32_t> line, start, end;
};

bool file_exists(const std::string& path) {
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fclose(f);
  return true;
}

void write_ints(std::FILE* f, const std::vector<int32_t>& v) {
  for (size_t i = 0; i < v.size(); i++) std::fprintf(f, "%s%d", i ? ", " : "", v[i]);
}

// The .tbc.json of the run.  Level and geometry keys follow the reference's writer:
// black/white 16-bit IRE from the output scale with a 10 % margin, the output sample rate,
// the field shape, the colour-burst window at the output rate and the active-video span of
// the system (NTSC 134..894, PAL 185..1107).
bool write_json(const std::string& path, const vhsdecode_hip_field_params_v1& fp, double output_rate_hz,
                const vhsdecode_hip_run_config_v1& cfg, const std::vector<FieldRecord>& fields) {
  const std::string tmp = path + ".tmp";
  std::FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return false;
  const bool pal = fp.is_pal != 0;
  const double margin = 0.1;
  const double black = ((0.0 - fp.vsync_ire) * fp.output_scale + fp.output_zero) * (1.0 - margin);
  const double white = ((100.0 - fp.vsync_ire) * fp.output_scale + fp.output_zero) * (1.0 + margin);
  const int burst_start = (int)(cfg.colour_burst_start_us * 1e-6 * output_rate_hz + 0.5);
  const int burst_end = (int)(cfg.colour_burst_end_us * 1e-6 * output_rate_hz + 0.5);
  std::fprintf(f, "{\n  \"videoParameters\": {\n");
  std::fprintf(f, "    \"system\": \"%s\",\n", cfg.system_name ? cfg.system_name : (pal ? "PAL" : "NTSC"));
  std::fprintf(f, "    \"isSubcarrierLocked\": false,\n");
  std::fprintf(f, "    \"isSourcePal\": %s,\n", pal ? "true" : "false");
  std::fprintf(f, "    \"numberOfSequentialFields\": %zu,\n", fields.size());
  std::fprintf(f, "    \"black16bIre\": %.1f,\n", black);
  std::fprintf(f, "    \"white16bIre\": %.1f,\n", white);
  std::fprintf(f, "    \"sampleRate\": %.0f,\n", output_rate_hz);
  std::fprintf(f, "    \"fieldWidth\": %d,\n", fp.output_line_length);
  std::fprintf(f, "    \"fieldHeight\": %d,\n", fp.field_lines);
  std::fprintf(f, "    \"colourBurstStart\": %d,\n", burst_start);
  std::fprintf(f, "    \"colourBurstEnd\": %d,\n", burst_end);
  std::fprintf(f, "    \"activeVideoStart\": %d,\n", pal ? 185 : 134);
  std::fprintf(f, "    \"activeVideoEnd\": %d,\n", pal ? 1107 : 894);
  std::fprintf(f, "    \"tapeFormat\": \"VHS\",\n");
  std::fprintf(f, "    \"isMapped\": false\n  },\n  \"fields\": [\n");
  for (size_t i = 0; i < fields.size(); i++) {
    const FieldRecord& r = fields[i];
    std::fprintf(f, "    {\n      \"isFirstField\": %s,\n", r.first ? "true" : "false");
    std::fprintf(f, "      \"seqNo\": %zu,\n", i + 1);
    std::fprintf(f, "      \"fileLoc\": %llu,\n", (unsigned long long)r.file_loc);
    std::fprintf(f, "      \"dropOuts\": {\n        \"fieldLine\": [");
    write_ints(f, r.line);
    std::fprintf(f, "],\n        \"startx\": [");
    write_ints(f, r.start);
    std::fprintf(f, "],\n        \"endx\": [");
    write_ints(f, r.end);
    std::fprintf(f, "]\n      }\n    }%s\n", i + 1 < fields.size() ? "," : "");
  }
  std::fprintf(f, "  ]\n}\n");
  if (std::fclose(f) != 0) return false;
  return std::rename(tmp.c_str(), path.c_str()) == 0;
}

int32_t run_impl(const vhsdecode_hip_run_config_v1& cfg, vhsdecode_hip_run_result_v1& res) {
  const vhsdecode_hip_serve_format_v1& fm = *cfg.format;
  const vhsdecode_hip_field_params_v1& fp = *fm.field;
  const vhsdecode_hip_vhs_sync_v1& sy = *fm.sync;
  const int stride = kBlock - fm.block_cut - fm.block_cut_end;
  if (fm.block_cut < 0 || fm.block_cut_end < 0 || stride <= 0)
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "block cuts leave no stride");
  if (sy.block_stride != stride || sy.block_cut != fm.block_cut || sy.block_length != kBlock)
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "sync block geometry differs from the format's block cuts");
  if (!(sy.line_period_us > 0.0) || fp.field_lines <= 0 || fp.output_line_length <= 0 || sy.read_samples <= 0 ||
      !(fp.sample_rate_mhz > 0.0))
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "field geometry must be positive");
  // Spans sit on multiples of the grid lcm(block length, stride), so a field's read
  // window (aligned down to a block, then to a stride) is the same window in every span
  // that holds it; consecutive spans overlap by the longest read window plus a grid step,
  // so the field the sequence reaches at the end of one span lies whole in the next.
  const int64_t grid = (int64_t)kBlock / std::gcd((int64_t)kBlock, (int64_t)stride) * stride;
  const double line_len = sy.line_period_us * fp.sample_rate_mhz;
  const double field_len = line_len * (fp.field_lines - 0.5);
  const int64_t window_max = sy.read_samples + 2 * (int64_t)stride;
  const int64_t overlap = window_max + grid;
  const int64_t min_span = overlap + grid + stride + kBlock;
  const int64_t span_req =
      cfg.span_samples > 0 ? cfg.span_samples : std::max(min_span, (int64_t)std::ceil(24.0 * field_len) + kBlock);
  if (span_req < min_span)
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT,
                     "span_samples must hold the span overlap (the longest read window and a grid step) and one "
                     "step of the span grid");
  if (span_req > (int64_t)INT32_MAX)
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "span_samples above 2^31 - 1");
  const int32_t nb_max = (int32_t)((span_req - kBlock) / stride + 1);
  const int64_t read_max = (int64_t)(nb_max - 1) * stride + kBlock;
  const int64_t plane_max = (int64_t)nb_max * stride;
  // records per call: every field of a span, its failures and the stop record
  const int32_t max_fields = (int32_t)std::ceil((double)plane_max / field_len) + 8;
  const int64_t advance = (plane_max - overlap) / grid * grid;
  const size_t per_field = (size_t)fp.field_lines * fp.output_line_length;
  const int32_t cap = cfg.dropouts ? cfg.dropouts->capacity : 0;
  if (cfg.dropouts && cap <= 0) return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "dropout capacity must be positive");

  const std::string base(cfg.output_base_utf8);
  const std::string tbc_path = base + ".tbc", json_path = base + ".tbc.json";
  const std::string chroma_path = base + "_chroma.tbc";
  const bool with_chroma = cfg.chroma != nullptr;
  if (!cfg.overwrite && (file_exists(tbc_path) || file_exists(json_path) || (with_chroma && file_exists(chroma_path))))
    return set_error(VHSDECODE_HIP_E_IO, "an output file exists and overwrite is off: " + tbc_path);

  // the cancel callback runs on this thread; the read wrapper latches a cancel it sees
  std::atomic<bool> cancelled{false};
  auto poll_cancel = [&]() -> bool {
    if (!cancelled.load() && cfg.cancel_callback && cfg.cancel_callback(cfg.user_data)) cancelled.store(true);
    return cancelled.load();
  };
  if (poll_cancel()) return set_error(VHSDECODE_HIP_E_CANCELLED, "the run was cancelled before it started");
  RUN_HIP(hipSetDevice(cfg.device_id));
  OwnedContext ctx;
  {
    vhsdecode_hip_demod_config_v1 dc{};
    dc.struct_size = sizeof(dc);
    dc.device_id = cfg.device_id;
    dc.block_cut = fm.block_cut;
    dc.block_cut_end = fm.block_cut_end;
    dc.max_blocks = nb_max;
    dc.filters = fm.filters;
    const int32_t rc = vhsdecode_hip_demod_create(&dc, &ctx.c);
    if (rc) return rc;
    if (fm.sinc_table) {
      const int32_t rs = vhsdecode_hip_set_sinc_table(ctx.c, fm.sinc_table, fm.sinc_count);
      if (rs) return rs;
    }
    if (with_chroma) {
      const int32_t rs = vhsdecode_hip_set_vhs_chroma(ctx.c, cfg.chroma);
      if (rs) return rs;
    }
  }
  OwnedStream st;
  RUN_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
  // device: RF, four planes, block flags, fields, records, count and next begin, dropout rows
  DeviceBuf d_rf, d_planes, d_flags, d_fields, d_cfields, d_records, d_meta, d_drop;
  RUN_HIP(hipMalloc(&d_rf.p, sizeof(int16_t) * (size_t)read_max));
  RUN_HIP(hipMalloc(&d_planes.p, sizeof(float) * 4 * (size_t)plane_max));
  RUN_HIP(hipMalloc(&d_flags.p, sizeof(uint32_t) * (size_t)nb_max));
  RUN_HIP(hipMalloc(&d_fields.p, sizeof(uint16_t) * per_field * (size_t)max_fields));
  if (with_chroma) RUN_HIP(hipMalloc(&d_cfields.p, sizeof(uint16_t) * per_field * (size_t)max_fields));
  const size_t rec_bytes = sizeof(vhsdecode_hip_vhs_field_v1) * (size_t)max_fields;
  RUN_HIP(hipMalloc(&d_records.p, rec_bytes));
  const size_t meta_bytes = 16;  // record count (int32, padded to 8), next begin (int64)
  RUN_HIP(hipMalloc(&d_meta.p, meta_bytes));
  // dropouts: row counts, then field line / start x / end x, [max_fields][cap] each
  const size_t drop_bytes = sizeof(int32_t) * (size_t)max_fields * (1 + 3 * (size_t)cap);
  if (cfg.dropouts) RUN_HIP(hipMalloc(&d_drop.p, drop_bytes));
  PinnedBuf h_rf[2], h_fields[2], h_cfields[2], h_records, h_meta, h_drop;
  for (PinnedBuf& b : h_rf) RUN_HIP(hipHostMalloc(&b.p, sizeof(int16_t) * (size_t)read_max, hipHostMallocDefault));
  for (PinnedBuf& b : h_fields)
    RUN_HIP(hipHostMalloc(&b.p, sizeof(uint16_t) * per_field * (size_t)max_fields, hipHostMallocDefault));
  if (with_chroma)
    for (PinnedBuf& b : h_cfields)
      RUN_HIP(hipHostMalloc(&b.p, sizeof(uint16_t) * per_field * (size_t)max_fields, hipHostMallocDefault));
  RUN_HIP(hipHostMalloc(&h_records.p, rec_bytes, hipHostMallocDefault));
  RUN_HIP(hipHostMalloc(&h_meta.p, meta_bytes, hipHostMallocDefault));
  if (cfg.dropouts) RUN_HIP(hipHostMalloc(&h_drop.p, drop_bytes, hipHostMallocDefault));
  std::vector<float> staging[2];
  if (cfg.input_sample_format == VHSDECODE_HIP_INPUT_FLOAT32)
    for (std::vector<float>& v : staging) v.resize((size_t)read_max);

  OwnedFile tbc, ctbc;
  tbc.f = std::fopen(tbc_path.c_str(), "wb");
  if (!tbc.f) return set_error(VHSDECODE_HIP_E_IO, "cannot create " + tbc_path);
  if (with_chroma) {
    ctbc.f = std::fopen(chroma_path.c_str(), "wb");
    if (!ctbc.f) return set_error(VHSDECODE_HIP_E_IO, "cannot create " + chroma_path);
  }

  const int32_t* h_count = static_cast<const int32_t*>(h_meta.p);
  const int64_t* h_next = reinterpret_cast<const int64_t*>(static_cast<unsigned char*>(h_meta.p) + 8);
  int32_t* d_count = static_cast<int32_t*>(d_meta.p);
  int64_t* d_next = reinterpret_cast<int64_t*>(static_cast<unsigned char*>(d_meta.p) + 8);
  const vhsdecode_hip_vhs_field_v1* recs = static_cast<const vhsdecode_hip_vhs_field_v1*>(h_records.p);
  int32_t* dd = static_cast<int32_t*>(d_drop.p);
  const int32_t* hd = static_cast<const int32_t*>(h_drop.p);
  const size_t rows = (size_t)max_fields * (size_t)cap;

  std::vector<FieldRecord> written;
  const uint64_t total = cfg.total_samples;
  uint32_t spans = 0, drop_rows = 0;
  bool done = false;

  // One span's RF into a pinned buffer (FP32 input rounded to int16); runs on the prefetch
  // thread for every span but the first.  `last`: no whole block of input after this span.
  struct SpanRead {
    int64_t pos = 0;
    int32_t nb = 0;
    int64_t n_read = 0;
    bool last = true;
    bool ok = false;
  };
  // The read callback, latching a cancel it reports through a short read.
  auto read_input = [&](void* dst, uint64_t p, size_t n) -> size_t {
    const size_t got = cfg.read_callback(cfg.user_data, dst, p, n);
    if (got < n && cfg.cancel_callback && cfg.cancel_callback(cfg.user_data)) cancelled.store(true);
    return got;
  };
  auto read_span = [&](int64_t p, int slot) -> SpanRead {
    SpanRead r;
    r.pos = p;
    if ((uint64_t)p + kBlock > total) return r;
    const int64_t avail = (int64_t)std::min<uint64_t>(total - (uint64_t)p, (uint64_t)read_max);
    int32_t nb = (int32_t)((avail - kBlock) / stride + 1);
    int64_t n = (int64_t)(nb - 1) * stride + kBlock;
    bool last = (uint64_t)p + (uint64_t)n + stride > total;
    int16_t* dst = static_cast<int16_t*>(h_rf[slot].p);
    size_t got;
    if (cfg.input_sample_format == VHSDECODE_HIP_INPUT_INT16) {
      got = read_input(dst, (uint64_t)p, (size_t)n);
    } else {
      std::vector<float>& stage = staging[slot];
      got = read_input(stage.data(), (uint64_t)p, (size_t)n);
      const size_t m = std::min(got, (size_t)n);
      for (size_t q = 0; q < m; q++) {
        const float v = stage[q];
        dst[q] = (int16_t)(std::isfinite(v) ? std::min(32767.0f, std::max(-32768.0f, std::nearbyint(v))) : 0.0f);
      }
    }
    if (got < (size_t)n) {  // a short read is the end of the input (or a cancel)
      if (got < (size_t)kBlock) return r;
      nb = (int32_t)(((int64_t)got - kBlock) / stride + 1);
      n = (int64_t)(nb - 1) * stride + kBlock;
      last = true;
    }
    r.nb = nb;
    r.n_read = n;
    r.last = last;
    r.ok = true;
    return r;
  };

  // .tbc writes on a writer thread, one job in flight, so the file keeps the field order;
  // a job writes the listed fields of a staging slot.
  std::future<bool> write_job;
  auto write_fields = [&](int slot, std::vector<int32_t> which) -> bool {
    const uint16_t* base_p = static_cast<const uint16_t*>(h_fields[slot].p);
    const uint16_t* cbase_p = static_cast<const uint16_t*>(h_cfields[slot].p);
    for (const int32_t f : which) {
      if (std::fwrite(base_p + (size_t)f * per_field, sizeof(uint16_t), per_field, tbc.f) != per_field) return false;
      if (with_chroma &&
          std::fwrite(cbase_p + (size_t)f * per_field, sizeof(uint16_t), per_field, ctbc.f) != per_field)
        return false;
    }
    return true;
  };
  auto finish_writes = [&]() -> bool { return !write_job.valid() || write_job.get(); };

  const auto t0 = std::chrono::steady_clock::now();
  SpanRead cur = read_span(0, 0);
  int64_t next_global = 0;  // where the field sequence continues (global plane coordinates)
  int64_t prev_pos = 0;
  bool fresh = true;
  int call = 0;
  for (int k = 0; cur.ok && !done; k++) {
    if (poll_cancel()) break;
    // the next span's read overlaps this span's device work (the future's destructor
    // waits for it on every exit path, before the buffers go)
    std::future<SpanRead> next;
    if (!cur.last) next = std::async(std::launch::async, read_span, cur.pos + advance, (k + 1) & 1);
    const int64_t pos = cur.pos;
    const int32_t nb = cur.nb;
    const int64_t plane = (int64_t)nb * stride;
    RUN_HIP(hipMemcpyAsync(d_rf.p, h_rf[k & 1].p, sizeof(int16_t) * (size_t)cur.n_read, hipMemcpyHostToDevice,
                           st.s));
    float* planes = static_cast<float*>(d_planes.p);
    vhsdecode_hip_span_v1 sp{};
    sp.struct_size = sizeof(sp);
    sp.video = planes;
    sp.video_lowpass = planes + plane;
    sp.envelope = planes + 2 * plane;
    sp.chroma = planes + 3 * plane;
    sp.block_flags = static_cast<uint32_t*>(d_flags.p);
    int32_t rc = vhsdecode_hip_demod_span_device(ctx.c, static_cast<const int16_t*>(d_rf.p), 0, nb, &sp, st.s);
    if (rc) return rc;
    spans++;
    if (next_global < pos) return set_error(VHSDECODE_HIP_E_INTERNAL, "the field sequence fell behind the span");
    int64_t begin = next_global - pos;
    int64_t shift = pos - prev_pos;
    prev_pos = pos;
    // calls on this span until the sequence's next read window leaves it
    for (int stall = 0;;) {
      if (with_chroma)
        rc = vhsdecode_hip_decode_vhs_fields_chroma_device(
            ctx.c, sp.video_lowpass, sp.video, sp.chroma, plane, begin, fresh ? 1 : 0, shift, &sy, &fp, max_fields,
            kRounds, static_cast<uint16_t*>(d_fields.p), static_cast<uint16_t*>(d_cfields.p), 0,
            static_cast<vhsdecode_hip_vhs_field_v1*>(d_records.p), d_count, d_next, st.s);
      else
        rc = vhsdecode_hip_decode_vhs_fields_device(ctx.c, sp.video_lowpass, sp.video, plane, begin, fresh ? 1 : 0,
                                                    shift, &sy, &fp, max_fields, kRounds,
                                                    static_cast<uint16_t*>(d_fields.p),
                                                    static_cast<vhsdecode_hip_vhs_field_v1*>(d_records.p), d_count,
                                                    d_next, st.s);
      if (rc) return rc;
      fresh = false;
      shift = 0;
      call++;
      if (cfg.dropouts) {
        const double* span_locs = nullptr;
        int32_t loc_stride = 0;
        rc = vhsdecode_hip_vhs_field_lines_device(ctx.c, 3, &span_locs, &loc_stride);
        if (rc) return rc;
        rc = vhsdecode_hip_detect_dropouts_device(
            ctx.c, sp.envelope, plane, span_locs, loc_stride,
            static_cast<const vhsdecode_hip_vhs_field_v1*>(d_records.p), d_count, max_fields, cfg.dropouts, dd,
            dd + max_fields, dd + max_fields + rows, dd + max_fields + 2 * rows, nullptr, nullptr, nullptr, st.s);
        if (rc) return rc;
        RUN_HIP(hipMemcpyAsync(h_drop.p, d_drop.p, drop_bytes, hipMemcpyDeviceToHost, st.s));
      }
      RUN_HIP(hipMemcpyAsync(h_meta.p, d_meta.p, meta_bytes, hipMemcpyDeviceToHost, st.s));
      RUN_HIP(hipMemcpyAsync(h_records.p, d_records.p, rec_bytes, hipMemcpyDeviceToHost, st.s));
      RUN_HIP(hipStreamSynchronize(st.s));
      const int32_t count = std::min(*h_count, max_fields);
      // staging slot `call & 1` was last written two calls ago; that job has finished
      const int slot = call & 1;
      if (count > 0) {
        RUN_HIP(hipMemcpyAsync(h_fields[slot].p, d_fields.p, sizeof(uint16_t) * per_field * (size_t)count,
                               hipMemcpyDeviceToHost, st.s));
        if (with_chroma)
          RUN_HIP(hipMemcpyAsync(h_cfields[slot].p, d_cfields.p, sizeof(uint16_t) * per_field * (size_t)count,
                                 hipMemcpyDeviceToHost, st.s));
      }
      RUN_HIP(hipStreamSynchronize(st.s));
      std::vector<int32_t> to_write;
      for (int32_t f = 0; f < count && !done; f++) {
        if (recs[f].status != 0) continue;  // a failed field: the reference writes nothing for it
        FieldRecord r;
        r.first = recs[f].is_first_field != 0;
        r.file_loc = (uint64_t)(pos + recs[f].window_start);
        if (cfg.dropouts) {
          const int32_t n = hd[f];
          if (n < 0 || n > cap)
            return set_error(VHSDECODE_HIP_E_CAPACITY, "a field has more dropout rows than the dropout capacity");
          const int32_t* fl = hd + max_fields + (size_t)f * cap;
          r.line.assign(fl, fl + n);
          r.start.assign(fl + rows, fl + rows + n);
          r.end.assign(fl + 2 * rows, fl + 2 * rows + n);
          drop_rows += (uint32_t)n;
        }
        to_write.push_back(f);
        written.push_back(std::move(r));
        if (cfg.maximum_output_fields && written.size() >= cfg.maximum_output_fields) done = true;
      }
      if (!finish_writes()) return set_error(VHSDECODE_HIP_E_IO, "cannot write " + tbc_path);
      if (!to_write.empty()) write_job = std::async(std::launch::async, write_fields, slot, std::move(to_write));
      const int64_t next_begin = *h_next;
      next_global = pos + next_begin;
      if (done) break;
      if (count >= max_fields) {
        begin = next_begin;
        continue;
      }
      const vhsdecode_hip_vhs_field_v1& stop = recs[count];
      if (stop.status == 5)
        return set_error(VHSDECODE_HIP_E_CAPACITY, "a read window holds more sync edges or pulses than the device lists");
      const bool fits = stop.window_start + stop.window_length <= plane;
      if (stop.status == 6 && !fits) break;  // the next window is the next span's
      if (count == 0 && next_begin == begin && ++stall > 2)
        return set_error(VHSDECODE_HIP_E_INTERNAL, "the field sequence made no progress");
      begin = next_begin;
    }
    if (cur.last || done) break;
    if (poll_cancel()) break;
    cur = next.get();
  }
  if (!finish_writes()) return set_error(VHSDECODE_HIP_E_IO, "cannot write " + tbc_path);
  if (poll_cancel()) return set_error(VHSDECODE_HIP_E_CANCELLED, "the run was cancelled");
  const auto t1 = std::chrono::steady_clock::now();
  if (std::fflush(tbc.f) != 0) return set_error(VHSDECODE_HIP_E_IO, "cannot write " + tbc_path);
  if (with_chroma && std::fflush(ctbc.f) != 0) return set_error(VHSDECODE_HIP_E_IO, "cannot write " + chroma_path);
  const double out_rate = sy.output_rate_mhz * 1e6;
  if (!write_json(json_path, fp, out_rate, cfg, written)) return set_error(VHSDECODE_HIP_E_IO, "cannot write " + json_path);
  res.fields_written = (uint32_t)written.size();
  res.output_line_length = (uint32_t)fp.output_line_length;
  res.output_field_lines = (uint32_t)fp.field_lines;
  res.elapsed_seconds = std::chrono::duration<double>(t1 - t0).count();
  res.spans = spans;
  res.dropout_rows = drop_rows;
  return VHSDECODE_HIP_OK;
}

}  // namespace

extern "C" int32_t vhsdecode_hip_run(const vhsdecode_hip_run_config_v1* config, vhsdecode_hip_run_result_v1* result) {
  if (!config || !result) return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "config and result are required");
  if (config->struct_size != sizeof(vhsdecode_hip_run_config_v1) ||
      result->struct_size != sizeof(vhsdecode_hip_run_result_v1))
    return set_error(VHSDECODE_HIP_E_STRUCT_SIZE, "run config or result struct_size mismatch");
  const vhsdecode_hip_serve_format_v1* fm = config->format;
  if (!fm || fm->struct_size != sizeof(vhsdecode_hip_serve_format_v1) || !fm->filters || !fm->field || !fm->sync)
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "run format with filters, field and sync params required");
  if (fm->field->struct_size != sizeof(vhsdecode_hip_field_params_v1) ||
      fm->sync->struct_size != sizeof(vhsdecode_hip_vhs_sync_v1) ||
      (config->dropouts && config->dropouts->struct_size != sizeof(vhsdecode_hip_dropout_params_v1)))
    return set_error(VHSDECODE_HIP_E_STRUCT_SIZE, "field, sync or dropout params struct_size mismatch");
  if (fm->filters->laserdisc)
    return set_error(VHSDECODE_HIP_E_UNSUPPORTED, "the whole run decodes tape formats; LaserDisc is block and span only");
  if (!config->read_callback || !config->output_base_utf8 || !config->output_base_utf8[0] ||
      config->total_samples == 0 || config->input_sample_format > VHSDECODE_HIP_INPUT_INT16)
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT,
                     "the run needs a read callback, an output base, samples and a known input format");
  *result = vhsdecode_hip_run_result_v1{};
  result->struct_size = sizeof(vhsdecode_hip_run_result_v1);
  try {
    return run_impl(*config, *result);
  } catch (const std::exception& e) {
    return set_error(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}
