// Whole-run decode behind vhsdecode_hip_run: the counterpart of vhsdecode_cuda_fast_run
// (src/VHSDecode.CudaFast.Native/src/vhsdecode_cuda_fast.cpp:238-370).  The RF is read
// through the caller's callback one span at a time into pinned host memory, each span runs
// the device chain (vhsdecode_hip_demod_span_device, vhsdecode_hip_decode_span_fields_device
// and, when asked, vhsdecode_hip_detect_dropouts_device) on one stream, and the fields are
// appended in order to <base>.tbc; <base>.tbc.json is written at the end.
//
// Span placement: span k starts at RF sample k * advance, a multiple of lcm(block length,
// block stride) (491520 samples for 1024/1024 cuts), and consecutive spans overlap by a
// field and a half plus the read-window lead (plus a dropout read window when detection is
// on).  Plane sample i of a span starting at RF sample p is RF sample p + cut + i, so the
// blocks two spans share are the same blocks, and each field's read-window origin (aligned
// down to a block length, then to a stride, in span coordinates) lands where it lands in
// one span from the start of the input: the TBC's FP32 source coordinates, which count from
// that origin, are then the same, and so is every output sample.  (A format whose lcm
// exceeds a field steps by the stride instead.)  A field whose line 0 lies within half a
// field of the last one written is the same field found again and is skipped.  Because the
// places are fixed, a prefetch thread reads span k + 1 through the callback while the
// device decodes span k (the callbacks run on that thread for every span but the first).
#include <hip/hip_runtime.h>

#include <algorithm>
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
  uint64_t file_loc = 0;  // RF sample of the field's line 0
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
// the field shape and the active-video span of the system (NTSC 134..894, PAL 185..1107).
bool write_json(const std::string& path, const vhsdecode_hip_field_params_v1& fp, double output_rate_hz,
                const std::vector<FieldRecord>& fields) {
  const std::string tmp = path + ".tmp";
  std::FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return false;
  const bool pal = fp.is_pal != 0;
  const double margin = 0.1;
  const double black = ((0.0 - fp.vsync_ire) * fp.output_scale + fp.output_zero) * (1.0 - margin);
  const double white = ((100.0 - fp.vsync_ire) * fp.output_scale + fp.output_zero) * (1.0 + margin);
  std::fprintf(f, "{\n  \"videoParameters\": {\n");
  std::fprintf(f, "    \"system\": \"%s\",\n", pal ? "PAL" : "NTSC");
  std::fprintf(f, "    \"isSubcarrierLocked\": false,\n");
  std::fprintf(f, "    \"isSourcePal\": %s,\n", pal ? "true" : "false");
  std::fprintf(f, "    \"numberOfSequentialFields\": %zu,\n", fields.size());
  std::fprintf(f, "    \"black16bIre\": %.1f,\n", black);
  std::fprintf(f, "    \"white16bIre\": %.1f,\n", white);
  std::fprintf(f, "    \"sampleRate\": %.0f,\n", output_rate_hz);
  std::fprintf(f, "    \"fieldWidth\": %d,\n", fp.output_line_length);
  std::fprintf(f, "    \"fieldHeight\": %d,\n", fp.field_lines);
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

// End of a field's dropout read window in span coordinates before the span clamps it: the
// window the dropout kernels place (DecodeReadWindowPlanner.Resolve,
// Decode/DecodeReadWindowPlanner.cs:37-63), read_samples from the block-aligned start,
// rounded up past a whole stride.
int64_t dropout_window_end(const vhsdecode_hip_dropout_params_v1& d, double line0) {
  const double intended = std::max(0.0, line0 - d.lead_lines * d.line_length);
  const int64_t read = std::max<int64_t>(0, (int64_t)intended - d.block_cut);
  const int64_t aligned = read / d.block_length * d.block_length;
  return ((aligned + d.read_samples) / d.block_stride + 1) * d.block_stride;
}

int32_t run_impl(const vhsdecode_hip_run_config_v1& cfg, vhsdecode_hip_run_result_v1& res) {
  const vhsdecode_hip_serve_format_v1& fm = *cfg.format;
  const vhsdecode_hip_field_params_v1& fp = *fm.field;
  const vhsdecode_hip_locator_v1& lz = *fm.locator;
  const int stride = kBlock - fm.block_cut - fm.block_cut_end;
  if (fm.block_cut < 0 || fm.block_cut_end < 0 || stride <= 0)
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "block cuts leave no stride");
  if (!(lz.line_length > 0.0) || fp.field_lines <= 0 || fp.output_line_length <= 0)
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "field geometry must be positive");
  // one field of input, and the span: at least three fields, so each span holds a whole
  // field beyond the half field of overlap with the previous one
  const double field_len = lz.line_length * (fp.field_lines - 0.5);
  const int64_t align = lz.block_length > 0 ? lz.block_length : kBlock;
  const int64_t grid = align / std::gcd(align, (int64_t)stride) * stride;
  const int64_t step = grid <= (int64_t)field_len ? grid : stride;
  // Spans sit at fixed places, span k at k * advance, so the next span's RF can be read
  // while the device decodes the current one.  Consecutive spans overlap by a field and a
  // half plus the read-window lead (and, with dropout detection, a whole read window: the
  // envelope mean runs over it) plus a step of the grid, so every field lies whole, with
  // its window, inside one span.
  const double lead = lz.lead_lines * lz.line_length;
  const int64_t window = cfg.dropouts ? cfg.dropouts->read_samples + 2 * (int64_t)stride : 0;
  const int64_t overlap = (int64_t)std::ceil(1.5 * field_len + lead) + fm.block_cut + window + step;
  const int64_t min_span = overlap + step + stride + kBlock;
  const int64_t span_req =
      cfg.span_samples > 0 ? cfg.span_samples : std::max(min_span, (int64_t)std::ceil(24.0 * field_len) + kBlock);
  if (span_req < min_span)
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT,
                     "span_samples must hold the span overlap (a field and a half, the lead and any dropout window) "
                     "and one step of the span grid");
  if (span_req > (int64_t)INT32_MAX)
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "span_samples above 2^31 - 1");
  const int32_t nb_max = (int32_t)((span_req - kBlock) / stride + 1);
  const int64_t read_max = (int64_t)(nb_max - 1) * stride + kBlock;
  const int64_t plane_max = (int64_t)nb_max * stride;
  const int32_t max_fields = (int32_t)std::ceil((double)plane_max / field_len) + 2;
  const int64_t advance = (plane_max - overlap) / step * step;
  const size_t per_field = (size_t)fp.field_lines * fp.output_line_length;
  const int32_t cap = cfg.dropouts ? cfg.dropouts->capacity : 0;
  if (cfg.dropouts && (cap <= 0 || cfg.dropouts->read_samples <= 0 || cfg.dropouts->block_length <= 0 ||
                       cfg.dropouts->block_stride <= 0))
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "dropout capacity, read window and block geometry must be positive");

  const std::string base(cfg.output_base_utf8);
  const std::string tbc_path = base + ".tbc", json_path = base + ".tbc.json";
  if (!cfg.overwrite && (file_exists(tbc_path) || file_exists(json_path)))
    return set_error(VHSDECODE_HIP_E_IO, "an output file exists and overwrite is off: " + tbc_path);

  if (cfg.cancel_callback && cfg.cancel_callback(cfg.user_data))
    return set_error(VHSDECODE_HIP_E_CANCELLED, "the run was cancelled before it started");
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
  }
  OwnedStream st;
  RUN_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
  // device: RF, four planes, block flags, fields, field metadata, dropout rows
  DeviceBuf d_rf, d_planes, d_flags, d_fields, d_meta, d_drop;
  RUN_HIP(hipMalloc(&d_rf.p, sizeof(int16_t) * (size_t)read_max));
  RUN_HIP(hipMalloc(&d_planes.p, sizeof(float) * 4 * (size_t)plane_max));
  RUN_HIP(hipMalloc(&d_flags.p, sizeof(uint32_t) * (size_t)nb_max));
  RUN_HIP(hipMalloc(&d_fields.p, sizeof(uint16_t) * per_field * (size_t)max_fields));
  // meta: count (int32, padded to 8), line 0 (double[max_fields]), first (int32[max_fields])
  const size_t meta_bytes = 8 + (size_t)max_fields * 12;
  RUN_HIP(hipMalloc(&d_meta.p, meta_bytes));
  // dropouts: row counts, then field line / start x / end x, [max_fields][cap] each
  const size_t drop_bytes = sizeof(int32_t) * (size_t)max_fields * (1 + 3 * (size_t)cap);
  if (cfg.dropouts) RUN_HIP(hipMalloc(&d_drop.p, drop_bytes));
  PinnedBuf h_rf[2], h_fields[2], h_meta, h_drop;
  for (PinnedBuf& b : h_rf) RUN_HIP(hipHostMalloc(&b.p, sizeof(int16_t) * (size_t)read_max, hipHostMallocDefault));
  for (PinnedBuf& b : h_fields)
    RUN_HIP(hipHostMalloc(&b.p, sizeof(uint16_t) * per_field * (size_t)max_fields, hipHostMallocDefault));
  RUN_HIP(hipHostMalloc(&h_meta.p, meta_bytes, hipHostMallocDefault));
  if (cfg.dropouts) RUN_HIP(hipHostMalloc(&h_drop.p, drop_bytes, hipHostMallocDefault));
  std::vector<float> staging[2];
  if (cfg.input_sample_format == VHSDECODE_HIP_INPUT_FLOAT32)
    for (std::vector<float>& v : staging) v.resize((size_t)read_max);

  OwnedFile tbc;
  tbc.f = std::fopen(tbc_path.c_str(), "wb");
  if (!tbc.f) return set_error(VHSDECODE_HIP_E_IO, "cannot create " + tbc_path);

  unsigned char* meta = static_cast<unsigned char*>(h_meta.p);
  const int32_t* h_count = reinterpret_cast<const int32_t*>(meta);
  const double* h_line0 = reinterpret_cast<const double*>(meta + 8);
  const int32_t* h_first = reinterpret_cast<const int32_t*>(meta + 8 + 8 * (size_t)max_fields);
  unsigned char* dm = static_cast<unsigned char*>(d_meta.p);
  int32_t* d_count = reinterpret_cast<int32_t*>(dm);
  double* d_line0 = reinterpret_cast<double*>(dm + 8);
  int32_t* d_first = reinterpret_cast<int32_t*>(dm + 8 + 8 * (size_t)max_fields);
  int32_t* dd = static_cast<int32_t*>(d_drop.p);
  const int32_t* hd = static_cast<const int32_t*>(h_drop.p);
  const size_t rows = (size_t)max_fields * (size_t)cap;

  std::vector<FieldRecord> written;
  double last_abs = -1.0;  // RF position of the last written field's line 0
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
      got = cfg.read_callback(cfg.user_data, dst, (uint64_t)p, (size_t)n);
    } else {
      std::vector<float>& stage = staging[slot];
      got = cfg.read_callback(cfg.user_data, stage.data(), (uint64_t)p, (size_t)n);
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

  // .tbc writes of one span's fields on a writer thread, one job in flight, so the file
  // keeps the field order; a job writes fields [first, first + n) of a staging slot.
  std::future<bool> write_job;
  auto write_fields = [&](int slot, std::vector<int32_t> which) -> bool {
    const uint16_t* base_p = static_cast<const uint16_t*>(h_fields[slot].p);
    for (const int32_t f : which)
      if (std::fwrite(base_p + (size_t)f * per_field, sizeof(uint16_t), per_field, tbc.f) != per_field) return false;
    return true;
  };
  auto finish_writes = [&]() -> bool { return !write_job.valid() || write_job.get(); };

  const auto t0 = std::chrono::steady_clock::now();
  SpanRead cur = read_span(0, 0);
  for (int k = 0; cur.ok && !done; k++) {
    if (cfg.cancel_callback && cfg.cancel_callback(cfg.user_data))
      return set_error(VHSDECODE_HIP_E_CANCELLED, "the run was cancelled");
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
    rc = vhsdecode_hip_decode_span_fields_device(ctx.c, sp.video_lowpass, sp.video, plane, &fp, &lz, max_fields,
                                                 static_cast<uint16_t*>(d_fields.p), d_count, d_line0, d_first,
                                                 st.s);
    if (rc) return rc;
    if (cfg.dropouts) {
      const double* locs = nullptr;
      int32_t nloc = 0;
      rc = vhsdecode_hip_field_locations_device(ctx.c, &locs, &nloc);
      if (rc) return rc;
      rc = vhsdecode_hip_detect_dropouts_device(ctx.c, sp.envelope, plane, locs, nloc, d_line0, d_first, d_count,
                                                max_fields, cfg.dropouts, dd, dd + max_fields, dd + max_fields + rows,
                                                dd + max_fields + 2 * rows, nullptr, nullptr, nullptr, st.s);
      if (rc) return rc;
      RUN_HIP(hipMemcpyAsync(h_drop.p, d_drop.p, drop_bytes, hipMemcpyDeviceToHost, st.s));
    }
    RUN_HIP(hipMemcpyAsync(h_meta.p, d_meta.p, meta_bytes, hipMemcpyDeviceToHost, st.s));
    RUN_HIP(hipStreamSynchronize(st.s));
    const int32_t count = std::min(*h_count, max_fields);
    // slot k & 1 was last written from two spans ago; that job has finished (one in flight)
    if (count > 0)
      RUN_HIP(hipMemcpyAsync(h_fields[k & 1].p, d_fields.p, sizeof(uint16_t) * per_field * (size_t)count,
                             hipMemcpyDeviceToHost, st.s));
    RUN_HIP(hipStreamSynchronize(st.s));
    spans++;
    std::vector<int32_t> to_write;
    for (int32_t f = 0; f < count; f++) {
      const double abs0 = (double)pos + fm.block_cut + h_line0[f];
      if (last_abs >= 0.0 && abs0 < last_abs + 0.5 * field_len) continue;  // written from the last span
      // a field whose read window leaves the span is written from the next span, which
      // holds it whole; so are the fields after it
      if (cfg.dropouts && !cur.last && dropout_window_end(*cfg.dropouts, h_line0[f]) > plane) break;
      FieldRecord r;
      r.first = h_first[f] != 0;
      r.file_loc = (uint64_t)std::max(0.0, std::floor(abs0));
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
      last_abs = abs0;
      if (cfg.maximum_output_fields && written.size() >= cfg.maximum_output_fields) {
        done = true;
        break;
      }
    }
    if (!finish_writes()) return set_error(VHSDECODE_HIP_E_IO, "cannot write " + tbc_path);
    if (!to_write.empty()) write_job = std::async(std::launch::async, write_fields, k & 1, std::move(to_write));
    if (cur.last || done) break;
    cur = next.get();
  }
  if (!finish_writes()) return set_error(VHSDECODE_HIP_E_IO, "cannot write " + tbc_path);
  const auto t1 = std::chrono::steady_clock::now();
  if (std::fflush(tbc.f) != 0) return set_error(VHSDECODE_HIP_E_IO, "cannot write " + tbc_path);
  const double out_rate = fp.output_line_length * (fp.sample_rate_mhz * 1e6 / lz.line_length);
  if (!write_json(json_path, fp, out_rate, written)) return set_error(VHSDECODE_HIP_E_IO, "cannot write " + json_path);
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
  if (!fm || fm->struct_size != sizeof(vhsdecode_hip_serve_format_v1) || !fm->filters || !fm->field ||
      !fm->locator)
    return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "run format with filters, field and locator params required");
  if (fm->field->struct_size != sizeof(vhsdecode_hip_field_params_v1) ||
      fm->locator->struct_size != sizeof(vhsdecode_hip_locator_v1) ||
      (config->dropouts && config->dropouts->struct_size != sizeof(vhsdecode_hip_dropout_params_v1)))
    return set_error(VHSDECODE_HIP_E_STRUCT_SIZE, "field, locator or dropout params struct_size mismatch");
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
