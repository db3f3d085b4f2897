// Host side of VHSDecode.Hip.Native: context management, filter-table staging,
// launch orchestration and the C ABI declared in include/vhsdecode_hip.h.
#include "vhsdecode_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "fft16k.hpp"
#include "sos_scan.hpp"

#include "vhsd_args.hpp"
#include "vhs_sync.hpp"
#include "chroma.hpp"

#include "host_common.hpp"

namespace {

thread_local std::string g_last_error;

int32_t fail(int32_t code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define VHSD_HIP(call)                                                                             \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess) return fail(VHSDECODE_HIP_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t count) {
    if (count <= n && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) n = count;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// SosFilter.DefaultPadLength (Dsp/SosFilter.cs:540-545)
int default_pad_length(const double* sos, int sections) {
  int zn = 0, zd = 0;
  for (int s = 0; s < sections; s++) {
    if (sos[s * 6 + 2] == 0.0) zn++;
    if (sos[s * 6 + 5] == 0.0) zd++;
  }
  return 3 * ((2 * sections) + 1 - std::min(zn, zd));
}

// FP32 steady-state initial conditions exactly as
// SosFilter.SteadyStateInitialConditionsFloat32 (Dsp/SosFilter.cs:924-951).
#pragma clang fp contract(off)
void steady_state_f32(const float (*c)[6], int sections, float* zi) {
  float scale = 1.0f;
  for (int i = 0; i < sections; i++) {
    const float b0 = c[i][0], b1 = c[i][1], b2 = c[i][2], a0 = c[i][3], a1 = c[i][4], a2 = c[i][5];
    float t1 = b1 - (a1 * b0);
    float t2 = b2 - (a2 * b0);
    float num = (0.0f + t1) + t2;
    float den = (1.0f + a1) + a2;
    float z0 = num / den;
    float z1 = ((1.0f + a1) * z0) - t1;
    zi[2 * i] = scale * z0;
    zi[2 * i + 1] = scale * z1;
    float ndc = ((0.0f + b0) + b1) + b2;
    float ddc = ((0.0f + a0) + a1) + a2;
    scale *= ndc / ddc;
  }
}

// Build SosScanArgs for an N-sample block: FP32 coefficients, zi, and per section
// the state-transition matrices A^(C 2^d) and the chunk response (FP64, stored FP32).
int32_t prepare_sos(const double* sos, int sections, int n, vhsd::SosScanArgs* a, std::vector<float>* trans) {
  if (sections < 1 || (sections > 4 && sections != vhsd::kMaxSections))
    return fail(VHSDECODE_HIP_E_UNSUPPORTED, "SOS filters of 1 to 4 or 10 sections are supported");
  std::memset(a, 0, sizeof(*a));
  a->sections = sections;
  a->pad = default_pad_length(sos, sections);
  const int L = n + 2 * a->pad;
  // every thread runs exactly kChunk samples; the extended block plus its LDS margins
  // must fit the 33792-float block buffer
  a->chunk = vhsd::kChunk;
  if (a->pad > 64 || L > vhsd::kScanThreads * vhsd::kChunk ||
      vhsd::kSosMargin + L + vhsd::kChunk + 64 > 2 * vhsd::kFftBufEntries)
    return fail(VHSDECODE_HIP_E_UNSUPPORTED, "SOS padding too long");
  float c32[vhsd::kMaxSections][6];
  for (int s = 0; s < sections; s++) {
    for (int q = 0; q < 6; q++) c32[s][q] = (float)sos[s * 6 + q];
    a->coef[s][0] = c32[s][0];
    a->coef[s][1] = c32[s][1];
    a->coef[s][2] = c32[s][2];
    a->coef[s][3] = c32[s][4];
    a->coef[s][4] = c32[s][5];
  }
  steady_state_f32(c32, sections, a->zi);
  // per section: the one-step transition with zero input (state (z0, z1): output
  // f = z0, z0' = z1 - a1 f, z1' = -a2 f), its powers A^(C 2^d) in FP64, then the
  // chunk's zero-input response (output t per unit entry state)
  const int L2 = vhsd::kScanLevels * 4;
  trans->assign((size_t)sections * L2 + (size_t)sections * vhsd::kChunk * 2, 0.0f);
  auto mul2 = [](const double* x, const double* y, double* o) {
    const double r0 = x[0] * y[0] + x[1] * y[2], r1 = x[0] * y[1] + x[1] * y[3];
    const double r2 = x[2] * y[0] + x[3] * y[2], r3 = x[2] * y[1] + x[3] * y[3];
    o[0] = r0;
    o[1] = r1;
    o[2] = r2;
    o[3] = r3;
  };
  for (int k = 0; k < sections; k++) {
    const double a1 = c32[k][4], a2 = c32[k][5];
    const double A[4] = {-a1, 1.0, -a2, 0.0};
    double P[4] = {1.0, 0.0, 0.0, 1.0}, base[4] = {A[0], A[1], A[2], A[3]};
    for (int e = a->chunk; e > 0; e >>= 1) {
      if (e & 1) mul2(P, base, P);
      mul2(base, base, base);
    }
    float* lev = trans->data() + (size_t)k * L2;
    for (int d = 0; d < vhsd::kScanLevels; d++) {
      for (int i = 0; i < 4; i++) lev[d * 4 + i] = (float)P[i];
      mul2(P, P, P);
    }
    float* resp = trans->data() + (size_t)sections * L2 + (size_t)k * vhsd::kChunk * 2;
    for (int col = 0; col < 2; col++) {
      double z0 = col == 0 ? 1.0 : 0.0, z1 = col == 1 ? 1.0 : 0.0;
      for (int t = 0; t < vhsd::kChunk; t++) {
        const double f = z0;
        resp[t * 2 + col] = (float)f;
        z0 = z1 - a1 * f;
        z1 = -a2 * f;
      }
    }
  }
  return VHSDECODE_HIP_OK;
}
#pragma clang fp contract(on)

std::vector<float2> octant_table() {
  std::vector<float2> t(vhsd::kOctantEntries);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int u = 0; u < vhsd::kOctantEntries; u++) {
    long double a = two_pi * (long double)u / 16384.0L;
    t[u] = make_float2((float)cosl(a), (float)sinl(a));
  }
  return t;
}

// TbcLineResampler.BuildKaiserSincLookup (Tbc/TbcLineResampler.cs:1297-1330)
double bessel_i0(double x) {
  double sum = 1.0, y = (x * x) / 4.0, term = 1.0;
  for (int k = 1; k < 100; k++) {
    term *= y / (double)(k * k);
    double next = sum + term;
    if (next == sum) break;
    sum = next;
  }
  return sum;
}

std::vector<float> kaiser_sinc_table() {
  const int taps = 16, phases = 65536, half = 8;
  const double beta = 5.0, i0b = bessel_i0(beta);
  std::vector<float> lut((size_t)(phases + 1) * taps);
  double w[16];
  for (int p = 0; p < phases; p++) {
    double ph = (double)p / phases, sum = 0.0;
    for (int t = 0; t < taps; t++) {
      double x = (half - 1 - t) + ph;
      double s = x == 0.0 ? 1.0 : std::sin(M_PI * x) / (M_PI * x);
      double r = x / half;
      double k = (r < -1.0 || r > 1.0) ? 0.0 : bessel_i0(beta * std::sqrt(1.0 - r * r)) / i0b;
      w[t] = s * k;
      sum += w[t];
    }
    for (int t = 0; t < taps; t++) lut[(size_t)p * taps + t] = (float)(w[t] / sum);
  }
  for (int t = 0; t < taps; t++) lut[(size_t)phases * taps + t] = lut[(size_t)(phases - 1) * taps + t];
  return lut;
}

double median(std::vector<double> v) {
  if (v.empty()) return NAN;
  for (double x : v)
    if (std::isnan(x)) return x;
  std::sort(v.begin(), v.end());
  size_t m = v.size() / 2;
  return (v.size() & 1) ? v[m] : (v[m - 1] + v[m]) / 2.0;
}

}  // namespace

int32_t vhsd_host::set_error(int32_t code, const std::string& msg) { return fail(code, msg); }

struct vhsdecode_hip_demod_context {
  std::mutex mu;
  int device = 0;
  hipStream_t stream = nullptr;
  int n = 0, cut = 0, cut_end = 0, stride = 0, max_blocks = 0;
  double fs = 0.0;
  int vlp_offset = 0, chroma_offset = 0;
  double diff_max = 0.0;
  double bias = 0.0;
  float video_corr = 0.0f, vlp_corr = 0.0f;
  bool has_chroma = false;
  vhsd::SosScanArgs env{}, chroma{};
  DevBuf<float2> rfv, vid, vlp, oct;
  DevBuf<float> env_trans, chroma_trans;
  // RF high boost and fsc notch (Betamax)
  bool has_boost = false, has_notch = false;
  vhsd::SosScanArgs boost{}, notch{};
  float boost_mult = 0.0f;
  DevBuf<float> boost_trans, notch_trans;
  DevBuf<float2> aux;
  // LaserDisc complex path
  bool ld = false;
  float clip_lo = 0.0f, clip_hi = 0.0f;   // offsets from the bias
  int nref = 0;                           // burst then pilot
  bool ref_is_pilot[2] = {false, false};
  DevBuf<float2> ref_filt[2];
  float ref_corr[2] = {0.0f, 0.0f};
  int ref_shift[2] = {0, 0};
  bool has_audio = false;
  int audio_bins = 0, audio_log2 = 0, audio_low[2] = {0, 0}, audio_dec = 0;
  DevBuf<float2> audio_st[2];
  double audio_rate = 0.0, audio_low_freq[2] = {0.0, 0.0};
  // sub-deemphasis
  bool has_subdeemph = false;
  vhsd::SosScanArgs sd_amp{};
  DevBuf<float2> sd_vh;
  DevBuf<float> sd_trans;
  float sd_params[7] = {};
  DevBuf<long long> block_start, out_offset;
  DevBuf<unsigned> flags;
  DevBuf<float2> rep_a;
  DevBuf<int> dbg_list;
  DevBuf<float2> stash;
  DevBuf<float> rep_dm;
  // pulse scratch
  DevBuf<int> tile_counts, counters, p_starts, p_lens;
  DevBuf<long long> edge_pos;
  DevBuf<unsigned char> edge_kind;
  // hsync scratch
  DevBuf<double> h_loc, h_ref;
  DevBuf<unsigned char> h_fil, h_fo;
  DevBuf<unsigned char> h_ptrs;
  DevBuf<unsigned char> h_work;
  // tbc
  DevBuf<float> lut;
  bool lut_ready = false;
  DevBuf<double> t_loc, t_adj, t_f64;
  DevBuf<unsigned short> t_u16;
  // batched fields
  DevBuf<double> f_loc, f_ref, f_x, f_y0, f_pow;
  int f_ref_nloc = 0;  // locations per field of the last field decode
  DevBuf<float> d_part;               // dropout mean partial sums
  DevBuf<unsigned long long> d_mask;  // dropout low/high bit masks
  DevBuf<unsigned> d_sum;             // dropout mask summaries
  DevBuf<vhsd::DropoutField> d_info;  // dropout per-field state
  DevBuf<unsigned char> f_fil, f_fo, f_tab;
  // VHS field sequence (vhs_sync.hip): one allocation carved into the per-field lists,
  // records and tables; the sequence state (level moving averages and sync history) lives
  // at its start and carries from call to call
  DevBuf<unsigned char> vs_mem, vs_state;
  bool vs_state_ready = false;
  // colour-under chroma (vhsdecode_hip_set_vhs_chroma): the format, its tables, the state
  // carried with the field sequence and the per-call work buffers
  bool chroma_on = false;
  vhsd::chroma::ChromaCfg ccfg{};
  DevBuf<float> c_het, c_sin, c_cos;
  DevBuf<float2> ld2_filt[2];
  double ld2_center[2] = {0.0, 0.0};
  bool ld2_on = false;
  DevBuf<unsigned char> c_state, c_work;
  bool c_state_ready = false;
  const double* vs_lines = nullptr;    // line locations, hsync-refined and render locations of the
  const double* vs_refined = nullptr;  // last vhs field decode ([fields][stride], window coordinates)
  const double* vs_render = nullptr;
  const double* vs_span = nullptr;     // render locations from the output first line, span coordinates
  const long long* vs_origins = nullptr;
  const int* vs_first_lines = nullptr;
  int vs_line_stride = 0;
  // the field sequence's launches captured as a HIP graph (replayed while every argument
  // of every launch is the same), and those arguments
  hipGraph_t vs_graph = nullptr;
  hipGraphExec_t vs_exec = nullptr;
  std::vector<unsigned char> vs_key;
  double f_pow_alpha = -1.0;
  int f_pow_len = 0;
  void* f_stage = nullptr;   // pinned host staging for the location upload
  size_t f_stage_bytes = 0;
  hipEvent_t f_stage_done = nullptr;
  // chroma on a caller stream (span_v1.chroma_stream): fork point and completion
  hipEvent_t chroma_fork = nullptr, chroma_done = nullptr;
  bool chroma_pending = false;
  // host-variant staging
  DevBuf<int16_t> s_rf;
  DevBuf<float> s_planes;
  DevBuf<unsigned> s_flags;
  // profiling: event pairs per kernel kind
  bool profiling = false;
  struct Rec {
    int kind;
    hipEvent_t a, b;
    long long units;
  };
  std::vector<Rec> recs;
  long long prof_launches[VHSDECODE_HIP_KERNEL_COUNT] = {};
  double prof_ms[VHSDECODE_HIP_KERNEL_COUNT] = {};
  long long prof_units[VHSDECODE_HIP_KERNEL_COUNT] = {};

  ~vhsdecode_hip_demod_context() {
    (void)hipSetDevice(device);
    rfv.release(); vid.release(); vlp.release(); oct.release(); sd_vh.release(); sd_trans.release();
    ref_filt[0].release(); ref_filt[1].release(); audio_st[0].release(); audio_st[1].release();
    boost_trans.release(); notch_trans.release(); aux.release();
    env_trans.release(); chroma_trans.release(); block_start.release(); out_offset.release(); flags.release();
    rep_a.release(); rep_dm.release(); stash.release(); dbg_list.release(); tile_counts.release(); counters.release(); p_starts.release();
    p_lens.release(); edge_pos.release(); edge_kind.release(); h_loc.release(); h_ref.release();
    h_fil.release(); h_fo.release(); h_ptrs.release(); h_work.release(); lut.release(); t_loc.release(); t_adj.release();
    t_f64.release(); t_u16.release(); s_rf.release(); s_planes.release(); s_flags.release();
    vs_mem.release(); vs_state.release();
    c_het.release(); c_sin.release(); c_cos.release(); c_state.release(); c_work.release();
    ld2_filt[0].release(); ld2_filt[1].release();
    f_loc.release(); f_ref.release(); f_x.release(); f_y0.release(); f_pow.release(); f_fil.release();
    f_fo.release(); f_tab.release();
    if (vs_exec) (void)hipGraphExecDestroy(vs_exec);
    if (vs_graph) (void)hipGraphDestroy(vs_graph);
    if (f_stage) (void)hipHostFree(f_stage);
    if (f_stage_done) (void)hipEventDestroy(f_stage_done);
    if (chroma_fork) (void)hipEventDestroy(chroma_fork);
    if (chroma_done) (void)hipEventDestroy(chroma_done);
    for (auto& r : recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// RAII bracket of one launch with HIP events when profiling is enabled.
struct ProfScope {
  vhsdecode_hip_demod_context* c;
  hipStream_t st;
  int kind;
  long long units;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(vhsdecode_hip_demod_context* ctx, hipStream_t s, int k, long long u) : c(ctx), st(s), kind(k), units(u) {
    if (!c->profiling) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      a = b = nullptr;
      return;
    }
    (void)hipEventRecord(a, st);
  }
  ~ProfScope() {
    if (!a) return;
    (void)hipEventRecord(b, st);
    c->recs.push_back({kind, a, b, units});
  }
};

hipStream_t pick_stream(vhsdecode_hip_demod_context* c, void* stream) {
  return stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
}

int32_t upload_half(DevBuf<float2>& dst, const double* full, int n) {
  const int m = n / 2 + 1;
  std::vector<float2> h(m);
  for (int k = 0; k < m; k++) h[k] = make_float2((float)full[2 * k], (float)full[2 * k + 1]);
  VHSD_HIP(dst.ensure(m));
  VHSD_HIP(hipMemcpy(dst.p, h.data(), sizeof(float2) * m, hipMemcpyHostToDevice));
  return VHSDECODE_HIP_OK;
}

// The half spectrum of the Hermitian part (H[k] + conj(H[N-k])) / 2 of a whole-circle
// response: Re(ifft(X H)) of a real signal's spectrum X is irfft of X times it, which is
// what the LaserDisc path's complex inverse followed by the real part computes.
int32_t upload_half_hermitian(DevBuf<float2>& dst, const double* full, int n) {
  const int m = n / 2 + 1;
  std::vector<double> h((size_t)2 * n, 0.0);
  for (int k = 0; k < m; k++) {
    const int q = (n - k) % n;
    h[2 * k] = 0.5 * (full[2 * k] + full[2 * q]);
    h[2 * k + 1] = 0.5 * (full[2 * k + 1] - full[2 * q + 1]);
  }
  return upload_half(dst, h.data(), n);
}

int32_t ensure_lut(vhsdecode_hip_demod_context* c) {
  if (c->lut_ready) return VHSDECODE_HIP_OK;
  std::vector<float> t = kaiser_sinc_table();
  VHSD_HIP(c->lut.ensure(t.size()));
  VHSD_HIP(hipMemcpy(c->lut.p, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice));
  c->lut_ready = true;
  return VHSDECODE_HIP_OK;
}

// Argument records of one context's block chain, and which kernels it runs.
struct ChainArgs {
  vhsd::DemodArgs d{};
  vhsd::RepairArgs r{};
  vhsd::SubDeemphArgs sd{};
  vhsd::ChromaArgs ch{};
  int demod_kind = vhsd::kTableDemod;  // plain, RF boost or LaserDisc
  bool repair = false, subdeemph = false, chroma = false;
  int chroma_sections = 0;
};

// Fill the records for blocks whose input offsets (into rf) and stitched output
// offsets (into the span planes) are the device arrays starts/offs, with per-block
// flags and scratch indexed by the same block index.
void build_chain(vhsdecode_hip_demod_context* c, const int16_t* rf, const long long* starts, const long long* offs,
                 const vhsdecode_hip_span_v1* out, unsigned* flags, float2* stash, float2* aux, float2* rep_a,
                 float* rep_dm, ChainArgs* ca) {
  vhsd::DemodArgs& a = ca->d;
  a = vhsd::DemodArgs{};
  a.rf = rf;
  a.block_start = starts;
  a.out_offset = offs;
  a.out_begin = c->cut;
  a.out_count = c->stride;
  a.rf_video = c->rfv.p;
  a.video = c->vid.p;
  a.vlp05 = c->vlp.p;
  a.oct = c->oct.p;
  a.vlp_offset = c->vlp_offset;
  a.fs = (float)c->fs;
  a.diff_max = c->diff_max;
  a.bias = (float)c->bias;
  a.video_corr = c->video_corr;
  a.vlp_corr = c->vlp_corr;
  a.env = c->env;
  a.env.trans = c->env_trans.p;
  a.out_video = out->video;
  a.out_vlp = out->video_lowpass;
  a.out_env = out->envelope;
  a.flags = flags;
  a.stash = stash;
  a.aux = aux;
  a.stash_sub = a.bias;
  if (c->has_boost) {
    a.boost = c->boost;
    a.boost.trans = c->boost_trans.p;
    a.boost_mult = c->boost_mult;
  }
  ca->demod_kind = c->ld ? vhsd::kTableDemodLd : (c->has_boost ? vhsd::kTableDemodBoost : vhsd::kTableDemod);
  if (c->ld) {
    // LaserDisc: the demod is stashed as offsets from the bias and clipped for video;
    // burst and pilot are extra passes of the video kernel
    a.stash_sub = 0.0f;
    a.clip_lo = c->clip_lo;
    a.clip_hi = c->clip_hi;
    a.fs_d = c->fs;
    a.diff_max = 0.0;
    int nref = 0;
    for (int r = 0; r < c->nref; r++) {
      float* dst = c->ref_is_pilot[r] ? out->pilot : out->chroma;
      if (!dst) continue;
      a.ref_filt[nref] = c->ref_filt[r].p;
      a.ref_out[nref] = dst;
      a.ref_shift[nref] = c->ref_shift[r];
      a.ref_corr[nref] = c->ref_corr[r];
      nref++;
    }
    a.nref = nref;
    if (c->has_audio && out->audio_left && out->audio_right) {
      a.audio = 1;
      a.audio_bins = c->audio_bins;
      a.audio_log2 = c->audio_log2;
      for (int ch = 0; ch < 2; ch++) {
        a.audio_low[ch] = c->audio_low[ch];
        a.audio_stage1[ch] = c->audio_st[ch].p;
        a.audio_low_freq[ch] = c->audio_low_freq[ch];
      }
      a.audio_scale = c->audio_rate / 6.283185307179586;
      a.audio_cut = c->cut / c->audio_dec;
      a.audio_stride = c->stride / c->audio_dec;
      a.audio_decimation = c->audio_dec;
      a.out_audio[0] = out->audio_left;
      a.out_audio[1] = out->audio_right;
    }
    return;
  }
  ca->repair = c->diff_max > 0.0;
  ca->r = vhsd::RepairArgs{a, rep_a, rep_dm};
  ca->subdeemph = c->has_subdeemph;
  if (c->has_subdeemph) {
    vhsd::SubDeemphArgs& sd = ca->sd;
    sd = vhsd::SubDeemphArgs{};
    sd.out_offset = offs;
    sd.out_begin = c->cut;
    sd.out_count = c->stride;
    sd.vh = c->sd_vh.p;
    sd.oct = c->oct.p;
    sd.stash = stash;
    sd.amp = c->sd_amp;
    sd.amp.trans = c->sd_trans.p;
    sd.half_deviation = c->sd_params[0];
    sd.exp_scaling = c->sd_params[1];
    sd.scaling1 = c->sd_params[2];
    sd.scaling2 = c->sd_params[3];
    sd.logistic_mid = c->sd_params[4];
    sd.logistic_rate = c->sd_params[5];
    sd.static_factor = c->sd_params[6];
    sd.out_video = out->video;
    sd.notch = c->has_notch ? 1 : 0;
    if (c->has_notch) {
      sd.notch_sos = c->notch;
      sd.notch_sos.trans = c->notch_trans.p;
    }
    sd.video = c->vid.p;
    sd.video_corr = c->video_corr;
    sd.aux = aux;
  }
  ca->chroma = c->has_chroma && out->chroma;
  if (ca->chroma) {
    vhsd::ChromaArgs& ch = ca->ch;
    ch = vhsd::ChromaArgs{};
    ch.rf = rf;
    ch.block_start = starts;
    ch.out_offset = offs;
    ch.out_begin = c->cut;
    ch.out_count = c->stride;
    ch.chroma_offset = c->chroma_offset;
    ch.sos = c->chroma;
    ch.sos.trans = c->chroma_trans.p;
    ch.out_chroma = out->chroma;
    ca->chroma_sections = c->chroma.sections;
  }
}

// The block chain of one context over nblocks blocks (kernel launches only, so a caller
// may capture the call into a graph).
int32_t demod_blocks_locked(vhsdecode_hip_demod_context* c, const int16_t* rf, const long long* starts,
                            const long long* offs, int32_t nblocks, const vhsdecode_hip_span_v1* out,
                            unsigned* flags, float2* stash, float2* aux, float2* rep_a, float* rep_dm,
                            hipStream_t st) {
  ChainArgs ca;
  build_chain(c, rf, starts, offs, out, flags, stash, aux, rep_a, rep_dm, &ca);
  // chroma only reads the RF: on a caller-supplied stream it overlaps the rest of the
  // chain and whatever the caller enqueues next (the field decode)
  hipStream_t cs = out->chroma_stream ? reinterpret_cast<hipStream_t>(out->chroma_stream) : nullptr;
  if (ca.chroma && cs) {
    if (!c->chroma_fork) VHSD_HIP(hipEventCreateWithFlags(&c->chroma_fork, hipEventDisableTiming));
    if (!c->chroma_done) VHSD_HIP(hipEventCreateWithFlags(&c->chroma_done, hipEventDisableTiming));
    VHSD_HIP(hipEventRecord(c->chroma_fork, st));
    VHSD_HIP(hipStreamWaitEvent(cs, c->chroma_fork, 0));
    ProfScope ps(c, cs, VHSDECODE_HIP_KERNEL_CHROMA, nblocks);
    VHSD_HIP(vhsd_launch_block_chroma(&ca.ch, nblocks, cs));
    VHSD_HIP(hipEventRecord(c->chroma_done, cs));
    c->chroma_pending = true;
  }
  VHSD_HIP(vhsd_launch_fill_u32(flags, nblocks, 0u, st));
  {
    ProfScope ps(c, st, VHSDECODE_HIP_KERNEL_BLOCK_DEMOD, nblocks);
    if (ca.demod_kind == vhsd::kTableDemodLd)
      VHSD_HIP(vhsd_launch_block_demod_ld(&ca.d, nblocks, st));
    else if (ca.demod_kind == vhsd::kTableDemodBoost)
      VHSD_HIP(vhsd_launch_block_demod_boost(&ca.d, nblocks, st));
    else
      VHSD_HIP(vhsd_launch_block_demod(&ca.d, nblocks, st));
  }
  if (ca.repair) {
    ProfScope ps(c, st, VHSDECODE_HIP_KERNEL_SPIKE_REPAIR, nblocks);
    VHSD_HIP(vhsd_launch_block_repair(&ca.r, nblocks, st));
  }
  {
    ProfScope ps(c, st, VHSDECODE_HIP_KERNEL_BLOCK_VIDEO, nblocks);
    VHSD_HIP(vhsd_launch_block_video(&ca.d, nblocks, st));
  }
  if (ca.subdeemph) {
    ProfScope ps(c, st, VHSDECODE_HIP_KERNEL_SUB_DEEMPHASIS, nblocks);
    VHSD_HIP(vhsd_launch_block_subdeemph(&ca.sd, nblocks, st));
  }
  if (ca.chroma && !cs) {
    ProfScope ps(c, st, VHSDECODE_HIP_KERNEL_CHROMA, nblocks);
    VHSD_HIP(vhsd_launch_block_chroma(&ca.ch, nblocks, st));
  }
  return VHSDECODE_HIP_OK;
}

// Per-block scratch of the chain for nblocks blocks (allocates outside any capture).
int32_t reserve_block_scratch(vhsdecode_hip_demod_context* c, int32_t nblocks) {
  VHSD_HIP(c->stash.ensure((size_t)nblocks * vhsd::kStashEntries));
  if (c->has_boost || c->has_notch || c->has_audio) VHSD_HIP(c->aux.ensure((size_t)nblocks * vhsd::kAuxEntries));
  if (c->diff_max > 0.0) {
    VHSD_HIP(c->rep_a.ensure((size_t)nblocks * 32768));
    VHSD_HIP(c->rep_dm.ensure((size_t)nblocks * 65536));
  }
  return VHSDECODE_HIP_OK;
}

int32_t demod_span_locked(vhsdecode_hip_demod_context* c, const int16_t* rf, int64_t first_sample, int32_t nblocks,
                          const vhsdecode_hip_span_v1* out, hipStream_t st) {
  if (nblocks <= 0) return VHSDECODE_HIP_OK;
  if (nblocks > c->max_blocks)
    return fail(VHSDECODE_HIP_E_CAPACITY, "nblocks exceeds the context's max_blocks");
  if (!rf || !out || !out->video || !out->video_lowpass || !out->envelope)
    return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "rf and the video, video_lowpass and envelope planes are required");
  if (first_sample < 0 || (first_sample & 1))
    return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "first_sample must be even and non-negative");
  std::vector<long long> starts(nblocks), offs(nblocks);
  for (int b = 0; b < nblocks; b++) {
    starts[b] = first_sample + (long long)b * c->stride;
    offs[b] = (long long)b * c->stride;
  }
  int32_t rc = reserve_block_scratch(c, nblocks);
  if (rc != VHSDECODE_HIP_OK) return rc;
  if (c->chroma_pending) {
    // a chroma launched on another stream may still read the descriptor arrays
    VHSD_HIP(hipStreamWaitEvent(st, c->chroma_done, 0));
    c->chroma_pending = false;
  }
  VHSD_HIP(hipMemcpyAsync(c->block_start.p, starts.data(), sizeof(long long) * nblocks, hipMemcpyHostToDevice, st));
  VHSD_HIP(hipMemcpyAsync(c->out_offset.p, offs.data(), sizeof(long long) * nblocks, hipMemcpyHostToDevice, st));
  unsigned* flags = out->block_flags ? out->block_flags : c->flags.p;
  return demod_blocks_locked(c, rf, c->block_start.p, c->out_offset.p, nblocks, out, flags, c->stash.p, c->aux.p,
                             c->rep_a.p, c->rep_dm.p, st);
}

}  // namespace

extern "C" {

int32_t vhsdecode_hip_get_abi_version(void) { return VHSDECODE_HIP_ABI_VERSION; }

const char* vhsdecode_hip_get_last_error(void) { return g_last_error.c_str(); }

int32_t vhsdecode_hip_get_runtime_info(int32_t device_id, vhsdecode_hip_runtime_info_v1* info) {
  try {
    if (!info) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "info is NULL");
    if (info->struct_size != sizeof(vhsdecode_hip_runtime_info_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "runtime_info struct_size mismatch");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
      info->device_count = 0;
      return fail(VHSDECODE_HIP_E_NO_DEVICE, "no HIP device is available");
    }
    if (device_id < 0 || device_id >= count) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "device_id out of range");
    hipDeviceProp_t prop;
    VHSD_HIP(hipGetDeviceProperties(&prop, device_id));
    info->device_count = count;
    info->device_id = device_id;
    info->compute_units = prop.multiProcessorCount;
    info->hbm_bytes = prop.totalGlobalMem;
    info->lds_bytes_per_workgroup = (int32_t)prop.sharedMemPerBlock;
    std::snprintf(info->device_name, sizeof(info->device_name), "%s", prop.name);
    std::snprintf(info->gcn_arch, sizeof(info->gcn_arch), "%s", prop.gcnArchName);
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

namespace {

// LaserDisc tables: RfVideo x RfMtf (FMA complex product, as the reference's
// NumpyComplexMultiply) for the analytic signal, Hermitian parts of Video,
// VideoLowPass05, burst and pilot for the real-output passes, and the audio slices.
int32_t setup_laserdisc(vhsdecode_hip_demod_context* c, const vhsdecode_hip_filters_v1* f,
                        const vhsdecode_hip_laserdisc_v1* ld) {
  const int n = c->n;
  c->ld = true;
  std::vector<double> fr((size_t)2 * n);
  for (int k = 0; k < n; k++) {
    double ar = f->rf_video[2 * k], ai = f->rf_video[2 * k + 1];
    if (ld->rf_mtf) {
      const double br = ld->rf_mtf[2 * k], bi = ld->rf_mtf[2 * k + 1];
      const double r = std::fma(ar, br, -(ai * bi));
      const double i = std::fma(ar, bi, ai * br);
      ar = r;
      ai = i;
    }
    fr[2 * k] = ar;
    fr[2 * k + 1] = ai;
  }
  int32_t rc;
  if ((rc = upload_half(c->rfv, fr.data(), n)) != 0) return rc;
  if ((rc = upload_half_hermitian(c->vid, f->video, n)) != 0) return rc;
  if ((rc = upload_half_hermitian(c->vlp, f->video_lowpass05, n)) != 0) return rc;
  if (ld->clip_high_hz > ld->clip_low_hz) {
    c->clip_lo = (float)(ld->clip_low_hz - c->bias);
    c->clip_hi = (float)(ld->clip_high_hz - c->bias);
  }
  const double* refs[2] = {ld->video_burst, ld->video_pilot};
  for (int r = 0; r < 2; r++) {
    if (!refs[r]) continue;
    const int shift = r == 0 ? ld->video_burst_offset : 0;
    if ((shift & 1) || shift < 0 || shift > c->cut_end)
      return fail(VHSDECODE_HIP_E_UNSUPPORTED, "video_burst_offset must be even and within the tail cut");
    const int i = c->nref++;
    if ((rc = upload_half_hermitian(c->ref_filt[i], refs[r], n)) != 0) return rc;
    c->ref_is_pilot[i] = r == 1;
    c->ref_shift[i] = shift;
    c->ref_corr[i] = (float)(c->bias * refs[r][0]);
  }
  if (const vhsdecode_hip_ld_audio_v1* au = ld->analog_audio) {
    if (au->struct_size != sizeof(vhsdecode_hip_ld_audio_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "analog_audio struct_size mismatch");
    const int bins = au->bin_count;
    int lg = 0;
    while ((1 << lg) < bins) lg++;
    if (bins < 2 || bins > 2048 || (1 << lg) != bins || n % bins)
      return fail(VHSDECODE_HIP_E_UNSUPPORTED, "analog audio slices must be powers of two from 2 to 2048 bins");
    const int dec = n / bins;
    if (c->cut % dec || c->stride % dec)
      return fail(VHSDECODE_HIP_E_UNSUPPORTED, "block cut and stride must be multiples of the audio decimation");
    if (!au->stage1[0] || !au->stage1[1] || !(au->slice_rate_hz > 0.0))
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "analog audio needs both stage-1 responses and a slice rate");
    for (int ch = 0; ch < 2; ch++) {
      if (au->low_bin[ch] < 0 || au->low_bin[ch] + bins / 2 >= n / 2)
        return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "analog audio slice is outside the positive half spectrum");
      std::vector<float2> h(bins);
      for (int i = 0; i < bins; i++) h[i] = make_float2((float)au->stage1[ch][2 * i], (float)au->stage1[ch][2 * i + 1]);
      VHSD_HIP(c->audio_st[ch].ensure(bins));
      VHSD_HIP(hipMemcpy(c->audio_st[ch].p, h.data(), sizeof(float2) * bins, hipMemcpyHostToDevice));
      c->audio_low[ch] = au->low_bin[ch];
      c->audio_low_freq[ch] = au->low_frequency_hz[ch];
    }
    c->audio_bins = bins;
    c->audio_log2 = lg;
    c->audio_dec = dec;
    c->audio_rate = au->slice_rate_hz;
    c->has_audio = true;
  }
  return VHSDECODE_HIP_OK;
}

}  // namespace

int32_t vhsdecode_hip_demod_create(const vhsdecode_hip_demod_config_v1* cfg, vhsdecode_hip_demod_context** out) {
  try {
    if (!cfg || !out) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "config and context must not be NULL");
    *out = nullptr;
    if (cfg->struct_size != sizeof(vhsdecode_hip_demod_config_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "demod_config struct_size mismatch");
    const vhsdecode_hip_filters_v1* f = cfg->filters;
    if (!f || f->struct_size != sizeof(vhsdecode_hip_filters_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "filters missing or struct_size mismatch");
    if (f->block_length != 32768)
      return fail(VHSDECODE_HIP_E_UNSUPPORTED, "this build demodulates 32768-sample blocks");
    const vhsdecode_hip_laserdisc_v1* ld = f->laserdisc;
    if (ld && ld->struct_size != sizeof(vhsdecode_hip_laserdisc_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "laserdisc struct_size mismatch");
    if (!f->rf_video || !f->video || !f->video_lowpass05)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "rf_video, video and video_lowpass05 are required");
    if (!ld && (!f->envelope_sos || f->envelope_sections != 1))
      return fail(VHSDECODE_HIP_E_UNSUPPORTED, "the envelope filter must be a single section (first-order low-pass)");
    if (ld && f->diff_demod_max > 0.0)
      return fail(VHSDECODE_HIP_E_UNSUPPORTED, "the LaserDisc path has no diff-demod repair");
    if (!(f->sample_rate_hz > 0.0)) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "sample_rate_hz must be positive");
    if (cfg->block_cut < 0 || cfg->block_cut_end < 0 || cfg->block_cut + cfg->block_cut_end >= f->block_length ||
        (cfg->block_cut & 1) || (cfg->block_cut_end & 1))
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "block cuts must be even and leave decoded samples");
    if (cfg->max_blocks <= 0) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "max_blocks must be positive");
    if ((f->video_lowpass05_offset & 1) || f->video_lowpass05_offset < 0 ||
        f->video_lowpass05_offset > cfg->block_cut_end)
      return fail(VHSDECODE_HIP_E_UNSUPPORTED, "video_lowpass05_offset must be even and within the tail cut");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
      return fail(VHSDECODE_HIP_E_NO_DEVICE, "no HIP device is available");
    if (cfg->device_id < 0 || cfg->device_id >= count)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "device_id out of range");
    VHSD_HIP(hipSetDevice(cfg->device_id));
    auto* c = new (std::nothrow) vhsdecode_hip_demod_context();
    if (!c) return fail(VHSDECODE_HIP_E_OUT_OF_MEMORY, "context allocation failed");
    c->device = cfg->device_id;
    c->n = f->block_length;
    c->cut = cfg->block_cut;
    c->cut_end = cfg->block_cut_end;
    c->stride = c->n - c->cut - c->cut_end;
    c->max_blocks = cfg->max_blocks;
    c->fs = f->sample_rate_hz;
    c->vlp_offset = f->video_lowpass05_offset;
    c->chroma_offset = f->chroma_offset;
    c->diff_max = f->diff_demod_max;
    c->bias = f->plane_bias_hz;
    if (!std::isfinite(c->bias) || (double)(float)c->bias != c->bias)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "plane_bias_hz must be exactly representable in FP32");
    // DC restored after the biased transforms: bias * (Re H[0] - 1)
    c->video_corr = (float)(c->bias * (f->video[0] - 1.0));
    c->vlp_corr = (float)(c->bias * (f->video_lowpass05[0] - 1.0));
    auto bail = [&](int32_t code) {
      delete c;
      return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
      return bail(fail(VHSDECODE_HIP_E_HIP, "hipStreamCreate failed"));
    int32_t rc;
    if (ld) {
      if ((rc = setup_laserdisc(c, f, ld)) != 0) return bail(rc);
    } else {
      if ((rc = upload_half(c->rfv, f->rf_video, c->n)) != 0) return bail(rc);
      if ((rc = upload_half(c->vid, f->video, c->n)) != 0) return bail(rc);
      if ((rc = upload_half(c->vlp, f->video_lowpass05, c->n)) != 0) return bail(rc);
    }
    std::vector<float2> oct = octant_table();
    if (c->oct.ensure(oct.size()) != hipSuccess ||
        hipMemcpy(c->oct.p, oct.data(), sizeof(float2) * oct.size(), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(VHSDECODE_HIP_E_HIP, "twiddle upload failed"));
    std::vector<float> tr;
    if (ld) {
      if (c->block_start.ensure(c->max_blocks) != hipSuccess || c->out_offset.ensure(c->max_blocks) != hipSuccess ||
          c->flags.ensure(c->max_blocks) != hipSuccess)
        return bail(fail(VHSDECODE_HIP_E_OUT_OF_MEMORY, "descriptor allocation failed"));
      *out = c;
      return VHSDECODE_HIP_OK;
    }
    if ((rc = prepare_sos(f->envelope_sos, f->envelope_sections, c->n, &c->env, &tr)) != 0) return bail(rc);
    if (c->env_trans.ensure(tr.size()) != hipSuccess ||
        hipMemcpy(c->env_trans.p, tr.data(), sizeof(float) * tr.size(), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(VHSDECODE_HIP_E_HIP, "envelope table upload failed"));
    if (f->chroma_sos && f->chroma_sections > 0) {
      if ((rc = prepare_sos(f->chroma_sos, f->chroma_sections, c->n, &c->chroma, &tr)) != 0) return bail(rc);
      if (c->chroma_trans.ensure(tr.size()) != hipSuccess ||
          hipMemcpy(c->chroma_trans.p, tr.data(), sizeof(float) * tr.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(VHSDECODE_HIP_E_HIP, "chroma table upload failed"));
      c->has_chroma = true;
    }
    if (const vhsdecode_hip_subdeemph_v1* sd = f->sub_deemphasis) {
      if (sd->struct_size != sizeof(vhsdecode_hip_subdeemph_v1))
        return bail(fail(VHSDECODE_HIP_E_STRUCT_SIZE, "sub_deemphasis struct_size mismatch"));
      if (!sd->high_pass || !sd->amplitude_sos || !(sd->deviation > 0.0))
        return bail(fail(VHSDECODE_HIP_E_INVALID_ARGUMENT,
                         "sub_deemphasis needs high_pass, amplitude_sos and a positive deviation"));
      // Video x HighPass in FP64, stored as an FP32 half spectrum
      std::vector<double> vh((size_t)2 * c->n);
      for (int k = 0; k <= c->n / 2; k++) {
        const double ar = f->video[2 * k], ai = f->video[2 * k + 1];
        const double br = sd->high_pass[2 * k], bi = sd->high_pass[2 * k + 1];
        vh[2 * k] = std::fma(ar, br, -(ai * bi));
        vh[2 * k + 1] = std::fma(ar, bi, ai * br);
      }
      if ((rc = upload_half(c->sd_vh, vh.data(), c->n)) != 0) return bail(rc);
      if ((rc = prepare_sos(sd->amplitude_sos, 1, c->n, &c->sd_amp, &tr)) != 0) return bail(rc);
      if (c->sd_trans.ensure(tr.size()) != hipSuccess ||
          hipMemcpy(c->sd_trans.p, tr.data(), sizeof(float) * tr.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(VHSDECODE_HIP_E_HIP, "sub-deemphasis table upload failed"));
      c->sd_params[0] = (float)(sd->deviation / 2.0);
      c->sd_params[1] = (float)sd->exponential_scaling;
      c->sd_params[2] = (float)sd->scaling1;
      c->sd_params[3] = (float)sd->scaling2;
      c->sd_params[4] = (float)sd->logistic_mid;
      c->sd_params[5] = (float)(sd->logistic_rate > 0.0 ? sd->logistic_rate : 0.0);
      c->sd_params[6] = (float)sd->static_factor;
      c->has_subdeemph = true;
    }
    if (f->rf_boost_sos && f->rf_boost_multiplier != 0.0) {
      if ((rc = prepare_sos(f->rf_boost_sos, 1, c->n, &c->boost, &tr)) != 0) return bail(rc);
      if (c->boost_trans.ensure(tr.size()) != hipSuccess ||
          hipMemcpy(c->boost_trans.p, tr.data(), sizeof(float) * tr.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(VHSDECODE_HIP_E_HIP, "RF boost table upload failed"));
      c->boost_mult = (float)f->rf_boost_multiplier;
      c->has_boost = true;
    }
    if (f->video_notch) {
      if (!c->has_subdeemph)
        return bail(fail(VHSDECODE_HIP_E_UNSUPPORTED, "the fsc notch is supported together with sub-deemphasis"));
      if ((rc = prepare_sos(f->video_notch, 1, c->n, &c->notch, &tr)) != 0) return bail(rc);
      if (c->notch_trans.ensure(tr.size()) != hipSuccess ||
          hipMemcpy(c->notch_trans.p, tr.data(), sizeof(float) * tr.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(VHSDECODE_HIP_E_HIP, "notch table upload failed"));
      c->has_notch = true;
    }
    if (c->block_start.ensure(c->max_blocks) != hipSuccess || c->out_offset.ensure(c->max_blocks) != hipSuccess ||
        c->flags.ensure(c->max_blocks) != hipSuccess)
      return bail(fail(VHSDECODE_HIP_E_OUT_OF_MEMORY, "descriptor allocation failed"));
    *out = c;
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

void vhsdecode_hip_demod_destroy(vhsdecode_hip_demod_context* c) {
  if (!c) return;
  {
    std::lock_guard<std::mutex> lock(c->mu);
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
  }
  delete c;
}

int32_t vhsdecode_hip_demod_span_device(vhsdecode_hip_demod_context* c, const int16_t* rf, int64_t first_sample,
                                        int32_t nblocks, const vhsdecode_hip_span_v1* out, void* stream) {
  try {
    if (!c) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "context is NULL");
    if (!out || out->struct_size != sizeof(vhsdecode_hip_span_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "span struct_size mismatch");
    std::lock_guard<std::mutex> lock(c->mu);
    VHSD_HIP(hipSetDevice(c->device));
    return demod_span_locked(c, rf, first_sample, nblocks, out, pick_stream(c, stream));
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

// ---------------------------------------------------------------- batched segments
// A batch decodes the whole blocks of many segments of mixed formats and sample rates.
// Every block gets a global index; per-block arrays (input and output offsets, flags,
// stash and scratch) are indexed by it, and each context contributes one argument
// record per kernel kind pointing at those shared arrays.  The whole batch is two
// launches whatever the mix of formats: the first stage runs the demod (plain, RF
// boost or LaserDisc by the record) of every block and the chroma filter (by the
// record's section count) of every block that has one; the second runs, per block,
// the repair (flagged blocks), the video stage and the sub-deemphasis.  The launch
// sequence is captured once into a HIP graph and replayed while the RF and plane
// pointers stay the same; the records are rebuilt (outside any capture) when they
// change.
struct vhsdecode_hip_batch {
  std::mutex mu;
  int device = 0;
  std::vector<vhsdecode_hip_demod_context*> groups;  // one per context
  std::vector<int> h_rec;                            // block -> group
  std::vector<long long> h_starts, h_offs;
  DevBuf<int> rec;
  DevBuf<long long> starts, offs;
  DevBuf<unsigned> flags;
  DevBuf<float2> stash, aux, rep_a;
  DevBuf<float> rep_dm;
  DevBuf<vhsd::DemodArgs> r_demod;
  DevBuf<vhsd::RepairArgs> r_repair;
  DevBuf<vhsd::SubDeemphArgs> r_sd;
  DevBuf<vhsd::ChromaArgs> r_chroma;
  // first stage: every demod kind and chroma section count in one launch
  DevBuf<int> s1_demod_ids, s1_chroma_ids, rec_kind, rec_sections;
  // second stage: repair, video and sub-deemphasis of every block in one launch
  DevBuf<int> rec_repair, rec_subdeemph;
  int n_demod = 0, n_chroma = 0;
  long long blocks = 0, out_end = 0, rf_end = 0;
  bool use_graph = false;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* key_rf = nullptr;
  vhsdecode_hip_span_v1 key_out{};
  bool have_key = false;

  ~vhsdecode_hip_batch() {
    (void)hipSetDevice(device);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    rec.release(); starts.release(); offs.release(); flags.release(); stash.release(); aux.release();
    rep_a.release(); rep_dm.release(); r_demod.release(); r_repair.release(); r_sd.release(); r_chroma.release();
    s1_demod_ids.release(); s1_chroma_ids.release(); rec_kind.release(); rec_sections.release();
    rec_repair.release(); rec_subdeemph.release();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

bool same_span(const vhsdecode_hip_span_v1& a, const vhsdecode_hip_span_v1& b) {
  return a.video == b.video && a.video_lowpass == b.video_lowpass && a.envelope == b.envelope &&
         a.chroma == b.chroma && a.pilot == b.pilot && a.audio_left == b.audio_left && a.audio_right == b.audio_right;
}

// Rebuild the records for these pointers and the per-kind block lists (synchronous).
int32_t batch_prepare(vhsdecode_hip_batch* b, const int16_t* rf, const vhsdecode_hip_span_v1* out) {
  const size_t g = b->groups.size();
  std::vector<ChainArgs> ca(g);
  std::vector<vhsd::DemodArgs> d(g);
  std::vector<vhsd::RepairArgs> r(g);
  std::vector<vhsd::SubDeemphArgs> sd(g);
  std::vector<vhsd::ChromaArgs> ch(g);
  for (size_t i = 0; i < g; i++) {
    build_chain(b->groups[i], rf, b->starts.p, b->offs.p, out, b->flags.p, b->stash.p, b->aux.p, b->rep_a.p,
                b->rep_dm.p, &ca[i]);
    d[i] = ca[i].d;
    r[i] = ca[i].r;
    sd[i] = ca[i].sd;
    ch[i] = ca[i].ch;
  }
  VHSD_HIP(hipMemcpy(b->r_demod.p, d.data(), sizeof(vhsd::DemodArgs) * g, hipMemcpyHostToDevice));
  VHSD_HIP(hipMemcpy(b->r_repair.p, r.data(), sizeof(vhsd::RepairArgs) * g, hipMemcpyHostToDevice));
  VHSD_HIP(hipMemcpy(b->r_sd.p, sd.data(), sizeof(vhsd::SubDeemphArgs) * g, hipMemcpyHostToDevice));
  VHSD_HIP(hipMemcpy(b->r_chroma.p, ch.data(), sizeof(vhsd::ChromaArgs) * g, hipMemcpyHostToDevice));
  // first stage: the demod of every block and the chroma of every block that has it
  std::vector<int> s1_demod, s1_chroma, rkind(g), rsec(g);
  for (size_t i = 0; i < g; i++) {
    rkind[i] = ca[i].demod_kind;
    rsec[i] = ca[i].chroma_sections;
  }
  for (long long blk = 0; blk < b->blocks; blk++) {
    s1_demod.push_back((int)blk);
    if (ca[b->h_rec[blk]].chroma) s1_chroma.push_back((int)blk);
  }
  b->n_demod = (int)s1_demod.size();
  b->n_chroma = (int)s1_chroma.size();
  VHSD_HIP(b->s1_demod_ids.ensure(std::max<size_t>(s1_demod.size(), 1)));
  VHSD_HIP(b->s1_chroma_ids.ensure(std::max<size_t>(s1_chroma.size(), 1)));
  VHSD_HIP(b->rec_kind.ensure(g));
  VHSD_HIP(b->rec_sections.ensure(g));
  if (!s1_demod.empty())
    VHSD_HIP(hipMemcpy(b->s1_demod_ids.p, s1_demod.data(), sizeof(int) * s1_demod.size(), hipMemcpyHostToDevice));
  if (!s1_chroma.empty())
    VHSD_HIP(hipMemcpy(b->s1_chroma_ids.p, s1_chroma.data(), sizeof(int) * s1_chroma.size(),
                       hipMemcpyHostToDevice));
  VHSD_HIP(hipMemcpy(b->rec_kind.p, rkind.data(), sizeof(int) * g, hipMemcpyHostToDevice));
  VHSD_HIP(hipMemcpy(b->rec_sections.p, rsec.data(), sizeof(int) * g, hipMemcpyHostToDevice));
  // second stage: repair (flagged blocks of records that have it), video, sub-deemphasis
  std::vector<int> rrep(g), rsd(g);
  for (size_t i = 0; i < g; i++) {
    rrep[i] = ca[i].repair ? 1 : 0;
    rsd[i] = ca[i].subdeemph ? 1 : 0;
  }
  VHSD_HIP(b->rec_repair.ensure(g));
  VHSD_HIP(b->rec_subdeemph.ensure(g));
  VHSD_HIP(hipMemcpy(b->rec_repair.p, rrep.data(), sizeof(int) * g, hipMemcpyHostToDevice));
  VHSD_HIP(hipMemcpy(b->rec_subdeemph.p, rsd.data(), sizeof(int) * g, hipMemcpyHostToDevice));
  return VHSDECODE_HIP_OK;
}

// Kernel launches only: safe to capture.
int32_t batch_enqueue(vhsdecode_hip_batch* b, hipStream_t st) {
  VHSD_HIP(vhsd_launch_fill_u32(b->flags.p, b->blocks, 0u, st));
  const vhsd::StageTable t1{b->s1_demod_ids.p, b->n_demod, b->s1_chroma_ids.p, b->rec.p, b->rec_kind.p,
                            b->rec_sections.p};
  VHSD_HIP(vhsd_launch_stage1_table(b->r_demod.p, b->r_chroma.p, t1, b->n_chroma, st));
  const vhsd::StageTable2 t2{b->rec.p, b->rec_repair.p, b->rec_subdeemph.p};
  VHSD_HIP(vhsd_launch_stage2_table(b->r_repair.p, b->r_demod.p, b->r_sd.p, t2, (int)b->blocks, st));
  return VHSDECODE_HIP_OK;
}

}  // namespace

int32_t vhsdecode_hip_batch_create(const vhsdecode_hip_batch_config_v1* cfg, vhsdecode_hip_batch** out) {
  try {
    if (!cfg || !out) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "config and batch must not be NULL");
    *out = nullptr;
    if (cfg->struct_size != sizeof(vhsdecode_hip_batch_config_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "batch_config struct_size mismatch");
    if (cfg->count < 0 || (cfg->count > 0 && !cfg->segments))
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "segments missing");
    auto b = std::make_unique<vhsdecode_hip_batch>();
    b->use_graph = cfg->use_graph != 0;
    std::vector<std::pair<long long, long long>> ranges;
    bool need_aux = false, need_rep = false;
    for (int i = 0; i < cfg->count; i++) {
      const vhsdecode_hip_batch_segment_v1& sg = cfg->segments[i];
      if (sg.struct_size != sizeof(vhsdecode_hip_batch_segment_v1))
        return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "batch_segment struct_size mismatch");
      vhsdecode_hip_demod_context* c = sg.context;
      if (!c || sg.rf_offset < 0 || (sg.rf_offset & 1) || sg.length < 0 || sg.out_offset < 0)
        return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "segment needs a context, an even rf_offset and non-negative sizes");
      if (!b->groups.empty() && c->device != b->device)
        return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "contexts on different devices");
      // whole blocks only, as RfBlockStreamDecoder reads them (a short read ends the stream)
      const long long nb = sg.length >= c->n ? (sg.length - c->n) / c->stride + 1 : 0;
      if (nb == 0) continue;
      if (c->has_audio && (sg.out_offset % c->audio_dec))
        return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "LaserDisc audio needs out_offset divisible by the decimation");
      b->device = c->device;
      size_t gi = 0;
      while (gi < b->groups.size() && b->groups[gi] != c) gi++;
      if (gi == b->groups.size()) b->groups.push_back(c);
      need_aux |= c->has_boost || c->has_notch || c->has_audio;
      need_rep |= !c->ld && c->diff_max > 0.0;
      for (long long k = 0; k < nb; k++) {
        b->h_starts.push_back(sg.rf_offset + k * c->stride);
        b->h_offs.push_back(sg.out_offset + k * c->stride);
        b->h_rec.push_back((int)gi);
      }
      ranges.emplace_back(sg.out_offset, sg.out_offset + nb * c->stride);
      b->out_end = std::max(b->out_end, sg.out_offset + nb * c->stride);
      b->rf_end = std::max(b->rf_end, sg.rf_offset + (nb - 1) * c->stride + c->n);
    }
    b->blocks = (long long)b->h_starts.size();
    if (b->blocks > (1LL << 30)) return fail(VHSDECODE_HIP_E_CAPACITY, "too many blocks in one batch");
    std::sort(ranges.begin(), ranges.end());
    for (size_t i = 1; i < ranges.size(); i++)
      if (ranges[i].first < ranges[i - 1].second)
        return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "segment output ranges overlap");
    VHSD_HIP(hipSetDevice(b->device));
    VHSD_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    const size_t nb = (size_t)std::max<long long>(b->blocks, 1), ng = std::max<size_t>(b->groups.size(), 1);
    VHSD_HIP(b->starts.ensure(nb));
    VHSD_HIP(b->offs.ensure(nb));
    VHSD_HIP(b->rec.ensure(nb));
    VHSD_HIP(b->flags.ensure(nb));
    VHSD_HIP(b->stash.ensure(nb * vhsd::kStashEntries));
    if (need_aux) VHSD_HIP(b->aux.ensure(nb * vhsd::kAuxEntries));
    if (need_rep) {
      VHSD_HIP(b->rep_a.ensure(nb * 32768));
      VHSD_HIP(b->rep_dm.ensure(nb * 65536));
    }
    VHSD_HIP(b->r_demod.ensure(ng));
    VHSD_HIP(b->r_repair.ensure(ng));
    VHSD_HIP(b->r_sd.ensure(ng));
    VHSD_HIP(b->r_chroma.ensure(ng));
    if (b->blocks > 0) {
      VHSD_HIP(hipMemcpy(b->starts.p, b->h_starts.data(), sizeof(long long) * b->blocks, hipMemcpyHostToDevice));
      VHSD_HIP(hipMemcpy(b->offs.p, b->h_offs.data(), sizeof(long long) * b->blocks, hipMemcpyHostToDevice));
      VHSD_HIP(hipMemcpy(b->rec.p, b->h_rec.data(), sizeof(int) * b->blocks, hipMemcpyHostToDevice));
    }
    *out = b.release();
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_batch_info(const vhsdecode_hip_batch* b, int64_t* blocks, int64_t* output_samples,
                                 int64_t* rf_samples) {
  if (!b) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "batch is NULL");
  if (blocks) *blocks = b->blocks;
  if (output_samples) *output_samples = b->out_end;
  if (rf_samples) *rf_samples = b->rf_end;
  return VHSDECODE_HIP_OK;
}

int32_t vhsdecode_hip_batch_run_device(vhsdecode_hip_batch* b, const int16_t* rf, int64_t rf_count,
                                       const vhsdecode_hip_span_v1* out, void* stream) {
  try {
    if (!b || !rf || !out) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "batch, rf and span are required");
    if (out->struct_size != sizeof(vhsdecode_hip_span_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "span struct_size mismatch");
    if (!out->video || !out->video_lowpass || !out->envelope)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "the video, video_lowpass and envelope planes are required");
    if (rf_count < b->rf_end) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "rf buffer is shorter than the segments");
    std::lock_guard<std::mutex> lock(b->mu);
    if (b->blocks == 0) return VHSDECODE_HIP_OK;
    VHSD_HIP(hipSetDevice(b->device));
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : b->stream;
    if (!b->have_key || b->key_rf != rf || !same_span(b->key_out, *out)) {
      // the previous run may still read the records: let it finish before rewriting them
      VHSD_HIP(hipDeviceSynchronize());
      const int32_t rc = batch_prepare(b, rf, out);
      if (rc != VHSDECODE_HIP_OK) return rc;
      if (b->exec) (void)hipGraphExecDestroy(b->exec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
      b->key_rf = rf;
      b->key_out = *out;
      b->have_key = true;
      if (b->use_graph) {
        VHSD_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int32_t erc = batch_enqueue(b, b->stream);
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(b->stream, &graph);
        if (erc != VHSDECODE_HIP_OK) {
          if (graph) (void)hipGraphDestroy(graph);
          b->have_key = false;
          return erc;
        }
        VHSD_HIP(ce);
        b->graph = graph;
        VHSD_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      }
    }
    if (b->use_graph) {
      VHSD_HIP(hipGraphLaunch(b->exec, st));
      return VHSDECODE_HIP_OK;
    }
    return batch_enqueue(b, st);
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

void vhsdecode_hip_batch_destroy(vhsdecode_hip_batch* b) {
  if (!b) return;
  {
    std::lock_guard<std::mutex> lock(b->mu);
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    (void)hipDeviceSynchronize();
  }
  delete b;
}

int32_t vhsdecode_hip_demod_span_host(vhsdecode_hip_demod_context* c, const int16_t* rf, int64_t rf_count,
                                      int64_t first_sample, int32_t nblocks, const vhsdecode_hip_span_v1* out) {
  try {
    if (!c) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "context is NULL");
    if (!out || out->struct_size != sizeof(vhsdecode_hip_span_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "span struct_size mismatch");
    if (!rf || nblocks <= 0 || first_sample < 0) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid span request");
    std::lock_guard<std::mutex> lock(c->mu);
    VHSD_HIP(hipSetDevice(c->device));
    const int64_t need = first_sample + (int64_t)(nblocks - 1) * c->stride + c->n;
    if (need > rf_count) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "rf buffer is shorter than the requested blocks");
    const int64_t in_count = need - first_sample;
    const size_t plane = (size_t)nblocks * c->stride;
    VHSD_HIP(c->s_rf.ensure((size_t)in_count));
    const size_t audio_plane = c->has_audio ? plane / c->audio_dec : 0;
    VHSD_HIP(c->s_planes.ensure(plane * 5 + 2 * audio_plane));
    VHSD_HIP(c->s_flags.ensure((size_t)nblocks));
    hipStream_t st = c->stream;
    VHSD_HIP(hipMemcpyAsync(c->s_rf.p, rf + first_sample, sizeof(int16_t) * in_count, hipMemcpyHostToDevice, st));
    vhsdecode_hip_span_v1 dev{};
    dev.struct_size = sizeof(dev);
    dev.video = c->s_planes.p;
    dev.video_lowpass = c->s_planes.p + plane;
    dev.envelope = c->s_planes.p + 2 * plane;
    dev.chroma = out->chroma ? c->s_planes.p + 3 * plane : nullptr;
    dev.block_flags = c->s_flags.p;
    dev.pilot = out->pilot ? c->s_planes.p + 4 * plane : nullptr;
    if (c->has_audio && out->audio_left && out->audio_right) {
      dev.audio_left = c->s_planes.p + 5 * plane;
      dev.audio_right = dev.audio_left + audio_plane;
    }
    int32_t rc = demod_span_locked(c, c->s_rf.p, 0, nblocks, &dev, st);
    if (rc != 0) return rc;
    if (out->video)
      VHSD_HIP(hipMemcpyAsync(out->video, dev.video, sizeof(float) * plane, hipMemcpyDeviceToHost, st));
    if (out->video_lowpass)
      VHSD_HIP(hipMemcpyAsync(out->video_lowpass, dev.video_lowpass, sizeof(float) * plane, hipMemcpyDeviceToHost, st));
    if (out->envelope)
      VHSD_HIP(hipMemcpyAsync(out->envelope, dev.envelope, sizeof(float) * plane, hipMemcpyDeviceToHost, st));
    if (out->chroma && dev.chroma)
      VHSD_HIP(hipMemcpyAsync(out->chroma, dev.chroma, sizeof(float) * plane, hipMemcpyDeviceToHost, st));
    if (out->block_flags)
      VHSD_HIP(hipMemcpyAsync(out->block_flags, dev.block_flags, sizeof(unsigned) * nblocks, hipMemcpyDeviceToHost, st));
    if (out->pilot && dev.pilot)
      VHSD_HIP(hipMemcpyAsync(out->pilot, dev.pilot, sizeof(float) * plane, hipMemcpyDeviceToHost, st));
    if (dev.audio_left) {
      VHSD_HIP(hipMemcpyAsync(out->audio_left, dev.audio_left, sizeof(float) * audio_plane, hipMemcpyDeviceToHost, st));
      VHSD_HIP(hipMemcpyAsync(out->audio_right, dev.audio_right, sizeof(float) * audio_plane, hipMemcpyDeviceToHost, st));
    }
    VHSD_HIP(hipStreamSynchronize(st));
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_find_pulses(vhsdecode_hip_demod_context* c, const float* sync_ref, int64_t length,
                                  double threshold, int32_t min_length, int32_t max_length, int32_t* starts,
                                  int32_t* lengths, int32_t capacity, int32_t* count) {
  try {
    if (!c || !sync_ref || !count || capacity < 0 || length < 0 || ((!starts || !lengths) && capacity > 0))
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid pulse request");
    std::lock_guard<std::mutex> lock(c->mu);
    VHSD_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int tiles = (int)((length + 4095) / 4096);
    const size_t edge_cap = (size_t)(length / 2 + 2);
    VHSD_HIP(c->tile_counts.ensure(std::max(tiles, 1)));
    VHSD_HIP(c->counters.ensure(2));
    VHSD_HIP(c->edge_pos.ensure(edge_cap));
    VHSD_HIP(c->edge_kind.ensure(edge_cap));
    VHSD_HIP(c->p_starts.ensure(std::max(capacity, 1)));
    VHSD_HIP(c->p_lens.ensure(std::max(capacity, 1)));
    ProfScope ps(c, st, VHSDECODE_HIP_KERNEL_PULSES, length);
    VHSD_HIP(vhsd_launch_pulses(sync_ref, length, c->bias, threshold, min_length, max_length, c->tile_counts.p, tiles,
                                c->edge_pos.p, c->edge_kind.p, (int)edge_cap, c->counters.p, c->p_starts.p,
                                c->p_lens.p, capacity, st));
    int host_counters[2] = {0, 0};
    VHSD_HIP(hipMemcpyAsync(host_counters, c->counters.p, sizeof(host_counters), hipMemcpyDeviceToHost, st));
    VHSD_HIP(hipStreamSynchronize(st));
    const int n = std::min(host_counters[1], capacity);
    if (n > 0) {
      VHSD_HIP(hipMemcpy(starts, c->p_starts.p, sizeof(int) * n, hipMemcpyDeviceToHost));
      VHSD_HIP(hipMemcpy(lengths, c->p_lens.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    }
    *count = host_counters[1];
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_refine_hsync(vhsdecode_hip_demod_context* c, const float* sync_ref, int64_t length,
                                   const double* locations, const uint8_t* filled, int32_t lines,
                                   const vhsdecode_hip_hsync_params_v1* params, double* refined, uint8_t* filled_out) {
  try {
    if (!c || !sync_ref || !locations || !filled || !params || !refined || !filled_out || lines < 0)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid hsync request");
    if (params->struct_size != sizeof(vhsdecode_hip_hsync_params_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "hsync_params struct_size mismatch");
    if (lines == 0) return VHSDECODE_HIP_OK;
    std::lock_guard<std::mutex> lock(c->mu);
    VHSD_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    VHSD_HIP(c->h_loc.ensure(lines));
    VHSD_HIP(c->h_ref.ensure(lines));
    VHSD_HIP(c->h_fil.ensure(lines));
    VHSD_HIP(c->h_fo.ensure(lines));
    VHSD_HIP(c->h_ptrs.ensure(6 * sizeof(void*) + 2 * sizeof(long long)));
    VHSD_HIP(hipMemcpyAsync(c->h_loc.p, locations, sizeof(double) * lines, hipMemcpyHostToDevice, st));
    VHSD_HIP(hipMemcpyAsync(c->h_fil.p, filled, lines, hipMemcpyHostToDevice, st));
    // pointer tables for the batched kernel (one field here)
    struct Table {
      const float* span;
      const double* loc;
      const unsigned char* fil;
      double* ref;
      unsigned char* fo;
      long long len;
      int lines;
      int pad;
    } t{sync_ref, c->h_loc.p, c->h_fil.p, c->h_ref.p, c->h_fo.p, (long long)length, lines, 0};
    static_assert(sizeof(Table) <= 6 * sizeof(void*) + 2 * sizeof(long long), "table size");
    VHSD_HIP(hipMemcpyAsync(c->h_ptrs.p, &t, sizeof(t), hipMemcpyHostToDevice, st));
    Table* dt = reinterpret_cast<Table*>(c->h_ptrs.p);
    vhsd::HsyncParams p{params->threshold, params->sample_rate_mhz, params->hsync_pulse_us, params->ire0,
                        params->hz_ire, params->is_pal, params->use_right_hsync};
    {
      ProfScope ps(c, st, VHSDECODE_HIP_KERNEL_HSYNC, 1);
      VHSD_HIP(c->h_work.ensure(vhsd_hsync_work_bytes(1, lines)));
      VHSD_HIP(vhsd_launch_hsync(&dt->span, &dt->len, &dt->loc, &dt->fil, &dt->lines, &p, &dt->ref, &dt->fo, 1, lines,
                                 c->bias, c->h_work.p, nullptr, nullptr, st));
    }
    VHSD_HIP(hipMemcpyAsync(refined, c->h_ref.p, sizeof(double) * lines, hipMemcpyDeviceToHost, st));
    VHSD_HIP(hipMemcpyAsync(filled_out, c->h_fo.p, lines, hipMemcpyDeviceToHost, st));
    VHSD_HIP(hipStreamSynchronize(st));
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_set_sinc_table(vhsdecode_hip_demod_context* c, const float* table, int64_t count) {
  try {
    if (!c || !table || count != (int64_t)65537 * 16)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "sinc table must hold 65537 x 16 floats");
    std::lock_guard<std::mutex> lock(c->mu);
    VHSD_HIP(hipSetDevice(c->device));
    VHSD_HIP(c->lut.ensure((size_t)count));
    VHSD_HIP(hipMemcpy(c->lut.p, table, sizeof(float) * count, hipMemcpyHostToDevice));
    c->lut_ready = true;
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

static int32_t tbc_params(const vhsdecode_hip_tbc_params_v1* p, double nominal, vhsd::TbcParams* t) {
  if (p->output_line_length <= 0 || p->first_line < 0 || p->line_count < 0)
    return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid TBC geometry");
  t->out_len = p->output_line_length;
  t->first_line = p->first_line;
  t->line_count = p->line_count;
  t->nominal = nominal;
  t->has_conv = p->convert != 0;
  t->ire0 = p->ire0;
  t->out_scale = p->output_scale;
  t->inv_hz_ire = 1.0 / p->hz_ire;
  t->conv_offset = (p->output_zero + 0.5) - (p->vsync_ire * p->output_scale);
  return VHSDECODE_HIP_OK;
}

int32_t vhsdecode_hip_tbc_resample(vhsdecode_hip_demod_context* c, const float* source, int64_t source_length,
                                   const double* locations, int32_t location_count,
                                   const vhsdecode_hip_tbc_params_v1* params, uint16_t* out_u16, double* out_f64) {
  try {
    if (!c || !source || !locations || !params || source_length <= 0)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid TBC request");
    if (params->struct_size != sizeof(vhsdecode_hip_tbc_params_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "tbc_params struct_size mismatch");
    if (location_count < 2) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "at least two line locations are required");
    for (int i = 0; i < location_count; i++)
      if (!std::isfinite(locations[i])) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "line locations must be finite");
    if (params->first_line < 0 || params->first_line + params->line_count >= location_count)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "first_line/line_count outside the line locations");
    std::lock_guard<std::mutex> lock(c->mu);
    VHSD_HIP(hipSetDevice(c->device));
    int32_t rc = ensure_lut(c);
    if (rc) return rc;
    // nominal line length and per-line level adjusts (TbcLineResampler.cs:508-509, 818-878, 997-1011)
    double nominal = params->nominal_line_length;
    if (!(nominal > 0.0)) {
      std::vector<double> l(location_count - 1);
      for (int i = 0; i + 1 < location_count; i++) l[i] = locations[i + 1] - locations[i];
      nominal = std::max(1.0, median(l));
    }
    const int L = params->output_line_length;
    const int lines = params->first_line + params->line_count;
    std::vector<double> f(lines), dev(lines);
    const double inv = 1.0 / nominal;
    for (int line = 0; line < lines; line++) {
      int left = std::clamp(line, 0, location_count - 2);
      f[line] = (locations[left + 1] * inv) - (locations[left] * inv);
    }
    std::vector<double> adj_line(lines, 0.0);
    if (lines > 0) {
      double med = median(f);
      for (int i = 0; i < lines; i++) dev[i] = std::fabs(f[i] - med);
      double mad = median(dev);
      double thr = mad > 0.0 ? 15.0 * mad : 0.001;
      for (int i = 0; i < lines; i++) adj_line[i] = std::fabs(f[i] - med) > thr ? med : f[i];
    }
    const bool per_sample = params->wow_level_smoothing > 0.0;
    std::vector<double> adj;
    if (per_sample) {
      adj.assign((size_t)lines * L, 0.0);
      for (int i = 0; i < lines; i++) std::fill(adj.begin() + (size_t)i * L, adj.begin() + (size_t)(i + 1) * L, adj_line[i]);
      const double alpha = 1.0 / (params->wow_level_smoothing * L);
      for (size_t i = 1; i < adj.size(); i++) adj[i] = std::fma(adj[i] - adj[i - 1], alpha, adj[i - 1]);
    } else {
      adj = adj_line;
    }
    vhsd::TbcParams t{};
    if ((rc = tbc_params(params, nominal, &t)) != 0) return rc;
    const size_t dest = (size_t)L * params->line_count;
    hipStream_t st = c->stream;
    VHSD_HIP(c->t_loc.ensure(location_count));
    VHSD_HIP(c->t_adj.ensure(std::max<size_t>(adj.size(), 1)));
    VHSD_HIP(c->t_u16.ensure(std::max<size_t>(dest, 1)));
    VHSD_HIP(c->t_f64.ensure(std::max<size_t>(dest, 1)));
    VHSD_HIP(hipMemcpyAsync(c->t_loc.p, locations, sizeof(double) * location_count, hipMemcpyHostToDevice, st));
    if (!adj.empty())
      VHSD_HIP(hipMemcpyAsync(c->t_adj.p, adj.data(), sizeof(double) * adj.size(), hipMemcpyHostToDevice, st));
    {
      ProfScope ps(c, st, VHSDECODE_HIP_KERNEL_TBC, (long long)dest);
      VHSD_HIP(vhsd_launch_tbc(source, source_length, c->bias, c->t_loc.p, c->t_adj.p, per_sample ? 0 : 1, &t, c->lut.p,
                               out_u16 ? c->t_u16.p : nullptr, out_f64 ? c->t_f64.p : nullptr, st));
    }
    if (out_u16 && dest)
      VHSD_HIP(hipMemcpyAsync(out_u16, c->t_u16.p, sizeof(uint16_t) * dest, hipMemcpyDeviceToHost, st));
    if (out_f64 && dest)
      VHSD_HIP(hipMemcpyAsync(out_f64, c->t_f64.p, sizeof(double) * dest, hipMemcpyDeviceToHost, st));
    VHSD_HIP(hipStreamSynchronize(st));
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_tbc_resample_device(vhsdecode_hip_demod_context* c, const float* source, int64_t source_length,
                                          const double* locations, const double* level_adjust, int32_t per_sample,
                                          const vhsdecode_hip_tbc_params_v1* params, uint16_t* out_u16,
                                          double* out_f64, void* stream) {
  try {
    if (!c || !source || !locations || !level_adjust || !params || source_length <= 0)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid TBC request");
    if (params->struct_size != sizeof(vhsdecode_hip_tbc_params_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "tbc_params struct_size mismatch");
    if (!(params->nominal_line_length > 0.0))
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "the device variant needs an explicit nominal line length");
    std::lock_guard<std::mutex> lock(c->mu);
    VHSD_HIP(hipSetDevice(c->device));
    int32_t rc = ensure_lut(c);
    if (rc) return rc;
    vhsd::TbcParams t{};
    if ((rc = tbc_params(params, params->nominal_line_length, &t)) != 0) return rc;
    hipStream_t st = pick_stream(c, stream);
    ProfScope ps(c, st, VHSDECODE_HIP_KERNEL_TBC, (long long)params->output_line_length * params->line_count);
    VHSD_HIP(vhsd_launch_tbc(source, source_length, c->bias, locations, level_adjust, per_sample ? 0 : 1, &t, c->lut.p,
                             out_u16, out_f64, st));
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_decode_fields_device(vhsdecode_hip_demod_context* c, const float* sync_ref, const float* video,
                                           int64_t span_length, const double* line_locations, const uint8_t* filled,
                                           const int64_t* origins, int32_t nfields,
                                           const vhsdecode_hip_field_params_v1* p, uint16_t* out,
                                           double* refined_out, void* stream) {
  try {
    if (!c || !sync_ref || !video || !line_locations || !filled || !p || !out || nfields < 0 || span_length <= 0)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid field decode request");
    if (p->struct_size != sizeof(vhsdecode_hip_field_params_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "field_params struct_size mismatch");
    if (p->field_lines <= 0 || p->field_lines > 511 || p->output_line_length <= 0 || !(p->nominal_line_length > 0.0))
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid field geometry");
    if (nfields == 0) return VHSDECODE_HIP_OK;
    const int nloc = p->field_lines + 1;
    const size_t nl = (size_t)nfields * nloc;
    for (size_t i = 0; i < nl; i++)
      if (!std::isfinite(line_locations[i]))
        return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "line locations must be finite");
    if (origins)
      for (int f = 0; f < nfields; f++)
        if (origins[f] < 0 || origins[f] >= span_length)
          return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "field origins must lie inside the span");
    std::lock_guard<std::mutex> lock(c->mu);
    VHSD_HIP(hipSetDevice(c->device));
    int32_t rc = ensure_lut(c);
    if (rc) return rc;
    hipStream_t st = pick_stream(c, stream);
    VHSD_HIP(c->f_loc.ensure(nl));
    VHSD_HIP(c->f_ref.ensure(nl));
    c->f_ref_nloc = nloc;
    VHSD_HIP(c->f_fil.ensure(nl));
    VHSD_HIP(c->f_fo.ensure(nl));
    VHSD_HIP(c->f_x.ensure((size_t)nfields * p->field_lines));
    VHSD_HIP(c->f_y0.ensure((size_t)nfields * p->field_lines));
    // per-field pointer tables for the batched hsync kernel
    const size_t tab_bytes = (size_t)nfields * 60;
    VHSD_HIP(c->f_tab.ensure(tab_bytes));
    // smoothing powers table: pow[s] = (1 - alpha)^(s + 1), s < output_line_length
    const bool smooth = p->wow_level_smoothing > 0.0;
    const double alpha = smooth ? 1.0 / (p->wow_level_smoothing * p->output_line_length) : 0.0;
    if (smooth && (c->f_pow_alpha != alpha || c->f_pow_len != p->output_line_length)) {
      std::vector<double> pw(p->output_line_length);
      for (int s2 = 0; s2 < p->output_line_length; s2++) pw[s2] = std::pow(1.0 - alpha, (double)(s2 + 1));
      VHSD_HIP(c->f_pow.ensure(pw.size()));
      VHSD_HIP(hipMemcpy(c->f_pow.p, pw.data(), sizeof(double) * pw.size(), hipMemcpyHostToDevice));
      c->f_pow_alpha = alpha;
      c->f_pow_len = p->output_line_length;
    }
    const double line_decay = smooth ? std::pow(1.0 - alpha, (double)p->output_line_length) : 0.0;
    // stage locations, flags and pointer tables in pinned memory, one async upload
    const size_t bytes = nl * sizeof(double) + nl + tab_bytes + 64;
    if (c->f_stage_done) VHSD_HIP(hipEventSynchronize(c->f_stage_done));
    if (bytes > c->f_stage_bytes) {
      if (c->f_stage) (void)hipHostFree(c->f_stage);
      c->f_stage = nullptr;
      VHSD_HIP(hipHostMalloc(&c->f_stage, bytes, hipHostMallocDefault));
      c->f_stage_bytes = bytes;
    }
    if (!c->f_stage_done) VHSD_HIP(hipEventCreateWithFlags(&c->f_stage_done, hipEventDisableTiming));
    unsigned char* stg = reinterpret_cast<unsigned char*>(c->f_stage);
    std::memcpy(stg, line_locations, nl * sizeof(double));
    std::memcpy(stg + nl * sizeof(double), filled, nl);
    unsigned char* tab = stg + nl * sizeof(double) + nl;
    tab = reinterpret_cast<unsigned char*>((reinterpret_cast<uintptr_t>(tab) + 15) & ~uintptr_t(15));
    const size_t tab_off = (size_t)(tab - stg);
    // table layout (byte offsets, 8-byte slots): spans, lens, locs, fil, ref, fo, then int line counts
    const size_t o_sp = 0, o_ln = 8 * (size_t)nfields, o_lc = 16 * (size_t)nfields, o_fi = 24 * (size_t)nfields,
                 o_rf = 32 * (size_t)nfields, o_fo = 40 * (size_t)nfields, o_org = 48 * (size_t)nfields,
                 o_cnt = 56 * (size_t)nfields;
    const size_t used_tab = o_cnt + 4 * (size_t)nfields;
    if (tab_off + used_tab > bytes) return fail(VHSDECODE_HIP_E_INTERNAL, "staging overflow");
    for (int f = 0; f < nfields; f++) {
      const void* vp[6] = {sync_ref, nullptr, c->f_loc.p + (size_t)f * nloc, c->f_fil.p + (size_t)f * nloc,
                           c->f_ref.p + (size_t)f * nloc, c->f_fo.p + (size_t)f * nloc};
      const long long len = span_length;
      std::memcpy(tab + o_sp + 8 * (size_t)f, &vp[0], 8);
      std::memcpy(tab + o_ln + 8 * (size_t)f, &len, 8);
      std::memcpy(tab + o_lc + 8 * (size_t)f, &vp[2], 8);
      std::memcpy(tab + o_fi + 8 * (size_t)f, &vp[3], 8);
      std::memcpy(tab + o_rf + 8 * (size_t)f, &vp[4], 8);
      std::memcpy(tab + o_fo + 8 * (size_t)f, &vp[5], 8);
      const long long org = origins ? (long long)origins[f] : 0;
      std::memcpy(tab + o_org + 8 * (size_t)f, &org, 8);
      std::memcpy(tab + o_cnt + 4 * (size_t)f, &nloc, 4);
    }
    VHSD_HIP(hipMemcpyAsync(c->f_loc.p, stg, nl * sizeof(double), hipMemcpyHostToDevice, st));
    VHSD_HIP(hipMemcpyAsync(c->f_fil.p, stg + nl * sizeof(double), nl, hipMemcpyHostToDevice, st));
    VHSD_HIP(hipMemcpyAsync(c->f_tab.p, tab, used_tab, hipMemcpyHostToDevice, st));
    VHSD_HIP(hipEventRecord(c->f_stage_done, st));
    unsigned char* dtab = c->f_tab.p;
    const float* const* d_spans = reinterpret_cast<const float* const*>(dtab + o_sp);
    const long long* d_lens = reinterpret_cast<const long long*>(dtab + o_ln);
    const double* const* d_locs = reinterpret_cast<const double* const*>(dtab + o_lc);
    const unsigned char* const* d_fils = reinterpret_cast<const unsigned char* const*>(dtab + o_fi);
    double* const* d_refs = reinterpret_cast<double* const*>(dtab + o_rf);
    unsigned char* const* d_fos = reinterpret_cast<unsigned char* const*>(dtab + o_fo);
    const int* d_lcount = reinterpret_cast<const int*>(dtab + o_cnt);
    const long long* d_org = reinterpret_cast<const long long*>(dtab + o_org);
    vhsd::HsyncParams hp{p->threshold, p->sample_rate_mhz, p->hsync_pulse_us, p->ire0, p->hz_ire, p->is_pal,
                         p->use_right_hsync};
    {
      ProfScope ps(c, st, VHSDECODE_HIP_KERNEL_HSYNC, nfields);
      VHSD_HIP(c->h_work.ensure(vhsd_hsync_work_bytes(nfields, nloc)));
      VHSD_HIP(vhsd_launch_hsync(d_spans, d_lens, d_locs, d_fils, d_lcount, &hp, d_refs, d_fos, nfields, nloc,
                                 c->bias, c->h_work.p, nullptr, nullptr, st));
    }
    const double out_scale = p->output_scale;
    const double conv_off = (p->output_zero + 0.5) - (p->vsync_ire * out_scale);
    {
      ProfScope ps(c, st, VHSDECODE_HIP_KERNEL_TBC, (long long)nfields * p->field_lines * p->output_line_length);
      VHSD_HIP(vhsd_launch_tbc_fields(video, span_length, c->bias, d_org, c->f_ref.p, nloc, p->field_lines, nullptr, 0, 0,
                                      nfields,
                                      p->nominal_line_length, line_decay, smooth ? c->f_pow.p : nullptr, c->f_x.p,
                                      c->f_y0.p, p->output_line_length, smooth ? 1 : 0, p->ire0, out_scale,
                                      1.0 / p->hz_ire, conv_off, c->lut.p, out, nullptr, nullptr, st));
    }
    if (refined_out) {
      VHSD_HIP(hipMemcpyAsync(refined_out, c->f_ref.p, nl * sizeof(double), hipMemcpyDeviceToHost, st));
      VHSD_HIP(hipStreamSynchronize(st));
    }
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

namespace {
// NumpyTrigFloat32 (Decode/VhsChromaDecoder.cs:3834-3893): numpy's float32 sin/cos, a
// Cody-Waite reduction by pi/2 and minimax polynomials in FMA form; arguments past the
// reduction's range take the C library's float sin/cos.
#pragma clang fp contract(off)
float numpy_trig_f32(float x, bool cosine) {
  auto bits = [](uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
  };
  if (std::isnan(x)) return NAN;
  if (std::fabs(x) > (cosine ? 71476.0625f : 117435.992f)) return cosine ? std::cos(x) : std::sin(x);
  float q = x * bits(0x3F22F983u);
  q = (q + 12582912.0f) - 12582912.0f;
  float r = std::fma(q, bits(0xBFC90FD8u), x);
  r = std::fma(q, bits(0xB4A8885Au), r);
  r = std::fma(q, bits(0xA7C234C4u), r);
  const float r2 = r * r;
  float cv = std::fma(bits(0x37CC730Bu), r2, bits(0xBAB6036Eu));
  cv = std::fma(cv, r2, bits(0x3D2AAA9Eu));
  cv = std::fma(cv, r2, -0.5f);
  cv = std::fma(cv, r2, 1.0f);
  float sv = std::fma(bits(0x363E9DDEu), r2, bits(0xB95035DDu));
  sv = std::fma(sv, r2, bits(0x3C0888CDu));
  sv = std::fma(sv, r2, -1.0f / 6.0f);
  sv = std::fma(sv, r2, 0.0f);
  sv = std::fma(sv, r, r);
  int iq = (int)std::nearbyint(q);
  if (cosine) iq++;
  const float v = (iq & 1) == 0 ? sv : cv;
  return (iq & 2) != 0 ? -v : v;
}
}  // namespace

int32_t vhsdecode_hip_set_vhs_chroma(vhsdecode_hip_demod_context* c, const vhsdecode_hip_chroma_params_v1* p) {
#pragma clang fp contract(off)
  try {
    if (!c) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid context");
    std::lock_guard<std::mutex> lock(c->mu);
    if (!p) {
      c->chroma_on = false;
      return VHSDECODE_HIP_OK;
    }
    if (p->struct_size != sizeof(vhsdecode_hip_chroma_params_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "chroma_params struct_size mismatch");
    const int N = p->field_lines, L = p->output_line_length;
    if (N <= 20 || L <= 0 || (long long)N * L > (1LL << 26) || p->sections < 1 ||
        p->sections > vhsd::chroma::kMaxSections || !(p->fsc_mhz > 0.0) || !(p->color_under_hz > 0.0) ||
        !(p->output_rate_hz > 0.0) || p->burst_start < 0 || p->burst_end <= p->burst_start ||
        p->burst_end > L || p->burst_end + 5 >= L || p->input_line_length <= 0)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid chroma geometry");
    VHSD_HIP(hipSetDevice(c->device));
    vhsd::chroma::ChromaCfg& cc = c->ccfg;
    cc = vhsd::chroma::ChromaCfg{};
    cc.is_ntsc = p->is_ntsc ? 1 : 0;
    cc.lines = N;
    cc.line_len = L;
    cc.burst_start = p->burst_start;
    cc.burst_end = p->burst_end;
    cc.win_max = p->burst_end + (p->burst_end - p->burst_start);
    cc.pad = default_pad_length(&p->final_sos[0][0], p->sections);
    cc.probe_ext = cc.win_max + 2 * cc.pad;
    // the whole-field band-pass's chunks and warm-up: twice the samples its slowest pole
    // takes to forget a state to 1e-8, plus a margin (measured fastest on decoded fields:
    // shorter warm-ups leave many chunks to run again, longer ones lengthen every chunk;
    // exactness does not depend on it: a chunk whose warm-up fell short is run again)
    {
      double rmax = 0.0;
      for (int s = 0; s < p->sections; s++) {
        const double a0 = p->final_sos[s][3], a1 = p->final_sos[s][4] / a0, a2 = p->final_sos[s][5] / a0;
        const double disc = a1 * a1 - 4.0 * a2;
        double r;
        if (disc < 0.0) {
          r = std::sqrt(a2);
        } else {
          const double q = std::sqrt(disc);
          r = std::max(std::fabs((-a1 + q) / 2.0), std::fabs((-a1 - q) / 2.0));
        }
        rmax = std::max(rmax, r);
      }
      if (!(rmax < 1.0)) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "the final chroma band-pass is not stable");
      const double decay = std::log(1e-8) / std::log(std::max(rmax, 1e-3));
      cc.iir_warm = (int)std::min(4096.0, std::max(64.0, std::ceil(2.0 * decay) + 64.0));
      cc.iir_chunk = vhsd::chroma::kIirChunk;
    }
    if ((size_t)cc.probe_ext * vhsd::chroma::kProbeThreads * sizeof(float) > 160 * 1024)
      return fail(VHSDECODE_HIP_E_UNSUPPORTED, "burst window too long for the probe kernel");
    cc.sections = p->sections;
    float c32[vhsd::kMaxSections][6];
    for (int s = 0; s < p->sections; s++)
      for (int q = 0; q < 6; q++) c32[s][q] = cc.sos[s][q] = (float)p->final_sos[s][q];
    float zi[2 * vhsd::kMaxSections];
    steady_state_f32(c32, p->sections, zi);
    for (int s = 0; s < p->sections; s++) {
      cc.zi[s][0] = zi[2 * s];
      cc.zi[s][1] = zi[2 * s + 1];
    }
    cc.rot[0] = p->rotation[0];
    cc.rot[1] = p->rotation[1];
    cc.input_line_length = p->input_line_length;
    cc.fsc_ratio = (p->output_rate_hz / 1000000.0) / p->fsc_mhz;
    cc.cuc_hz = p->color_under_hz;
    cc.fsc_mhz = p->fsc_mhz;
    cc.burst_abs_ref = p->burst_abs_ref;
    // BuildHeterodyneTable and BuildCarrierTables at 4 fsc over one field (C library cos,
    // as Math.Cos on Linux)
    const size_t n = (size_t)N * L;
    const double rate = p->fsc_mhz * 4.0;
    const double het_scale = (p->fsc_mhz + p->color_under_hz / 1000000.0) / rate;
    const double wave_scale = p->fsc_mhz / rate;
    std::vector<float> het(4 * n), bs(n), bc(n);
    for (int ph = 0; ph < 4; ph++) {
      const double off = ((M_PI / 2.0) * ph) + 0.0;
      for (size_t i = 0; i < n; i++) het[ph * n + i] = (float)-std::cos(((2.0 * M_PI) * het_scale * (double)i) + off);
    }
    for (size_t i = 0; i < n; i++) {
      const float theta = (float)((2.0 * M_PI) * wave_scale * (double)i);
      bs[i] = numpy_trig_f32(theta, false);
      bc[i] = numpy_trig_f32(theta, true);
    }
    VHSD_HIP(c->c_het.ensure(4 * n));
    VHSD_HIP(c->c_sin.ensure(n));
    VHSD_HIP(c->c_cos.ensure(n));
    VHSD_HIP(hipMemcpy(c->c_het.p, het.data(), 4 * n * sizeof(float), hipMemcpyHostToDevice));
    VHSD_HIP(hipMemcpy(c->c_sin.p, bs.data(), n * sizeof(float), hipMemcpyHostToDevice));
    VHSD_HIP(hipMemcpy(c->c_cos.p, bc.data(), n * sizeof(float), hipMemcpyHostToDevice));
    VHSD_HIP(c->c_state.ensure(sizeof(vhsd::chroma::ChromaState)));
    cc.het = c->c_het.p;
    cc.bsin = c->c_sin.p;
    cc.bcos = c->c_cos.p;
    c->chroma_on = true;
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_decode_vhs_fields_chroma_device(
    vhsdecode_hip_demod_context* c, const float* sync_ref, const float* video, const float* chroma,
    int64_t span_length, int64_t first_begin, int32_t reset_state, int64_t state_shift,
    const vhsdecode_hip_vhs_sync_v1* sy, const vhsdecode_hip_field_params_v1* p, int32_t max_fields,
    int32_t rounds, uint16_t* out, uint16_t* out_chroma, int32_t field_number_base,
    vhsdecode_hip_vhs_field_v1* fields, int32_t* field_count, int64_t* next_begin, void* stream) {
  if (!chroma || !out_chroma || field_number_base < 0)
    return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid chroma plane, output or field number");
  const vhsd_chroma_io io{chroma, out_chroma, field_number_base};
  return vhsd_decode_vhs_fields_from(c, sync_ref, video, span_length, first_begin, reset_state, state_shift, sy, p,
                                     max_fields, rounds, out, fields, field_count, next_begin, stream, 0, &io);
}

int32_t vhsdecode_hip_decode_vhs_fields_device(vhsdecode_hip_demod_context* c, const float* sync_ref,
                                               const float* video, int64_t span_length, int64_t first_begin,
                                               int32_t reset_state, int64_t state_shift,
                                               const vhsdecode_hip_vhs_sync_v1* sy,
                                               const vhsdecode_hip_field_params_v1* p, int32_t max_fields,
                                               int32_t rounds, uint16_t* out, vhsdecode_hip_vhs_field_v1* fields,
                                               int32_t* field_count, int64_t* next_begin, void* stream) {
  return vhsd_decode_vhs_fields_from(c, sync_ref, video, span_length, first_begin, reset_state, state_shift, sy, p,
                                     max_fields, rounds, out, fields, field_count, next_begin, stream, 0);
}

int32_t vhsd_decode_vhs_fields_from(vhsdecode_hip_demod_context* c, const float* sync_ref, const float* video,
                                    int64_t span_length, int64_t first_begin, int32_t reset_state,
                                    int64_t state_shift, const vhsdecode_hip_vhs_sync_v1* sy,
                                    const vhsdecode_hip_field_params_v1* p, int32_t max_fields, int32_t rounds,
                                    uint16_t* out, vhsdecode_hip_vhs_field_v1* fields, int32_t* field_count,
                                    int64_t* next_begin, void* stream, int32_t first_record,
                                    const vhsd_chroma_io* chroma) {
  using namespace vhsd::vsync;
  try {
    if (!c || !sync_ref || !video || !sy || !p || !out || !field_count || max_fields <= 0 || span_length <= 0 ||
        first_begin < 0 || rounds < 1 || rounds > 64 || first_record < 0 || first_record >= max_fields)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid vhs field decode request");
    if (sy->struct_size != sizeof(vhsdecode_hip_vhs_sync_v1) || p->struct_size != sizeof(vhsdecode_hip_field_params_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "vhs_sync or field_params struct_size mismatch");
    if (span_length > INT32_MAX)
      return fail(VHSDECODE_HIP_E_UNSUPPORTED, "spans are limited to 2^31 - 1 samples (int32 pulse positions)");
    const double fs_mhz = c->fs / 1000000.0;
    const int out_lines = p->field_lines;
    const int line_stride = out_lines + 10;
    const int max_first = 4;
    if (out_lines <= 0 || out_lines + max_first + 1 > line_stride || line_stride > 511 ||
        p->output_line_length <= 0 || !(sy->line_period_us > 0.0) || sy->num_pulses <= 0 ||
        sy->block_length <= 0 || sy->block_stride <= 0 || sy->block_cut < 0 || sy->read_samples <= 0 ||
        !(sy->frames_per_second > 0.0) || !(sy->frame_lines > 0.0) || !(p->hz_ire != 0.0))
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid vhs sync or field geometry");
    if ((3.75 * fs_mhz) + 2.0 > kHistBins)
      return fail(VHSDECODE_HIP_E_UNSUPPORTED, "sample rate too high for the hsync-length histogram");
    std::lock_guard<std::mutex> lock(c->mu);
    if (!reset_state && !c->vs_state_ready)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "no vhs field sequence to continue on this context");
    VHSD_HIP(hipSetDevice(c->device));
    int32_t rc = ensure_lut(c);
    if (rc) return rc;
    hipStream_t st = pick_stream(c, stream);
    if (chroma) {
      if (!c->chroma_on) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "no chroma format set on this context");
      if (c->ccfg.lines != out_lines || c->ccfg.line_len != p->output_line_length)
        return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "chroma format and field geometry differ");
      if (!reset_state && !c->c_state_ready)
        return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "no chroma state to continue on this context");
    }
    // nominal-threshold pulses of the span: their vertical-sync groups place the read windows
    const int capacity = (int)std::min<int64_t>(span_length / 16 + 1024, INT32_MAX / 2);
    const int tiles = (int)((span_length + 4095) / 4096);
    const size_t edge_cap_span = (size_t)(span_length / 2 + 2);
    VHSD_HIP(c->tile_counts.ensure(std::max(tiles, 1)));
    VHSD_HIP(c->counters.ensure(2));
    VHSD_HIP(c->edge_pos.ensure(edge_cap_span));
    VHSD_HIP(c->edge_kind.ensure(edge_cap_span));
    VHSD_HIP(c->p_starts.ensure(capacity));
    VHSD_HIP(c->p_lens.ensure(capacity));
    // SyncAnalyzer.FromParameters (Dsp/SyncAnalysis.cs:88-114) and the vhs detector geometry
    SyncCfg cfg{};
    cfg.fs_mhz = fs_mhz;
    double lp_us = sy->line_period_us;
    if (fs_mhz > 0.0) lp_us = std::nearbyint(lp_us * fs_mhz) / fs_mhz;
    cfg.nominal = lp_us * fs_mhz;
    cfg.hsync_us = sy->hsync_pulse_us;
    cfg.eq_us = sy->eq_pulse_us;
    cfg.vsync_us = sy->vsync_pulse_us;
    cfg.hsync_tol = sy->hsync_tolerance_us;
    cfg.eq_tol = sy->eq_tolerance_us;
    cfg.num_pulses = sy->num_pulses;
    cfg.is_pal = sy->is_pal ? 1 : 0;
    cfg.ire0 = p->ire0;
    cfg.hz_ire = p->hz_ire;
    cfg.vsync_ire = p->vsync_ire;
    cfg.field_lines0 = cfg.is_pal ? out_lines - 1 : out_lines;
    cfg.field_lines1 = cfg.is_pal ? out_lines : out_lines - 1;
    cfg.output_lines = out_lines;
    cfg.block_length = sy->block_length;
    cfg.block_cut = sy->block_cut;
    cfg.stride = sy->block_stride;
    cfg.read_samples = sy->read_samples;
    const double vf = sy->frames_per_second * 2.0, hf = sy->frames_per_second * sy->frame_lines;
    cfg.ser_eq_len = (int)std::nearbyint(c->fs * sy->eq_pulse_us * 1e-6);
    cfg.ser_line_len = (int)std::nearbyint(c->fs / hf);
    (void)vf;
    cfg.pulse_min = std::max(0, (int)std::ceil((sy->eq_pulse_us * fs_mhz) / 8.0));
    cfg.pulse_max = std::max(1, (int)(cfg.nominal * 5.0));
    const double field_rate = sy->frames_per_second * 2.0;
    cfg.fs_window = (int)std::nearbyint(field_rate < 60.0 ? field_rate / 5.0 : field_rate / 6.0);
    if (cfg.fs_window < 1 || cfg.fs_window > kAvgCap)
      return fail(VHSDECODE_HIP_E_UNSUPPORTED, "field-state moving-average window out of range");
    cfg.bias = c->bias;
    // the most pulses a read window can hold: each is at least the shortest length the
    // pulse searches keep, and is followed by a sample above the threshold
    const int64_t window_max = sy->read_samples + 2 * (int64_t)sy->block_stride;
    const int min_kept = std::max(1, std::min(cfg.pulse_min, (int)std::ceil(cfg.ser_eq_len / 8.0)));
    const int64_t pulse_cap = window_max / (min_kept + 1) + 16;
    if (pulse_cap > INT32_MAX / 8) return fail(VHSDECODE_HIP_E_UNSUPPORTED, "read windows too long");
    cfg.pulse_cap = (int)pulse_cap;
    cfg.line_stride = line_stride;
    cfg.max_records = max_fields;
    // one allocation for every list, record and table
    const size_t F = (size_t)max_fields;
    size_t off = 0;
    auto take = [&](size_t bytes) {
      const size_t o = off;
      off += (bytes + 255) & ~(size_t)255;
      return o;
    };
    const size_t o_ctl = take(sizeof(SeqCtl)), o_spec = take(4), o_ws = take(8 * F), o_wl = take(8 * F);
    const size_t PC = (size_t)cfg.pulse_cap;
    const size_t o_pst = take(4 * F * PC), o_pln = take(4 * F * PC), o_ovf = take(4 * F);
    const size_t o_search = take(sizeof(LevelSearch) * F), o_levels = take(sizeof(FieldLevels) * F);
    const size_t o_before = take(sizeof(LevelStateDev) * F), o_after = take(sizeof(LevelStateDev));
    const size_t o_raw = take(4 * F), o_logic = take(sizeof(FieldLogic) * F);
    const size_t o_rk = take(4 * F * PC), o_rs = take(4 * F * PC), o_aux = take((size_t)kLogicAux * F * PC);
    const size_t o_rec = take(sizeof(FieldRecord) * F);
    const size_t o_lines = take(8 * F * line_stride);
    const size_t o_tp = take(8 * F * 6), o_lc = take(4 * F), o_org = take(8 * F), o_fl = take(4 * F);
    const size_t o_thr = take(8 * F), o_dc = take(8 * F), o_zero = take(line_stride), o_oi = take(4 * F);
    const size_t o_refb = take(8 * F * line_stride), o_fob = take(F * line_stride), o_rend = take(8 * F * line_stride);
    const size_t o_acc = take(4 * 8 * F * line_stride);
    // segmented pulse searches: segments of the longest read window
    const int segs = (int)((window_max + kSegLen - 1) / kSegLen);
    const size_t FS = F * (size_t)segs;
    const size_t o_smin = take(8 * FS), o_snan = take(4 * FS), o_sthr = take(8 * F * kSpecR);
    const size_t o_sint = take(4 * F * 9), o_sdbl = take(8 * F * 3);
    const size_t o_ssum = take(sizeof(SegSummary) * FS * kSpecR), o_scar = take(4 * FS), o_soff = take(4 * FS);
    VHSD_HIP(c->vs_mem.ensure(off));
    unsigned char* m = c->vs_mem.p;
    const size_t state_hist = (sizeof(LevelStateDev) + 255) & ~(size_t)255;
    if (!c->vs_state.p) VHSD_HIP(c->vs_state.ensure(state_hist + sizeof(SyncHistoryDev)));
    VHSD_HIP(c->f_x.ensure(F * line_stride));
    VHSD_HIP(c->f_y0.ensure(F * line_stride));
    LevelStateDev* level_state = reinterpret_cast<LevelStateDev*>(c->vs_state.p);
    SyncHistoryDev* history = reinterpret_cast<SyncHistoryDev*>(c->vs_state.p + state_hist);
    if (reset_state) {
      VHSD_HIP(vhsd_launch_vhs_sync_reset(level_state, history, st));
      c->vs_state_ready = true;
    }
    if (chroma && reset_state) {
      VHSD_HIP(vhsd_launch_chroma_state(reinterpret_cast<vhsd::chroma::ChromaState*>(c->c_state.p), 1,
                                        chroma->field_number_base, nullptr, 0, st));
      c->c_state_ready = true;
    }
    SeqCtl* ctl = reinterpret_cast<SeqCtl*>(m + o_ctl);
    VHSD_HIP(vhsd_launch_vhs_ctl_init(ctl, first_begin, first_record, st));
    SyncWork w{};
    w.sync_ref = sync_ref;
    w.span_len = span_length;
    w.ctl = ctl;
    w.state_shift = reset_state ? 0 : state_shift;
    w.spec_count = reinterpret_cast<int*>(m + o_spec);
    w.win_start = reinterpret_cast<long long*>(m + o_ws);
    w.win_len = reinterpret_cast<long long*>(m + o_wl);
    w.pulse_start = reinterpret_cast<int*>(m + o_pst);
    w.pulse_len = reinterpret_cast<int*>(m + o_pln);
    w.overflow = reinterpret_cast<int*>(m + o_ovf);
    w.search = reinterpret_cast<LevelSearch*>(m + o_search);
    w.levels = reinterpret_cast<FieldLevels*>(m + o_levels);
    w.level_before = reinterpret_cast<LevelStateDev*>(m + o_before);
    w.level_state_in = level_state;
    w.level_state_out = level_state;
    w.level_state_after_all = reinterpret_cast<LevelStateDev*>(m + o_after);
    w.raw_count = reinterpret_cast<int*>(m + o_raw);
    w.logic = reinterpret_cast<FieldLogic*>(m + o_logic);
    w.ref_kind = reinterpret_cast<int*>(m + o_rk);
    w.ref_start = reinterpret_cast<int*>(m + o_rs);
    w.logic_aux = m + o_aux;
    w.records = fields ? reinterpret_cast<FieldRecord*>(fields) : reinterpret_cast<FieldRecord*>(m + o_rec);
    w.history_in = history;
    w.history_out = history;
    w.lines = reinterpret_cast<double*>(m + o_lines);
    {
      SegBuffers& b = w.seg;
      b.segs = segs;
      b.seg_min = reinterpret_cast<double*>(m + o_smin);
      b.seg_nan = reinterpret_cast<int*>(m + o_snan);
      b.thr = reinterpret_cast<double*>(m + o_sthr);
      int* ints = reinterpret_cast<int*>(m + o_sint);
      b.choice = ints;
      b.fallback = ints + F;
      b.active = ints + 2 * F;
      b.count = ints + 3 * F;
      b.active2 = ints + 4 * F;
      b.count2 = ints + 5 * F;
      b.count3 = ints + 6 * F;
      b.pending = ints + 7 * F;
      double* dbls = reinterpret_cast<double*>(m + o_sdbl);
      b.thr2 = dbls;
      b.thr3 = dbls + F;
      b.dc3 = dbls + 2 * F;
      b.sum = reinterpret_cast<SegSummary*>(m + o_ssum);
      b.seg_carry = reinterpret_cast<int*>(m + o_scar);
      b.seg_off = reinterpret_cast<int*>(m + o_soff);
    }
    FieldTail t{};
    t.spans = reinterpret_cast<const float**>(m + o_tp);
    t.span_len = reinterpret_cast<long long*>(m + o_tp + 8 * F);
    t.locs = reinterpret_cast<const double**>(m + o_tp + 16 * F);
    t.filled = reinterpret_cast<const unsigned char**>(m + o_tp + 24 * F);
    t.refined = reinterpret_cast<double**>(m + o_tp + 32 * F);
    t.filled_out = reinterpret_cast<unsigned char**>(m + o_tp + 40 * F);
    t.line_counts = reinterpret_cast<int*>(m + o_lc);
    t.origins = reinterpret_cast<long long*>(m + o_org);
    t.first_lines = reinterpret_cast<int*>(m + o_fl);
    t.threshold = reinterpret_cast<double*>(m + o_thr);
    t.dc = reinterpret_cast<double*>(m + o_dc);
    t.zeros = m + o_zero;
    t.refined_base = reinterpret_cast<double*>(m + o_refb);
    t.filled_out_base = m + o_fob;
    t.render = reinterpret_cast<double*>(m + o_rend);
    t.out_index = reinterpret_cast<int*>(m + o_oi);
    const size_t acc_n = F * line_stride;
    t.acc_lines = reinterpret_cast<double*>(m + o_acc);
    t.acc_refined = t.acc_lines + acc_n;
    t.acc_render = t.acc_refined + acc_n;
    t.acc_span = t.acc_render + acc_n;
    // ApplyNtscFscPhaseShiftCore with fsc = TbcFrameSpec.OutputSampleRateHz / 4
    // (TbcFieldDecodePipeline.cs:1394-1400, 4022-4038)
    const double fsc_mhz = (sy->output_rate_mhz * 1000000.0) / 4000000.0;
    t.fsc_shift = (cfg.is_pal || !(fsc_mhz > 0.0)) ? 0.0 : ((117.25 / 360.0) / fsc_mhz) * fs_mhz;
    // OutputFirstLine (TbcFieldDecodePipeline.cs:2475-2483)
    t.out_first_line[0] = cfg.is_pal ? 3 : 1;
    t.out_first_line[1] = cfg.is_pal ? 4 : 1;
    // chroma work buffers (one allocation): probe windows and results, the per-field sequence
    // choice, analysis level adjusts, the chroma field and the final filter's two buffers
    vhsd::chroma::ChromaWork cw{};
    if (chroma) {
      const vhsd::chroma::ChromaCfg& cc = c->ccfg;
      const size_t nf = (size_t)cc.lines * cc.line_len;
      const int lvl_stride = out_lines + max_first;
      const size_t mext = nf + 2 * (size_t)cc.pad;
      const int chunks = (int)((mext + cc.iir_chunk - 1) / cc.iir_chunk);
      size_t co = 0;
      auto ctake = [&](size_t bytes) {
        const size_t o = co;
        co += (bytes + 255) & ~(size_t)255;
        return o;
      };
      const size_t o_win = ctake(4 * F * cc.lines * cc.win_max), o_probe = ctake(8 * F * 2 * cc.lines * 5);
      const size_t o_flip = ctake(4 * F * 2), o_summ = ctake(8 * F * 2 * 4), o_ch = ctake(4 * F);
      const size_t o_fno = ctake(4 * F), o_ones = ctake(4 * F), o_ax = ctake(8 * F * lvl_stride);
      const size_t o_ay = ctake(8 * F * lvl_stride), o_fld = ctake(4 * F * nf);
      const size_t o_fa = ctake(4 * F * (size_t)chunks * cc.iir_chunk), o_fb = ctake(4 * F * (size_t)chunks * cc.iir_chunk);
      const size_t o_th = ctake(cc.is_ntsc ? 8 * F * nf : 8);
      const size_t o_ist = ctake(4 * F * (size_t)chunks * 2 * 2 * vhsd::chroma::kMaxSections);
      VHSD_HIP(c->c_work.ensure(co));
      unsigned char* cm = c->c_work.p;
      cw.plane = chroma->plane;
      cw.plane_len = span_length;
      cw.win = reinterpret_cast<float*>(cm + o_win);
      cw.probe = reinterpret_cast<double*>(cm + o_probe);
      cw.flip = reinterpret_cast<int*>(cm + o_flip);
      cw.summ = reinterpret_cast<double*>(cm + o_summ);
      cw.chosen = reinterpret_cast<int*>(cm + o_ch);
      cw.field_no = reinterpret_cast<int*>(cm + o_fno);
      cw.ones = reinterpret_cast<int*>(cm + o_ones);
      cw.an_x = reinterpret_cast<double*>(cm + o_ax);
      cw.an_y0 = reinterpret_cast<double*>(cm + o_ay);
      cw.lvl_stride = lvl_stride;
      cw.field = reinterpret_cast<float*>(cm + o_fld);
      cw.fa = reinterpret_cast<float*>(cm + o_fa);
      cw.fb = reinterpret_cast<float*>(cm + o_fb);
      cw.iir_state = reinterpret_cast<float*>(cm + o_ist);
      cw.theta = reinterpret_cast<double*>(cm + o_th);
      cw.iir_chunks = chunks;
      cw.state = reinterpret_cast<vhsd::chroma::ChromaState*>(c->c_state.p);
      cw.out = chroma->out;
    }
    vhsd::HsyncParams hp{p->threshold, fs_mhz, sy->hsync_pulse_us, p->ire0, p->hz_ire, cfg.is_pal,
                         p->use_right_hsync};
    VHSD_HIP(c->h_work.ensure(vhsd_hsync_work_bytes(max_fields, line_stride)));
    const bool smooth = p->wow_level_smoothing > 0.0;
    const double alpha = smooth ? 1.0 / (p->wow_level_smoothing * p->output_line_length) : 0.0;
    if (smooth && (c->f_pow_alpha != alpha || c->f_pow_len != p->output_line_length)) {
      std::vector<double> pw(p->output_line_length);
      for (int s2 = 0; s2 < p->output_line_length; s2++) pw[s2] = std::pow(1.0 - alpha, (double)(s2 + 1));
      VHSD_HIP(c->f_pow.ensure(pw.size()));
      VHSD_HIP(hipMemcpy(c->f_pow.p, pw.data(), sizeof(double) * pw.size(), hipMemcpyHostToDevice));
      c->f_pow_alpha = alpha;
      c->f_pow_len = p->output_line_length;
    }
    const double line_decay = smooth ? std::pow(1.0 - alpha, (double)p->output_line_length) : 0.0;
    const double out_scale = p->output_scale;
    const double conv_off = (p->output_zero + 0.5) - (p->vsync_ire * out_scale);
    // Everything after the state reset and the start position: the nominal-threshold
    // pulses of the span, whose vertical-sync groups place the read windows, then the
    // rounds.  Each round speculates the read windows from where the sequence stands,
    // decodes them in parallel and verifies them in order; a round after a mispredicted
    // window or a failed field carries on from there, and a halted call's rounds do nothing.
    auto enqueue = [&](hipStream_t q) -> int32_t {
      // the zero `filled` row (BuildUpstreamLineLocations fills no line); written by a fill kernel
      VHSD_HIP(vhsd_launch_fill_u32(reinterpret_cast<unsigned*>(m + o_zero), (line_stride + 3) / 4, 0u, q));
      if (chroma) VHSD_HIP(vhsd_launch_chroma_state(cw.state, 0, -1, cw.ones, max_fields, q));
      {
        ProfScope ps(c, q, VHSDECODE_HIP_KERNEL_PULSES, span_length);
        VHSD_HIP(vhsd_launch_pulses(sync_ref, span_length, c->bias, p->threshold, cfg.pulse_min, cfg.pulse_max,
                                    c->tile_counts.p, tiles, c->edge_pos.p, c->edge_kind.p, (int)edge_cap_span,
                                    c->counters.p, c->p_starts.p, c->p_lens.p, capacity, q));
      }
      for (int r = 0; r < rounds; r++) {
        SyncWork wr = w;
        if (r > 0) wr.state_shift = 0;
        {
          ProfScope ps(c, q, VHSDECODE_HIP_KERNEL_LOCATE, max_fields);
          VHSD_HIP(vhsd_launch_vhs_sync(&cfg, &wr, c->p_starts.p, c->p_lens.p, c->counters.p + 1, capacity,
                                        span_length, max_fields, q));
        }
        VHSD_HIP(vhsd_launch_vhs_tail_tables(&cfg, &wr, &t, max_fields, q));
        {
          ProfScope ps(c, q, VHSDECODE_HIP_KERNEL_HSYNC, max_fields);
          VHSD_HIP(vhsd_launch_hsync(t.spans, t.span_len, t.locs, t.filled, t.line_counts, &hp, t.refined,
                                     t.filled_out, max_fields, line_stride, c->bias, c->h_work.p, t.threshold, t.dc,
                                     q));
        }
        if (chroma) {
          // the burst phase analysis on the hsync-refined lines, whose burst-refined
          // locations are the render locations (luma and chroma)
          ProfScope ps(c, q, VHSDECODE_HIP_KERNEL_CHROMA_FIELD, max_fields);
          VHSD_HIP(vhsd_launch_tbc_levels(t.refined_base, line_stride, out_lines, cw.ones, cw.lvl_stride,
                                          p->nominal_line_length, line_decay, cw.an_x, cw.an_y0, &ctl->decoded,
                                          max_fields, q));
          VHSD_HIP(vhsd_launch_chroma_analysis(&c->ccfg, &cw, t.refined_base, t.render, line_stride, line_stride,
                                               t.origins, smooth ? c->f_pow.p : nullptr, p->nominal_line_length,
                                               smooth ? 1 : 0, c->lut.p, t.fsc_shift, max_fields, &ctl->decoded, q));
        } else {
          VHSD_HIP(vhsd_launch_vhs_render(&cfg, &wr, &t, max_fields, q));
        }
        VHSD_HIP(vhsd_launch_vhs_export(&cfg, &wr, &t, max_fields, q));
        {
          ProfScope ps(c, q, VHSDECODE_HIP_KERNEL_TBC, (long long)max_fields * out_lines * p->output_line_length);
          VHSD_HIP(vhsd_launch_tbc_fields(video, span_length, c->bias, t.origins, t.render, line_stride, out_lines,
                                          t.first_lines, max_first, 1, max_fields, p->nominal_line_length,
                                          line_decay, smooth ? c->f_pow.p : nullptr, c->f_x.p, c->f_y0.p,
                                          p->output_line_length, smooth ? 1 : 0, p->ire0, out_scale,
                                          1.0 / p->hz_ire, conv_off, c->lut.p, out, t.out_index, &ctl->decoded, q));
        }
        if (chroma) {
          ProfScope ps(c, q, VHSDECODE_HIP_KERNEL_CHROMA_FIELD, max_fields);
          VHSD_HIP(vhsd_launch_chroma_fields(&c->ccfg, &cw, t.render, line_stride, t.origins, t.first_lines, c->f_x.p,
                                             c->f_y0.p, smooth ? c->f_pow.p : nullptr, p->nominal_line_length,
                                             smooth ? 1 : 0, c->lut.p, w.records, ctl, t.out_index, max_fields,
                                             &ctl->decoded, q));
        }
      }
      VHSD_HIP(vhsd_launch_vhs_results(ctl, field_count, next_begin, q));
      return VHSDECODE_HIP_OK;
    };
    if (c->profiling) {
      // kernel timings need the launches themselves
      const int32_t erc = enqueue(st);
      if (erc != VHSDECODE_HIP_OK) return erc;
    } else {
      // the launches' arguments: a call that repeats them replays the captured graph
      std::vector<unsigned char> key;
      auto put = [&](const void* v, size_t bytes) {
        const unsigned char* b = static_cast<const unsigned char*>(v);
        key.insert(key.end(), b, b + bytes);
      };
      const void* ptrs[] = {sync_ref, video, out, field_count, next_begin, c->tile_counts.p, c->edge_pos.p,
                            c->edge_kind.p, c->counters.p, c->p_starts.p, c->p_lens.p, c->h_work.p, c->f_pow.p,
                            c->f_x.p, c->f_y0.p, c->lut.p, m};
      put(ptrs, sizeof(ptrs));
      const long long scal[] = {span_length, capacity, tiles, (long long)edge_cap_span, max_fields, rounds,
                                line_stride, out_lines, smooth ? 1 : 0};
      put(scal, sizeof(scal));
      const double dbl[] = {c->bias, p->threshold, line_decay, out_scale, conv_off, p->ire0, p->hz_ire,
                            p->nominal_line_length};
      put(dbl, sizeof(dbl));
      put(&cfg, sizeof(cfg));
      put(&w, sizeof(w));
      put(&t, sizeof(t));
      put(&hp, sizeof(hp));
      put(p, sizeof(*p));
      const int chroma_on = chroma ? 1 : 0;
      put(&chroma_on, sizeof(chroma_on));
      if (chroma) {
        put(&cw, sizeof(cw));
        put(&c->ccfg, sizeof(c->ccfg));
      }
      if (!c->vs_exec || key != c->vs_key) {
        if (c->vs_exec) (void)hipGraphExecDestroy(c->vs_exec);
        if (c->vs_graph) (void)hipGraphDestroy(c->vs_graph);
        c->vs_exec = nullptr;
        c->vs_graph = nullptr;
        c->vs_key.clear();
        VHSD_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int32_t erc = enqueue(c->stream);
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
        if (erc != VHSDECODE_HIP_OK) {
          if (graph) (void)hipGraphDestroy(graph);
          return erc;
        }
        VHSD_HIP(ce);
        c->vs_graph = graph;
        VHSD_HIP(hipGraphInstantiate(&c->vs_exec, graph, nullptr, nullptr, 0));
        c->vs_key = std::move(key);
      }
      VHSD_HIP(hipGraphLaunch(c->vs_exec, st));
    }

    c->f_ref_nloc = 0;
    c->vs_lines = t.acc_lines;
    c->vs_refined = t.acc_refined;
    c->vs_render = t.acc_render;
    c->vs_span = t.acc_span;
    c->vs_line_stride = line_stride;
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_vhs_field_lines_device(vhsdecode_hip_demod_context* c, int32_t which, const double** table,
                                             int32_t* stride) {
  if (!c || !table || !stride || which < 0 || which > 3)
    return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid arguments");
  std::lock_guard<std::mutex> lock(c->mu);
  if (!c->vs_render) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "no vhs field decode yet");
  *table = which == 0 ? c->vs_lines : which == 1 ? c->vs_refined : which == 2 ? c->vs_render : c->vs_span;
  *stride = c->vs_line_stride;
  return VHSDECODE_HIP_OK;
}

int32_t vhsdecode_hip_field_locations_device(vhsdecode_hip_demod_context* c, const double** locations,
                                             int32_t* location_count) {
  if (!c || !locations || !location_count) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid arguments");
  std::lock_guard<std::mutex> lock(c->mu);
  if (!c->f_ref.p || c->f_ref_nloc <= 0) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "no field decode yet");
  *locations = c->f_ref.p;
  *location_count = c->f_ref_nloc;
  return VHSDECODE_HIP_OK;
}

int32_t vhsdecode_hip_detect_dropouts_device(vhsdecode_hip_demod_context* c, const float* envelope,
                                             int64_t span_length, const double* locations, int32_t location_count,
                                             const vhsdecode_hip_vhs_field_v1* fields, const int32_t* field_count,
                                             int32_t max_fields,
                                             const vhsdecode_hip_dropout_params_v1* p, int32_t* row_count,
                                             int32_t* field_line, int32_t* start_x, int32_t* end_x,
                                             int32_t* range_count, int32_t* range_start, int32_t* range_end,
                                             void* stream) {
  try {
    if (!c || !envelope || !locations || !fields || !field_count || !p || !row_count ||
        !field_line || !start_x || !end_x || max_fields < 0 || span_length <= 0)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid dropout detection request");
    if (p->struct_size != sizeof(vhsdecode_hip_dropout_params_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "dropout_params struct_size mismatch");
    const bool ranges = range_count || range_start || range_end;
    if (ranges && !(range_count && range_start && range_end))
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "range outputs are given all three or none");
    if (location_count < 2 || location_count > vhsd::kDropMaxLocations)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "location_count must be 2..512");
    if (p->capacity <= 0 || p->output_line_length <= 0 || p->merge_threshold < 0 || p->minimum_length < 0 ||
        !(p->hysteresis > 0.0) || !(p->threshold_fraction >= 0.0) || p->first_field_lines <= 0 ||
        p->second_field_lines <= 0)
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid dropout parameters");
    if (span_length > INT32_MAX)
      return fail(VHSDECODE_HIP_E_UNSUPPORTED, "spans are limited to 2^31 - 1 samples (int32 range positions)");
    if (max_fields == 0) return VHSDECODE_HIP_OK;
    std::lock_guard<std::mutex> lock(c->mu);
    VHSD_HIP(hipSetDevice(c->device));
    hipStream_t st = pick_stream(c, stream);
    const int words_cap = (int)std::min<int64_t>(span_length / 64 + 2, (int64_t)vhsd::kDropSummaryWords * 32);
    VHSD_HIP(c->d_part.ensure((size_t)max_fields * 2 * 128));
    VHSD_HIP(c->d_mask.ensure((size_t)max_fields * 2 * words_cap));
    VHSD_HIP(c->d_sum.ensure((size_t)max_fields * 2 * vhsd::kDropSummaryWords));
    VHSD_HIP(c->d_info.ensure((size_t)max_fields));
    vhsd::DropoutArgs a{};
    a.envelope = envelope;
    a.span_len = span_length;
    a.locations = locations;
    a.nloc = location_count;
    a.fields = fields;
    a.field_count = field_count;
    a.fraction = p->threshold_fraction;
    a.absolute_threshold = p->absolute_threshold;
    a.hysteresis = p->hysteresis;
    a.merge = p->merge_threshold;
    a.min_len = p->minimum_length;
    a.out_len = p->output_line_length;
    a.first_lines = p->first_field_lines;
    a.second_lines = p->second_field_lines;
    a.capacity = p->capacity;
    a.partials = c->d_part.p;
    a.group_min = c->d_part.p + (size_t)max_fields * 128;
    a.info = c->d_info.p;
    a.masks = c->d_mask.p;
    a.summaries = c->d_sum.p;
    a.words_cap = words_cap;
    a.row_count = row_count;
    a.field_line = field_line;
    a.start_x = start_x;
    a.end_x = end_x;
    a.range_count = range_count;
    a.range_start = range_start;
    a.range_end = range_end;
    ProfScope ps(c, st, VHSDECODE_HIP_KERNEL_DROPOUTS, max_fields);
    VHSD_HIP(vhsd_launch_dropouts(&a, max_fields, st));
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_set_profiling(vhsdecode_hip_demod_context* c, int32_t enable) {
  if (!c) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "context is NULL");
  std::lock_guard<std::mutex> lock(c->mu);
  c->profiling = enable != 0;
  return VHSDECODE_HIP_OK;
}

int32_t vhsdecode_hip_get_profile(vhsdecode_hip_demod_context* c, vhsdecode_hip_profile_v1* p) {
  try {
    if (!c || !p) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "context and profile are required");
    if (p->struct_size != sizeof(vhsdecode_hip_profile_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "profile struct_size mismatch");
    std::lock_guard<std::mutex> lock(c->mu);
    VHSD_HIP(hipSetDevice(c->device));
    for (auto& r : c->recs) {
      float ms = 0.0f;
      VHSD_HIP(hipEventSynchronize(r.b));
      VHSD_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      c->prof_launches[r.kind] += 1;
      c->prof_ms[r.kind] += ms;
      c->prof_units[r.kind] += r.units;
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    c->recs.clear();
    for (int k = 0; k < VHSDECODE_HIP_KERNEL_COUNT; k++) {
      p->launches[k] = c->prof_launches[k];
      p->total_ms[k] = c->prof_ms[k];
      p->units[k] = c->prof_units[k];
      c->prof_launches[k] = 0;
      c->prof_ms[k] = 0.0;
      c->prof_units[k] = 0;
    }
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_debug_replace_spikes(vhsdecode_hip_demod_context* c, const float* demod, const float* diffed,
                                           double max_value, float* out, void* stream) {
  try {
    if (!c || !demod || !diffed || !out) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid spike request");
    std::lock_guard<std::mutex> lock(c->mu);
    VHSD_HIP(hipSetDevice(c->device));
    VHSD_HIP(c->dbg_list.ensure(32768));
    VHSD_HIP(vhsd_launch_debug_replace_spikes(demod, diffed, c->dbg_list.p, max_value, out, pick_stream(c, stream)));
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_set_ld_audio_phase2(vhsdecode_hip_demod_context* c, const vhsdecode_hip_ld_audio_phase2_v1* p) {
  try {
    if (!c) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid context");
    std::lock_guard<std::mutex> lock(c->mu);
    if (!p) {
      c->ld2_on = false;
      return VHSDECODE_HIP_OK;
    }
    if (p->struct_size != sizeof(vhsdecode_hip_ld_audio_phase2_v1))
      return fail(VHSDECODE_HIP_E_STRUCT_SIZE, "ld_audio_phase2 struct_size mismatch");
    if (!p->stage2[0] || !p->stage2[1] || !std::isfinite(p->center_hz[0]) || !std::isfinite(p->center_hz[1]))
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid stage-2 filters or carriers");
    if (p->block_length != 2 * vhsd::kFftM)
      return fail(VHSDECODE_HIP_E_UNSUPPORTED, "stage-2 block length must be 32768");
    VHSD_HIP(hipSetDevice(c->device));
    const int M = vhsd::kFftM;
    std::vector<float2> h(M + 1);
    for (int ch = 0; ch < 2; ch++) {
      for (int k = 0; k <= M; k++)
        h[k] = make_float2((float)p->stage2[ch][2 * k], (float)p->stage2[ch][2 * k + 1]);
      VHSD_HIP(c->ld2_filt[ch].ensure(M + 1));
      VHSD_HIP(hipMemcpy(c->ld2_filt[ch].p, h.data(), (M + 1) * sizeof(float2), hipMemcpyHostToDevice));
      c->ld2_center[ch] = p->center_hz[ch];
    }
    c->ld2_on = true;
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_ld_audio_phase2_device(vhsdecode_hip_demod_context* c, const float* left, const float* right,
                                             int64_t count, float* out_left, float* out_right, void* stream) {
  try {
    if (!c || count < 0 || (count > 0 && (!left || !right || !out_left || !out_right)))
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid stage-2 audio request");
    if (count > 0 && (out_left == left || out_left == right || out_right == left || out_right == right))
      return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "stage-2 output must not overlay its input");
    std::lock_guard<std::mutex> lock(c->mu);
    if (!c->ld2_on) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "vhsdecode_hip_set_ld_audio_phase2 not called");
    if (count == 0) return VHSDECODE_HIP_OK;
    VHSD_HIP(hipSetDevice(c->device));
    const long long N = 2 * vhsd::kFftM, J = N - 512;
    int nmid = 0;
    if (count > N)
      for (long long sample = J; sample < count - J; sample += J) nmid++;
    const int nblocks = count <= N ? 1 : nmid + 2;
    vhsd::LdAudio2Args a{};
    a.in[0] = left;
    a.in[1] = right;
    a.out[0] = out_left;
    a.out[1] = out_right;
    a.count = count;
    a.nmid = nmid;
    a.center[0] = c->ld2_center[0];
    a.center[1] = c->ld2_center[1];
    a.filt[0] = c->ld2_filt[0].p;
    a.filt[1] = c->ld2_filt[1].p;
    a.oct = c->oct.p;
    const hipStream_t st = pick_stream(c, stream);
    VHSD_HIP(vhsd_launch_ld_audio_phase2(&a, nblocks, st));
    if (!stream) VHSD_HIP(hipStreamSynchronize(st));
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_debug_fft16k(vhsdecode_hip_demod_context* c, float* data, int32_t count, int32_t inverse,
                                   void* stream) {
  try {
    if (!c || !data || count <= 0) return fail(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid FFT request");
    std::lock_guard<std::mutex> lock(c->mu);
    VHSD_HIP(hipSetDevice(c->device));
    VHSD_HIP(vhsd_launch_debug_fft(data, count, inverse, c->oct.p, pick_stream(c, stream)));
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return fail(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

}  // extern "C"
