// Host side of the HiFi Hilbert discriminator (vhsdecode_hip_hifi_*): per-context
// device scratch for one block's transform and phase, the twiddle tables, and the
// launches of csrc/hifi_kernels.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <exception>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "host_common.hpp"
#include "vhsdecode_hip.h"

extern "C" hipError_t vhsd_launch_hifi(const float* in, float* out, float2* work, float* phase, const float2* oct,
                                       const float2* fine, int n1, double carrier_scale, double rate, double deviation,
                                       hipStream_t stream);

struct vhsdecode_hip_hifi {
  std::mutex mu;
  int device = 0;
  hipStream_t stream = nullptr;
  long long n = 0;
  int n1 = 0;
  double carrier_scale = 0.0, rate = 0.0, deviation = 0.0;
  float2* work = nullptr;
  float* phase = nullptr;
  float2* oct = nullptr;
  float2* fine = nullptr;
  float* s_in = nullptr;   // host-variant staging
  float* s_out = nullptr;
  size_t s_count = 0;

  ~vhsdecode_hip_hifi() {
    (void)hipSetDevice(device);
    for (void* p : {(void*)work, (void*)phase, (void*)oct, (void*)fine, (void*)s_in, (void*)s_out})
      if (p) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

using vhsd_host::set_error;

#define HIFI_HIP(call)                                                                               \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess) return set_error(VHSDECODE_HIP_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int32_t run_blocks(vhsdecode_hip_hifi* h, const float* in, float* out, int32_t nblocks, hipStream_t st) {
  for (int b = 0; b < nblocks; b++) {
    const long long off = (long long)b * h->n;
    HIFI_HIP(vhsd_launch_hifi(in + off, out + off, h->work, h->phase, h->oct, h->fine, h->n1, h->carrier_scale,
                              h->rate, h->deviation, st));
  }
  return VHSDECODE_HIP_OK;
}

}  // namespace

extern "C" {

int32_t vhsdecode_hip_hifi_create(const vhsdecode_hip_hifi_config_v1* cfg, vhsdecode_hip_hifi** out) {
  try {
    if (!cfg || !out) return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "config and context must not be NULL");
    *out = nullptr;
    if (cfg->struct_size != sizeof(vhsdecode_hip_hifi_config_v1))
      return set_error(VHSDECODE_HIP_E_STRUCT_SIZE, "hifi_config struct_size mismatch");
    // HiFiHilbertDiscriminator constructor checks (HiFiHilbertDiscriminator.cs:9-36)
    if (cfg->sample_rate_hz <= 0) return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "sample_rate_hz must be positive");
    if (!std::isfinite(cfg->carrier_hz) || cfg->carrier_hz <= 0.0 || cfg->carrier_hz > 2147483647.0)
      return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "carrier_hz out of range");
    if (!std::isfinite(cfg->deviation_hz) || cfg->deviation_hz <= 0.0 || cfg->deviation_hz > 2147483647.0)
      return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "deviation_hz out of range");
    const long long n = cfg->block_length;
    int n1 = (int)(n / 16384);
    if (n < 16384 || n % 16384 || n1 > 256 || (n1 & (n1 - 1)))
      return set_error(VHSDECODE_HIP_E_UNSUPPORTED, "block_length must be 16384 * 2^k with k <= 8");
    const double carrier = std::nearbyint(cfg->carrier_hz);  // round half to even
    const double deviation = std::trunc(cfg->deviation_hz);
    if (deviation <= 0.0) return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "deviation_hz truncates to zero");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
      return set_error(VHSDECODE_HIP_E_NO_DEVICE, "no HIP device is available");
    if (cfg->device_id < 0 || cfg->device_id >= count)
      return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "device_id out of range");
    HIFI_HIP(hipSetDevice(cfg->device_id));
    auto* h = new (std::nothrow) vhsdecode_hip_hifi();
    if (!h) return set_error(VHSDECODE_HIP_E_OUT_OF_MEMORY, "context allocation failed");
    auto bail = [&](int32_t code) {
      delete h;
      return code;
    };
    h->device = cfg->device_id;
    h->n = n;
    h->n1 = n1;
    h->carrier_scale = carrier / deviation;
    h->rate = (double)(float)cfg->sample_rate_hz;
    h->deviation = deviation;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
      return bail(set_error(VHSDECODE_HIP_E_HIP, "hipStreamCreate failed"));
    if (hipMalloc((void**)&h->work, sizeof(float2) * n) != hipSuccess ||
        hipMalloc((void**)&h->phase, sizeof(float) * n) != hipSuccess ||
        hipMalloc((void**)&h->oct, sizeof(float2) * 2049) != hipSuccess ||
        hipMalloc((void**)&h->fine, sizeof(float2) * 256) != hipSuccess)
      return bail(set_error(VHSDECODE_HIP_E_OUT_OF_MEMORY, "HiFi scratch allocation failed"));
    const long double two_pi = 6.283185307179586476925286766559005768L;
    std::vector<float2> oct(2049), fine(256);
    for (int u = 0; u < 2049; u++) {
      const long double a = two_pi * (long double)u / 16384.0L;
      oct[u] = make_float2((float)cosl(a), (float)sinl(a));
    }
    for (int r = 0; r < 256; r++) {
      const long double a = -two_pi * (long double)r / (long double)n;
      fine[r] = make_float2((float)cosl(a), (float)sinl(a));
    }
    if (hipMemcpy(h->oct, oct.data(), sizeof(float2) * oct.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->fine, fine.data(), sizeof(float2) * fine.size(), hipMemcpyHostToDevice) != hipSuccess)
      return bail(set_error(VHSDECODE_HIP_E_HIP, "twiddle upload failed"));
    *out = h;
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return set_error(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

void vhsdecode_hip_hifi_destroy(vhsdecode_hip_hifi* h) {
  if (!h) return;
  {
    std::lock_guard<std::mutex> lock(h->mu);
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
  }
  delete h;
}

int32_t vhsdecode_hip_hifi_demodulate_device(vhsdecode_hip_hifi* h, const float* in, float* out, int32_t nblocks,
                                             void* stream) {
  try {
    if (!h || !in || !out || nblocks < 0) return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid HiFi request");
    std::lock_guard<std::mutex> lock(h->mu);
    HIFI_HIP(hipSetDevice(h->device));
    // one block's scratch is shared by the blocks of a call: they run in stream order
    return run_blocks(h, in, out, nblocks, stream ? reinterpret_cast<hipStream_t>(stream) : h->stream);
  } catch (const std::exception& e) {
    return set_error(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

int32_t vhsdecode_hip_hifi_demodulate_host(vhsdecode_hip_hifi* h, const float* in, float* out, int32_t nblocks) {
  try {
    if (!h || !in || !out || nblocks < 0) return set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "invalid HiFi request");
    std::lock_guard<std::mutex> lock(h->mu);
    HIFI_HIP(hipSetDevice(h->device));
    const size_t count = (size_t)nblocks * (size_t)h->n;
    if (count == 0) return VHSDECODE_HIP_OK;
    if (count > h->s_count) {
      if (h->s_in) (void)hipFree(h->s_in);
      if (h->s_out) (void)hipFree(h->s_out);
      h->s_in = h->s_out = nullptr;
      h->s_count = 0;
      HIFI_HIP(hipMalloc((void**)&h->s_in, sizeof(float) * count));
      HIFI_HIP(hipMalloc((void**)&h->s_out, sizeof(float) * count));
      h->s_count = count;
    }
    hipStream_t st = h->stream;
    HIFI_HIP(hipMemcpyAsync(h->s_in, in, sizeof(float) * count, hipMemcpyHostToDevice, st));
    // the untouched last sample of each block keeps the caller's value
    HIFI_HIP(hipMemcpyAsync(h->s_out, out, sizeof(float) * count, hipMemcpyHostToDevice, st));
    const int32_t rc = run_blocks(h, h->s_in, h->s_out, nblocks, st);
    if (rc != VHSDECODE_HIP_OK) return rc;
    HIFI_HIP(hipMemcpyAsync(out, h->s_out, sizeof(float) * count, hipMemcpyDeviceToHost, st));
    HIFI_HIP(hipStreamSynchronize(st));
    return VHSDECODE_HIP_OK;
  } catch (const std::exception& e) {
    return set_error(VHSDECODE_HIP_E_INTERNAL, e.what());
  }
}

}  // extern "C"
