// Host-side helpers shared by the library's translation units.
#pragma once
#include <stdint.h>

#include <string>

namespace vhsd_host {
// Records `msg` as the calling thread's last error (vhsdecode_hip_get_last_error) and
// returns `code`.
int32_t set_error(int32_t code, const std::string& msg);
}  // namespace vhsd_host

struct vhsdecode_hip_demod_context;
struct vhsdecode_hip_vhs_sync_v1;
struct vhsdecode_hip_field_params_v1;
struct vhsdecode_hip_vhs_field_v1;

// vhsdecode_hip_decode_vhs_fields_device continuing a sequence whose first `first_record`
// records (and their output fields) are already in `fields` / `out`: this call's records
// follow them, so every call of a sequence uses the same buffers (and its launches can be
// replayed).  field_count receives the records of the sequence so far; max_fields bounds
// them.
// The chroma side of a vhs field-sequence call (vhsdecode_hip_decode_vhs_fields_chroma_device).
struct vhsd_chroma_io {
  const float* plane;          // the span's chroma plane
  uint16_t* out;               // [max_fields][field samples]
  int32_t field_number_base;   // fields decoded before the sequence (used on reset)
};

extern "C" __attribute__((visibility("hidden"))) int32_t vhsd_decode_vhs_fields_from(
    vhsdecode_hip_demod_context* c, const float* sync_ref, const float* video, int64_t span_length,
    int64_t first_begin, int32_t reset_state, int64_t state_shift, const vhsdecode_hip_vhs_sync_v1* sy,
    const vhsdecode_hip_field_params_v1* p, int32_t max_fields, int32_t rounds, uint16_t* out,
    vhsdecode_hip_vhs_field_v1* fields, int32_t* field_count, int64_t* next_begin, void* stream,
    int32_t first_record, const vhsd_chroma_io* chroma = nullptr);
