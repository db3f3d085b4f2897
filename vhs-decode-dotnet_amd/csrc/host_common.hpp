// Host-side helpers shared by the library's translation units.
#pragma once
#include <stdint.h>

#include <string>

namespace vhsd_host {
// Records `msg` as the calling thread's last error (vhsdecode_hip_get_last_error) and
// returns `code`.
int32_t set_error(int32_t code, const std::string& msg);
}  // namespace vhsd_host
