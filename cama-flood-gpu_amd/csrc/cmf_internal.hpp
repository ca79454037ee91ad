// Internal declarations shared by the host and device sources of libcmfhip.
#pragma once

#include "../../include/cmf_hip.h"

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

namespace cmf {

// Thread-local last-error message; set_error returns its code so that
// callers can write `return set_error(...)`.
int set_error(int code, const char* fmt, ...);

}  // namespace cmf
