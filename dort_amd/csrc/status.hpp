// status.hpp -- the library's error reporting, shared by the HIP and the host-only sources.
#pragma once
#include "../../include/dortgpu.h"

namespace dgpu {

// Records a printf-style message as the calling thread's last error (dgpu_last_error)
// and returns st.
dgpu_status fail(dgpu_status st, const char* fmt, ...);

}  // namespace dgpu
