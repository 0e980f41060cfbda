// One process-wide lock for the library's device allocations, frees and graph captures.
// Handles are independent and may be driven from different host threads at once, but a
// free or an allocation (the runtime may synchronise the device for them) in one thread
// while another thread captures a handle's step graph invalidates that capture.  Every
// entry point that allocates, frees or captures holds this lock; the steps themselves
// (graph launches on the handle's own non-blocking stream) do not.
#pragma once
#include <mutex>

namespace mflbm {

inline std::recursive_mutex& device_lock() {
    static std::recursive_mutex m;
    return m;
}

}  // namespace mflbm
