// abi_guard.hpp — no C++ exception leaves an extern "C" entry point.  Each
// gc_status entry point runs its body inside GC_ABI_BEGIN / GC_ABI_END: a
// failed host allocation returns GC_ERR_OUT_OF_MEMORY, any other exception
// GC_ERR_STATE, with the message in gc_last_error(NULL).
#pragma once

#include <new>
#include <string>

#include "gcsim.h"

namespace gc {

gc_status set_error(const std::string& msg, gc_status code);

inline gc_status abi_failure(const char* name, bool oom) noexcept {
    const gc_status code = oom ? GC_ERR_OUT_OF_MEMORY : GC_ERR_STATE;
    try {
        return set_error(std::string(name) + (oom ? ": out of host memory" : ": unexpected C++ exception"),
                         code);
    } catch (...) {
        return code;
    }
}

}  // namespace gc

#define GC_ABI_BEGIN try {
#define GC_ABI_END(name)                                 \
    }                                                    \
    catch (const std::bad_alloc&) {                      \
        return gc::abi_failure(name, true);              \
    }                                                    \
    catch (...) {                                        \
        return gc::abi_failure(name, false);             \
    }
