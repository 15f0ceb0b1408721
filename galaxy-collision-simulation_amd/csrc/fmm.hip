// fmm.hip — hip-fmm engine.
#include "gc_fmm.hpp"

namespace gc {

struct FmmEngine::Impl {};

FmmEngine::FmmEngine() : impl_(new Impl()) {}
FmmEngine::~FmmEngine() { delete impl_; }

int FmmEngine::compute(const DeviceState&, const gc_params&, size_t, size_t, bool, double,
                       hipStream_t, std::string* err) {
    if (err) *err = "the FMM solver is not available in this build";
    return GC_ERR_UNSUPPORTED;
}

int FmmEngine::time_kernels(const DeviceState&, const gc_params&, int, hipStream_t, double*,
                            double*, std::string* err) {
    if (err) *err = "the FMM solver is not available in this build";
    return GC_ERR_UNSUPPORTED;
}

}  // namespace gc
