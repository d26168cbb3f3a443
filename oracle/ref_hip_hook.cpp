// ref_hip_hook.cpp — test plumbing for oracle/_ref/libps26119_ref_hip.so: the reference core
// built with its GPU hook switched on (src/pdhg/cpu_backend.cpp compiled with
// PS26119_HAVE_CUDA, so make_backend(true, …) calls make_cuda_backend(), reference
// src/pdhg/backend.h:83-85), with that hook answered by the MI355X backend of
// ps26119-solver_amd/csrc/adapter/hip_backend.cpp. The reference's own drivers then run
// unchanged over the psh_* kernels; the GPU tests compare them with the same drivers on the
// reference CPU backend. Test infrastructure only: the product never links this.
#include <memory>

#include "pdhg/backend.h"

namespace ps26119::pdhg {

std::unique_ptr<Backend> make_hip_backend();  // hip_backend.cpp

std::unique_ptr<Backend> make_cuda_backend() { return make_hip_backend(); }

}  // namespace ps26119::pdhg
