// hm_ctx.hpp -- the opaque context behind the C-ABI.
#pragma once

#include <memory>
#include <string>

#include "hm_adjoint.hpp"
#include "hm_diffusive.hpp"
#include "hm_reconstructor.hpp"

namespace hm {
// Resolves, once per device, the device address of the allocation-free entry
// point's scratch array (hm_noalloc.hip); hm_ctx_create calls it.
int noalloc_setup(int device, std::string& err);
}  // namespace hm

struct hm_ctx {
    std::unique_ptr<hm::ConvectiveFluxReconstructorGPU> rec;
    std::unique_ptr<hm::AdjointSingle> adj;  // created on the first VJP call; destroyed before rec
    std::unique_ptr<hm::DiffusiveFluxReconstructorGPU> diff;  // created by hm_set_diffusive_transport
    std::string err;
    int levels = 0;  // live hm_level objects built on this context
};
