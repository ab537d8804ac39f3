// hm_ctx.hpp -- the opaque context behind the C-ABI.
#pragma once

#include <memory>
#include <string>

#include "hm_reconstructor.hpp"

struct hm_ctx {
    std::unique_ptr<hm::ConvectiveFluxReconstructorGPU> rec;
    std::string err;
    int levels = 0;  // live hm_level objects built on this context
};
