// hm_kernels_five.hip -- five-equation (Allaire) instantiation of the
// multi-species RHS pipeline (kernels in hm_multi.hpp).
#include "hm_multi.hpp"

namespace hm {

namespace {

template <int D, class W>
hipError_t launch_five_d(const Plan<W>& plan, const KParams& k, int scheme, hipStream_t s) {
    switch (k.num_species) {
        case 2: return launch_multi_dn<D, 2, false>(plan, k, scheme, s);
        case 3: return launch_multi_dn<D, 3, false>(plan, k, scheme, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

template <class W>
hipError_t launch_rhs_five(int dim, int scheme, const Plan<W>& plan, const KParams& k, hipStream_t s) {
    switch (dim) {
        case 1: return launch_five_d<1>(plan, k, scheme, s);
        case 2: return launch_five_d<2>(plan, k, scheme, s);
        case 3: return launch_five_d<3>(plan, k, scheme, s);
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_rhs_five<SegSingle>(int, int, const Plan<SegSingle>&, const KParams&, hipStream_t);
template hipError_t launch_rhs_five<SegTable>(int, int, const Plan<SegTable>&, const KParams&, hipStream_t);

}  // namespace hm
