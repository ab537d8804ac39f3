// hm_kernels_four.hip -- four-equation conservative instantiation of the
// multi-species RHS pipeline (kernels in hm_multi.hpp).
#include "hm_multi.hpp"

namespace hm {

namespace {

template <int D, class W>
hipError_t launch_four_d(const Plan<W>& plan, const KParams& k, int scheme, hipStream_t s) {
    switch (k.num_species) {
        case 1: return launch_multi_dn<D, 1, true>(plan, k, scheme, s);
        case 2: return launch_multi_dn<D, 2, true>(plan, k, scheme, s);
        case 3: return launch_multi_dn<D, 3, true>(plan, k, scheme, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

template <class W>
hipError_t launch_rhs_four(int dim, int scheme, const Plan<W>& plan, const KParams& k, hipStream_t s) {
    switch (dim) {
        case 1: return launch_four_d<1>(plan, k, scheme, s);
        case 2: return launch_four_d<2>(plan, k, scheme, s);
        case 3: return launch_four_d<3>(plan, k, scheme, s);
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_rhs_four<SegSingle>(int, int, const Plan<SegSingle>&, const KParams&, hipStream_t);
template hipError_t launch_rhs_four<SegTable>(int, int, const Plan<SegTable>&, const KParams&, hipStream_t);

}  // namespace hm
