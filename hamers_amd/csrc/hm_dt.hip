// hm_dt.hip -- spectral radiuses and stable time step on the GPU (see
// hm_dt.hpp), and the C-ABI entry point that returns them for a patch list.
#include <vector>

#include "hm_cell.hpp"
#include "hm_ctx.hpp"
#include "hm_dt.hpp"
#include "hm_multi.hpp"
#include "hm_tables.hpp"

namespace hm {

namespace {

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

// One lane per interior cell; lanes past the segment's end contribute 0.
// MODEL: HM_MODEL_*; NS species of the multi-species models.
template <int D, int MODEL, int NS>
__global__ __launch_bounds__(kBlock) void k_spectral(SegTableT<DevPatch> w, KParams k, unsigned long long* acc) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const DevPatch* P;
    int ax;
    int64_t loc;
    double s[3] = {0.0, 0.0, 0.0}, sum = 0.0;
    if (w.find(gid, P, ax, loc)) {
        const Boxes<D> B(*P);
        // 32-bit index arithmetic (a segment has fewer than 2^31 items)
        const uint32_t r = uint32_t(loc) / uint32_t(B.n[0]);
        const int i = int(uint32_t(loc) - r * uint32_t(B.n[0]));
        const int j = (D > 1) ? int((D > 2) ? r % uint32_t(B.n[1]) : r) : 0;
        const int kk = (D > 2) ? int(r / uint32_t(B.n[1])) : 0;
        const uint32_t iq = uint32_t(q_idx<D>(*P, i, j, kk));
        auto q = [&](int e, uint32_t at) { return ldg(P->Q[e], at); };
        double u[3], c;
        if constexpr (MODEL == HM_MODEL_SINGLE_SPECIES) {
            double rho, p;
            cell_state_single<D>(q, iq, k.gamma[0], rho, u, p, c);
        } else {
            double Zr[NS], Z[NS], rho, p, c2;
            cell_state_multi<D, NS, MODEL == HM_MODEL_FOUR_EQN_CONSERVATIVE>(k, q, iq, Zr, Z, rho, u, p, c2);
            c = sqrt(c2);
        }
        // Euler.cpp:592-598 (1D), 713-721 (2D), 861-871 (3D)
#pragma unroll
        for (int a = 0; a < D; a++) s[a] = div_rn(fabs(u[a]) + c, P->dx[a]);
        if constexpr (D == 1)
            sum = s[0];
        else if constexpr (D == 2)
            sum = s[0] + s[1];
        else
            sum = s[0] + s[1] + s[2];
    }
    // fmax over the wavefront (every lane takes part), then one atomic per wavefront
#pragma unroll
    for (int a = 0; a < D; a++) s[a] = wave_max(s[a]);
    sum = wave_max(sum);
    if ((threadIdx.x & 63u) == 0u) {
#pragma unroll
        for (int a = 0; a < D; a++)
            if (s[a] > 0.0) atomicMax(acc + a, static_cast<unsigned long long>(__double_as_longlong(s[a])));
        if (sum > 0.0) atomicMax(acc + 3, static_cast<unsigned long long>(__double_as_longlong(sum)));
    }
}

__global__ void k_dt_clear(unsigned long long* acc) {
    if (threadIdx.x < kDtAcc) acc[threadIdx.x] = 0ull;
}

__global__ void k_dt_reset_time(double* res) {
    if (threadIdx.x == 0) res[kDtTime] = 0.0;
}

// RungeKuttaLevelIntegrator.cpp:1213-1239 and 1321-1329: dt is the smaller of
// the patches' dt and 1 / the sum of the axis maxima (tbox::MathUtilities::Min,
// a < b ? a : b), times cfl.
template <int D>
__global__ void k_dt_final(const unsigned long long* acc, double cfl, double* res) {
    if (threadIdx.x != 0) return;
    double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < D; a++) s[a] = __longlong_as_double(static_cast<long long>(acc[a]));
    const double maxsum = __longlong_as_double(static_cast<long long>(acc[3]));
    double ssum = 0.0;
#pragma unroll
    for (int a = 0; a < D; a++) ssum += s[a];
    const double dt_patch = double(1) / maxsum;
    const double dt_sum = 1.0 / ssum;
    const double dt_level = dt_patch < dt_sum ? dt_patch : dt_sum;
    const double dt = dt_level * cfl;
    res[kDtS] = s[0];
    res[kDtS + 1] = s[1];
    res[kDtS + 2] = s[2];
    res[kDtPatch] = dt_patch;
    res[kDtLevel] = dt_level;
    res[kDtStep] = dt;
    res[kDtTime] = res[kDtTime] + dt;
}

template <int D, int MODEL, int NS>
void launch_spectral_k(const SegTableT<DevPatch>& w, int64_t items, const KParams& k, unsigned long long* acc,
                       hipStream_t s) {
    const unsigned nb = unsigned((items + kBlock - 1) / kBlock);
    if (nb) hipLaunchKernelGGL((k_spectral<D, MODEL, NS>), dim3(nb), dim3(kBlock), 0, s, w, k, acc);
}

template <int D>
hipError_t launch_stable_dt_d(const hm_config& cfg, const SegTableT<DevPatch>& w, int64_t items, const KParams& k,
                              double cfl, const DtBuffers& buf, hipStream_t s) {
    hipLaunchKernelGGL(k_dt_clear, dim3(1), dim3(64), 0, s, buf.acc);
    const int ns = cfg.num_species;
    switch (cfg.model) {
        case HM_MODEL_SINGLE_SPECIES: launch_spectral_k<D, HM_MODEL_SINGLE_SPECIES, 1>(w, items, k, buf.acc, s); break;
        case HM_MODEL_FOUR_EQN_CONSERVATIVE:
            if (ns == 1) launch_spectral_k<D, HM_MODEL_FOUR_EQN_CONSERVATIVE, 1>(w, items, k, buf.acc, s);
            else if (ns == 2) launch_spectral_k<D, HM_MODEL_FOUR_EQN_CONSERVATIVE, 2>(w, items, k, buf.acc, s);
            else launch_spectral_k<D, HM_MODEL_FOUR_EQN_CONSERVATIVE, 3>(w, items, k, buf.acc, s);
            break;
        default:
            if (ns == 2) launch_spectral_k<D, HM_MODEL_FIVE_EQN_ALLAIRE, 2>(w, items, k, buf.acc, s);
            else launch_spectral_k<D, HM_MODEL_FIVE_EQN_ALLAIRE, 3>(w, items, k, buf.acc, s);
            break;
    }
    hipLaunchKernelGGL(k_dt_final<D>, dim3(1), dim3(64), 0, s, buf.acc, cfl, buf.res);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_stable_dt(const hm_config& cfg, const SegTableT<DevPatch>& work, int64_t items, const KParams& k,
                            double cfl, const DtBuffers& buf, hipStream_t s) {
    switch (cfg.dim) {
        case 1: return launch_stable_dt_d<1>(cfg, work, items, k, cfl, buf, s);
        case 2: return launch_stable_dt_d<2>(cfg, work, items, k, cfl, buf, s);
        case 3: return launch_stable_dt_d<3>(cfg, work, items, k, cfl, buf, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_dt_reset_time(const DtBuffers& buf, hipStream_t s) {
    hipLaunchKernelGGL(k_dt_reset_time, dim3(1), dim3(64), 0, s, buf.res);
    return hipGetLastError();
}

}  // namespace hm

extern "C" {

int hm_compute_spectral_radiuses_and_stable_dt_on_patches(hm_ctx* ctx, const hm_patch_view* patches,
                                                          int num_patches, double* out, void* stream) {
    using namespace hm;
    if (!ctx || !ctx->rec) return HM_ERR_INVALID_ARGUMENT;
    ConvectiveFluxReconstructorGPU& rec = *ctx->rec;
    if (num_patches < 1 || !patches || !out) return rec.fail(HM_ERR_INVALID_ARGUMENT, "invalid patch list or output");
    const hm_config& cfg = rec.config();
    const int d = cfg.dim, nq = rec.numConservativeArrays();
    TableImage<DevPatch> tab;
    tab.items.resize(num_patches);
    for (int i = 0; i < num_patches; i++) {
        const hm_patch_view& v = patches[i];
        DevPatch& p = tab.items[i];
        std::memset(&p, 0, sizeof p);
        double box = 1.0;
        for (int a = 0; a < 3; a++) {
            const bool used = a < d;
            p.n[a] = used ? v.n[a] : 1;
            p.gq[a] = used ? v.ghost[a] : 0;
            p.dx[a] = used ? v.dx[a] : 1.0;
            p.wrec[a] = -1;
            if (used && (v.n[a] < 1 || v.ghost[a] < 0 || !(v.dx[a] > 0.0)))
                return rec.fail(HM_ERR_INVALID_ARGUMENT, "patch " + std::to_string(i) + ": bad extent, ghost width or dx");
            if (used) box *= double(v.n[a]) + 2.0 * double(v.ghost[a]);
        }
        if (box >= 536870912.0)
            return rec.fail(HM_ERR_INVALID_ARGUMENT, "patch " + std::to_string(i) + ": ghost box of 2^29 cells or more");
        for (int e = 0; e < nq; e++) {
            if (!v.Q[e])
                return rec.fail(HM_ERR_INVALID_ARGUMENT,
                                "patch " + std::to_string(i) + ": conservative variable " + std::to_string(e) + " is NULL");
            p.Q[e] = v.Q[e];
        }
        int64_t cells = 1;
        for (int a = 0; a < d; a++) cells *= v.n[a];
        tab.add(i, 0, cells);
    }
    if (int st = rec.selectDevice()) return st;
    const size_t tb = tab.bytes();
    const size_t bytes = tb + up256(sizeof(unsigned long long) * kDtAcc) + up256(sizeof(double) * kDtSlots);
    void* dev = nullptr;
    if (hipMalloc(&dev, bytes) != hipSuccess) return rec.fail(HM_ERR_OUT_OF_MEMORY, "dt work tables");
    std::vector<char> img(tb, 0);
    const SegTableT<DevPatch> w = tab.write(img.data(), static_cast<char*>(dev));
    DtBuffers buf;
    buf.acc = reinterpret_cast<unsigned long long*>(static_cast<char*>(dev) + tb);
    buf.res = reinterpret_cast<double*>(static_cast<char*>(dev) + tb + up256(sizeof(unsigned long long) * kDtAcc));
    hipStream_t s = static_cast<hipStream_t>(stream);
    double res[kDtSlots] = {};
    int st = HM_OK;
    if (hipMemcpyAsync(dev, img.data(), tb, hipMemcpyHostToDevice, s) != hipSuccess ||
        launch_dt_reset_time(buf, s) != hipSuccess ||
        launch_stable_dt(cfg, w, tab.total, rec.kernelParams(), 1.0, buf, s) != hipSuccess ||
        hipMemcpyAsync(res, buf.res, sizeof res, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        st = rec.fail(HM_ERR_DEVICE, "stable dt computation failed");
    (void)hipFree(dev);
    if (st) return st;
    for (int a = 0; a < d; a++) out[a] = res[kDtS + a];
    out[d] = res[kDtPatch];
    out[d + 1] = res[kDtLevel];
    return HM_OK;
}

}  // extern "C"
