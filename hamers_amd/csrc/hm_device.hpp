// hm_device.hpp -- device-side patch descriptors and work-list addressing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hamers_convective.h"

namespace hm {

constexpr int kGc = 4;     // convective ghost width (WCNS56-HLLC-HLL.cpp:22)
constexpr int kGs = 2;     // ghost width of dilatation / vorticity (WCNS56-HLLC-HLL.cpp:116-123)
constexpr int kBlock = 256;

// Element i of a per-patch array (one component of a patch's field), read as
// global memory at a 32-bit byte offset from a wave-uniform base: the load
// then takes its base in scalar registers and one 32-bit vector offset
// instead of a 64-bit flat address per lane.  Every entry point keeps a
// patch's ghost box below 2^29 cells, and callers index one component
// (cells or faces of one box) from that component's own base, so i * 8 fits
// 32 bits.  i may be formed with wrap-around (base + i - k, i >= k) as
// unsigned arithmetic.
__device__ __forceinline__ double ldg(const double* base, uint32_t i) {
    typedef __attribute__((address_space(1))) const char gchar;
    typedef __attribute__((address_space(1))) const double gdouble;
    return *(gdouble*)((gchar*)base + i * 8u);
}

// Block number a kernel works on for hardware block blockIdx.x of a grid of
// nb blocks.  The eight XCDs of the chip (each with its own L2) take blocks
// round-robin: hardware blocks b, b + 8, b + 16, ... run on XCD b % 8.  The
// map hands each XCD one contiguous range of block numbers instead, so that
// neighbouring blocks of a patch, which re-read each other's stencil rows,
// share an L2.  A bijection on [0, nb).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t xcd = b & 7u, local = b >> 3, q = nb >> 3, r = nb & 7u;
    return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// Store counterpart of ldg.
__device__ __forceinline__ void stg(double* base, uint32_t i, double v) {
    typedef __attribute__((address_space(1))) char gchar;
    typedef __attribute__((address_space(1))) double gdouble;
    *(gdouble*)((gchar*)base + i * 8u) = v;
}

// Per-patch descriptor as the kernels see it: caller arrays plus this
// library's scratch for the patch.
struct DevPatch {
    int n[3];                     // interior extents (1 on unused axes)
    int gq[3];                    // ghost width of Q (0 on unused axes)
    double dx[3];
    const double* Q[HM_MAX_EQN];  // conservative variables, per equation, over the Q ghost box
    double* F[3];                 // output face flux per axis (component-major)
    double* S;                    // output source (component-major interior)
    double* prim;                 // scratch: V[0..neq) and c over the 4-ghost box
    double* sens;                 // scratch: theta, Omega over the 2-ghost box
    double* fmid[3];              // scratch: midpoint flux per axis, neq comps over (n_a+3) x n_b box
    double* ustar[3];             // scratch: HLLC normal velocity per axis (advective models)
    int64_t wrec[3];              // offset of this patch's axis-a chunk in a face-state slot (KParams::wrec), -1: none
};

// Extents of the boxes the kernels walk, for a patch of dimension D.
template <int D>
struct Boxes {
    int n[3];
    __device__ __forceinline__ explicit Boxes(const DevPatch& P) {
#pragma unroll
        for (int a = 0; a < 3; a++) n[a] = (a < D) ? P.n[a] : 1;
    }
    // Box grown by g on the used axes.
    __device__ __forceinline__ int ext(int a, int g) const { return (a < D) ? n[a] + 2 * g : 1; }
    __device__ __forceinline__ int64_t vol(int g) const {
        return int64_t(ext(0, g)) * ext(1, g) * ext(2, g);
    }
    // Linear index of cell (i,j,k) in the box grown by g.
    __device__ __forceinline__ int64_t idx(int g, int i, int j, int k) const {
        const int gx = g, gy = (D > 1) ? g : 0, gz = (D > 2) ? g : 0;
        return (i + gx) + int64_t(j + gy) * ext(0, g) + int64_t(k + gz) * ext(0, g) * ext(1, g);
    }
    // Midpoint faces of axis A: i_A in [-1, n_A+1], others interior.
    __device__ __forceinline__ int mid_ext(int A, int a) const { return (a == A) ? n[a] + 3 : n[a]; }
    __device__ __forceinline__ int64_t mid_vol(int A) const {
        return int64_t(mid_ext(A, 0)) * mid_ext(A, 1) * mid_ext(A, 2);
    }
    __device__ __forceinline__ int64_t mid_idx(int A, int i, int j, int k) const {
        const int ci = i + (A == 0), cj = j + (A == 1), ck = k + (A == 2);
        return ci + int64_t(cj) * mid_ext(A, 0) + int64_t(ck) * mid_ext(A, 0) * mid_ext(A, 1);
    }
    // Output faces of axis A: (n_A+1) x n_b.
    __device__ __forceinline__ int face_ext(int A, int a) const { return (a == A) ? n[a] + 1 : n[a]; }
    __device__ __forceinline__ int64_t face_vol(int A) const {
        return int64_t(face_ext(A, 0)) * face_ext(A, 1) * face_ext(A, 2);
    }
    __device__ __forceinline__ int64_t face_idx(int A, int i, int j, int k) const {
        return i + int64_t(j) * face_ext(A, 0) + int64_t(k) * face_ext(A, 0) * face_ext(A, 1);
    }
};

// Q ghost box addressing (caller's ghost width).
template <int D>
__device__ __forceinline__ int64_t q_idx(const DevPatch& P, int i, int j, int k) {
    const int e0 = P.n[0] + 2 * P.gq[0];
    const int e1 = (D > 1) ? P.n[1] + 2 * P.gq[1] : 1;
    const int gy = (D > 1) ? P.gq[1] : 0, gz = (D > 2) ? P.gq[2] : 0;
    return (i + P.gq[0]) + int64_t(j + gy) * e0 + int64_t(k + gz) * e0 * e1;
}

// A work segment is a contiguous run of items (cells or faces) of one patch
// and one axis.  Items of all segments are numbered consecutively and every
// segment starts on a multiple of the wavefront size (kSegAlign), so one
// wavefront never spans two segments: the per-axis dispatch inside a kernel
// is uniform per wavefront, while one block may still span several small
// patches.  Numbers in [start+cnt, next start) are holes.
constexpr int kSegAlign = 64;

template <class T>
struct SegSingleT {  // one patch, up to 4 segments (one per axis, one extra), all by value
    T p;
    int64_t start[5];
    int64_t cnt[4];
    int nseg;
    __device__ __forceinline__ bool find(int64_t gid, const T*& P, int& seg_axis, int64_t& loc) const {
        if (gid >= start[nseg]) return false;
        int s = 0;
        while (s + 1 < nseg && gid >= start[s + 1]) s++;
        loc = gid - start[s];
        if (loc >= cnt[s]) return false;
        P = &p;
        seg_axis = s;
        return true;
    }
};

// The same 64-bit value in every lane of the wavefront (lane 0's).
__device__ __forceinline__ int64_t wave_uniform(int64_t v) {
    const int lo = __builtin_amdgcn_readfirstlane(int(uint64_t(v) & 0xffffffffu));
    const int hi = __builtin_amdgcn_readfirstlane(int(uint64_t(v) >> 32));
    return int64_t((uint64_t(uint32_t(hi)) << 32) | uint32_t(lo));
}

template <class T>
struct SegTableT {  // many patches: device arrays
    const T* patches;
    const int64_t* start;  // nseg + 1 entries
    const int64_t* cnt;    // nseg entries
    const int* seg;        // per segment: patch * 4 + axis
    int nseg;
    // gid = blockIdx.x * (a multiple of 64) + threadIdx.x: a wavefront's items
    // are 64 consecutive numbers from a multiple of 64, which lie in one
    // segment.  So the segment is looked up once per wavefront, on the
    // wavefront's first item: the search, the patch pointer and the patch's
    // fields are then wave-uniform (scalar loads and registers).
    __device__ __forceinline__ bool find(int64_t gid, const T*& P, int& seg_axis, int64_t& loc) const {
        const int64_t wb = wave_uniform(gid & ~int64_t(kSegAlign - 1));
        if (wb >= start[nseg]) return false;
        int lo = 0, hi = nseg - 1;  // last segment with start <= wb
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (start[mid] <= wb)
                lo = mid;
            else
                hi = mid - 1;
        }
        loc = gid - start[lo];
        if (loc >= cnt[lo]) return false;
        const int pd = seg[lo];
        P = patches + (pd >> 2);
        seg_axis = pd & 3;
        return true;
    }
};

using SegSingle = SegSingleT<DevPatch>;
using SegTable = SegTableT<DevPatch>;

}  // namespace hm
