// Shared by the FWI step kernels: per-shot kernel parameters passed by value
// (one shot) or as a kernel-argument array indexed by blockIdx.z (a batch of
// shots in one launch), and the temporally blocked FWI entry points.
#pragma once

#include "trtm_common.hpp"

namespace trtm {

template <class T>
struct ByValue {
    T v;
    __device__ __forceinline__ const T& get() const { return v; }
};
constexpr int kShotsPerLaunch = 8;  // per-shot entries fit the 4 KB kernel-argument limit
template <class T>
struct ByArray {
    T v[kShotsPerLaunch];
    __device__ __forceinline__ const T& get() const { return v[blockIdx.z]; }
};

// float only; descriptors already checked by the caller
int fwi_block_forward(const trtm_fwi_desc* d, int n, int it_begin, int it_end, hipStream_t s);
int fwi_block_adjoint(const trtm_fwi_desc* d, int n, int it_begin, int it_end, hipStream_t s);
int fwi_block_configs(int systems, int pass, int32_t* out, int max_n);
int fwi_scat_forward(const trtm_fwi_desc* d, int n, int it_begin, int it_end, hipStream_t s);

}  // namespace trtm
