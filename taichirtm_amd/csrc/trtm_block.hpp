// Shared by the temporally blocked kernels (LDS-tiled and one-wavefront-per-tile).
#pragma once

#include "trtm_common.hpp"

namespace trtm {

int fail(int code, const char* fmt, ...);

struct ShotDev {
    float* in[8];   // u, v, w, sxx, szz, sxz, syx, syz read at launch start
    float* out[8];  // written at launch end
    int nt, it_end, save, isnap;
    int num_snaps, snap_base, check_interval, n_groups;
    int n_rec, it_lo, it_off, snaps_regular;  // isnaps[k] == (k + 1 + snap_base) * isnap
    const int *loc, *order, *start;
    const float *su, *sv, *sw;
    const int* rec_loc;
    float *rec_u, *rec_v, *rec_w;
    const int *inj_tstart, *inj_titems, *rec_tstart, *rec_titems;
    // per list item (blocked kernel): region cell and group code / receiver
    const int *inj_tcells, *rec_tcells;
    // wavefront kernel: per-tile slot (-1: none) of a per-lane map of the
    // single-entry injection cells (wavelet row) and of the receivers (index)
    const int *imap_slot, *imap, *rmap_slot, *rmap;
    float *snap_u, *snap_v, *snap_w;
    long long snap_stride;  // floats per snapshot slot
    int snap_sponge, pad3;  // 1: stored values lack the sponge factor (backward applies it)
    const int* isnaps;
    float *image_u, *image_v, *image_w;
    int* check_flags;
};

__device__ __forceinline__ int xcd_tile_b(int bid, int tiles) {
    const int per = (tiles + 7) / 8;
    return (bid % 8) * per + bid / 8;
}

// Pointers read from the shot table are generic; cast them to the global
// address space so the compiler emits global_* (not flat_*) memory
// instructions, which do not share the LDS wait counter.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> G(T* p) {
    return (gptr<T>)p;
}

// Tables no thread writes during a launch (tile lists, wavelets, observed
// data), read through the constant address space: uniform reads become scalar
// loads.
template <class T>
using cptr = __attribute__((address_space(4))) const T*;
template <class T>
__device__ __forceinline__ cptr<T> C(const T* p) {
    return (cptr<T>)p;
}

static __global__ void k_fill_int_b(int* p, int n, int value) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = value;
}

// exclusive scan of count[0..n) into start[0..n], one workgroup
static __global__ __launch_bounds__(1024) void k_scan_b(const int* count, int* start, int n) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int chunk = (n + 1023) / 1024;
    const int b = t * chunk, e = min(n, b + chunk);
    int s = 0;
    for (int k = b; k < e; k++) s += count[k];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = t > 0 ? part[t - 1] : 0;
    for (int k = b; k < e; k++) {
        start[k] = run;
        run += count[k];
    }
    if (t == 1023) start[n] = part[1023];
}

// A launch configuration: tile rows, tile columns, threads per workgroup,
// steps per launch.  nt == 64 selects the one-wavefront-per-tile kernel, whose
// tile is tx x (64 - 2k) cells.
struct BlockCfg {
    int tx, tz, nt, k;
};

constexpr int kWaveThreads = 64;

inline bool is_wave(const BlockCfg& c) { return c.nt == kWaveThreads; }

// per-tile list sizes of a configuration (buckets are tiles, or tile rows for
// the wavefront kernel)
struct ListSizes {
    size_t inj_buckets, rec_buckets, inj_item_bytes, rec_item_bytes, map_bytes;
};

// per-lane maps of the wavefront kernel, carved from one workspace region
struct WaveMaps {
    int *imap_slot, *imap, *rmap_slot, *rmap, *tmp0, *tmp1, *dup;
};

ListSizes wave_list_sizes(const trtm_step_desc* d, BlockCfg c);
WaveMaps wave_maps(const trtm_step_desc* d, BlockCfg c, void* region);
int wave_build_lists(const trtm_step_desc* d, BlockCfg c, int* istart, int* icur, void* iitems,
                     int* rstart, int* rcur, void* ritems, void* maps, hipStream_t s);
bool launch_wave(bool fwd, int systems, BlockCfg c, const KArgs& m, const ShotDev* shots,
                 int n_shots, int it0, int tiles_x, int tiles_z, hipStream_t s);
void wave_configs(int systems, const BlockCfg** list, int* n);

}  // namespace trtm
