// Host side of the one-wavefront-per-tile kernel: compiled shapes, per
// (tile, row) injection / receiver lists, launch dispatch.  The kernel itself
// is in trtm_wave_kernel.hpp, instantiated per wave system and direction.
#include <algorithm>

#include "trtm_wave_kernel.hpp"

namespace trtm {

// ---------------------------------------------------------------------------
// host side

// configurations compiled per wave system: {tile rows, 64 - 2K, 64, K}
static const BlockCfg kWaveSH[] = {{16, 56, 64, 4}, {8, 56, 64, 4}, {24, 56, 64, 4},
                                   {16, 60, 64, 2}};
static const BlockCfg kWavePSV[] = {{16, 56, 64, 4}, {8, 56, 64, 4}, {16, 60, 64, 2},
                                    {8, 60, 64, 2}};
static const BlockCfg kWaveBoth[] = {{8, 56, 64, 4}, {8, 60, 64, 2}};

void wave_configs(int systems, const BlockCfg** list, int* n) {
    if (systems == TRTM_SH) {
        *list = kWaveSH;
        *n = sizeof(kWaveSH) / sizeof(BlockCfg);
    } else if (systems == TRTM_PSV) {
        *list = kWavePSV;
        *n = sizeof(kWavePSV) / sizeof(BlockCfg);
    } else {
        *list = kWaveBoth;
        *n = sizeof(kWaveBoth) / sizeof(BlockCfg);
    }
}

static int tiles_of(int n, int t) { return (n + t - 1) / t; }

ListSizes wave_list_sizes(const trtm_step_desc* d, BlockCfg c) {
    const int rx = c.tx + 2 * c.k;
    const size_t tiles = (size_t)tiles_of(d->nx, c.tx) * tiles_of(d->nz, c.tz);
    // a cell lies in at most ceil(2K / TX) + 1 regions along x, likewise along z
    const size_t mult = (size_t)(tiles_of(2 * c.k, c.tx) + 1) * (tiles_of(2 * c.k, c.tz) + 1);
    ListSizes L;
    L.inj_buckets = tiles * rx + 1;
    L.rec_buckets = tiles * c.tx + 1;
    L.inj_item_bytes = mult * (size_t)std::max(d->n_inj_groups, 0) * sizeof(WaveInj);
    L.rec_item_bytes = (size_t)std::max(d->n_rec, 0) * sizeof(WaveRec);
    return L;
}

// mode 0 counts, mode 1 fills.  A group goes to the (tile, row) bucket of
// every region (tile + halo K) holding its cell; a receiver to the (tile, row)
// bucket of the tile owning its cell.
__global__ void k_wave_lists(int mode, int tx, int tz, int K, int nz, int tiles_x, int tiles_z,
                             int n_groups, const int* loc, const int* order, const int* start,
                             int* icount, const int* istart, WaveInj* iitems, int n_rec,
                             const int* rloc, int* rcount, const int* rstart, WaveRec* ritems) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int rx = tx + 2 * K;
    if (q < n_groups) {
        const int k0 = order[start[q]];
        const int ci = loc[2 * k0], cj = loc[2 * k0 + 1];
        for (int ti = std::max(0, (ci - K) / tx - 1); ti <= std::min(tiles_x - 1, (ci + K) / tx); ti++)
            for (int tk = std::max(0, (cj - K) / tz - 1); tk <= std::min(tiles_z - 1, (cj + K) / tz);
                 tk++) {
                const int ri0 = ti * tx - K, rj0 = tk * tz - K;
                if (ci < ri0 || ci >= ri0 + rx || cj < rj0 || cj >= rj0 + 64) continue;
                const int bucket = (ti * tiles_z + tk) * rx + (ci - ri0);
                const int pos = atomicAdd(icount + bucket, 1);
                if (mode == 1)
                    iitems[istart[bucket] + pos] = WaveInj{cj - rj0, start[q], start[q + 1], ci * nz + cj};
            }
    }
    if (q < n_rec) {
        const int ci = rloc[2 * q], cj = rloc[2 * q + 1];
        const int ti = ci / tx, tk = cj / tz;
        const int bucket = (ti * tiles_z + tk) * tx + (ci - ti * tx);
        const int pos = atomicAdd(rcount + bucket, 1);
        if (mode == 1) ritems[rstart[bucket] + pos] = WaveRec{cj - (tk * tz - K), q};
    }
}

int wave_build_lists(const trtm_step_desc* d, BlockCfg c, int* istart, int* icur, void* iitems,
                     int* rstart, int* rcur, void* ritems, hipStream_t s) {
    const ListSizes L = wave_list_sizes(d, c);
    const int tiles_x = tiles_of(d->nx, c.tx), tiles_z = tiles_of(d->nz, c.tz);
    const int nib = (int)L.inj_buckets - 1, nrb = (int)L.rec_buckets - 1;
    const int ng = d->n_inj_groups;
    const int nr = (d->rec_u || d->rec_v) ? d->n_rec : 0;
    const dim3 tb(256), lg((std::max(ng, nr) + 255) / 256);
    hipLaunchKernelGGL(k_fill_int_b, dim3((nib + 255) / 256), tb, 0, s, icur, nib, 0);
    hipLaunchKernelGGL(k_fill_int_b, dim3((nrb + 255) / 256), tb, 0, s, rcur, nrb, 0);
    if (ng > 0 || nr > 0)
        hipLaunchKernelGGL(k_wave_lists, lg, tb, 0, s, 0, c.tx, c.tz, c.k, d->nz, tiles_x, tiles_z,
                           ng, d->inj_loc, d->inj_order, d->inj_group_start, icur,
                           (const int*)nullptr, (WaveInj*)nullptr, nr, d->rec_loc, rcur,
                           (const int*)nullptr, (WaveRec*)nullptr);
    hipLaunchKernelGGL(k_scan_b, dim3(1), dim3(1024), 0, s, icur, istart, nib);
    hipLaunchKernelGGL(k_scan_b, dim3(1), dim3(1024), 0, s, rcur, rstart, nrb);
    hipLaunchKernelGGL(k_fill_int_b, dim3((nib + 255) / 256), tb, 0, s, icur, nib, 0);
    hipLaunchKernelGGL(k_fill_int_b, dim3((nrb + 255) / 256), tb, 0, s, rcur, nrb, 0);
    if (ng > 0 || nr > 0)
        hipLaunchKernelGGL(k_wave_lists, lg, tb, 0, s, 1, c.tx, c.tz, c.k, d->nz, tiles_x, tiles_z,
                           ng, d->inj_loc, d->inj_order, d->inj_group_start, icur,
                           (const int*)istart, static_cast<WaveInj*>(iitems), nr, d->rec_loc, rcur,
                           (const int*)rstart, static_cast<WaveRec*>(ritems));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TRTM_OK : fail(TRTM_ERR_HIP, "wave tile lists: %s", hipGetErrorString(e));
}

bool launch_wave(bool fwd, int systems, BlockCfg c, const KArgs& m, const ShotDev* shots,
                 int n_shots, int it0, int tiles_x, int tiles_z, hipStream_t s) {
    const bool psv = systems & TRTM_PSV, sh = systems & TRTM_SH;
    auto* f = psv && sh ? (fwd ? launch_wave_both_fwd : launch_wave_both_bwd)
              : psv     ? (fwd ? launch_wave_psv_fwd : launch_wave_psv_bwd)
                        : (fwd ? launch_wave_sh_fwd : launch_wave_sh_bwd);
    return f(c.tx, c.k, m, shots, n_shots, it0, tiles_x, tiles_z, s);
}

}  // namespace trtm
