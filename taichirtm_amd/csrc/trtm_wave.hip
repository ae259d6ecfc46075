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

static size_t al256(size_t x) { return (x + 255) & ~size_t(255); }

struct WaveCounts {
    size_t tiles, rx, mult, n_islots, n_rslots;
};

static WaveCounts wave_counts(const trtm_step_desc* d, BlockCfg c) {
    WaveCounts w;
    w.rx = c.tx + 2 * c.k;
    w.tiles = (size_t)tiles_of(d->nx, c.tx) * tiles_of(d->nz, c.tz);
    // a cell lies in at most ceil(2K / TX) + 1 regions along x, likewise along z
    w.mult = (size_t)(tiles_of(2 * c.k, c.tx) + 1) * (tiles_of(2 * c.k, c.tz) + 1);
    w.n_islots = std::min(w.tiles, w.mult * (size_t)std::max(d->n_inj_groups, 0));
    w.n_rslots = std::min(w.tiles, (size_t)std::max(d->n_rec, 0));
    return w;
}

ListSizes wave_list_sizes(const trtm_step_desc* d, BlockCfg c) {
    const WaveCounts w = wave_counts(d, c);
    ListSizes L;
    L.inj_buckets = w.tiles * w.rx + 1;
    L.rec_buckets = w.tiles * c.tx + 1;
    L.inj_item_bytes = w.mult * (size_t)std::max(d->n_inj_groups, 0) * sizeof(WaveInj);
    L.rec_item_bytes = (size_t)std::max(d->n_rec, 0) * sizeof(WaveRec);
    L.map_bytes = 4 * al256((w.tiles + 1) * sizeof(int)) +
                  al256(w.n_islots * w.rx * 64 * sizeof(int)) +
                  al256(w.n_rslots * c.tx * 64 * sizeof(int)) +
                  al256((size_t)std::max(d->n_rec, 0) * sizeof(int));
    return L;
}

WaveMaps wave_maps(const trtm_step_desc* d, BlockCfg c, void* region) {
    const WaveCounts w = wave_counts(d, c);
    char* p = static_cast<char*>(region);
    WaveMaps m;
    auto take = [&](size_t bytes) {
        int* q = reinterpret_cast<int*>(p);
        p += al256(bytes);
        return q;
    };
    m.imap_slot = take((w.tiles + 1) * sizeof(int));
    m.rmap_slot = take((w.tiles + 1) * sizeof(int));
    m.tmp0 = take((w.tiles + 1) * sizeof(int));
    m.tmp1 = take((w.tiles + 1) * sizeof(int));
    m.imap = take(w.n_islots * w.rx * 64 * sizeof(int));
    m.rmap = take(w.n_rslots * c.tx * 64 * sizeof(int));
    m.dup = take((size_t)std::max(d->n_rec, 0) * sizeof(int));
    return m;
}

// Injection groups of one entry go to a per-lane map of every region holding
// their cell (the wavelet row); groups of several entries to the (tile, row)
// buckets.  Receivers: the first to claim its cell goes to the per-lane map of
// the owning tile, further receivers on the same cell to the buckets.
struct WaveListArgs {
    int tx, tz, K, nz, tiles_x, tiles_z;
    int n_groups;
    const int *loc, *order, *start;
    int n_rec;
    const int* rloc;
    int *icount, *istart, *rcount, *rstart;
    WaveInj* iitems;
    WaveRec* ritems;
    WaveMaps maps;
};

// calls f(tile, row in region, lane) for every region holding cell (ci, cj)
template <class F>
__device__ void for_regions(const WaveListArgs& a, int ci, int cj, F f) {
    const int rx = a.tx + 2 * a.K;
    for (int ti = max(0, (ci - a.K) / a.tx - 1); ti <= min(a.tiles_x - 1, (ci + a.K) / a.tx); ti++)
        for (int tk = max(0, (cj - a.K) / a.tz - 1); tk <= min(a.tiles_z - 1, (cj + a.K) / a.tz);
             tk++) {
            const int ri0 = ti * a.tx - a.K, rj0 = tk * a.tz - a.K;
            if (ci < ri0 || ci >= ri0 + rx || cj < rj0 || cj >= rj0 + 64) continue;
            f(ti * a.tiles_z + tk, ci - ri0, cj - rj0);
        }
}

// pass 0: flag tiles needing maps, count multi-entry groups per bucket
__global__ void k_wave_mark(WaveListArgs a) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int rx = a.tx + 2 * a.K;
    if (q < a.n_groups) {
        const int e0 = a.start[q], e1 = a.start[q + 1];
        const int k0 = a.order[e0];
        for_regions(a, a.loc[2 * k0], a.loc[2 * k0 + 1], [&](int tile, int r, int) {
            if (e1 - e0 == 1)
                a.maps.imap_slot[tile] = 1;
            else
                atomicAdd(a.icount + tile * rx + r, 1);
        });
    }
    if (q < a.n_rec) {
        const int ci = a.rloc[2 * q], cj = a.rloc[2 * q + 1];
        a.maps.rmap_slot[(ci / a.tx) * a.tiles_z + cj / a.tz] = 1;
    }
}

// flags -> compact slot indices (exclusive prefix of the flags), -1 = none
__global__ void k_wave_slots(int* islot, const int* ipre, int* rslot, const int* rpre, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    islot[k] = islot[k] ? ipre[k] : -1;
    rslot[k] = rslot[k] ? rpre[k] : -1;
}

// pass 1: maps and multi-entry buckets; receivers claim their cell
__global__ void k_wave_fill(WaveListArgs a) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int rx = a.tx + 2 * a.K;
    if (q < a.n_groups) {
        const int e0 = a.start[q], e1 = a.start[q + 1];
        const int k0 = a.order[e0];
        const int ci = a.loc[2 * k0], cj = a.loc[2 * k0 + 1];
        for_regions(a, ci, cj, [&](int tile, int r, int lane) {
            if (e1 - e0 == 1) {
                const int slot = a.maps.imap_slot[tile];
                a.maps.imap[((size_t)slot * rx + r) * 64 + lane] = k0;
            } else {
                const int b = tile * rx + r;
                a.iitems[a.istart[b] + atomicAdd(a.icount + b, 1)] =
                    WaveInj{lane, e0, e1, ci * a.nz + cj};
            }
        });
    }
    if (q < a.n_rec) {
        const int ci = a.rloc[2 * q], cj = a.rloc[2 * q + 1];
        const int ti = ci / a.tx, tk = cj / a.tz, t = ci - ti * a.tx;
        const int tile = ti * a.tiles_z + tk, lane = cj - (tk * a.tz - a.K);
        const int slot = a.maps.rmap_slot[tile];
        int* cell = a.maps.rmap + ((size_t)slot * a.tx + t) * 64 + lane;
        const bool first = atomicCAS(cell, -1, q) == -1;
        a.maps.dup[q] = first ? 0 : 1;
        if (!first) atomicAdd(a.rcount + tile * a.tx + t, 1);
    }
}

// pass 2: receivers sharing a cell with an earlier one go to the buckets
__global__ void k_wave_dups(WaveListArgs a) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.n_rec || !a.maps.dup[q]) return;
    const int ci = a.rloc[2 * q], cj = a.rloc[2 * q + 1];
    const int ti = ci / a.tx, tk = cj / a.tz, t = ci - ti * a.tx;
    const int b = (ti * a.tiles_z + tk) * a.tx + t;
    a.ritems[a.rstart[b] + atomicAdd(a.rcount + b, 1)] = WaveRec{cj - (tk * a.tz - a.K), q};
}

static void fill(int* p, size_t n, int v, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_fill_int_b, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p,
                           (int)n, v);
}

int wave_build_lists(const trtm_step_desc* d, BlockCfg c, int* istart, int* icur, void* iitems,
                     int* rstart, int* rcur, void* ritems, void* maps, hipStream_t s) {
    const ListSizes L = wave_list_sizes(d, c);
    const WaveCounts w = wave_counts(d, c);
    WaveListArgs a;
    a.tx = c.tx;
    a.tz = c.tz;
    a.K = c.k;
    a.nz = d->nz;
    a.tiles_x = tiles_of(d->nx, c.tx);
    a.tiles_z = tiles_of(d->nz, c.tz);
    a.n_groups = d->n_inj_groups;
    a.loc = d->inj_loc;
    a.order = d->inj_order;
    a.start = d->inj_group_start;
    a.n_rec = (d->rec_u || d->rec_v) ? d->n_rec : 0;
    a.rloc = d->rec_loc;
    a.icount = icur;
    a.istart = istart;
    a.rcount = rcur;
    a.rstart = rstart;
    a.iitems = static_cast<WaveInj*>(iitems);
    a.ritems = static_cast<WaveRec*>(ritems);
    a.maps = wave_maps(d, c, maps);
    const int nib = (int)L.inj_buckets - 1, nrb = (int)L.rec_buckets - 1;
    const int T = (int)w.tiles;
    const dim3 tb(256), lg((std::max(std::max(a.n_groups, a.n_rec), 1) + 255) / 256);
    fill(icur, nib, 0, s);
    fill(rcur, nrb, 0, s);
    fill(a.maps.imap_slot, T, 0, s);
    fill(a.maps.rmap_slot, T, 0, s);
    hipLaunchKernelGGL(k_wave_mark, lg, tb, 0, s, a);
    hipLaunchKernelGGL(k_scan_b, dim3(1), dim3(1024), 0, s, a.maps.imap_slot, a.maps.tmp0, T);
    hipLaunchKernelGGL(k_scan_b, dim3(1), dim3(1024), 0, s, a.maps.rmap_slot, a.maps.tmp1, T);
    hipLaunchKernelGGL(k_wave_slots, dim3((T + 255) / 256), tb, 0, s, a.maps.imap_slot,
                       (const int*)a.maps.tmp0, a.maps.rmap_slot, (const int*)a.maps.tmp1, T);
    fill(a.maps.imap, w.n_islots * w.rx * 64, -1, s);
    fill(a.maps.rmap, w.n_rslots * c.tx * 64, -1, s);
    hipLaunchKernelGGL(k_scan_b, dim3(1), dim3(1024), 0, s, icur, istart, nib);
    fill(icur, nib, 0, s);
    hipLaunchKernelGGL(k_wave_fill, lg, tb, 0, s, a);
    hipLaunchKernelGGL(k_scan_b, dim3(1), dim3(1024), 0, s, rcur, rstart, nrb);
    fill(rcur, nrb, 0, s);
    hipLaunchKernelGGL(k_wave_dups, lg, tb, 0, s, a);
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
