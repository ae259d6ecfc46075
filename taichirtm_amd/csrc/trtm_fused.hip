// Fused time step: one launch per step for the whole reference kernel group
// BC -> V -> inject -> S -> sponge -> record -> snapshot / image -> finite check
// (forward_modeling.py:539-572, backward_modeling.py:539-576).
//
// Each 256-thread workgroup owns a tile of TX x 64 cells (z contiguous, one
// wave64 per tile row).  Phase A computes the new velocities on the tile plus
// the one-cell halo the stress stencil needs (+x/+z forward, -x/-z backward)
// into LDS, reading the old state; injections are added to the LDS values.
// Phase B updates the tile's stresses from LDS, applies the sponge and writes
// the new state.  State ping-pongs between two buffers, so no tile reads a
// value another tile is writing in the same launch.  The tile's receivers are
// recorded from LDS.  Blocks are renumbered so that each XCD gets a contiguous
// band of tiles and neighbouring tiles share that XCD's L2.
#include <algorithm>

#include "trtm_common.hpp"

namespace trtm {

int fail(int code, const char* fmt, ...);

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

static int pick_tx(const trtm_step_desc* d) {
    (void)d;
    return 8;
}

struct WsLayout {
    size_t alt, inj_start, inj_cursor, inj_items, rec_start, rec_cursor, rec_items, total;
    int tiles_x, tiles_z;
};

static WsLayout ws_layout(const trtm_step_desc* d, int tx) {
    WsLayout L;
    const size_t n = (size_t)d->nx * d->nz;
    L.tiles_x = (d->nx + tx - 1) / tx;
    L.tiles_z = (d->nz + kTileZ - 1) / kTileZ;
    const size_t tiles = (size_t)L.tiles_x * L.tiles_z;
    size_t off = 0;
    L.alt = off;        off = align256(off + 8 * n * sizeof(float));
    L.inj_start = off;  off = align256(off + (tiles + 1) * sizeof(int));
    L.inj_cursor = off; off = align256(off + tiles * sizeof(int));
    L.inj_items = off;  off = align256(off + 4 * (size_t)std::max(d->n_inj_groups, 0) * sizeof(int));
    L.rec_start = off;  off = align256(off + (tiles + 1) * sizeof(int));
    L.rec_cursor = off; off = align256(off + tiles * sizeof(int));
    L.rec_items = off;  off = align256(off + (size_t)std::max(d->n_rec, 0) * sizeof(int));
    L.total = off;
    return L;
}

int64_t fused_workspace_bytes(const trtm_step_desc* d) {
    return (int64_t)ws_layout(d, pick_tx(d)).total;
}

__device__ __forceinline__ bool in_vregion(int ci, int cj, int tile_i, int tile_k, int tx,
                                           bool fwd) {
    const int ilo = tile_i * tx - (fwd ? 0 : 1), jlo = tile_k * kTileZ - (fwd ? 0 : 1);
    return ci >= ilo && ci <= ilo + tx && cj >= jlo && cj <= jlo + kTileZ;
}

// mode 0: count, mode 1: fill.  Injection groups go to every tile whose
// velocity region (tile plus halo) holds their cell.
__global__ void k_tiles_inj(int mode, int tx, int tiles_x, int tiles_z, bool fwd, int n_groups,
                            const int* loc, const int* order, const int* start, int* count,
                            const int* list_start, int* items) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    const int k0 = order[start[g]];
    const int ci = loc[2 * k0], cj = loc[2 * k0 + 1];
    for (int di = -1; di <= 1; di++)
        for (int dk = -1; dk <= 1; dk++) {
            const int ti = ci / tx + di, tk = cj / kTileZ + dk;
            if (ti < 0 || ti >= tiles_x || tk < 0 || tk >= tiles_z) continue;
            if (!in_vregion(ci, cj, ti, tk, tx, fwd)) continue;
            const int tile = ti * tiles_z + tk;
            const int pos = atomicAdd(count + tile, 1);
            if (mode == 1) items[list_start[tile] + pos] = g;
        }
}

// Receivers go to the one tile that holds their cell.
__global__ void k_tiles_rec(int mode, int tx, int tiles_z, int n_rec, const int* loc, int* count,
                            const int* list_start, int* items) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n_rec) return;
    const int tile = (loc[2 * l] / tx) * tiles_z + loc[2 * l + 1] / kTileZ;
    const int pos = atomicAdd(count + tile, 1);
    if (mode == 1) items[list_start[tile] + pos] = l;
}

__global__ void k_fill_int(int* p, int n, int value) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = value;
}

// Exclusive scan of count[0..n) into start[0..n], single workgroup of 1024.
__global__ __launch_bounds__(1024) void k_scan(const int* count, int* start, int n) {
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

int fused_prepare(const trtm_step_desc* d, bool fwd, hipStream_t s, TilePlan* plan) {
    const int tx = pick_tx(d);
    const WsLayout L = ws_layout(d, tx);
    char* ws = static_cast<char*>(d->workspace);
    const int tiles = L.tiles_x * L.tiles_z;
    int* inj_start = reinterpret_cast<int*>(ws + L.inj_start);
    int* inj_cursor = reinterpret_cast<int*>(ws + L.inj_cursor);
    int* inj_items = reinterpret_cast<int*>(ws + L.inj_items);
    int* rec_start = reinterpret_cast<int*>(ws + L.rec_start);
    int* rec_cursor = reinterpret_cast<int*>(ws + L.rec_cursor);
    int* rec_items = reinterpret_cast<int*>(ws + L.rec_items);
    const dim3 tb(256), tg((tiles + 255) / 256);
    hipLaunchKernelGGL(k_fill_int, tg, tb, 0, s, inj_cursor, tiles, 0);
    hipLaunchKernelGGL(k_fill_int, tg, tb, 0, s, rec_cursor, tiles, 0);
    const int ng = d->n_inj_groups;
    const int nr = (d->rec_u || d->rec_v) ? d->n_rec : 0;
    if (ng > 0)
        hipLaunchKernelGGL(k_tiles_inj, dim3((ng + 255) / 256), tb, 0, s, 0, tx, L.tiles_x,
                           L.tiles_z, fwd, ng, d->inj_loc, d->inj_order, d->inj_group_start,
                           inj_cursor, (const int*)nullptr, (int*)nullptr);
    if (nr > 0)
        hipLaunchKernelGGL(k_tiles_rec, dim3((nr + 255) / 256), tb, 0, s, 0, tx, L.tiles_z, nr,
                           d->rec_loc, rec_cursor, (const int*)nullptr, (int*)nullptr);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, inj_cursor, inj_start, tiles);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, rec_cursor, rec_start, tiles);
    hipLaunchKernelGGL(k_fill_int, tg, tb, 0, s, inj_cursor, tiles, 0);
    hipLaunchKernelGGL(k_fill_int, tg, tb, 0, s, rec_cursor, tiles, 0);
    if (ng > 0)
        hipLaunchKernelGGL(k_tiles_inj, dim3((ng + 255) / 256), tb, 0, s, 1, tx, L.tiles_x,
                           L.tiles_z, fwd, ng, d->inj_loc, d->inj_order, d->inj_group_start,
                           inj_cursor, (const int*)inj_start, inj_items);
    if (nr > 0)
        hipLaunchKernelGGL(k_tiles_rec, dim3((nr + 255) / 256), tb, 0, s, 1, tx, L.tiles_z, nr,
                           d->rec_loc, rec_cursor, (const int*)rec_start, rec_items);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(TRTM_ERR_HIP, "tile plan: %s", hipGetErrorString(e));
    plan->tx = tx;
    plan->tiles_x = L.tiles_x;
    plan->tiles_z = L.tiles_z;
    plan->inj_start = inj_start;
    plan->inj_items = inj_items;
    plan->rec_start = rec_start;
    plan->rec_items = rec_items;
    return TRTM_OK;
}

// Block id -> tile, so that XCD x (blocks x, x+8, x+16, ...) gets tiles
// [x*per, (x+1)*per): a contiguous band of tile rows.
__device__ __forceinline__ int xcd_tile(int bid, int tiles) {
    const int per = (tiles + 7) / 8;
    return (bid % 8) * per + bid / 8;
}

template <bool FWD, bool PSV, bool SH, int TX>
__global__ __launch_bounds__(256) void k_fused(KArgs a, KArgs o, InjArgs q, StepIO p,
                                               TilePlan pl) {
    constexpr int RW = kTileZ + 1;  // velocity region: tile + 1 halo column
    constexpr int RH = TX + 1;      //                  tile + 1 halo row
    __shared__ float sU[PSV ? RH : 1][RW];
    __shared__ float sW[PSV ? RH : 1][RW];
    __shared__ float sV[SH ? RH : 1][RW];

    const int tiles = pl.tiles_x * pl.tiles_z;
    const int tile = xcd_tile(blockIdx.x, tiles);
    if (tile >= tiles) return;  // whole block leaves together
    const int i0 = (tile / pl.tiles_z) * TX, j0 = (tile % pl.tiles_z) * kTileZ;
    const int ilo = FWD ? i0 : i0 - 1, jlo = FWD ? j0 : j0 - 1;
    const int nx = a.nx, nz = a.nz;
    const size_t NZ = (size_t)nz;
    const int tid = threadIdx.x;
    const float* __restrict__ sm = a.sm;
    auto masked = [&](const float* f, size_t c) { return sm ? f[c] * sm[c] : f[c]; };

    // Phase A: new velocities (before the sponge) on the region
    for (int e = tid; e < RH * RW; e += 256) {
        const int r = e / RW, cz = e - r * RW;
        const int i = ilo + r, j = jlo + cz;
        float nu = 0.0f, nw = 0.0f, nv = 0.0f;
        if (i >= 0 && i < nx && j >= 0 && j < nz) {
            const size_t c = (size_t)i * NZ + j;
            const bool inner = i >= 1 && i < nx - 1 && j >= 1 && j < nz - 1;
            if (PSV) {
                nu = a.u[c];
                nw = a.w[c];
                if (inner) {
                    float sxx_x, szz_z, sxz_x, sxz_z;
                    if (FWD) {
                        sxx_x = (a.sxx[c] - a.sxx[c - NZ]) * a.idx;
                        szz_z = (masked(a.szz, c) - masked(a.szz, c - 1)) * a.idz;
                        sxz_x = (masked(a.sxz, c) - masked(a.sxz, c - NZ)) * a.idx;
                        sxz_z = (masked(a.sxz, c) - masked(a.sxz, c - 1)) * a.idz;
                        nu = nu + ((sxx_x + sxz_z) * a.dt) * a.iru[c];
                        nw = nw + ((sxz_x + szz_z) * a.dt) * a.irw[c];
                    } else {
                        sxx_x = (a.sxx[c + NZ] - a.sxx[c]) * a.idx;
                        szz_z = (masked(a.szz, c + 1) - masked(a.szz, c)) * a.idz;
                        sxz_x = (masked(a.sxz, c + NZ) - masked(a.sxz, c)) * a.idx;
                        sxz_z = (masked(a.sxz, c + 1) - masked(a.sxz, c)) * a.idz;
                        nu = nu + ((-(sxx_x + sxz_z)) * a.dt) * a.iru[c];
                        nw = nw + ((-(sxz_x + szz_z)) * a.dt) * a.irw[c];
                    }
                }
            }
            if (SH) {
                nv = a.v[c];
                if (inner) {
                    float syx_x, syz_z;
                    if (FWD) {
                        syx_x = (a.syx[c] - a.syx[c - NZ]) * a.idx;
                        syz_z = (masked(a.syz, c) - masked(a.syz, c - 1)) * a.idz;
                        nv = nv + ((syx_x + syz_z) * a.dt) * a.ir[c];
                    } else {
                        syx_x = (a.syx[c + NZ] - a.syx[c]) * a.idx;
                        syz_z = (masked(a.syz, c + 1) - masked(a.syz, c)) * a.idz;
                        nv = nv + ((-(syx_x + syz_z)) * a.dt) * a.ir[c];
                    }
                }
            }
        }
        if (PSV) {
            sU[r][cz] = nu;
            sW[r][cz] = nw;
        }
        if (SH) sV[r][cz] = nv;
    }
    __syncthreads();

    // injection into the region, one thread per cell, entries in order
    if (q.n_groups > 0) {
        const int b = pl.inj_start[tile], end = pl.inj_start[tile + 1];
        const size_t nt = (size_t)a.nt;
        for (int k = b + tid; k < end; k += 256) {
            const int g = pl.inj_items[k];
            const int e0 = q.start[g], e1 = q.start[g + 1];
            const int k0 = q.order[e0];
            const int ci = q.loc[2 * k0], cj = q.loc[2 * k0 + 1];
            const int r = ci - ilo, cz = cj - jlo;
            const size_t c = (size_t)ci * NZ + cj;
            if (PSV) {
                float x = sU[r][cz], z = sW[r][cz];
                for (int e = e0; e < e1; e++) {
                    const size_t off = (size_t)q.order[e] * nt + q.col;
                    if (FWD) {
                        x = x + (q.su[off] * a.factor) * a.iru[c];
                        z = z + (q.sw[off] * a.factor) * a.irw[c];
                    } else {
                        x = x + q.su[off];
                        z = z + q.sw[off];
                    }
                }
                sU[r][cz] = x;
                sW[r][cz] = z;
            }
            if (SH) {
                float y = sV[r][cz];
                for (int e = e0; e < e1; e++) {
                    const size_t off = (size_t)q.order[e] * nt + q.col;
                    y = FWD ? y + (q.sv[off] * a.factor) * a.ir[c] : y + q.sv[off];
                }
                sV[r][cz] = y;
            }
        }
        __syncthreads();
    }

    // Phase B: stresses on the tile, sponge, new state, snapshot / image / check
    const int lane = tid & 63, wave = tid >> 6;
    const size_t n = (size_t)nx * NZ;
    const float sdt = FWD ? a.dt : -a.dt;
    int code = 0;
    for (int r = wave; r < TX; r += 4) {
        const int i = i0 + r, j = j0 + lane;
        if (i >= nx || j >= nz) continue;
        const size_t c = (size_t)i * NZ + j;
        const float A = sponge_at(i, j, nx, nz, a.fw, a.tab);
        const bool inner = i >= 1 && i < nx - 1 && j >= 1 && j < nz - 1;
        const int lr = i - ilo, lc = j - jlo;
        if (PSV) {
            float sxx = a.sxx[c], szz = masked(a.szz, c), sxz = masked(a.sxz, c);
            if (inner) {
                float u_x, u_z, w_x, w_z;
                if (FWD) {
                    u_x = (sU[lr + 1][lc] - sU[lr][lc]) * a.idx;
                    u_z = (sU[lr][lc + 1] - sU[lr][lc]) * a.idz;
                    w_x = (sW[lr + 1][lc] - sW[lr][lc]) * a.idx;
                    w_z = (sW[lr][lc + 1] - sW[lr][lc]) * a.idz;
                } else {
                    u_x = (sU[lr][lc] - sU[lr - 1][lc]) * a.idx;
                    u_z = (sU[lr][lc] - sU[lr][lc - 1]) * a.idz;
                    w_x = (sW[lr][lc] - sW[lr - 1][lc]) * a.idx;
                    w_z = (sW[lr][lc] - sW[lr][lc - 1]) * a.idz;
                }
                const float div_uw = u_x + w_z;
                const float lam = a.lam[c], mu = a.mu[c];
                sxx = sxx + sdt * (lam * div_uw + (2.0f * mu) * u_x);
                szz = szz + sdt * (lam * div_uw + (2.0f * mu) * w_z);
                sxz = sxz + (sdt * a.mxz[c]) * (u_z + w_x);
            }
            o.sxx[c] = sxx * A;
            o.szz[c] = szz * A;
            o.sxz[c] = sxz * A;
            const float u = sU[lr][lc] * A, w = sW[lr][lc] * A;
            o.u[c] = u;
            o.w[c] = w;
            if (p.snap_slot >= 0) {
                p.snap_u[(size_t)p.snap_slot * n + c] = u;
                p.snap_w[(size_t)p.snap_slot * n + c] = w;
            }
            if (p.image_slot >= 0) {
                p.image_u[c] = p.image_u[c] + p.snap_u[(size_t)p.image_slot * n + c] * u;
                p.image_w[c] = p.image_w[c] + p.snap_w[(size_t)p.image_slot * n + c] * w;
            }
            if (p.check_slot >= 0) {
                if (not_finite_or_big(u) && code < p.check_base + 1) code = p.check_base + 1;
                if (not_finite_or_big(w)) code = p.check_base + 3;
            }
        }
        if (SH) {
            float syx = a.syx[c], syz = masked(a.syz, c);
            if (inner) {
                float v_x, v_z;
                if (FWD) {
                    v_x = (sV[lr + 1][lc] - sV[lr][lc]) * a.idx;
                    v_z = (sV[lr][lc + 1] - sV[lr][lc]) * a.idz;
                } else {
                    v_x = (sV[lr][lc] - sV[lr - 1][lc]) * a.idx;
                    v_z = (sV[lr][lc] - sV[lr][lc - 1]) * a.idz;
                }
                syx = syx + (sdt * a.myx[c]) * v_x;
                syz = syz + (sdt * a.myz[c]) * v_z;
            }
            o.syx[c] = syx * A;
            o.syz[c] = syz * A;
            const float v = sV[lr][lc] * A;
            o.v[c] = v;
            if (p.snap_slot >= 0) p.snap_v[(size_t)p.snap_slot * n + c] = v;
            if (p.image_slot >= 0)
                p.image_v[c] = p.image_v[c] + p.snap_v[(size_t)p.image_slot * n + c] * v;
            if (p.check_slot >= 0 && not_finite_or_big(v) && code < p.check_base + 2)
                code = p.check_base + 2;
        }
    }
    if (code) atomicMax(p.check_flags + p.check_slot, code);

    // receivers of this tile, after the sponge (LDS is not written in phase B)
    if (p.n_rec > 0 && p.rec_row >= 0) {
        const int b = pl.rec_start[tile], end = pl.rec_start[tile + 1];
        for (int k = b + tid; k < end; k += 256) {
            const int l = pl.rec_items[k];
            const int ci = p.rec_loc[2 * l], cj = p.rec_loc[2 * l + 1];
            const float A = sponge_at(ci, cj, nx, nz, a.fw, a.tab);
            const size_t off = (size_t)p.rec_row * p.n_rec + l;
            const int lr = ci - ilo, lc = cj - jlo;
            if (PSV) {
                p.rec_u[off] = sU[lr][lc] * A;
                p.rec_w[off] = sW[lr][lc] * A;
            }
            if (SH) p.rec_v[off] = sV[lr][lc] * A;
        }
    }
}

template <bool FWD, bool PSV, bool SH>
static hipError_t fused_step(const KArgs& a, const KArgs& o, const InjArgs& q, const StepIO& p,
                             const TilePlan& pl, hipStream_t s) {
    const int tiles = pl.tiles_x * pl.tiles_z;
    const int grid = 8 * ((tiles + 7) / 8);
    hipLaunchKernelGGL((k_fused<FWD, PSV, SH, 8>), dim3(grid), dim3(256), 0, s, a, o, q, p, pl);
    return hipGetLastError();
}

hipError_t launch_fused_step(bool fwd, int systems, const KArgs& a, const KArgs& out,
                             const InjArgs& q, const StepIO& p, const TilePlan& plan,
                             hipStream_t s) {
    const bool psv = systems & TRTM_PSV, sh = systems & TRTM_SH;
    if (fwd) {
        if (psv && sh) return fused_step<true, true, true>(a, out, q, p, plan, s);
        if (psv) return fused_step<true, true, false>(a, out, q, p, plan, s);
        if (sh) return fused_step<true, false, true>(a, out, q, p, plan, s);
    } else {
        if (psv && sh) return fused_step<false, true, true>(a, out, q, p, plan, s);
        if (psv) return fused_step<false, true, false>(a, out, q, p, plan, s);
        if (sh) return fused_step<false, false, true>(a, out, q, p, plan, s);
    }
    return hipSuccess;
}

}  // namespace trtm
