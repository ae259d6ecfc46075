// Shared definitions for the gfx950 elastic FD / RTM kernels.
//
// Floating-point contract: every kernel evaluates the reference's float32
// expressions operand by operand in the reference's order (built with
// -ffp-contract=off, no reassociation), so results are bitwise reproducible
// against the CPU restatement of the reference built the same way.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "taichirtm_hip.h"

namespace trtm {

constexpr int kBlockZ = 64;  // one wave64 spans 64 consecutive z cells
constexpr int kBlockX = 4;   // 4 waves per 256-thread workgroup

// Everything a step kernel reads or writes, passed by value as kernel args.
struct KArgs {
    int nx, nz, nt, fw;
    float dt, idx, idz, factor;
    const float* tab;  // sponge table, fw+1 entries
    const float *lam, *mu, *mxz, *myx, *myz, *ir, *iru, *irw;
    const float* sm;  // surface matrix or nullptr
    float *u, *v, *w, *sxx, *szz, *sxz, *syx, *syz;
};

// Cerjan sponge coefficient of cell (i, j), forward_modeling.py:203-241.
// The reference fills three strips in sequence (left, right, bottom) with
// later writes winning; evaluating the strips in reverse order and returning
// the first hit gives the same value for every fw <= min(nx, nz).
__device__ __forceinline__ float sponge_at(int i, int j, int nx, int nz, int fw,
                                           const float* __restrict__ tab) {
    const int jb = nz - 1 - j;  // distance to the bottom edge
    if (jb < fw && i >= jb && i < nx - jb) return tab[fw - jb];
    const int ir = nx - 1 - i;  // distance to the right edge
    if (ir < fw && j < nz - ir - 1) return tab[fw - ir];
    if (i < fw && j < nz - i - 1) return tab[fw - i];
    return 1.0f;
}

__device__ __forceinline__ bool not_finite_or_big(float x) {
    return (x != x) || (fabsf(x) > 1e30f);
}

// Per-step extras for the post-update kernels: which snapshot slot to write
// (forward) or correlate (backward), which finite-check slot to fill; -1 = none.
struct StepIO {
    int snap_slot;
    int image_slot;
    int check_slot;
    int check_base;  // 0 forward, 3 backward
    int rec_row;     // row of the time-major record buffers
    float *snap_u, *snap_v, *snap_w;
    float *image_u, *image_v, *image_w;
    int* check_flags;
    int n_rec;
    const int* rec_loc;
    float *rec_u, *rec_v, *rec_w;
};

struct InjArgs {
    int n_groups;
    int col;  // wavelet / observed-data column of this step
    const int* loc;
    const int* order;
    const int* start;
    const float *su, *sv, *sw;
};

// Per-call tiling of the fused kernel and the per-tile lists of injection
// groups (cells inside the tile's velocity region, halo included) and of
// receivers (cells inside the tile), built on the device in the workspace.
struct TilePlan {
    int tx;  // rows (x) per tile; a tile is tx x kTileZ cells
    int tiles_x, tiles_z;
    const int* inj_start;  // tiles+1
    const int* inj_items;  // group indices
    const int* rec_start;  // tiles+1
    const int* rec_items;  // receiver indices
};

constexpr int kTileZ = 64;

}  // namespace trtm
