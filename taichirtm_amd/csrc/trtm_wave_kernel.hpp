// Temporally blocked, multi-shot time stepping with one wavefront per tile
// and the tile's state in registers.
//
// A wavefront owns TX x (64 - 2K) cells of one shot.  Lane l holds the column
// z = z0 - K + l; the region's TX + 2K rows of every state field are
// registers of that lane.  The wavefront advances K time steps with no LDS
// traffic and no barriers: x-neighbours are other registers of the same lane
// and z-neighbours come from the adjacent lane through DPP wavefront shifts
// (the two edge lanes read 0).  Each half-step leaves one more edge row and
// one more edge lane stale, so after K steps exactly the owned cells are
// valid, and stale cells are never read by valid ones.
//
// Per step (forward_modeling.py:539-572, backward_modeling.py:539-576): the
// topography mask on the stresses, the velocity update, the injections, the
// stress update with the stress sponge, the velocity sponge, then from the
// post-sponge velocities of the owned cells: snapshots, the image (accumulated
// in registers, written once per launch), the finite check and the receivers.
// Every expression is the reference's, operand by operand, so results are
// bit-identical to the one-step-per-launch kernels.
//
// Injection groups and receivers are bucketed per (tile, row) on the device,
// so that every register index is a compile-time constant.
#pragma once

#include <type_traits>

#include "trtm_block.hpp"

namespace trtm {

// value held by lane l + 1 (lane 63 reads 0): DPP wave_shl:1
__device__ __forceinline__ float lane_next(float x) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x130, 0xF, 0xF, true));
}

// value held by lane l - 1 (lane 0 reads 0): DPP wave_shr:1
__device__ __forceinline__ float lane_prev(float x) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x138, 0xF, 0xF, true));
}

// injection item: target lane, entry range [e0, e1) of the group in order[],
// cell index; receiver item: target lane, receiver index
struct WaveInj {
    int lane, e0, e1, cell;
};
struct WaveRec {
    int lane, rec;
};

// A value every lane of the wavefront holds (loaded from a uniform address),
// moved to a scalar register so that branches on it stay uniform.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uni(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}

// An LDS value made opaque to the optimiser, so that a select between an
// updated and an old value stays a select instead of becoming a branch around
// the read (which would serialise the LDS latency row by row).
__device__ __forceinline__ float opaque(float x) {
    asm("" : "+v"(x));
    return x;
}

// Per-lane select with a uniform lane mask (a VOP3 v_cndmask): lanes whose
// bit is set take the updated value.  Written out so the update of every row
// stays in one basic block, where the scheduler interleaves independent rows,
// instead of one exec-masked block per row (a serial dependency chain each).
__device__ __forceinline__ float pick(uint64_t mask, float updated, float old) {
    float r;
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(old), "v"(updated), "s"(mask));
    return r;
}

template <bool FWD, bool PSV, bool SH, int TX, int K>
__global__ __launch_bounds__(kWaveThreads) void k_wave(KArgs m, const ShotDev* __restrict__ shots,
                                                       int it0, int tiles_x, int tiles_z) {
    constexpr int RX = TX + 2 * K, ZU = 64 - 2 * K;
    static_assert(RX <= 64, "row masks are 64 bits");
    // Per-lane columns of the region's P-SV medium and sponge factors: lane l
    // only ever reads what lane l wrote, so no barrier is needed.  Fields:
    // iru, irw, lam, mu, mxz (P-SV), sponge factor.  The SH medium stays in
    // registers.
    constexpr int FA = PSV ? 5 : 0, NF = FA + 1;
    __shared__ float sC[NF][RX][64];

    const int lane = threadIdx.x;
    const gptr<const ShotDev> sh = G(shots) + blockIdx.y;
    const int it_end = sh->it_end, it_lo = sh->it_lo, it_off = sh->it_off;
    const int tiles = tiles_x * tiles_z;
    const int tile = xcd_tile_b(blockIdx.x, tiles);
    if (tile >= tiles || it0 + it_off >= it_end || it0 + it_off + K <= it_lo) return;

    const int nx = m.nx, nz = m.nz, nt = sh->nt, fw = m.fw;
    const int ti = tile / tiles_z, tk = tile - ti * tiles_z;
    const int ri0 = ti * TX - K, rj0 = tk * ZU - K;
    const int j = rj0 + lane;
    const bool jg = j >= 0 && j < nz;
    const bool jin = j >= 1 && j < nz - 1;
    const bool jown = lane >= K && lane < K + ZU && j < nz;
    const int jc = min(max(j, 0), nz - 1);  // column used for addresses (always in the grid)
    // rows of the region inside the grid [glo, ghi) and inside its interior [rlo, rhi)
    const int glo = max(0, -ri0), ghi = min(RX, nx - ri0);
    const int rlo = max(0, 1 - ri0), rhi = min(RX, nx - 1 - ri0);
    auto rin = [&](int r) { return r >= glo && r < ghi; };
    // uniform pointer to the start of grid row ri0 + r of a field
    auto row = [&](auto* base, int r) { return G(base) + (ptrdiff_t)(ri0 + r) * nz; };
    const size_t n = (size_t)nx * nz;
    const float dt = m.dt, idx = m.idx, idz = m.idz;
    const float sdt = FWD ? m.dt : -m.dt;
    const bool frame = fw > 0 && (ri0 < fw || ri0 + RX > nx - fw || rj0 + 64 > nz - fw);
    // general tiles touch the sponge or the grid's x-edges; the others update
    // every row where the column is interior and skip the sponge
    const bool general = frame || rlo > 0 || rhi < RX;
    const uint64_t jin_mask = __builtin_amdgcn_ballot_w64(jin);

    // state, SH medium into registers; P-SV medium and sponge columns into LDS.
    // Loads use a uniform in-grid row and the clamped column, so every lane
    // loads; cells outside the grid are then set to 0.
    float U[RX], W[RX], XX[RX], ZZ[RX], XZ[RX], V[RX], YX[RX], YZ[RX];
    float cI[RX], cX[RX], cZ[RX];  // inv_rho, sdt * myx, sdt * myz
#pragma unroll
    for (int r = 0; r < RX; r++) {
        const int rr = rin(r) ? r : K;
        const bool g = jg && rin(r);
        auto ld = [&](auto* base) { const float x = row(base, rr)[jc]; return g ? x : 0.0f; };
        if (PSV) {
            sC[0][r][lane] = ld(m.iru);
            sC[1][r][lane] = ld(m.irw);
            sC[2][r][lane] = ld(m.lam);
            sC[3][r][lane] = ld(m.mu);
            sC[4][r][lane] = ld(m.mxz);
            U[r] = ld(sh->in[0]);
            W[r] = ld(sh->in[2]);
            XX[r] = ld(sh->in[3]);
            ZZ[r] = ld(sh->in[4]);
            XZ[r] = ld(sh->in[5]);
        }
        if (SH) {
            cI[r] = ld(m.ir);
            cX[r] = sdt * ld(m.myx);
            cZ[r] = sdt * ld(m.myz);
            V[r] = ld(sh->in[1]);
            YX[r] = ld(sh->in[6]);
            YZ[r] = ld(sh->in[7]);
        }
    }
    if (general) {  // factor 1 outside the frame (x-edge tiles without a sponge)
#pragma unroll
        for (int r = 0; r < RX; r++)
            sC[FA][r][lane] = (frame && jg && rin(r)) ? sponge_at(ri0 + r, j, nx, nz, fw, m.tab)
                                                      : 1.0f;
    }

    // image accumulators of the owned rows
    const bool imaging = !FWD && sh->num_snaps > 0;
    float aU[TX], aW[TX], aV[TX];
#pragma unroll
    for (int t = 0; t < TX; t++) aU[t] = aW[t] = aV[t] = 0.0f;
    if (imaging) {
#pragma unroll
        for (int t = 0; t < TX; t++) {
            const int r = K + t;
            const int rr = rin(r) ? r : K;
            const bool g = jown && rin(r);
            if (PSV) {
                const float x = row(sh->image_u, rr)[jc], z = row(sh->image_w, rr)[jc];
                aU[t] = g ? x : 0.0f;
                aW[t] = g ? z : 0.0f;
            }
            if (SH) {
                const float y = row(sh->image_v, rr)[jc];
                aV[t] = g ? y : 0.0f;
            }
        }
    }

    // rows of the region holding injection cells, owned rows holding receivers
    const int n_groups = sh->n_groups, n_rec = sh->n_rec;
    const cptr<int> istart = C(sh->inj_tstart) + (size_t)tile * RX;
    const cptr<int> rstart = C(sh->rec_tstart) + (size_t)tile * TX;
    uint64_t imask = 0, rmask = 0;
    if (n_groups > 0) {
#pragma unroll
        for (int r = 0; r < RX; r++)
            if (uni(istart[r]) < uni(istart[r + 1])) imask |= 1ull << r;
    }
    if (n_rec > 0) {
#pragma unroll
        for (int t = 0; t < TX; t++)
            if (uni(rstart[t]) < uni(rstart[t + 1])) rmask |= 1ull << t;
    }

    // per-lane maps: the wavelet row of a single-entry injection cell in each
    // region row, the receiver of each owned row (-1: none); loaded once
    int io[RX], ro[TX];
    uint64_t iomask = 0, romask = 0;
    {
        const int islot = n_groups > 0 ? uni(C(sh->imap_slot)[tile]) : -1;
        const int rslot = n_rec > 0 ? uni(C(sh->rmap_slot)[tile]) : -1;
#pragma unroll
        for (int r = 0; r < RX; r++) {
            io[r] = islot >= 0 ? G(sh->imap)[((size_t)islot * RX + r) * 64 + lane] : -1;
            if (__builtin_amdgcn_ballot_w64(io[r] >= 0)) iomask |= 1ull << r;
        }
#pragma unroll
        for (int t = 0; t < TX; t++) {
            ro[t] = rslot >= 0 ? G(sh->rmap)[((size_t)rslot * TX + t) * 64 + lane] : -1;
            if (__builtin_amdgcn_ballot_w64(ro[t] >= 0)) romask |= 1ull << t;
        }
    }

    const bool has_sm = m.sm != nullptr;
    const int save = sh->save, isnap = sh->isnap, num_snaps = sh->num_snaps;
    const int snap_base = sh->snap_base, cint = sh->check_interval;
    const int coff = ri0 * nz + jc;  // cell of (row r, this lane) = coff + r * nz, rows in the grid

    // Rows live at step s (valid inputs, results needed later): the velocity
    // update reads one stale row more on one side, the stress update on the
    // other, so the valid rows shrink by one on each side per step.  With the
    // steps unrolled these ranges are compile-time constants.
    auto velocity = [&](int r, auto gen) {
        const int rm = r > 0 ? r - 1 : 0, rp = r < RX - 1 ? r + 1 : RX - 1;
        uint64_t upd = jin_mask;
        if constexpr (decltype(gen)::value) upd = (r >= rlo && r < rhi) ? jin_mask : 0ull;
        if (PSV) {
            const float iru = opaque(sC[0][r][lane]), irw = opaque(sC[1][r][lane]);
            float sxx_x, szz_z, sxz_x, sxz_z, nu, nw;
            if (FWD) {
                const float zzp = lane_prev(ZZ[r]), xzp = lane_prev(XZ[r]);
                sxx_x = (XX[r] - XX[rm]) * idx;
                szz_z = (ZZ[r] - zzp) * idz;
                sxz_x = (XZ[r] - XZ[rm]) * idx;
                sxz_z = (XZ[r] - xzp) * idz;
                nu = U[r] + ((sxx_x + sxz_z) * dt) * iru;
                nw = W[r] + ((sxz_x + szz_z) * dt) * irw;
            } else {
                const float zzn = lane_next(ZZ[r]), xzn = lane_next(XZ[r]);
                sxx_x = (XX[rp] - XX[r]) * idx;
                szz_z = (zzn - ZZ[r]) * idz;
                sxz_x = (XZ[rp] - XZ[r]) * idx;
                sxz_z = (xzn - XZ[r]) * idz;
                nu = U[r] + ((-(sxx_x + sxz_z)) * dt) * iru;
                nw = W[r] + ((-(sxz_x + szz_z)) * dt) * irw;
            }
            U[r] = pick(upd, nu, U[r]);
            W[r] = pick(upd, nw, W[r]);
        }
        if (SH) {
            float syx_x, syz_z, nv;
            if (FWD) {
                const float yzp = lane_prev(YZ[r]);
                syx_x = (YX[r] - YX[rm]) * idx;
                syz_z = (YZ[r] - yzp) * idz;
                nv = V[r] + ((syx_x + syz_z) * dt) * cI[r];
            } else {
                const float yzn = lane_next(YZ[r]);
                syx_x = (YX[rp] - YX[r]) * idx;
                syz_z = (yzn - YZ[r]) * idz;
                nv = V[r] + ((-(syx_x + syz_z)) * dt) * cI[r];
            }
            V[r] = pick(upd, nv, V[r]);
        }
    };
    // stress update of row r; general tiles: row range and stress sponge
    auto stress = [&](int r, auto sp) {
        const int rm = r > 0 ? r - 1 : 0, rp = r < RX - 1 ? r + 1 : RX - 1;
        uint64_t upd = jin_mask;
        if constexpr (decltype(sp)::value) upd = (r >= rlo && r < rhi) ? jin_mask : 0ull;
        float a = 1.0f;
        if constexpr (decltype(sp)::value) a = opaque(sC[FA][r][lane]);
        if (PSV) {
            float u_x, u_z, w_x, w_z;
            if (FWD) {
                const float un = lane_next(U[r]), wn = lane_next(W[r]);
                u_x = (U[rp] - U[r]) * idx;
                u_z = (un - U[r]) * idz;
                w_x = (W[rp] - W[r]) * idx;
                w_z = (wn - W[r]) * idz;
            } else {
                const float up = lane_prev(U[r]), wp = lane_prev(W[r]);
                u_x = (U[r] - U[rm]) * idx;
                u_z = (U[r] - up) * idz;
                w_x = (W[r] - W[rm]) * idx;
                w_z = (W[r] - wp) * idz;
            }
            const float div_uw = u_x + w_z;
            const float lam = opaque(sC[2][r][lane]), mu = opaque(sC[3][r][lane]);
            const float mxz = opaque(sC[4][r][lane]);
            const float nxx = XX[r] + sdt * (lam * div_uw + (2.0f * mu) * u_x);
            const float nzz = ZZ[r] + sdt * (lam * div_uw + (2.0f * mu) * w_z);
            const float nxz = XZ[r] + (sdt * mxz) * (u_z + w_x);
            float sxx = pick(upd, nxx, XX[r]), szz = pick(upd, nzz, ZZ[r]);
            float sxz = pick(upd, nxz, XZ[r]);
            if constexpr (decltype(sp)::value) {
                sxx = sxx * a;
                szz = szz * a;
                sxz = sxz * a;
            }
            XX[r] = sxx;
            ZZ[r] = szz;
            XZ[r] = sxz;
        }
        if (SH) {
            float v_x, v_z;
            if (FWD) {
                const float vn = lane_next(V[r]);
                v_x = (V[rp] - V[r]) * idx;
                v_z = (vn - V[r]) * idz;
            } else {
                const float vp = lane_prev(V[r]);
                v_x = (V[r] - V[rm]) * idx;
                v_z = (V[r] - vp) * idz;
            }
            const float nyx = YX[r] + cX[r] * v_x;
            const float nyz = YZ[r] + cZ[r] * v_z;
            float syx = pick(upd, nyx, YX[r]), syz = pick(upd, nyz, YZ[r]);
            if constexpr (decltype(sp)::value) {
                syx = syx * a;
                syz = syz * a;
            }
            YX[r] = syx;
            YZ[r] = syz;
        }
    };

    // snapshot / image / check cadence: one division at the first step this
    // launch runs, then counters advanced by one per step
    bool started = false;
    int f_rem = 0, f_q = 0, b_rem = 0, b_q = 0, c_rem = 0, c_q = 0;
#pragma unroll
    for (int s = 0; s < K; s++) {
        const int it = it0 + it_off + s;
        if (it < it_lo || it >= it_end) continue;
        const int tb = nt - 1 - it;  // backward time index
        if (!started) {
            started = true;
            if (FWD) {
                f_rem = it % isnap;
                f_q = it / isnap;
            } else {
                b_rem = tb % isnap;
                b_q = tb / isnap;
            }
            if (cint > 0) {
                c_rem = it % cint;
                c_q = it / cint;
            }
        } else {
            if (FWD) {
                if (++f_rem == isnap) {
                    f_rem = 0;
                    f_q++;
                }
            } else if (--b_rem < 0) {
                b_rem = isnap - 1;
                b_q--;
            }
            if (cint > 0 && ++c_rem == cint) {
                c_rem = 0;
                c_q++;
            }
        }
        int snap = -1, img = -1, chk = -1;
        if (FWD && save && f_rem == 0 && it != 0) {
            const int sl = f_q - 1 - snap_base;
            if (sl >= 0 && sl < num_snaps) snap = sl;
        }
        if (imaging && tb > 0 && b_rem == 0) {
            const int sl = b_q - 1 - snap_base;
            if (sl >= 0 && sl < num_snaps && uni(C(sh->isnaps)[sl]) == tb) img = sl;
        }
        if (cint > 0 && c_rem == 0) chk = c_q;

        // forward snapshot values the image of this step correlates with
        float pU[TX], pW[TX], pV[TX];
        if (img >= 0) {
            const size_t so = (size_t)img * n;
            const gptr<const float> fu = G((const float*)sh->snap_u) + so;
            const gptr<const float> fw_ = G((const float*)sh->snap_w) + so;
            const gptr<const float> fv = G((const float*)sh->snap_v) + so;
#pragma unroll
            for (int t = 0; t < TX; t++) {
                const int c = coff + (rin(K + t) ? K + t : K) * nz;
                if (PSV) {
                    pU[t] = fu[c];
                    pW[t] = fw_[c];
                }
                if (SH) pV[t] = fv[c];
            }
        }

        // samples injected into the map cells this step (issued before the
        // velocity update, which hides their latency)
        float qU[RX], qW[RX], qV[RX];
        if (iomask) {
            const size_t col = (size_t)(FWD ? it : tb);
#pragma unroll
            for (int r = 0; r < RX; r++) {
                if (!((iomask >> r) & 1ull)) continue;
                const size_t o = (size_t)max(io[r], 0) * nt + col;
                if (PSV) {
                    qU[r] = G(sh->su)[o];
                    qW[r] = G(sh->sw)[o];
                }
                if (SH) qV[r] = G(sh->sv)[o];
            }
        }

        // topography: the stresses the velocity update reads are masked in place
        if (has_sm) {
            const gptr<const float> smp = G(m.sm);
#pragma unroll
            for (int r = 0; r < RX; r++) {
                const float q = smp[coff + (rin(r) ? r : K) * nz];
                if (PSV) {
                    ZZ[r] = ZZ[r] * q;
                    XZ[r] = XZ[r] * q;
                }
                if (SH) YZ[r] = YZ[r] * q;
            }
        }

        // ---- velocity update (live rows)
        constexpr int v_lo = FWD ? 1 : 0, v_hi = FWD ? 0 : 1;
        if (general) {
#pragma unroll
            for (int r = 0; r < RX; r++)
                if (r >= s + v_lo && r < RX - s - v_hi) velocity(r, std::true_type{});
        } else {
#pragma unroll
            for (int r = 0; r < RX; r++)
                if (r >= s + v_lo && r < RX - s - v_hi) velocity(r, std::false_type{});
        }

        // ---- injections: single-entry cells from the map ...
        if (iomask) {
#pragma unroll
            for (int r = 0; r < RX; r++) {
                if (!((iomask >> r) & 1ull)) continue;
                const uint64_t mk = __builtin_amdgcn_ballot_w64(io[r] >= 0);
                if (PSV) {
                    const float x = FWD ? U[r] + (qU[r] * m.factor) * opaque(sC[0][r][lane])
                                        : U[r] + qU[r];
                    const float z = FWD ? W[r] + (qW[r] * m.factor) * opaque(sC[1][r][lane])
                                        : W[r] + qW[r];
                    U[r] = pick(mk, x, U[r]);
                    W[r] = pick(mk, z, W[r]);
                }
                if (SH) {
                    const float y = FWD ? V[r] + (qV[r] * m.factor) * cI[r] : V[r] + qV[r];
                    V[r] = pick(mk, y, V[r]);
                }
            }
        }
        // ... and cells with several entries, added in entry order
        if (imask) {
            const cptr<WaveInj> items = C(reinterpret_cast<const WaveInj*>(sh->inj_titems));
            const cptr<int> order = C(sh->order);
            const cptr<float> su = C(sh->su), sv = C(sh->sv), sw = C(sh->sw);
            const cptr<float> iru = C(m.iru), irw = C(m.irw), ir = C(m.ir);
            const size_t col = (size_t)(FWD ? it : tb);
            const size_t ntt = (size_t)nt;
#pragma unroll
            for (int r = 0; r < RX; r++) {
                if (!((imask >> r) & 1ull)) continue;
                const int b = uni(istart[r]), e = uni(istart[r + 1]);
                for (int q = b; q < e; q++) {
                    const int gl = uni(items[q].lane), e0 = uni(items[q].e0);
                    const int e1 = uni(items[q].e1), gc = uni(items[q].cell);
                    const bool me = lane == gl;
                    if (PSV) {
                        float x = U[r], z = W[r];
                        for (int q2 = e0; q2 < e1; q2++) {
                            const size_t o = (size_t)uni(order[q2]) * ntt + col;
                            if (FWD) {
                                x = x + (uni(su[o]) * m.factor) * uni(iru[gc]);
                                z = z + (uni(sw[o]) * m.factor) * uni(irw[gc]);
                            } else {
                                x = x + uni(su[o]);
                                z = z + uni(sw[o]);
                            }
                        }
                        U[r] = me ? x : U[r];
                        W[r] = me ? z : W[r];
                    }
                    if (SH) {
                        float y = V[r];
                        for (int q2 = e0; q2 < e1; q2++) {
                            const size_t o = (size_t)uni(order[q2]) * ntt + col;
                            y = FWD ? y + (uni(sv[o]) * m.factor) * uni(ir[gc]) : y + uni(sv[o]);
                        }
                        V[r] = me ? y : V[r];
                    }
                }
            }
        }

        // ---- stress update + stress sponge, then the velocity sponge (live rows)
        if (general) {
#pragma unroll
            for (int r = 0; r < RX; r++)
                if (r >= s + 1 && r < RX - s - 1) stress(r, std::true_type{});
#pragma unroll
            for (int r = 0; r < RX; r++) {
                if (r < s + 1 || r >= RX - s - 1) continue;
                const float a = opaque(sC[FA][r][lane]);
                if (PSV) {
                    U[r] = U[r] * a;
                    W[r] = W[r] * a;
                }
                if (SH) V[r] = V[r] * a;
            }
        } else {
#pragma unroll
            for (int r = 0; r < RX; r++)
                if (r >= s + 1 && r < RX - s - 1) stress(r, std::false_type{});
        }

        // ---- snapshots / image / finite check of the owned cells (post-sponge)
        if (snap >= 0) {
            const size_t so = (size_t)snap * n;
            const gptr<float> fu = G(sh->snap_u) + so, fw_ = G(sh->snap_w) + so;
            const gptr<float> fv = G(sh->snap_v) + so;
            if (jown) {
#pragma unroll
                for (int t = 0; t < TX; t++) {
                    const int r = K + t;
                    if (!rin(r)) continue;
                    const int c = coff + r * nz;
                    if (PSV) {
                        fu[c] = U[r];
                        fw_[c] = W[r];
                    }
                    if (SH) fv[c] = V[r];
                }
            }
        }
        if (img >= 0) {
#pragma unroll
            for (int t = 0; t < TX; t++) {
                const int r = K + t;
                if (PSV) {
                    aU[t] = aU[t] + pU[t] * U[r];
                    aW[t] = aW[t] + pW[t] * W[r];
                }
                if (SH) aV[t] = aV[t] + pV[t] * V[r];
            }
        }
        if (chk >= 0) {
            int code = 0;
#pragma unroll
            for (int t = 0; t < TX; t++) {
                const int r = K + t;
                if (!rin(r) || !jown) continue;
                if (PSV) {
                    if (not_finite_or_big(U[r]) && code < 1) code = 1;
                    if (not_finite_or_big(W[r])) code = 3;
                }
                if (SH && not_finite_or_big(V[r]) && code < 2) code = 2;
            }
            const int wc = __ballot(code == 3) ? 3 : __ballot(code == 2) ? 2 : __ballot(code == 1) ? 1 : 0;
            if (wc && lane == 0) atomicMax(sh->check_flags + chk, wc + (FWD ? 0 : 3));
        }

        // ---- receivers of the owned rows, after the sponge: the first receiver
        // of each cell from the map ...
        if (romask) {
            const size_t rrow = (size_t)(FWD ? it : tb) * n_rec;
            const gptr<float> ru = G(sh->rec_u) + rrow, rv = G(sh->rec_v) + rrow;
            const gptr<float> rw = G(sh->rec_w) + rrow;
#pragma unroll
            for (int t = 0; t < TX; t++) {
                if (!((romask >> t) & 1ull)) continue;
                if (ro[t] >= 0) {
                    if (PSV) {
                        ru[ro[t]] = U[K + t];
                        rw[ro[t]] = W[K + t];
                    }
                    if (SH) rv[ro[t]] = V[K + t];
                }
            }
        }
        // ... and further receivers on the same cell
        if (rmask) {
            const cptr<WaveRec> ritems = C(reinterpret_cast<const WaveRec*>(sh->rec_titems));
            const size_t rrow = (size_t)(FWD ? it : tb) * n_rec;
            const gptr<float> ru = G(sh->rec_u) + rrow, rv = G(sh->rec_v) + rrow;
            const gptr<float> rw = G(sh->rec_w) + rrow;
#pragma unroll
            for (int t = 0; t < TX; t++) {
                if (!((rmask >> t) & 1ull)) continue;
                const int b = uni(rstart[t]), e = uni(rstart[t + 1]);
                for (int q = b; q < e; q++) {
                    const int gl = uni(ritems[q].lane), l = uni(ritems[q].rec);
                    if (lane == gl) {
                        if (PSV) {
                            ru[l] = U[K + t];
                            rw[l] = W[K + t];
                        }
                        if (SH) rv[l] = V[K + t];
                    }
                }
            }
        }
    }

    // ---- owned cells: state (post-sponge) and image
    if (jown) {
#pragma unroll
        for (int t = 0; t < TX; t++) {
            const int r = K + t;
            if (!rin(r)) continue;
            const int c = coff + r * nz;
            if (PSV) {
                G(sh->out[0])[c] = U[r];
                G(sh->out[2])[c] = W[r];
                G(sh->out[3])[c] = XX[r];
                G(sh->out[4])[c] = ZZ[r];
                G(sh->out[5])[c] = XZ[r];
                if (imaging) {
                    G(sh->image_u)[c] = aU[t];
                    G(sh->image_w)[c] = aW[t];
                }
            }
            if (SH) {
                G(sh->out[1])[c] = V[r];
                G(sh->out[6])[c] = YX[r];
                G(sh->out[7])[c] = YZ[r];
                if (imaging) G(sh->image_v)[c] = aV[t];
            }
        }
    }
}

template <bool FWD, bool PSV, bool SH, int TX, int K>
void launch_wave_cfg(const KArgs& m, const ShotDev* shots, int n_shots, int it0, int tiles_x,
                     int tiles_z, hipStream_t s) {
    const int tiles = tiles_x * tiles_z;
    const dim3 grid(8 * ((tiles + 7) / 8), n_shots);
    hipLaunchKernelGGL((k_wave<FWD, PSV, SH, TX, K>), grid, dim3(kWaveThreads), 0, s, m, shots,
                       it0, tiles_x, tiles_z);
}

// one instantiation unit per wave system and direction (trtm_wave_<sys>_<dir>.hip)
#define TRTM_WAVE_UNIT(NAME) \
    bool NAME(int tx, int k, const KArgs& m, const ShotDev* shots, int n_shots, int it0, \
              int tiles_x, int tiles_z, hipStream_t s)
TRTM_WAVE_UNIT(launch_wave_sh_fwd);
TRTM_WAVE_UNIT(launch_wave_sh_bwd);
TRTM_WAVE_UNIT(launch_wave_psv_fwd);
TRTM_WAVE_UNIT(launch_wave_psv_bwd);
TRTM_WAVE_UNIT(launch_wave_both_fwd);
TRTM_WAVE_UNIT(launch_wave_both_bwd);

}  // namespace trtm
