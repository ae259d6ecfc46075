// Temporally blocked FWI time stepping (float): K steps per launch, up to
// kShotsPerLaunch shots per launch (shot = blockIdx.z).
//
// Forward (k_fwi_fwd_blk): the forward step of forward_modeling.py:539-551
// (flat free surface) with the layout of the blocked RTM kernel: a workgroup
// loads its tile plus a K-cell halo of the state into LDS, keeps the region's
// coefficients and sponge factors in registers and advances K steps; halo
// cells are recomputed redundantly with the same expressions, so every value
// is bit-identical to the one-step-per-launch FWI kernels.  Sources and
// receivers come from the per-cell group maps (source groups are injected at
// every region cell that holds one, records are written by the tile that owns
// the cell).  With HIST the tile also writes the step's history that the
// adjoint reads: the post-injection velocities of every step and the pre-step
// stresses of every hist_every-th step (the compact layout, hist_every > 0), or
// both of every step (hist_every = 0, the layout of the one-step-per-launch
// adjoint).
//
// Adjoint (k_fwi_adj_blk): the exact adjoint step (residual injection, sponge
// transpose, stress-update transpose with the lam / mu / mxz (myx / myz)
// gradients, source-term gradient, velocity-update transpose with the
// inv_rho_u / inv_rho_w (inv_rho) gradients), in gather form, K steps per
// launch in descending time.  Each thread keeps the adjoint state of its region
// cells in registers; only the values neighbours gather (the stress
// transposes q and the scaled velocity adjoints b) pass through LDS.  The
// gradients of the tile's cells are accumulated in registers in the same
// order as the per-step kernels and written once per launch.  It reads the
// compact history with hist_every = K: a launch lies inside one interval of
// K steps whose first step's stresses are stored, and the pre-step stresses of
// the interval's later steps are recomputed at the start of the launch, over
// the tile and its low ring, by the forward stress update and sponge from the
// stored velocities (the forward's expressions in its order, so bit for bit).
//
// State ping-pongs between the caller's buffers and a workspace copy (tiles
// read the halo that neighbouring tiles write).  A call issues an even number
// of launches, so the state ends in the caller's buffers; a shot whose record
// ends inside the range (and so skips launches) is copied back at the end.
#include <algorithm>
#include <vector>

#include "trtm_block.hpp"
#include "trtm_fwi.hpp"

namespace trtm {

struct FB {
    int nx, nz, nt, fw;
    int n_rec, hist_base, hist_len, flip;  // flip: the current state is in alt
    int hist_every;                        // compact history: stresses of every hist_every-th step
    float dt, idx, idz, factor;
    const float* tab;
    const float *lam, *mu, *mxz, *myx, *myz, *ir, *iru, *irw;
    float* st[8];   // caller's state: u v w sxx szz sxz syx syz (adjoint: pu ... psyz)
    float* alt[8];  // workspace copy
    const int *src_map, *rec_map;
    const int *src_order, *src_start, *rec_order, *rec_start;
    const float *wav_u, *wav_v, *wav_w;
    float *rec_u, *rec_v, *rec_w;        // forward records, time-major (nt, n_rec)
    const float *res_u, *res_v, *res_w;  // data adjoint, time-major (nt, n_rec)
    float* hist;
    float *g_lam, *g_mu, *g_mxz, *g_myx, *g_myz, *g_ir, *g_iru, *g_irw;
};
static_assert(sizeof(ByArray<FB>) + 32 <= 4096, "per-shot kernel arguments exceed 4 KB");

// history fields of one step, in the order sxx, szz, sxz, u, w, syx, syz, v
// over the active systems
template <bool PSV, bool SH>
struct HF {
    static constexpr int sxx = 0, szz = 1, sxz = 2, u = 3, w = 4;
    static constexpr int syx = PSV ? 5 : 0, syz = PSV ? 6 : 1, v = PSV ? 7 : 2;
    static constexpr int n = (PSV ? 5 : 0) + (SH ? 3 : 0);
};

// compact history of hist_len steps: the velocity slots (one per step: u, w, v
// of the active systems), then the stress slots (one per hist_every steps,
// from hist_base: sxx, szz, sxz, syx, syz of the active systems)
template <bool PSV, bool SH>
struct HC {
    static constexpr int u = 0, w = 1, v = PSV ? 2 : 0;
    static constexpr int nv = (PSV ? 2 : 0) + (SH ? 1 : 0);
    static constexpr int sxx = 0, szz = 1, sxz = 2, syx = PSV ? 3 : 0, syz = PSV ? 4 : 1;
    static constexpr int ns = (PSV ? 3 : 0) + (SH ? 2 : 0);
};

// Entry of the sponge table that sponge_at reads for cell (i, j), or -1 for
// the coefficient 1 (same strip order as sponge_at).
__device__ __forceinline__ int sponge_slot(int i, int j, int nx, int nz, int fw) {
    const int jb = nz - 1 - j;
    if (jb < fw && i >= jb && i < nx - jb) return fw - jb;
    const int ir = nx - 1 - i;
    if (ir < fw && j < nz - ir - 1) return fw - ir;
    if (i < fw && j < nz - i - 1) return fw - i;
    return -1;
}

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// Element c of a global array as a uniform base plus a 32-bit byte offset:
// the saddr form of the global loads and stores, no 64-bit address per lane
// (grids below 2^30 cells).
template <class T>
__device__ __forceinline__ T ld(const T* base, int c) {
    const gptr<const char> b = (gptr<const char>)(const char*)base;
    return *(gptr<const T>)(b + (unsigned)c * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ void st(T* base, int c, T v) {
    const gptr<char> b = (gptr<char>)(char*)base;
    *(gptr<T>)(b + (unsigned)c * (unsigned)sizeof(T)) = v;
}

// Workgroup barrier ordering LDS only: loads still in flight to LDS (the
// asynchronous history copies) are not waited for here but by vm_wait().
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}
// wait for every vector-memory load of this wave (vmcnt 0)
__device__ __forceinline__ void vm_wait() { __builtin_amdgcn_s_waitcnt(0x0F70); }
// wait until at most N vector-memory loads of this wave are in flight
// (vmcnt holds 6 bits: a larger N waits for 63, which only waits longer)
template <int N>
__device__ __forceinline__ void vm_wait_n() {
    constexpr int M = N < 63 ? N : 63;
    __builtin_amdgcn_s_waitcnt(0x0F70 | (M & 15) | ((M >> 4) << 14));
}
// wait until at most n * Q loads are in flight (n = 0 .. 8; n is a constant
// once the caller's loop is unrolled)
template <int Q>
__device__ __forceinline__ void vm_wait_windows(int n) {
    switch (n) {
    case 0: vm_wait_n<0>(); break;
    case 1: vm_wait_n<Q>(); break;
    case 2: vm_wait_n<2 * Q>(); break;
    case 3: vm_wait_n<3 * Q>(); break;
    case 4: vm_wait_n<4 * Q>(); break;
    case 5: vm_wait_n<5 * Q>(); break;
    case 6: vm_wait_n<6 * Q>(); break;
    case 7: vm_wait_n<7 * Q>(); break;
    default: vm_wait_n<8 * Q>(); break;
    }
}

// Adjoint history window of a tile: rows i0-1 .. i0+TX, columns j0-1 .. j0+TZ,
// in 64-cell chunks per field.  The velocity windows come in two buffers (the
// step being taken and the next); the stress windows in `slots` buffers: the
// recomputed pre-step stresses of the interval's later steps, and the stored
// stresses of its first step.
template <bool PSV, bool SH, int TX, int TZ>
struct HistWindow {
    static constexpr int HX = TX + 2, HZ = TZ + 2;
    static constexpr int chunks = (HX * HZ + 63) / 64;
    static constexpr int field = chunks * 64;  // floats per field
    static constexpr int vfloats = HC<PSV, SH>::nv * field;  // one velocity window
    static constexpr int sfloats = HC<PSV, SH>::ns * field;  // one stress window
};
// steps between stored stresses read by the adjoint of K steps per launch (a
// launch of 8 steps spans two such intervals)
constexpr int adj_stress_every(int K) { return K < 4 ? K : 4; }
template <int KS>
constexpr int adj_stress_slots() { return KS - 1 > 2 ? KS - 1 : 2; }

template <bool PSV, bool SH, int TX, int TZ, int K>
constexpr size_t adj_lds_bytes() {
    using W = HistWindow<PSV, SH, TX, TZ>;
    constexpr int RN = (TX + 2 * K) * (TZ + 2 * K);
    return 4 * (2 * (size_t)W::vfloats +
                (size_t)adj_stress_slots<adj_stress_every(K)>() * W::sfloats +
                (size_t)((PSV ? 5 : 0) + (SH ? 3 : 0)) * RN);
}
constexpr size_t kLdsLimit = 160 * 1024;

// Region set-up shared by both kernels.  Every global load of the prologue is
// unconditional (cells outside the grid read cell 0 and are masked by their
// flags), so the loads of all of a thread's cells are in flight together.
struct RegionCell {
    int c;       // global offset (0 outside the grid)
    unsigned f;  // bit0 in grid, bit1 interior, bit2 in tile, bit3 reads -1 ok, bit4 reads +1 ok
    int slot;    // sponge table entry, -1 for 1
};

template <int TX, int TZ, int K>
__device__ __forceinline__ RegionCell region_cell(int e, int i0, int j0, int nx, int nz, int fw) {
    constexpr int RX = TX + 2 * K, RZ = TZ + 2 * K, RN = RX * RZ;
    const int r = e / RZ, z = e - r * RZ;
    const int i = i0 - K + r, j = j0 - K + z;
    RegionCell rc{0, 0u, -1};
    if (e < RN && i >= 0 && i < nx && j >= 0 && j < nz) {
        rc.c = i * nz + j;
        rc.f = 1u;
        if (i >= 1 && i < nx - 1 && j >= 1 && j < nz - 1) rc.f |= 2u;
        if (i >= i0 && i < i0 + TX && j >= j0 && j < j0 + TZ) rc.f |= 4u;
        if (r >= 1 && z >= 1) rc.f |= 8u;
        if (r <= RX - 2 && z <= RZ - 2) rc.f |= 16u;
        rc.slot = sponge_slot(i, j, nx, nz, fw);
    }
    return rc;
}

// ---------------------------------------------------------------------------
// forward, steps t_a .. t_b-1 (t_b - t_a <= K)

// Waves per SIMD asked of the forward kernel: the 8-step shapes whose LDS
// leaves room for two workgroups per CU get the registers that allow two
// (24 x 64 with 768 threads: 6 waves, 85 VGPRs; 32 x 64 with 1024: 8 waves,
// 64 VGPRs); the rest keep the compiler's choice.
template <bool PSV, int TX, int TZ, int NT, int K>
constexpr int fwd_waves_per_eu() {
    return (PSV && K == 8 && TX == 24 && TZ == 64 && NT == 768)    ? 6
           : (PSV && K == 8 && TX == 32 && TZ == 64 && NT == 1024) ? 8
           : (PSV && K == 8 && TX == 48 && TZ == 48 && NT == 1024) ? 8
           : (PSV && K == 8 && TX == 40 && TZ == 56 && NT == 1024) ? 8
                                                                   : 1;
}

template <bool PSV, bool SH, int TX, int TZ, int NT, int K, bool HIST, class S>
__global__ __launch_bounds__(NT, (fwd_waves_per_eu<PSV && !SH, TX, TZ, NT, K>())) void
k_fwi_fwd_blk(S src, int t_a, int t_b, int tiles_x, int tiles_z) {
    constexpr int RX = TX + 2 * K, RZ = TZ + 2 * K, RN = RX * RZ;
    constexpr int NC = (RN + NT - 1) / NT;
    constexpr int NP = PSV ? RN : 1, NS = SH ? RN : 1;
    using H = HF<PSV, SH>;
    __shared__ float sU[NP], sW[NP], sXX[NP], sZZ[NP], sXZ[NP];
    __shared__ float sV[NS], sYX[NS], sYZ[NS];

    const FB& a = src.get();
    const int tiles = tiles_x * tiles_z;
    const int tile = xcd_tile_b(blockIdx.x, tiles);
    const int nt = a.nt;
    if (tile >= tiles || t_a >= nt) return;
    const int i0 = (tile / tiles_z) * TX, j0 = (tile % tiles_z) * TZ;
    const int ri0 = i0 - K, rj0 = j0 - K;
    const int nx = a.nx, nz = a.nz;
    const int t = threadIdx.x;
    const bool has_src = a.src_map != nullptr, has_rec = a.rec_map != nullptr;

    int gc[NC], sg[NC], rg[NC];
    unsigned fl[NC];  // bit0 in grid, bit1 interior, bit2 in tile, bit3 V ok, bit4 S ok
    float A[NC], c0[NC], c1[NC], c2[NC], c3[NC], c4[NC], d0[NC], d1[NC], d2[NC];
    int slot[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const RegionCell rc = region_cell<TX, TZ, K>(t + NT * k, i0, j0, nx, nz, a.fw);
        gc[k] = rc.c;
        fl[k] = rc.f;
        slot[k] = rc.slot;
    }
    {
        float* const* in = a.flip ? a.alt : a.st;
        const gptr<const float> tab = G(a.tab);
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const int e = t + NT * k;
            const int c = gc[k];
            const bool g = fl[k] & 1u;
            const float sa = tab[slot[k] > 0 ? slot[k] : 0];
            A[k] = slot[k] >= 0 ? sa : 1.0f;
            sg[k] = has_src ? ld(a.src_map, c) : -1;
            rg[k] = has_rec ? ld(a.rec_map, c) : -1;
            if (!g) sg[k] = -1;
            if ((fl[k] & 5u) != 5u) rg[k] = -1;
            if (PSV) {
                c0[k] = ld(a.iru, c);
                c1[k] = ld(a.irw, c);
                c2[k] = ld(a.lam, c);
                c3[k] = ld(a.mu, c);
                c4[k] = ld(a.mxz, c);
                const float u = ld(in[0], c), w = ld(in[2], c), xx = ld(in[3], c), zz = ld(in[4], c),
                            xz = ld(in[5], c);
                if (e < RN) {
                    sU[e] = g ? u : 0.0f;
                    sW[e] = g ? w : 0.0f;
                    sXX[e] = g ? xx : 0.0f;
                    sZZ[e] = g ? zz : 0.0f;
                    sXZ[e] = g ? xz : 0.0f;
                }
            }
            if (SH) {
                d0[k] = ld(a.ir, c);
                d1[k] = ld(a.myx, c);
                d2[k] = ld(a.myz, c);
                const float v = ld(in[1], c), yx = ld(in[6], c), yz = ld(in[7], c);
                if (e < RN) {
                    sV[e] = g ? v : 0.0f;
                    sYX[e] = g ? yx : 0.0f;
                    sYZ[e] = g ? yz : 0.0f;
                }
            }
        }
    }
    __syncthreads();

    const size_t n = (size_t)nx * nz;
    const float idx = a.idx, idz = a.idz, dt = a.dt, factor = a.factor;
    const int hi = t_b < nt ? t_b : nt;
    const gptr<const int> sorder = G(a.src_order), sstart = G(a.src_start);
    const gptr<const int> rorder = G(a.rec_order), rstart = G(a.rec_start);
    const gptr<const float> wu = G(a.wav_u), wv = G(a.wav_v), ww = G(a.wav_w);
    const gptr<float> ru = G(a.rec_u), rv = G(a.rec_v), rw = G(a.rec_w);

    // compact history: floor quotient and remainder of the step's slot
    // (it - hist_base) by hist_every, advanced by one per step
    int hq = 0, hr = 0;
    if (HIST && a.hist_every > 0) {
        const int s0 = t_a - a.hist_base;
        hq = s0 / a.hist_every;
        hr = s0 - hq * a.hist_every;
        if (hr < 0) {
            hr += a.hist_every;
            hq--;
        }
    }
    // this wave holds a source or a receiver cell (wave-uniform tests ahead
    // of the per-cell ones)
    bool any_src = false, any_rec = false;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        any_src = any_src || sg[k] >= 0;
        any_rec = any_rec || rg[k] >= 0;
    }
    const bool w_src = __builtin_amdgcn_ballot_w64(any_src) != 0;
    const bool w_rec = __builtin_amdgcn_ballot_w64(any_rec) != 0;
    bool started = false;  // a step ran: LDS velocities await the sponge
    for (int s = 0; s < K; s++) {
        const int it = t_a + s;
        if (it >= hi) break;
        if (HIST && s > 0 && ++hr == a.hist_every) {
            hr = 0;
            ++hq;
        }
        // this step's history: stress fields (hs) and velocity fields (hv)
        float *hs = nullptr, *hv = nullptr;
        if (HIST) {
            const int slot = it - a.hist_base;
            if (slot >= 0 && slot < a.hist_len) {
                if (a.hist_every > 0) {
                    hv = a.hist + (size_t)slot * (HC<PSV, SH>::nv * n);
                    if (hr == 0)
                        hs = a.hist + ((size_t)a.hist_len * HC<PSV, SH>::nv +
                                       (size_t)hq * HC<PSV, SH>::ns) * n;
                } else {
                    hs = hv = a.hist + (size_t)slot * (H::n * n);
                }
            }
        }
        const bool cmp = a.hist_every > 0;
        const int o_syx = cmp ? HC<PSV, SH>::syx : H::syx, o_syz = cmp ? HC<PSV, SH>::syz : H::syz;
        const int o_u = cmp ? HC<PSV, SH>::u : H::u, o_w = cmp ? HC<PSV, SH>::w : H::w;
        const int o_v = cmp ? HC<PSV, SH>::v : H::v;
        // ---- history: the stresses this step's velocity update reads
        if (HIST && hs) {
#pragma unroll
            for (int k = 0; k < NC; k++) {
                if ((fl[k] & 5u) != 5u) continue;
                const int e = t + NT * k;
                const int c = gc[k];
                if (PSV) {
                    st(hs + H::sxx * n, c, sXX[e]);
                    st(hs + H::szz * n, c, sZZ[e]);
                    st(hs + H::sxz * n, c, sXZ[e]);
                }
                if (SH) {
                    st(hs + o_syx * n, c, sYX[e]);
                    st(hs + o_syz * n, c, sYZ[e]);
                }
            }
        }
        // ---- velocity update + source injection (region); a slot outside
        // the grid holds 0 in LDS and keeps it (no update, no source, sponge
        // 1), so only spare slots past the region skip
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const int e = t + NT * k;
            if (NT * (k + 1) > RN && e >= RN) continue;
            const bool upd = (fl[k] & 2u) && (fl[k] & 8u);
            if (PSV) {
                float u = sU[e], w = sW[e];
                if (started) {
                    u = u * A[k];
                    w = w * A[k];
                }
                if (upd) {
                    const float sxx_x = (sXX[e] - sXX[e - RZ]) * idx;
                    const float szz_z = (sZZ[e] - sZZ[e - 1]) * idz;
                    const float sxz_x = (sXZ[e] - sXZ[e - RZ]) * idx;
                    const float sxz_z = (sXZ[e] - sXZ[e - 1]) * idz;
                    u = u + ((sxx_x + sxz_z) * dt) * c0[k];
                    w = w + ((sxz_x + szz_z) * dt) * c1[k];
                }
                if (w_src && sg[k] >= 0)
                    for (int q = sstart[sg[k]]; q < sstart[sg[k] + 1]; q++) {
                        const size_t o = (size_t)sorder[q] * nt + it;
                        u = u + (wu[o] * factor) * c0[k];
                        w = w + (ww[o] * factor) * c1[k];
                    }
                sU[e] = u;
                sW[e] = w;
                if (HIST && hv && (fl[k] & 4u)) {
                    st(hv + o_u * n, gc[k], u);
                    st(hv + o_w * n, gc[k], w);
                }
            }
            if (SH) {
                float v = sV[e];
                if (started) v = v * A[k];
                if (upd) {
                    const float syx_x = (sYX[e] - sYX[e - RZ]) * idx;
                    const float syz_z = (sYZ[e] - sYZ[e - 1]) * idz;
                    v = v + ((syx_x + syz_z) * dt) * d0[k];
                }
                if (w_src && sg[k] >= 0)
                    for (int q = sstart[sg[k]]; q < sstart[sg[k] + 1]; q++)
                        v = v + (wv[(size_t)sorder[q] * nt + it] * factor) * d0[k];
                sV[e] = v;
                if (HIST && hv && (fl[k] & 4u)) st(hv + o_v * n, gc[k], v);
            }
        }
        __syncthreads();

        // ---- stress update + sponge (region), records (tile); slots outside
        // the grid keep their 0 as above
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const int e = t + NT * k;
            if (NT * (k + 1) > RN && e >= RN) continue;
            const bool upd = (fl[k] & 2u) && (fl[k] & 16u);
            if (PSV) {
                float sxx = sXX[e], szz = sZZ[e], sxz = sXZ[e];
                if (upd) {
                    const float u_x = (sU[e + RZ] - sU[e]) * idx;
                    const float u_z = (sU[e + 1] - sU[e]) * idz;
                    const float w_x = (sW[e + RZ] - sW[e]) * idx;
                    const float w_z = (sW[e + 1] - sW[e]) * idz;
                    const float div_uw = u_x + w_z;
                    sxx = sxx + dt * (c2[k] * div_uw + (2.0f * c3[k]) * u_x);
                    szz = szz + dt * (c2[k] * div_uw + (2.0f * c3[k]) * w_z);
                    sxz = sxz + (dt * c4[k]) * (u_z + w_x);
                }
                sXX[e] = sxx * A[k];
                sZZ[e] = szz * A[k];
                sXZ[e] = sxz * A[k];
            }
            if (SH) {
                float syx = sYX[e], syz = sYZ[e];
                if (upd) {
                    const float v_x = (sV[e + RZ] - sV[e]) * idx;
                    const float v_z = (sV[e + 1] - sV[e]) * idz;
                    syx = syx + (dt * d1[k]) * v_x;
                    syz = syz + (dt * d2[k]) * v_z;
                }
                sYX[e] = syx * A[k];
                sYZ[e] = syz * A[k];
            }
            if (w_rec && rg[k] >= 0) {
                const size_t row = (size_t)it * a.n_rec;
                for (int q = rstart[rg[k]]; q < rstart[rg[k] + 1]; q++) {
                    const int l = rorder[q];
                    if (PSV && a.rec_u) {
                        ru[row + l] = sU[e] * A[k];
                        rw[row + l] = sW[e] * A[k];
                    }
                    if (SH && a.rec_v) rv[row + l] = sV[e] * A[k];
                }
            }
        }
        started = true;
        __syncthreads();
    }

    // ---- the tile's state (post-sponge velocities)
    float* const* out = a.flip ? a.st : a.alt;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        if ((fl[k] & 5u) != 5u) continue;
        const int e = t + NT * k;
        const int c = gc[k];
        if (PSV) {
            st(out[0], c, sU[e] * A[k]);
            st(out[2], c, sW[e] * A[k]);
            st(out[3], c, sXX[e]);
            st(out[4], c, sZZ[e]);
            st(out[5], c, sXZ[e]);
        }
        if (SH) {
            st(out[1], c, sV[e] * A[k]);
            st(out[6], c, sYX[e]);
            st(out[7], c, sYZ[e]);
        }
    }
}

// ---------------------------------------------------------------------------
// adjoint, steps t_b-1 down to t_a (t_b - t_a <= K, inside K / KS intervals of
// KS steps from hist_base, KS = min(K, 4)), interval by interval from the top;
// steps at or beyond the shot's record length leave its state unchanged

template <bool PSV, bool SH, int TX, int TZ, int NT, int K, class S>
__global__ __launch_bounds__(NT) void k_fwi_adj_blk(S src, int t_a, int t_b, int tiles_x,
                                                    int tiles_z) {
    constexpr int RX = TX + 2 * K, RZ = TZ + 2 * K, RN = RX * RZ;
    constexpr int NC = (RN + NT - 1) / NT;
    constexpr int NP = PSV ? RN : 1, NS = SH ? RN : 1;
    using W = HistWindow<PSV, SH, TX, TZ>;
    using HV = HC<PSV, SH>;
    constexpr int KS = adj_stress_every(K);
    constexpr int NSL = adj_stress_slots<KS>();
    constexpr int FW_ = W::field;
    static_assert(adj_lds_bytes<PSV, SH, TX, TZ, K>() <= kLdsLimit, "LDS");
    __shared__ float qX[NP], qZ[NP], qS[NP], bU[NP], bW[NP];
    __shared__ float qY[NS], qV[NS], bV[NS];
    // velocity windows [field][chunk * 64 + lane] of alternate steps (two
    // arrays, so the LDS reads of one are not ordered after the loads into the
    // other), and the interval's pre-step stress windows: slot s - alpha - 1
    // for step s > alpha, slot 1 for step alpha (taken once step alpha + 2 is done)
    __shared__ float hwA[W::vfloats], hwB[W::vfloats];
    __shared__ float hsw[NSL * W::sfloats];

    const FB& a = src.get();
    const int tiles = tiles_x * tiles_z;
    const int tile = xcd_tile_b(blockIdx.x, tiles);
    const int nt = a.nt;
    if (tile >= tiles || t_a >= nt) return;
    const int i0 = (tile / tiles_z) * TX, j0 = (tile % tiles_z) * TZ;
    const int ri0 = i0 - K, rj0 = j0 - K;
    const int nx = a.nx, nz = a.nz;
    const int t = threadIdx.x;
    const bool has_src = a.src_map != nullptr, has_rec = a.rec_map != nullptr;
    float* const* in = a.flip ? a.alt : a.st;
    const size_t n = (size_t)nx * nz;

    // The gradient and history loads of a region cell outside the tile (or
    // off the interior) read a nearby tile cell (an interior one) instead:
    // always in bounds, and no cache lines beyond the tile's.
    auto tile_cell = [&](int e, bool interior) {
        const int r = e / RZ, z = e - r * RZ;
        int ti = min(clampi(ri0 + r, i0, i0 + TX - 1), nx - 1);
        int tj = min(clampi(rj0 + z, j0, j0 + TZ - 1), nz - 1);
        if (interior) {
            ti = clampi(ti, 1, nx - 2);
            tj = clampi(tj, 1, nz - 2);
        }
        return ti * nz + tj;
    };
    int gc[NC], sg[NC], rg[NC];
    // bit0 in grid, bit1 interior, bit2 in tile, bit3 phase 1 ok, bit4 phase 2
    // ok, bit5 in the stress window (the tile and its low ring, in the grid)
    unsigned fl[NC];
    float A[NC], c0[NC], c1[NC], c2[NC], c3[NC], c4[NC], d0[NC], d1[NC], d2[NC];
    float pu[NC], pw[NC], pxx[NC], pzz[NC], pxz[NC], pv[NC], pyx[NC], pyz[NC];
    float gl[NC], gm[NC], gx[NC], gu[NC], gw[NC], gyx[NC], gyz[NC], gr[NC];
    int slot[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const int e = t + NT * k;
        const RegionCell rc = region_cell<TX, TZ, K>(e, i0, j0, nx, nz, a.fw);
        gc[k] = rc.c;
        fl[k] = rc.f;
        slot[k] = rc.slot;
        const int r = e / RZ, z = e - r * RZ;
        if ((rc.f & 1u) && r >= K - 1 && r <= K - 1 + TX && z >= K - 1 && z <= K - 1 + TZ)
            fl[k] |= 32u;
        if (!(rc.f & 1u) && e < RN) {
            // outside the grid: the transposes read zero here (not interior)
            if (PSV) qX[e] = qZ[e] = qS[e] = bU[e] = bW[e] = 0.0f;
            if (SH) qY[e] = qV[e] = bV[e] = 0.0f;
        }
    }
    {
        const gptr<const float> tab = G(a.tab);
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const int c = gc[k], cg = tile_cell(t + NT * k, false);
            const float sa = tab[slot[k] > 0 ? slot[k] : 0];
            A[k] = slot[k] >= 0 ? sa : 1.0f;
            sg[k] = has_src ? ld(a.src_map, cg) : -1;
            rg[k] = has_rec ? ld(a.rec_map, c) : -1;
            if ((fl[k] & 5u) != 5u) sg[k] = -1;
            if (!(fl[k] & 1u)) rg[k] = -1;
            if (PSV) {
                c0[k] = ld(a.iru, c);
                c1[k] = ld(a.irw, c);
                c2[k] = ld(a.lam, c);
                c3[k] = ld(a.mu, c);
                c4[k] = ld(a.mxz, c);
                pu[k] = ld(in[0], c);
                pw[k] = ld(in[2], c);
                pxx[k] = ld(in[3], c);
                pzz[k] = ld(in[4], c);
                pxz[k] = ld(in[5], c);
                gl[k] = ld(a.g_lam, cg);
                gm[k] = ld(a.g_mu, cg);
                gx[k] = ld(a.g_mxz, cg);
                gu[k] = ld(a.g_iru, cg);
                gw[k] = ld(a.g_irw, cg);
            }
            if (SH) {
                d0[k] = ld(a.ir, c);
                d1[k] = ld(a.myx, c);
                d2[k] = ld(a.myz, c);
                pv[k] = ld(in[1], c);
                pyx[k] = ld(in[6], c);
                pyz[k] = ld(in[7], c);
                gyx[k] = ld(a.g_myx, cg);
                gyz[k] = ld(a.g_myz, cg);
                gr[k] = ld(a.g_ir, cg);
            }
        }
    }

    const float idx = a.idx, idz = a.idz, dt = a.dt, factor = a.factor;
    const gptr<const int> sorder = G(a.src_order), sstart = G(a.src_start);
    const gptr<const int> rorder = G(a.rec_order), rstart = G(a.rec_start);
    const gptr<const float> wu = G(a.wav_u), wv = G(a.wav_v), ww = G(a.wav_w);
    const gptr<const float> res_u = G(a.res_u), res_v = G(a.res_v), res_w = G(a.res_w);

    // History windows go global -> LDS asynchronously.  Wave w copies chunks
    // q = w, w + NW, ... of a window (q = field * chunks + chunk); the source
    // offset of this lane in each velocity chunk is the same every step, so
    // it is computed once.
    constexpr int NW = NT / 64;
    constexpr int NQV = HV::nv * W::chunks, QPV = (NQV + NW - 1) / NW;
    constexpr int NQS = HV::ns * W::chunks, QPS = (NQS + NW - 1) / NW;
    const int lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);
    auto win_off = [&](int q, int nq) {
        const int f = q / W::chunks, ch = q - f * W::chunks;
        const int w = ch * 64 + lane;
        const int r = w / W::HZ, z = w - r * W::HZ;
        const int i = clampi(i0 - 1 + r, 0, nx - 1), j = clampi(j0 - 1 + z, 0, nz - 1);
        return (unsigned)((q < nq ? f : 0) * (int)n + i * nz + j) * 4u;
    };
    // A wave's m-th load of a velocity window copies chunk q = wave + NW m,
    // or, past the last chunk, chunk q mod NQV again (the same bytes to the
    // same place), so every wave issues exactly QPV loads per window and can
    // wait for a window by count.
    auto vchunk = [&](int m) {
        const int q = wave + NW * m;  // wave-uniform
        return q < NQV ? q : q % NQV;
    };
    unsigned voff[QPV];  // byte offset from the step's velocity slot
#pragma unroll
    for (int m = 0; m < QPV; m++) voff[m] = win_off(vchunk(m), NQV);
    auto fetch_v = [&](int it, float* dst) {
        const char* base = (const char*)(a.hist + (size_t)(it - a.hist_base) * (HV::nv * n));
#pragma unroll
        for (int m = 0; m < QPV; m++)
            __builtin_amdgcn_global_load_lds(
                (const void*)(base + voff[m]),
                (__attribute__((address_space(3))) void*)&dst[vchunk(m) * 64], 4, 0, 0);
    };
    auto fetch_s = [&](const float* hstr, float* dst) {  // stored stresses of an interval
        const char* base = (const char*)hstr;
#pragma unroll
        for (int m = 0; m < QPS; m++) {
            const int q = wave + NW * m;
            if (q < NQS)
                __builtin_amdgcn_global_load_lds(
                    (const void*)(base + win_off(q, NQS)),
                    (__attribute__((address_space(3))) void*)&dst[q * 64], 4, 0, 0);
        }
    };
    // window offset of each region cell (own interior cells only use it)
    int wo[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const int e = t + NT * k;
        const int r = e / RZ, z = e - r * RZ;
        wo[k] = (r - K + 1) * W::HZ + (z - K + 1);
    }
    // Intervals from the top: interval iv takes steps top_i down to t_lo,
    // its pre-step stresses stored at step alpha_i.
    int top_i = (t_b < nt ? t_b : nt) - 1;
#pragma unroll
    for (int iv = 0; iv < K / KS; iv++) {
        if (top_i < t_a) break;
        const int alpha_i = a.hist_base + ((top_i - a.hist_base) / KS) * KS;
        const int t_lo = alpha_i > t_a ? alpha_i : t_a;
        const float* const hstr =
            a.hist + ((size_t)a.hist_len * HV::nv +
                      (size_t)((alpha_i - a.hist_base) / KS) * HV::ns) * n;
        // the previous interval's last step read the windows and slots reused here
        if (iv > 0) lds_barrier();
        // ---- pre-step stresses of steps alpha_i + 1 .. top_i over the stress
        // window: the forward stress update and sponge (k_fwi_fwd_blk) from the
        // stored stresses of step alpha_i and the velocities of steps alpha_i ..
        // top_i - 1.  Every velocity window of the interval is requested first:
        // v(alpha_i + r), r < top_i - 1 - alpha_i, parked in stress slot r + 1
        // (read by recomputation step r, overwritten by step r + 1), v(top_i - 1)
        // in hwB and v(top_i) in hwA, where the first two adjoint steps read
        // them.  Requests past top_i - 1 repeat v(top_i), so every wave issues
        // KS windows and waits by count.
        if (top_i > alpha_i) {
            float rxx[NC], rzz[NC], rxz[NC], ryx[NC], ryz[NC];
#pragma unroll
            for (int k = 0; k < NC; k++) {
                rxx[k] = rzz[k] = rxz[k] = ryx[k] = ryz[k] = 0.0f;
                if (!(fl[k] & 32u)) continue;
                const int c = gc[k];
                if (PSV) {
                    rxx[k] = ld(hstr + HV::sxx * n, c);
                    rzz[k] = ld(hstr + HV::szz * n, c);
                    rxz[k] = ld(hstr + HV::sxz * n, c);
                }
                if (SH) {
                    ryx[k] = ld(hstr + HV::syx * n, c);
                    ryz[k] = ld(hstr + HV::syz * n, c);
                }
            }
#pragma unroll
            for (int r = 0; r < KS - 1; r++) {
                const int s = alpha_i + r;
                if (s < top_i - 1)
                    fetch_v(s, hsw + (r + 1) * W::sfloats);
                else if (s == top_i - 1)
                    fetch_v(s, hwB);
                else
                    fetch_v(top_i, hwA);
            }
            fetch_v(top_i, hwA);
            vm_wait_windows<QPV>(KS);  // the loads before the windows
            // stresses of step s + 1 = alpha_i + r + 1 from step s, into slot r
            auto recompute = [&](int r, const float* buf) {
                float* const dst = hsw + r * W::sfloats;
#pragma unroll
                for (int k = 0; k < NC; k++) {
                    if (!(fl[k] & 32u)) continue;
                    const int w = wo[k];
                    if (PSV) {
                        float sxx = rxx[k], szz = rzz[k], sxz = rxz[k];
                        if (fl[k] & 2u) {
                            const float u0 = buf[HV::u * FW_ + w], w0 = buf[HV::w * FW_ + w];
                            const float u_x = (buf[HV::u * FW_ + w + W::HZ] - u0) * idx;
                            const float u_z = (buf[HV::u * FW_ + w + 1] - u0) * idz;
                            const float w_x = (buf[HV::w * FW_ + w + W::HZ] - w0) * idx;
                            const float w_z = (buf[HV::w * FW_ + w + 1] - w0) * idz;
                            const float div_uw = u_x + w_z;
                            sxx = sxx + dt * (c2[k] * div_uw + (2.0f * c3[k]) * u_x);
                            szz = szz + dt * (c2[k] * div_uw + (2.0f * c3[k]) * w_z);
                            sxz = sxz + (dt * c4[k]) * (u_z + w_x);
                        }
                        rxx[k] = sxx * A[k];
                        rzz[k] = szz * A[k];
                        rxz[k] = sxz * A[k];
                        dst[HV::sxx * FW_ + w] = rxx[k];
                        dst[HV::szz * FW_ + w] = rzz[k];
                        dst[HV::sxz * FW_ + w] = rxz[k];
                    }
                    if (SH) {
                        float syx = ryx[k], syz = ryz[k];
                        if (fl[k] & 2u) {
                            const float v0 = buf[HV::v * FW_ + w];
                            const float v_x = (buf[HV::v * FW_ + w + W::HZ] - v0) * idx;
                            const float v_z = (buf[HV::v * FW_ + w + 1] - v0) * idz;
                            syx = syx + (dt * d1[k]) * v_x;
                            syz = syz + (dt * d2[k]) * v_z;
                        }
                        ryx[k] = syx * A[k];
                        ryz[k] = syz * A[k];
                        dst[HV::syx * FW_ + w] = ryx[k];
                        dst[HV::syz * FW_ + w] = ryz[k];
                    }
                }
            };
#pragma unroll
            for (int r = 0; r < KS - 1; r++) {
                const int s = alpha_i + r;
                if (s >= top_i) break;
                vm_wait_windows<QPV>(KS - 1 - r);  // window r arrived
                lds_barrier();
                if (s == top_i - 1)
                    recompute(r, hwB);
                else
                    recompute(r, hsw + (r + 1) * W::sfloats);
            }
        } else {
            fetch_v(top_i, hwA);
            if (t_lo == alpha_i) fetch_s(hstr, hsw + W::sfloats);
        }
        // the stored stresses of step alpha_i for its adjoint step (when the launch
        // takes it after recomputed ones): loaded at step min(top_i, alpha_i + 2),
        // written to slot 1 at step alpha_i, once slot 1's last reader is done
        const bool late_alpha = t_lo == alpha_i && top_i > alpha_i;
        const int alpha_load = top_i < alpha_i + 2 ? top_i : alpha_i + 2;
        float mxx[NC], mzz[NC], mxz[NC], myx_[NC], myz_[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) mxx[k] = mzz[k] = mxz[k] = myx_[k] = myz_[k] = 0.0f;

#pragma unroll
        for (int s = 0; s < KS; s++) {
            const int it = top_i - s;
            if (it < t_lo) break;
            const float* const buf = (s & 1) ? hwB : hwA;  // v(top_i - s)
            float* const nxt = (s & 1) ? hwA : hwB;
            const float* const sw = hsw + (it == alpha_i ? 1 : it - alpha_i - 1) * W::sfloats;

            // ---- phase 1a: data adjoint at the receivers; stress transposes q
            // of every region cell from its own sponged stress adjoint
#pragma unroll
            for (int k = 0; k < NC; k++) {
                if (!(fl[k] & 1u)) continue;
                const int e = t + NT * k;
                if (rg[k] >= 0) {
                    const size_t row = (size_t)it * a.n_rec;
                    for (int q = rstart[rg[k]]; q < rstart[rg[k] + 1]; q++) {
                        const int l = rorder[q];
                        if (PSV) {
                            pu[k] += res_u[row + l];
                            pw[k] += res_w[row + l];
                        }
                        if (SH) pv[k] += res_v[row + l];
                    }
                }
                const bool in_ = fl[k] & 2u;
                if (PSV) {
                    const float axx = dt * (A[k] * pxx[k]), azz = dt * (A[k] * pzz[k]);
                    qX[e] = in_ ? (c2[k] + 2.0f * c3[k]) * axx + c2[k] * azz : 0.0f;
                    qZ[e] = in_ ? c2[k] * axx + (c2[k] + 2.0f * c3[k]) * azz : 0.0f;
                    qS[e] = in_ ? (c4[k] * dt) * (A[k] * pxz[k]) : 0.0f;
                }
                if (SH) {
                    qY[e] = in_ ? (d1[k] * dt) * (A[k] * pyx[k]) : 0.0f;
                    qV[e] = in_ ? (d2[k] * dt) * (A[k] * pyz[k]) : 0.0f;
                }
            }
            vm_wait();  // this step's windows (and the residuals) arrived
            lds_barrier();
            // the next step's velocity window (the one of top_i - 1 came with the
            // recomputation); the window it replaces was last read a step ago
            if (s > 0 && it - 1 >= t_lo) fetch_v(it - 1, nxt);
            if (late_alpha && it == alpha_load) {
#pragma unroll
                for (int k = 0; k < NC; k++) {
                    if (!(fl[k] & 32u)) continue;
                    const int c = gc[k];
                    if (PSV) {
                        mxx[k] = ld(hstr + HV::sxx * n, c);
                        mzz[k] = ld(hstr + HV::szz * n, c);
                        mxz[k] = ld(hstr + HV::sxz * n, c);
                    }
                    if (SH) {
                        myx_[k] = ld(hstr + HV::syx * n, c);
                        myz_[k] = ld(hstr + HV::syz * n, c);
                    }
                }
            }
            if (late_alpha && it == alpha_i) {
                float* const dst = hsw + W::sfloats;
#pragma unroll
                for (int k = 0; k < NC; k++) {
                    if (!(fl[k] & 32u)) continue;
                    const int w = wo[k];
                    if (PSV) {
                        dst[HV::sxx * FW_ + w] = mxx[k];
                        dst[HV::szz * FW_ + w] = mzz[k];
                        dst[HV::sxz * FW_ + w] = mxz[k];
                    }
                    if (SH) {
                        dst[HV::syx * FW_ + w] = myx_[k];
                        dst[HV::syz * FW_ + w] = myz_[k];
                    }
                }
            }

            // ---- phase 1b: new velocity adjoint; lam / mu / mxz (myx / myz) and
            // source-term gradients of the tile; b = dt * inv_rho * p for phase 2
#pragma unroll
            for (int k = 0; k < NC; k++) {
                if (!(fl[k] & 1u)) continue;
                const int e = t + NT * k;
                const bool in_ = fl[k] & 2u;
                const bool own_in = (fl[k] & 6u) == 6u;
                if (fl[k] & 8u) {
                    if (PSV) {
                        const float nu =
                            A[k] * pu[k] + idx * (qX[e - RZ] - qX[e]) + idz * (qS[e - 1] - qS[e]);
                        const float nw =
                            A[k] * pw[k] + idx * (qS[e - RZ] - qS[e]) + idz * (qZ[e - 1] - qZ[e]);
                        if (own_in) {
                            const float axx = dt * A[k] * pxx[k], azz = dt * A[k] * pzz[k];
                            const float axz = dt * A[k] * pxz[k];
                            const int w = wo[k];
                            const float u0 = buf[HV::u * FW_ + w], w0 = buf[HV::w * FW_ + w];
                            const float ux = (buf[HV::u * FW_ + w + W::HZ] - u0) * idx;
                            const float uz = (buf[HV::u * FW_ + w + 1] - u0) * idz;
                            const float wx = (buf[HV::w * FW_ + w + W::HZ] - w0) * idx;
                            const float wz = (buf[HV::w * FW_ + w + 1] - w0) * idz;
                            gl[k] += (axx + azz) * (ux + wz);
                            gm[k] += 2.0f * (axx * ux + azz * wz);
                            gx[k] += axz * (uz + wx);
                        }
                        pu[k] = nu;
                        pw[k] = nw;
                    }
                    if (SH) {
                        const float nv =
                            A[k] * pv[k] + idx * (qY[e - RZ] - qY[e]) + idz * (qV[e - 1] - qV[e]);
                        if (own_in) {
                            const int w = wo[k];
                            const float v0 = buf[HV::v * FW_ + w];
                            const float vx = (buf[HV::v * FW_ + w + W::HZ] - v0) * idx;
                            const float vz = (buf[HV::v * FW_ + w + 1] - v0) * idz;
                            gyx[k] += dt * A[k] * pyx[k] * vx;
                            gyz[k] += dt * A[k] * pyz[k] * vz;
                        }
                        pv[k] = nv;
                    }
                    if (sg[k] >= 0)
                        for (int q = sstart[sg[k]]; q < sstart[sg[k] + 1]; q++) {
                            const size_t o = (size_t)sorder[q] * nt + it;
                            if (PSV) {
                                gu[k] += pu[k] * wu[o] * factor;
                                gw[k] += pw[k] * ww[o] * factor;
                            }
                            if (SH) gr[k] += pv[k] * wv[o] * factor;
                        }
                }
                if (PSV) {
                    bU[e] = in_ ? dt * c0[k] * pu[k] : 0.0f;
                    bW[e] = in_ ? dt * c1[k] * pw[k] : 0.0f;
                }
                if (SH) bV[e] = in_ ? dt * d0[k] * pv[k] : 0.0f;
            }
            lds_barrier();

            // ---- phase 2: new stress adjoint; inv_rho_u / inv_rho_w (inv_rho)
            // gradients of the tile
#pragma unroll
            for (int k = 0; k < NC; k++) {
                if ((fl[k] & 17u) != 17u) continue;
                const int e = t + NT * k;
                const bool own_in = (fl[k] & 6u) == 6u;
                if (PSV) {
                    const float bu0 = bU[e], bw0 = bW[e];
                    pxx[k] = A[k] * pxx[k] + idx * (bu0 - bU[e + RZ]);
                    pxz[k] = A[k] * pxz[k] + idz * (bu0 - bU[e + 1]) + idx * (bw0 - bW[e + RZ]);
                    pzz[k] = A[k] * pzz[k] + idz * (bw0 - bW[e + 1]);
                    if (own_in) {
                        const int w = wo[k];
                        const float xx = sw[HV::sxx * FW_ + w], xz = sw[HV::sxz * FW_ + w];
                        const float zz = sw[HV::szz * FW_ + w];
                        const float sxx_x = (xx - sw[HV::sxx * FW_ + w - W::HZ]) * idx;
                        const float sxz_z = (xz - sw[HV::sxz * FW_ + w - 1]) * idz;
                        const float sxz_x = (xz - sw[HV::sxz * FW_ + w - W::HZ]) * idx;
                        const float szz_z = (zz - sw[HV::szz * FW_ + w - 1]) * idz;
                        gu[k] += pu[k] * dt * (sxx_x + sxz_z);
                        gw[k] += pw[k] * dt * (sxz_x + szz_z);
                    }
                }
                if (SH) {
                    const float bv0 = bV[e];
                    pyx[k] = A[k] * pyx[k] + idx * (bv0 - bV[e + RZ]);
                    pyz[k] = A[k] * pyz[k] + idz * (bv0 - bV[e + 1]);
                    if (own_in) {
                        const int w = wo[k];
                        const float yx = sw[HV::syx * FW_ + w], yz = sw[HV::syz * FW_ + w];
                        const float syx_x = (yx - sw[HV::syx * FW_ + w - W::HZ]) * idx;
                        const float syz_z = (yz - sw[HV::syz * FW_ + w - 1]) * idz;
                        gr[k] += pv[k] * dt * (syx_x + syz_z);
                    }
                }
            }
        }

        top_i = alpha_i - 1;
    }
    // ---- the tile's adjoint state and gradients
    float* const* out = a.flip ? a.st : a.alt;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        if ((fl[k] & 5u) != 5u) continue;
        const int c = gc[k];
        if (PSV) {
            st(out[0], c, pu[k]);
            st(out[2], c, pw[k]);
            st(out[3], c, pxx[k]);
            st(out[4], c, pzz[k]);
            st(out[5], c, pxz[k]);
            st(a.g_lam, c, gl[k]);
            st(a.g_mu, c, gm[k]);
            st(a.g_mxz, c, gx[k]);
            st(a.g_iru, c, gu[k]);
            st(a.g_irw, c, gw[k]);
        }
        if (SH) {
            st(out[1], c, pv[k]);
            st(out[6], c, pyx[k]);
            st(out[7], c, pyz[k]);
            st(a.g_myx, c, gyx[k]);
            st(a.g_myz, c, gyz[k]);
            st(a.g_ir, c, gr[k]);
        }
    }
}

// ---------------------------------------------------------------------------
// scattered-field forward, steps t_a .. t_b-1 (t_b - t_a <= K)
//
// Advances this model's state exactly as k_fwi_fwd_blk (same expressions, so
// its checkpoints are bit-identical) and, beside it, the scattered state
// ds = s(model) - s(reference): with D the difference operators and b, C the
// buoyancies and moduli of the velocity and stress updates,
//   dv' = A (dv + dt (b_r D ds) + dt (db D s)) + f factor db   (source cells)
//   ds' = A (ds + dt (C_r D dv') + dt (dC D v'))
// where db = b - b_r and dC = C - C_r cell by cell (given, b_r = b - db and
// C_r = C - dC), and s, v' are this model's pre-step stresses and
// post-update velocities.  The recurrence is
// linear in ds, so ds carries float32 rounding relative to itself: records of
// ds are the residual d(model) - d(reference) without the cancellation of
// two float32 records that agree to a few 1e-5.

struct FS {
    int nx, nz, nt, fw;
    int n_rec, flip, pad0, pad1;
    float dt, idx, idz, factor;
    const float* tab;
    const float *lam, *mu, *mxz, *myx, *myz, *ir, *iru, *irw;         // this model
    const float *elam, *emu, *emxz, *emyx, *emyz, *eir, *eiru, *eirw;  // minus the reference's
    float* st[8];    // this model's state u v w sxx szz sxz syx syz
    float* alt[8];   // its workspace copy
    float* dst[8];   // scattered state, same order
    float* dalt[8];  // its workspace copy
    const int *src_map, *rec_map;
    const int *src_order, *src_start, *rec_order, *rec_start;
    const float *wav_u, *wav_v, *wav_w;
    float *rec_u, *rec_v, *rec_w;  // scattered records, time-major (nt, n_rec)
};
constexpr int kScatShotsPerLaunch = 4;  // per-shot entries fit the 4 KB kernel-argument limit
template <class T>
struct ByArray4 {
    T v[kScatShotsPerLaunch];
    __device__ __forceinline__ const T& get() const { return v[blockIdx.z]; }
};
static_assert(sizeof(ByArray4<FS>) + 32 <= 4096, "scattered kernel arguments exceed 4 KB");

template <bool PSV, bool SH, int TX, int TZ, int NT, int K, class S>
__global__ __launch_bounds__(NT) void k_fwi_scat_blk(S src, int t_a, int t_b, int tiles_x,
                                                     int tiles_z) {
    constexpr int RX = TX + 2 * K, RZ = TZ + 2 * K, RN = RX * RZ;
    constexpr int NC = (RN + NT - 1) / NT;
    constexpr int NP = PSV ? RN : 1, NS = SH ? RN : 1;
    __shared__ float sU[NP], sW[NP], sXX[NP], sZZ[NP], sXZ[NP];
    __shared__ float tU[NP], tW[NP], tXX[NP], tZZ[NP], tXZ[NP];
    __shared__ float sV[NS], sYX[NS], sYZ[NS];
    __shared__ float tV[NS], tYX[NS], tYZ[NS];

    const FS& a = src.get();
    const int tiles = tiles_x * tiles_z;
    const int tile = xcd_tile_b(blockIdx.x, tiles);
    const int nt = a.nt;
    if (tile >= tiles || t_a >= nt) return;
    const int i0 = (tile / tiles_z) * TX, j0 = (tile % tiles_z) * TZ;
    const int nx = a.nx, nz = a.nz;
    const int t = threadIdx.x;
    const bool has_src = a.src_map != nullptr, has_rec = a.rec_map != nullptr;

    int gc[NC], sg[NC], rg[NC];
    unsigned fl[NC];  // bit0 in grid, bit1 interior, bit2 in tile, bit3 V ok, bit4 S ok
    float A[NC];
    // this model's coefficients (c, d), the reference's (r, q) and their
    // differences (e, h): P-SV iru irw lam mu mxz, SH ir myx myz
    float c0[NC], c1[NC], c2[NC], c3[NC], c4[NC], r0[NC], r1[NC], r2[NC], r3[NC], r4[NC];
    float e0[NC], e1[NC], e2[NC], e3[NC], e4[NC];
    float d0[NC], d1[NC], d2[NC], q0[NC], q1[NC], q2[NC], h0[NC], h1[NC], h2[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const RegionCell rc = region_cell<TX, TZ, K>(t + NT * k, i0, j0, nx, nz, a.fw);
        gc[k] = rc.c;
        fl[k] = rc.f;
        const int slot = rc.slot;
        const gptr<const float> tab = G(a.tab);
        const float sa = tab[slot > 0 ? slot : 0];
        A[k] = slot >= 0 ? sa : 1.0f;
    }
    {
        float* const* in = a.flip ? a.alt : a.st;
        float* const* din = a.flip ? a.dalt : a.dst;
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const int e = t + NT * k;
            const int c = gc[k];
            const bool g = fl[k] & 1u;
            sg[k] = has_src ? ld(a.src_map, c) : -1;
            rg[k] = has_rec ? ld(a.rec_map, c) : -1;
            if (!g) sg[k] = -1;
            if ((fl[k] & 5u) != 5u) rg[k] = -1;
            if (PSV) {
                c0[k] = ld(a.iru, c);
                c1[k] = ld(a.irw, c);
                c2[k] = ld(a.lam, c);
                c3[k] = ld(a.mu, c);
                c4[k] = ld(a.mxz, c);
                e0[k] = ld(a.eiru, c);
                e1[k] = ld(a.eirw, c);
                e2[k] = ld(a.elam, c);
                e3[k] = ld(a.emu, c);
                e4[k] = ld(a.emxz, c);
                r0[k] = c0[k] - e0[k];
                r1[k] = c1[k] - e1[k];
                r2[k] = c2[k] - e2[k];
                r3[k] = c3[k] - e3[k];
                r4[k] = c4[k] - e4[k];
                const float u = ld(in[0], c), w = ld(in[2], c), xx = ld(in[3], c), zz = ld(in[4], c),
                            xz = ld(in[5], c);
                const float du = ld(din[0], c), dw = ld(din[2], c), dxx = ld(din[3], c),
                            dzz = ld(din[4], c), dxz = ld(din[5], c);
                if (e < RN) {
                    sU[e] = g ? u : 0.0f;
                    sW[e] = g ? w : 0.0f;
                    sXX[e] = g ? xx : 0.0f;
                    sZZ[e] = g ? zz : 0.0f;
                    sXZ[e] = g ? xz : 0.0f;
                    tU[e] = g ? du : 0.0f;
                    tW[e] = g ? dw : 0.0f;
                    tXX[e] = g ? dxx : 0.0f;
                    tZZ[e] = g ? dzz : 0.0f;
                    tXZ[e] = g ? dxz : 0.0f;
                }
            }
            if (SH) {
                d0[k] = ld(a.ir, c);
                d1[k] = ld(a.myx, c);
                d2[k] = ld(a.myz, c);
                h0[k] = ld(a.eir, c);
                h1[k] = ld(a.emyx, c);
                h2[k] = ld(a.emyz, c);
                q0[k] = d0[k] - h0[k];
                q1[k] = d1[k] - h1[k];
                q2[k] = d2[k] - h2[k];
                const float v = ld(in[1], c), yx = ld(in[6], c), yz = ld(in[7], c);
                const float dv = ld(din[1], c), dyx = ld(din[6], c), dyz = ld(din[7], c);
                if (e < RN) {
                    sV[e] = g ? v : 0.0f;
                    sYX[e] = g ? yx : 0.0f;
                    sYZ[e] = g ? yz : 0.0f;
                    tV[e] = g ? dv : 0.0f;
                    tYX[e] = g ? dyx : 0.0f;
                    tYZ[e] = g ? dyz : 0.0f;
                }
            }
        }
    }
    __syncthreads();

    const float idx = a.idx, idz = a.idz, dt = a.dt, factor = a.factor;
    const int hi = t_b < nt ? t_b : nt;
    const gptr<const int> sorder = G(a.src_order), sstart = G(a.src_start);
    const gptr<const int> rorder = G(a.rec_order), rstart = G(a.rec_start);
    const gptr<const float> wu = G(a.wav_u), wv = G(a.wav_v), ww = G(a.wav_w);
    const gptr<float> ru = G(a.rec_u), rv = G(a.rec_v), rw = G(a.rec_w);

    bool started = false;  // a step ran: LDS velocities await the sponge
    for (int s = 0; s < K; s++) {
        const int it = t_a + s;
        if (it >= hi) break;
        // ---- velocity update + source injection (region)
#pragma unroll
        for (int k = 0; k < NC; k++) {
            if (!(fl[k] & 1u)) continue;
            const int e = t + NT * k;
            const bool upd = (fl[k] & 2u) && (fl[k] & 8u);
            if (PSV) {
                float u = sU[e], w = sW[e], du = tU[e], dw = tW[e];
                if (started) {
                    u = u * A[k];
                    w = w * A[k];
                    du = du * A[k];
                    dw = dw * A[k];
                }
                if (upd) {
                    const float sxx_x = (sXX[e] - sXX[e - RZ]) * idx;
                    const float szz_z = (sZZ[e] - sZZ[e - 1]) * idz;
                    const float sxz_x = (sXZ[e] - sXZ[e - RZ]) * idx;
                    const float sxz_z = (sXZ[e] - sXZ[e - 1]) * idz;
                    u = u + ((sxx_x + sxz_z) * dt) * c0[k];
                    w = w + ((sxz_x + szz_z) * dt) * c1[k];
                    const float dsxx_x = (tXX[e] - tXX[e - RZ]) * idx;
                    const float dszz_z = (tZZ[e] - tZZ[e - 1]) * idz;
                    const float dsxz_x = (tXZ[e] - tXZ[e - RZ]) * idx;
                    const float dsxz_z = (tXZ[e] - tXZ[e - 1]) * idz;
                    du = du + (((dsxx_x + dsxz_z) * dt) * r0[k] + ((sxx_x + sxz_z) * dt) * e0[k]);
                    dw = dw + (((dsxz_x + dszz_z) * dt) * r1[k] + ((sxz_x + szz_z) * dt) * e1[k]);
                }
                if (sg[k] >= 0)
                    for (int q = sstart[sg[k]]; q < sstart[sg[k] + 1]; q++) {
                        const size_t o = (size_t)sorder[q] * nt + it;
                        const float fu = wu[o] * factor, fw_ = ww[o] * factor;
                        u = u + fu * c0[k];
                        w = w + fw_ * c1[k];
                        du = du + fu * e0[k];
                        dw = dw + fw_ * e1[k];
                    }
                sU[e] = u;
                sW[e] = w;
                tU[e] = du;
                tW[e] = dw;
            }
            if (SH) {
                float v = sV[e], dv = tV[e];
                if (started) {
                    v = v * A[k];
                    dv = dv * A[k];
                }
                if (upd) {
                    const float syx_x = (sYX[e] - sYX[e - RZ]) * idx;
                    const float syz_z = (sYZ[e] - sYZ[e - 1]) * idz;
                    v = v + ((syx_x + syz_z) * dt) * d0[k];
                    const float dyx_x = (tYX[e] - tYX[e - RZ]) * idx;
                    const float dyz_z = (tYZ[e] - tYZ[e - 1]) * idz;
                    dv = dv + (((dyx_x + dyz_z) * dt) * q0[k] + ((syx_x + syz_z) * dt) * h0[k]);
                }
                if (sg[k] >= 0)
                    for (int q = sstart[sg[k]]; q < sstart[sg[k] + 1]; q++) {
                        const float fv = wv[(size_t)sorder[q] * nt + it] * factor;
                        v = v + fv * d0[k];
                        dv = dv + fv * h0[k];
                    }
                sV[e] = v;
                tV[e] = dv;
            }
        }
        __syncthreads();

        // ---- stress update + sponge (region), records (tile)
#pragma unroll
        for (int k = 0; k < NC; k++) {
            if (!(fl[k] & 1u)) continue;
            const int e = t + NT * k;
            const bool upd = (fl[k] & 2u) && (fl[k] & 16u);
            if (PSV) {
                float sxx = sXX[e], szz = sZZ[e], sxz = sXZ[e];
                float dxx = tXX[e], dzz = tZZ[e], dxz = tXZ[e];
                if (upd) {
                    const float u_x = (sU[e + RZ] - sU[e]) * idx;
                    const float u_z = (sU[e + 1] - sU[e]) * idz;
                    const float w_x = (sW[e + RZ] - sW[e]) * idx;
                    const float w_z = (sW[e + 1] - sW[e]) * idz;
                    const float div_uw = u_x + w_z;
                    sxx = sxx + dt * (c2[k] * div_uw + (2.0f * c3[k]) * u_x);
                    szz = szz + dt * (c2[k] * div_uw + (2.0f * c3[k]) * w_z);
                    sxz = sxz + (dt * c4[k]) * (u_z + w_x);
                    const float du_x = (tU[e + RZ] - tU[e]) * idx;
                    const float du_z = (tU[e + 1] - tU[e]) * idz;
                    const float dw_x = (tW[e + RZ] - tW[e]) * idx;
                    const float dw_z = (tW[e + 1] - tW[e]) * idz;
                    const float ddiv = du_x + dw_z;
                    dxx = dxx + dt * ((r2[k] * ddiv + (2.0f * r3[k]) * du_x) +
                                      (e2[k] * div_uw + (2.0f * e3[k]) * u_x));
                    dzz = dzz + dt * ((r2[k] * ddiv + (2.0f * r3[k]) * dw_z) +
                                      (e2[k] * div_uw + (2.0f * e3[k]) * w_z));
                    dxz = dxz + dt * (r4[k] * (du_z + dw_x) + e4[k] * (u_z + w_x));
                }
                sXX[e] = sxx * A[k];
                sZZ[e] = szz * A[k];
                sXZ[e] = sxz * A[k];
                tXX[e] = dxx * A[k];
                tZZ[e] = dzz * A[k];
                tXZ[e] = dxz * A[k];
            }
            if (SH) {
                float syx = sYX[e], syz = sYZ[e], dyx = tYX[e], dyz = tYZ[e];
                if (upd) {
                    const float v_x = (sV[e + RZ] - sV[e]) * idx;
                    const float v_z = (sV[e + 1] - sV[e]) * idz;
                    syx = syx + (dt * d1[k]) * v_x;
                    syz = syz + (dt * d2[k]) * v_z;
                    const float dv_x = (tV[e + RZ] - tV[e]) * idx;
                    const float dv_z = (tV[e + 1] - tV[e]) * idz;
                    dyx = dyx + dt * (q1[k] * dv_x + h1[k] * v_x);
                    dyz = dyz + dt * (q2[k] * dv_z + h2[k] * v_z);
                }
                sYX[e] = syx * A[k];
                sYZ[e] = syz * A[k];
                tYX[e] = dyx * A[k];
                tYZ[e] = dyz * A[k];
            }
            if (rg[k] >= 0) {
                const size_t row = (size_t)it * a.n_rec;
                for (int q = rstart[rg[k]]; q < rstart[rg[k] + 1]; q++) {
                    const int l = rorder[q];
                    if (PSV && a.rec_u) {
                        ru[row + l] = tU[e] * A[k];
                        rw[row + l] = tW[e] * A[k];
                    }
                    if (SH && a.rec_v) rv[row + l] = tV[e] * A[k];
                }
            }
        }
        started = true;
        __syncthreads();
    }

    // ---- the tile's states (post-sponge velocities)
    float* const* out = a.flip ? a.st : a.alt;
    float* const* dout = a.flip ? a.dst : a.dalt;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        if ((fl[k] & 5u) != 5u) continue;
        const int e = t + NT * k;
        const int c = gc[k];
        if (PSV) {
            st(out[0], c, sU[e] * A[k]);
            st(out[2], c, sW[e] * A[k]);
            st(out[3], c, sXX[e]);
            st(out[4], c, sZZ[e]);
            st(out[5], c, sXZ[e]);
            st(dout[0], c, tU[e] * A[k]);
            st(dout[2], c, tW[e] * A[k]);
            st(dout[3], c, tXX[e]);
            st(dout[4], c, tZZ[e]);
            st(dout[5], c, tXZ[e]);
        }
        if (SH) {
            st(out[1], c, sV[e] * A[k]);
            st(out[6], c, sYX[e]);
            st(out[7], c, sYZ[e]);
            st(dout[1], c, tV[e] * A[k]);
            st(dout[6], c, tYX[e]);
            st(dout[7], c, tYZ[e]);
        }
    }
}

// ---------------------------------------------------------------------------
// host side

template <bool PSV, bool SH, int TX, int TZ, int K>
constexpr bool adj_fits() {
    return adj_lds_bytes<PSV, SH, TX, TZ, K>() <= kLdsLimit;
}

template <bool PSV, bool SH, int TX, int TZ, int K>
constexpr bool fwd_fits() {
    return 4 * (size_t)(TX + 2 * K) * (TZ + 2 * K) * ((PSV ? 5 : 0) + (SH ? 3 : 0)) <= kLdsLimit;
}

static bool fwd_cfg_fits(int systems, const BlockCfg& c) {
    const int nf = ((systems & TRTM_PSV) ? 5 : 0) + ((systems & TRTM_SH) ? 3 : 0);
    return 4 * (size_t)(c.tx + 2 * c.k) * (c.tz + 2 * c.k) * nf <= kLdsLimit;
}

// adj_lds_bytes for a runtime shape
static bool adj_cfg_fits(int systems, const BlockCfg& c) {
    const bool psv = systems & TRTM_PSV, sh = systems & TRTM_SH;
    const size_t nf = (psv ? 5 : 0) + (sh ? 3 : 0), nv = (psv ? 2 : 0) + (sh ? 1 : 0);
    const size_t ns = (psv ? 3 : 0) + (sh ? 2 : 0);
    const size_t rn = (size_t)(c.tx + 2 * c.k) * (c.tz + 2 * c.k);
    const size_t field = (((size_t)(c.tx + 2) * (c.tz + 2) + 63) / 64) * 64;
    const int ks = adj_stress_every(c.k);
    const size_t slots = ks - 1 > 2 ? ks - 1 : 2;
    return 4 * (2 * nv * field + slots * ns * field + rn * nf) <= kLdsLimit;
}

// compiled shapes (tile rows, tile columns, threads, steps per launch)
static const BlockCfg kFwiCfg[] = {{24, 64, 768, 4},  {32, 64, 1024, 4}, {16, 64, 512, 4},
                                   {16, 64, 512, 8},  {32, 64, 1024, 2}, {16, 64, 256, 2},
                                   {16, 48, 512, 4},  {24, 32, 512, 4}, {20, 40, 512, 4}};
// (the 512-thread 16 x 48, 24 x 32 and 20 x 40 shapes keep the P-SV adjoint
// within half the LDS of a CU, so two workgroups share one and each one's
// barrier waits overlap the other's work; 20 x 40 has the fewest halo cells
// per tile cell of them with 40-cell columns, the fastest adjoint measured on
// the 4000 x 2000 grid, 5 shots, 4 % ahead of 16 x 48; 22 x 38, 28 x 30 and
// 26 x 32 fell between 20 x 40 and 16 x 48 or behind it; 20 x 40 with 448
// threads, three cells for every thread, was 13 % slower.  Measured slower:
// 16 x 48 and 24 x 32 with 8 steps per launch and 1024 threads, which read the
// state, medium and gradients half as often per step but leave one workgroup
// per CU (+22-25 %), and 16 x 32 with 320 threads, three per CU (+45 %))
// defaults: the forward passes {24, 64, 768, 4} (the fastest measured on
// MI355X for 8 P-SV shots of 600 x 300), the adjoint {16, 48, 512, 4} (its
// LDS fits every wave system); the first listed shape that fits otherwise
static const BlockCfg kFwiDefaultFwd = {24, 64, 768, 4}, kFwiDefaultAdj = {16, 48, 512, 4};
// forward-only shapes (not listed for the adjoint pass): 8 steps per launch
// halve the state and coefficient traffic per step of the checkpointing and
// history passes; the 1024-thread 48 x 48 and 40 x 56 tiles fill half the LDS
// of a CU with fewer halo cells per tile cell than 32 x 64 (the salt job's
// checkpointing pass 5 %, its history pass 2 % faster per launch; the tuner
// kept them over 34 x 64 and 36 x 60)
static const BlockCfg kFwiFwdOnly[] = {{32, 64, 1024, 8}, {24, 64, 768, 8}, {48, 64, 1024, 8},
                                       {48, 48, 1024, 8}, {40, 56, 1024, 8}};

static bool scat_cfg_fits(int systems, const BlockCfg& c);

// shapes of a pass (0 forward, 1 adjoint, 2 scattered forward) whose LDS fits
// the workgroup limit for these systems
int fwi_block_configs(int systems, int pass, int32_t* out, int max_n) {
    int n = 0;
    auto add = [&](const BlockCfg& c) {
        if (n < max_n) {
            out[4 * n] = c.tx;
            out[4 * n + 1] = c.tz;
            out[4 * n + 2] = c.nt;
            out[4 * n + 3] = c.k;
        }
        n++;
    };
    for (const BlockCfg& c : kFwiCfg)
        if (pass == 0 ? fwd_cfg_fits(systems, c)
                      : pass == 1 ? adj_cfg_fits(systems, c) : scat_cfg_fits(systems, c))
            add(c);
    if (pass != 1)
        for (const BlockCfg& c : kFwiFwdOnly)
            if (pass == 0 ? fwd_cfg_fits(systems, c) : scat_cfg_fits(systems, c)) add(c);
    return n;
}

// the pass's default, or the first listed shape that fits the systems and
// matches every nonzero request
static int fwi_choose(const trtm_fwi_desc* d, bool adjoint, BlockCfg* out) {
    const BlockCfg& def = adjoint ? kFwiDefaultAdj : kFwiDefaultFwd;
    auto fits = [&](const BlockCfg& c) {
        return adjoint ? adj_cfg_fits(d->systems, c) : fwd_cfg_fits(d->systems, c);
    };
    if (d->tile_x <= 0 && d->tile_z <= 0 && d->block_steps <= 0 && fits(def)) {
        *out = def;
        return TRTM_OK;
    }
    auto match = [&](const BlockCfg& c) {
        return (d->tile_x <= 0 || c.tx == d->tile_x) && (d->tile_z <= 0 || c.tz == d->tile_z) &&
               (d->block_steps <= 0 || c.k == d->block_steps);
    };
    for (const BlockCfg& c : kFwiCfg)
        if (fits(c) && match(c)) {
            *out = c;
            return TRTM_OK;
        }
    if (!adjoint)
        for (const BlockCfg& c : kFwiFwdOnly)
            if (fwd_cfg_fits(d->systems, c) && match(c)) {
                *out = c;
                return TRTM_OK;
            }
    return fail(TRTM_ERR_UNSUPPORTED, "blocked FWI %s kernel has no tile_x=%d tile_z=%d "
                "block_steps=%d build", adjoint ? "adjoint" : "forward", d->tile_x, d->tile_z,
                d->block_steps);
}

static FB make_fb(const trtm_fwi_desc* d, bool adjoint) {
    FB a;
    a.nx = d->nx; a.nz = d->nz; a.nt = d->nt; a.fw = d->fw;
    a.n_rec = d->n_rec; a.hist_base = d->hist_base; a.hist_len = d->hist_len; a.flip = 0;
    a.hist_every = d->hist_every;
    a.dt = (float)d->dt; a.idx = (float)d->inv_dx; a.idz = (float)d->inv_dz;
    a.factor = (float)d->src_factor;
    a.tab = (const float*)d->sponge_table;
    a.lam = (const float*)d->lam; a.mu = (const float*)d->mu; a.mxz = (const float*)d->mxz;
    a.myx = (const float*)d->myx; a.myz = (const float*)d->myz; a.ir = (const float*)d->inv_rho;
    a.iru = (const float*)d->inv_rho_u; a.irw = (const float*)d->inv_rho_w;
    void* const fwd[8] = {d->u, d->v, d->w, d->sxx, d->szz, d->sxz, d->syx, d->syz};
    void* const adj[8] = {d->pu, d->pv, d->pw, d->psxx, d->pszz, d->psxz, d->psyx, d->psyz};
    const size_t n = (size_t)d->nx * d->nz;
    for (int f = 0; f < 8; f++) {
        a.st[f] = (float*)(adjoint ? adj[f] : fwd[f]);
        a.alt[f] = (float*)d->workspace + f * n;
    }
    a.src_map = d->n_src_groups > 0 ? d->cell_map : nullptr;
    const bool records = !adjoint && (d->rec_u || d->rec_v);
    a.rec_map = d->n_rec_groups > 0 && (adjoint || records) ? d->cell_map + n : nullptr;
    a.src_order = d->src_order; a.src_start = d->src_group_start;
    a.rec_order = d->rec_order; a.rec_start = d->rec_group_start;
    a.wav_u = (const float*)d->wav_u; a.wav_v = (const float*)d->wav_v;
    a.wav_w = (const float*)d->wav_w;
    a.rec_u = (float*)d->rec_u; a.rec_v = (float*)d->rec_v; a.rec_w = (float*)d->rec_w;
    a.res_u = (const float*)d->res_u; a.res_v = (const float*)d->res_v;
    a.res_w = (const float*)d->res_w;
    a.hist = (float*)d->hist;
    a.g_lam = (float*)d->g_lam; a.g_mu = (float*)d->g_mu; a.g_mxz = (float*)d->g_mxz;
    a.g_myx = (float*)d->g_myx; a.g_myz = (float*)d->g_myz; a.g_ir = (float*)d->g_ir;
    a.g_iru = (float*)d->g_iru; a.g_irw = (float*)d->g_irw;
    return a;
}

template <bool PSV, bool SH, int TX, int TZ, int NT, int K, class S>
static void launch_fwd(bool hist, const S& src, int shots, int t_a, int t_b, int tiles_x,
                       int tiles_z, hipStream_t s) {
    const int tiles = tiles_x * tiles_z;
    const dim3 grid(8 * ((tiles + 7) / 8), 1, shots);
    if (hist)
        hipLaunchKernelGGL((k_fwi_fwd_blk<PSV, SH, TX, TZ, NT, K, true, S>), grid, dim3(NT), 0, s,
                           src, t_a, t_b, tiles_x, tiles_z);
    else
        hipLaunchKernelGGL((k_fwi_fwd_blk<PSV, SH, TX, TZ, NT, K, false, S>), grid, dim3(NT), 0,
                           s, src, t_a, t_b, tiles_x, tiles_z);
}

template <bool PSV, bool SH, int TX, int TZ, int NT, int K, class S>
static void launch_fb(bool adjoint, bool hist, const S& src, int shots, int t_a, int t_b,
                      int tiles_x, int tiles_z, hipStream_t s) {
    const int tiles = tiles_x * tiles_z;
    const dim3 grid(8 * ((tiles + 7) / 8), 1, shots);
    if (adjoint)
        hipLaunchKernelGGL((k_fwi_adj_blk<PSV, SH, TX, TZ, NT, K, S>), grid, dim3(NT), 0, s, src,
                           t_a, t_b, tiles_x, tiles_z);
    else if (hist)
        hipLaunchKernelGGL((k_fwi_fwd_blk<PSV, SH, TX, TZ, NT, K, true, S>), grid, dim3(NT), 0, s,
                           src, t_a, t_b, tiles_x, tiles_z);
    else
        hipLaunchKernelGGL((k_fwi_fwd_blk<PSV, SH, TX, TZ, NT, K, false, S>), grid, dim3(NT), 0,
                           s, src, t_a, t_b, tiles_x, tiles_z);
}

template <bool PSV, bool SH, class S>
static int launch_cfg(const BlockCfg& c, bool adjoint, bool hist, const S& src, int shots,
                      int t_a, int t_b, int nx, int nz, hipStream_t s) {
    const int tx_ = (nx + c.tx - 1) / c.tx, tz_ = (nz + c.tz - 1) / c.tz;
#define TRTM_FB(TX, TZ, NT, K)                                                                  \
    if (c.tx == TX && c.tz == TZ && c.nt == NT && c.k == K) {                                   \
        if constexpr (adj_fits<PSV, SH, TX, TZ, K>() && fwd_fits<PSV, SH, TX, TZ, K>()) {       \
            launch_fb<PSV, SH, TX, TZ, NT, K, S>(adjoint, hist, src, shots, t_a, t_b, tx_, tz_, \
                                                 s);                                            \
            return TRTM_OK;                                                                     \
        } else if constexpr (fwd_fits<PSV, SH, TX, TZ, K>()) {                                  \
            if (!adjoint) {                                                                     \
                launch_fwd<PSV, SH, TX, TZ, NT, K, S>(hist, src, shots, t_a, t_b, tx_, tz_, s); \
                return TRTM_OK;                                                                 \
            }                                                                                   \
        }                                                                                       \
    }
    TRTM_FB(16, 64, 512, 4)
    TRTM_FB(16, 64, 256, 2)
    TRTM_FB(32, 64, 1024, 4)
    TRTM_FB(16, 64, 512, 8)
    TRTM_FB(32, 64, 1024, 2)
    TRTM_FB(24, 64, 768, 4)
    TRTM_FB(16, 48, 512, 4)
    TRTM_FB(24, 32, 512, 4)
    TRTM_FB(20, 40, 512, 4)
#undef TRTM_FB
    // forward-only shapes
#define TRTM_FF(TX, TZ, NT, K)                                                                  \
    if constexpr (fwd_fits<PSV, SH, TX, TZ, K>()) {                                             \
        if (!adjoint && c.tx == TX && c.tz == TZ && c.nt == NT && c.k == K) {                   \
            const int tiles = tx_ * tz_;                                                        \
            const dim3 grid(8 * ((tiles + 7) / 8), 1, shots);                                   \
            if (hist)                                                                           \
                hipLaunchKernelGGL((k_fwi_fwd_blk<PSV, SH, TX, TZ, NT, K, true, S>), grid,      \
                                   dim3(NT), 0, s, src, t_a, t_b, tx_, tz_);                    \
            else                                                                                \
                hipLaunchKernelGGL((k_fwi_fwd_blk<PSV, SH, TX, TZ, NT, K, false, S>), grid,     \
                                   dim3(NT), 0, s, src, t_a, t_b, tx_, tz_);                    \
            return TRTM_OK;                                                                     \
        }                                                                                       \
    }
    TRTM_FF(32, 64, 1024, 8)
    TRTM_FF(24, 64, 768, 8)
    TRTM_FF(48, 64, 1024, 8)
    TRTM_FF(48, 48, 1024, 8)
    TRTM_FF(40, 56, 1024, 8)
#undef TRTM_FF
    return fail(TRTM_ERR_UNSUPPORTED, "blocked FWI shape %dx%d/%d/%d not compiled", c.tx, c.tz,
                c.nt, c.k);
}

template <class S>
static int launch_sys(int systems, const BlockCfg& c, bool adjoint, bool hist, const S& src,
                      int shots, int t_a, int t_b, int nx, int nz, hipStream_t s) {
    if (systems == (TRTM_PSV | TRTM_SH))
        return launch_cfg<true, true>(c, adjoint, hist, src, shots, t_a, t_b, nx, nz, s);
    if (systems == TRTM_PSV)
        return launch_cfg<true, false>(c, adjoint, hist, src, shots, t_a, t_b, nx, nz, s);
    return launch_cfg<false, true>(c, adjoint, hist, src, shots, t_a, t_b, nx, nz, s);
}

// Steps [it_begin, it_end) of n shots, ascending (forward) or descending
// (adjoint) in launches of K steps; shots in groups of kShotsPerLaunch.  Each
// shot's state alternates between its buffers and the workspace; a shot takes
// part in a launch while the launch's first step is below its record length.
static int fwi_block_run(const trtm_fwi_desc* d, int n, int it_begin, int it_end, bool adjoint,
                         hipStream_t s) {
    if (it_begin >= it_end) return TRTM_OK;
    BlockCfg c;
    int rc = fwi_choose(d, adjoint, &c);
    if (rc) return rc;
    for (int q = 0; q < n; q++)
        if (!d[q].workspace)
            return fail(TRTM_ERR_ARG, "shot %d: blocked FWI kernel needs the workspace", q);
    const bool hist = !adjoint && d->hist != nullptr;
    const int K = c.k, groups = (n + kShotsPerLaunch - 1) / kShotsPerLaunch;
    const int KS = adj_stress_every(K);
    if (adjoint && d->hist_every != KS)
        return fail(TRTM_ERR_ARG, "the blocked FWI adjoint of %d steps per launch reads the "
                    "compact history with hist_every = %d (got %d)", K, KS, d->hist_every);
    if (hist && d->hist_every < 0)
        return fail(TRTM_ERR_ARG, "hist_every must be >= 0");
    std::vector<FB> fb(n);
    for (int q = 0; q < n; q++) fb[q] = make_fb(&d[q], adjoint);
    // launches of up to K steps in stepping order (adjoint: whole intervals of
    // KS steps from hist_base, whose first steps' stresses are stored, at most
    // K / KS of them); an even number of them (the first one split in two when
    // needed) leaves the state in the caller's buffers for every shot that
    // runs to the end of the range
    std::vector<std::pair<int, int>> chunks;
    if (adjoint) {
        for (int hi = it_end; hi > it_begin;) {
            const int q = (hi - 1 - d->hist_base) / KS;  // interval of step hi - 1
            const int q_lo = std::max(q - (K / KS - 1), 0);
            const int lo = std::max(d->hist_base + q_lo * KS, it_begin);
            chunks.push_back({lo, hi});
            hi = lo;
        }
    } else {
        for (int t_a = it_begin; t_a < it_end; t_a += K) chunks.push_back({t_a, std::min(t_a + K, it_end)});
    }
    if (chunks.size() % 2 == 1 && chunks[0].second - chunks[0].first > 1) {
        const auto first = chunks[0];
        const int cut = adjoint ? first.second - 1 : first.first + 1;
        chunks[0] = adjoint ? std::make_pair(cut, first.second) : std::make_pair(first.first, cut);
        chunks.insert(chunks.begin() + 1, adjoint ? std::make_pair(first.first, cut)
                                                  : std::make_pair(cut, first.second));
    }
    // Each group runs the whole range before the next one starts: a launch
    // then reads the state that the group's previous launch has just written,
    // while it is still in the last-level cache (two 8-shot groups of state
    // and history between them would not fit).
    for (int g = 0; g < groups; g++) {
        const int q0 = g * kShotsPerLaunch, m = std::min(kShotsPerLaunch, n - q0);
        for (const auto& ch : chunks) {
            const int t_a = ch.first, t_b = ch.second;
            if (n == 1) {
                rc = launch_sys(d->systems, c, adjoint, hist, ByValue<FB>{fb[0]}, 1, t_a, t_b,
                                d->nx, d->nz, s);
            } else {
                ByArray<FB> arr;
                for (int q = 0; q < m; q++) arr.v[q] = fb[q0 + q];
                rc = launch_sys(d->systems, c, adjoint, hist, arr, m, t_a, t_b, d->nx, d->nz, s);
            }
            if (rc) return rc;
            for (int q = q0; q < q0 + m; q++)
                if (t_a < fb[q].nt) fb[q].flip ^= 1;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return fail(TRTM_ERR_HIP, "blocked FWI %s launch: %s", adjoint ? "adjoint" : "forward",
                        hipGetErrorString(e));
    }
    // state left in the workspace goes back to the caller's buffers
    const size_t bytes = (size_t)d->nx * d->nz * sizeof(float);
    for (int q = 0; q < n; q++) {
        if (!fb[q].flip) continue;
        for (int f = 0; f < 8; f++) {
            const bool psv_f = f == 0 || f == 2 || (f >= 3 && f <= 5);
            if (!(psv_f ? (d->systems & TRTM_PSV) : (d->systems & TRTM_SH))) continue;
            const hipError_t e =
                hipMemcpyAsync(fb[q].st[f], fb[q].alt[f], bytes, hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess)
                return fail(TRTM_ERR_HIP, "blocked FWI state copy: %s", hipGetErrorString(e));
        }
    }
    return TRTM_OK;
}

// ---- scattered-field forward

template <bool PSV, bool SH, int TX, int TZ, int K>
constexpr bool scat_fits() {
    return 4 * (size_t)(TX + 2 * K) * (TZ + 2 * K) * ((PSV ? 10 : 0) + (SH ? 6 : 0)) <= kLdsLimit;
}

static bool scat_cfg_fits(int systems, const BlockCfg& c) {
    const int nf = ((systems & TRTM_PSV) ? 10 : 0) + ((systems & TRTM_SH) ? 6 : 0);
    return 4 * (size_t)(c.tx + 2 * c.k) * (c.tz + 2 * c.k) * nf <= kLdsLimit;
}

// the forward default when it fits, else the first listed shape that fits
// and matches every nonzero request
static int scat_choose(const trtm_fwi_desc* d, BlockCfg* out) {
    const BlockCfg& def = kFwiDefaultFwd;
    if (d->tile_x <= 0 && d->tile_z <= 0 && d->block_steps <= 0 && scat_cfg_fits(d->systems, def)) {
        *out = def;
        return TRTM_OK;
    }
    for (const BlockCfg* list : {&kFwiCfg[0], &kFwiFwdOnly[0]}) {
        const int cnt = list == kFwiCfg ? (int)(sizeof(kFwiCfg) / sizeof(BlockCfg))
                                        : (int)(sizeof(kFwiFwdOnly) / sizeof(BlockCfg));
        for (int q = 0; q < cnt; q++) {
            const BlockCfg& c = list[q];
            if (scat_cfg_fits(d->systems, c) && (d->tile_x <= 0 || c.tx == d->tile_x) &&
                (d->tile_z <= 0 || c.tz == d->tile_z) &&
                (d->block_steps <= 0 || c.k == d->block_steps)) {
                *out = c;
                return TRTM_OK;
            }
        }
    }
    return fail(TRTM_ERR_UNSUPPORTED, "scattered FWI kernel has no tile_x=%d tile_z=%d "
                "block_steps=%d build for these wave systems", d->tile_x, d->tile_z,
                d->block_steps);
}

template <bool PSV, bool SH>
static int launch_scat_cfg(const BlockCfg& c, const ByArray4<FS>& src, int shots, int t_a,
                           int t_b, int nx, int nz, hipStream_t s) {
    const int tx_ = (nx + c.tx - 1) / c.tx, tz_ = (nz + c.tz - 1) / c.tz;
    const int tiles = tx_ * tz_;
    const dim3 grid(8 * ((tiles + 7) / 8), 1, shots);
#define TRTM_FS(TX, TZ, NT, K)                                                                  \
    if constexpr (scat_fits<PSV, SH, TX, TZ, K>()) {                                            \
        if (c.tx == TX && c.tz == TZ && c.nt == NT && c.k == K) {                               \
            hipLaunchKernelGGL((k_fwi_scat_blk<PSV, SH, TX, TZ, NT, K, ByArray4<FS>>), grid,    \
                               dim3(NT), 0, s, src, t_a, t_b, tx_, tz_);                        \
            return TRTM_OK;                                                                     \
        }                                                                                       \
    }
    TRTM_FS(24, 64, 768, 4)
    TRTM_FS(32, 64, 1024, 4)
    TRTM_FS(16, 64, 512, 4)
    TRTM_FS(16, 64, 512, 8)
    TRTM_FS(32, 64, 1024, 2)
    TRTM_FS(16, 64, 256, 2)
    TRTM_FS(16, 48, 512, 4)
    TRTM_FS(24, 32, 512, 4)
    TRTM_FS(20, 40, 512, 4)
    TRTM_FS(32, 64, 1024, 8)
    TRTM_FS(24, 64, 768, 8)
    TRTM_FS(48, 64, 1024, 8)
    TRTM_FS(48, 48, 1024, 8)
    TRTM_FS(40, 56, 1024, 8)
#undef TRTM_FS
    return fail(TRTM_ERR_UNSUPPORTED, "scattered FWI shape %dx%d/%d/%d not compiled", c.tx, c.tz,
                c.nt, c.k);
}

static FS make_fs(const trtm_fwi_desc* d) {
    FS a;
    a.nx = d->nx; a.nz = d->nz; a.nt = d->nt; a.fw = d->fw;
    a.n_rec = d->n_rec; a.flip = 0; a.pad0 = a.pad1 = 0;
    a.dt = (float)d->dt; a.idx = (float)d->inv_dx; a.idz = (float)d->inv_dz;
    a.factor = (float)d->src_factor;
    a.tab = (const float*)d->sponge_table;
    a.lam = (const float*)d->lam; a.mu = (const float*)d->mu; a.mxz = (const float*)d->mxz;
    a.myx = (const float*)d->myx; a.myz = (const float*)d->myz; a.ir = (const float*)d->inv_rho;
    a.iru = (const float*)d->inv_rho_u; a.irw = (const float*)d->inv_rho_w;
    a.elam = (const float*)d->dif_lam; a.emu = (const float*)d->dif_mu;
    a.emxz = (const float*)d->dif_mxz; a.emyx = (const float*)d->dif_myx;
    a.emyz = (const float*)d->dif_myz; a.eir = (const float*)d->dif_inv_rho;
    a.eiru = (const float*)d->dif_inv_rho_u; a.eirw = (const float*)d->dif_inv_rho_w;
    void* const fwd[8] = {d->u, d->v, d->w, d->sxx, d->szz, d->sxz, d->syx, d->syz};
    void* const sc[8] = {d->du, d->dv, d->dw, d->dsxx, d->dszz, d->dsxz, d->dsyx, d->dsyz};
    const size_t n = (size_t)d->nx * d->nz;
    for (int f = 0; f < 8; f++) {
        a.st[f] = (float*)fwd[f];
        a.alt[f] = (float*)d->workspace + f * n;
        a.dst[f] = (float*)sc[f];
        a.dalt[f] = (float*)d->scat_workspace + f * n;
    }
    a.src_map = d->n_src_groups > 0 ? d->cell_map : nullptr;
    a.rec_map = d->n_rec_groups > 0 && (d->rec_u || d->rec_v) ? d->cell_map + n : nullptr;
    a.src_order = d->src_order; a.src_start = d->src_group_start;
    a.rec_order = d->rec_order; a.rec_start = d->rec_group_start;
    a.wav_u = (const float*)d->wav_u; a.wav_v = (const float*)d->wav_v;
    a.wav_w = (const float*)d->wav_w;
    a.rec_u = (float*)d->rec_u; a.rec_v = (float*)d->rec_v; a.rec_w = (float*)d->rec_w;
    return a;
}

// Steps [it_begin, it_end) of n shots in launches of K steps (an even number,
// the first split when needed), up to kScatShotsPerLaunch shots per launch,
// each group of shots through the whole range before the next.
int fwi_scat_forward(const trtm_fwi_desc* d, int n, int it_begin, int it_end, hipStream_t s) {
    if (it_begin >= it_end) return TRTM_OK;
    BlockCfg c;
    int rc = scat_choose(d, &c);
    if (rc) return rc;
    std::vector<FS> fs(n);
    for (int q = 0; q < n; q++) fs[q] = make_fs(&d[q]);
    const int K = c.k, groups = (n + kScatShotsPerLaunch - 1) / kScatShotsPerLaunch;
    std::vector<std::pair<int, int>> chunks;
    for (int l = 0; l * K < it_end - it_begin; l++)
        chunks.push_back({it_begin + l * K, std::min(it_begin + (l + 1) * K, it_end)});
    if (chunks.size() % 2 == 1 && chunks[0].second - chunks[0].first > 1) {
        const auto first = chunks[0];
        chunks[0] = {first.first, first.first + 1};
        chunks.insert(chunks.begin() + 1, {first.first + 1, first.second});
    }
    for (int g = 0; g < groups; g++) {
        const int q0 = g * kScatShotsPerLaunch, m = std::min(kScatShotsPerLaunch, n - q0);
        for (const auto& ch : chunks) {
            ByArray4<FS> arr;
            for (int q = 0; q < m; q++) arr.v[q] = fs[q0 + q];
            for (int q = m; q < kScatShotsPerLaunch; q++) arr.v[q] = fs[q0];
            if (d->systems == (TRTM_PSV | TRTM_SH))
                rc = launch_scat_cfg<true, true>(c, arr, m, ch.first, ch.second, d->nx, d->nz, s);
            else if (d->systems == TRTM_PSV)
                rc = launch_scat_cfg<true, false>(c, arr, m, ch.first, ch.second, d->nx, d->nz, s);
            else
                rc = launch_scat_cfg<false, true>(c, arr, m, ch.first, ch.second, d->nx, d->nz, s);
            if (rc) return rc;
            for (int q = q0; q < q0 + m; q++)
                if (ch.first < fs[q].nt) fs[q].flip ^= 1;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return fail(TRTM_ERR_HIP, "scattered FWI launch: %s", hipGetErrorString(e));
    }
    const size_t bytes = (size_t)d->nx * d->nz * sizeof(float);
    for (int q = 0; q < n; q++) {
        if (!fs[q].flip) continue;
        for (int f = 0; f < 8; f++) {
            const bool psv_f = f == 0 || f == 2 || (f >= 3 && f <= 5);
            if (!(psv_f ? (d->systems & TRTM_PSV) : (d->systems & TRTM_SH))) continue;
            hipError_t e =
                hipMemcpyAsync(fs[q].st[f], fs[q].alt[f], bytes, hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(fs[q].dst[f], fs[q].dalt[f], bytes, hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess)
                return fail(TRTM_ERR_HIP, "scattered FWI state copy: %s", hipGetErrorString(e));
        }
    }
    return TRTM_OK;
}

int fwi_block_forward(const trtm_fwi_desc* d, int n, int it_begin, int it_end, hipStream_t s) {
    return fwi_block_run(d, n, it_begin, it_end, false, s);
}

int fwi_block_adjoint(const trtm_fwi_desc* d, int n, int it_begin, int it_end, hipStream_t s) {
    return fwi_block_run(d, n, it_begin, it_end, true, s);
}

}  // namespace trtm
