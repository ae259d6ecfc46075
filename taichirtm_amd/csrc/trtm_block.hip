// Temporally blocked, multi-shot time stepping (one launch per K steps).
//
// A workgroup owns a tile of TX x 64 cells of one shot.  It loads the tile
// plus a K-cell halo on every side (the region, RX x RZ cells) of the state
// into LDS once, keeps the region's medium coefficients and sponge factors in
// registers, and advances K time steps inside LDS: per step the velocity
// update, the injections, the stress update and the sponge
// (forward_modeling.py:539-572, backward_modeling.py:539-576).  Each half-step
// leaves one more halo row/column stale on one side (the velocity update reads
// i-1/j-1 forward, the stress update i+1/j+1), so after K steps exactly the
// tile is valid; halo cells are recomputed redundantly by the neighbouring
// tiles with the same expressions, so every valid value is bit-identical to
// the one-step-per-launch kernels.
//
// Per step the tile's receivers are recorded, snapshots written, the image
// accumulated in registers (written back once per launch) and the finite
// check taken, all from the post-sponge values.  Velocities stay unsponged in
// LDS and the sponge factor is applied when they are next used, which gives
// the same rounding as the reference's in-place multiply.
//
// The grid's second dimension is the shot: shots share the medium and differ
// in state, record length, injection and recording geometry (ShotDev).
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "trtm_block.hpp"

namespace trtm {

int fail(int code, const char* fmt, ...);

// (512-thread shapes: at least 4 waves per SIMD, i.e. at most 128 VGPRs, so
// two workgroups share a CU; a 32 x 64, 8-step P-SV shape with 1024 threads
// held to 64 VGPRs for two per CU spilled 35-62 registers)
template <bool FWD, bool PSV, bool SH, int TX, int TZ, int NT, int K>
__global__ __launch_bounds__(NT, NT == 512 ? 4 : 1) void k_block(KArgs m,
                                                                 const ShotDev* __restrict__ shots,
                                                                 int it0, int tiles_x,
                                                                 int tiles_z) {
    constexpr int RX = TX + 2 * K, RZ = TZ + 2 * K, RN = RX * RZ;
    constexpr int NC = (RN + NT - 1) / NT;
    static_assert(2 * K <= TX && 2 * K <= TZ,
                  "the per-tile injection lists hold a cell in at most 2 x 2 regions");
    // Region arrays with RZ + 1 cells of padding on each side and room for
    // every thread slot: without a surface matrix the updates run branch-free
    // on every slot (neighbour reads of edge and spare slots stay inside the
    // array) and the result is selected per cell.  SH-only kernels and the
    // P-SV-only shapes with at most four cells per thread run them; the other
    // P-SV kernels keep per-cell branches (their registers are the tighter
    // limit).
    constexpr bool BFP = PSV && !SH && NC <= 4;  // branch-free P-SV updates
    constexpr int PAD = (!PSV || BFP) ? RZ + 1 : 0;
    constexpr int NA = PAD == 0 ? RN : (RN > NT * NC ? RN : NT * NC) + 2 * PAD;
    constexpr int NP = PSV ? NA : 1, NS = SH ? NA : 1;
    __shared__ float sU_[NP], sW_[NP], sXX_[NP], sZZ_[NP], sXZ_[NP];
    __shared__ float sV_[NS], sYX_[NS], sYZ_[NS];
    constexpr int OP = PSV ? PAD : 0, OS = SH ? PAD : 0;
    float* const sU = sU_ + OP;
    float* const sW = sW_ + OP;
    float* const sXX = sXX_ + OP;
    float* const sZZ = sZZ_ + OP;
    float* const sXZ = sXZ_ + OP;
    float* const sV = sV_ + OS;
    float* const sYX = sYX_ + OS;
    float* const sYZ = sYZ_ + OS;

    // small shapes: injection group of each region cell, and the injected
    // values of the launch's K steps for up to kInjTab single-entry groups of
    // the tile (see below)
    constexpr bool HOIST = NC <= 2;
    constexpr int kInjTab = 64;
    __shared__ int gmap[HOIST ? RN : 1];
    __shared__ float tabU[HOIST && PSV ? kInjTab * K : 1], tabW[HOIST && PSV ? kInjTab * K : 1];
    __shared__ float tabV[HOIST && SH ? kInjTab * K : 1];

    const ShotDev* __restrict__ sh = shots + blockIdx.y;
    const int it_end = sh->it_end, it_lo = sh->it_lo, it_off = sh->it_off;
    const int tiles = tiles_x * tiles_z;
    const int tile = xcd_tile_b(blockIdx.x, tiles);
    // this launch covers the shot's steps it0 + it_off .. it0 + it_off + K - 1
    if (tile >= tiles || it0 + it_off >= it_end || it0 + it_off + K <= it_lo) return;
    const int nt = sh->nt;
    const int i0 = (tile / tiles_z) * TX, j0 = (tile % tiles_z) * TZ;
    const int ri0 = i0 - K, rj0 = j0 - K;
    const int nx = m.nx, nz = m.nz;
    const int t = threadIdx.x;
    const gptr<const float> sm = G(m.sm);
    const int it_first = it0 + it_off;

    // per owned cell: global offset, flags, sponge factor, coefficients
    int gc[NC];
    unsigned fl[NC];  // bit0 in grid, bit1 interior, bit2 in tile, bit3 V ok, bit4 S ok
    float A[NC], c0[NC], c1[NC], c2[NC], c3[NC], c4[NC], d0[NC], d1[NC], d2[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const int e = t + NT * k;
        const int r = e / RZ, z = e - r * RZ;
        const int i = ri0 + r, j = rj0 + z;
        unsigned f = 0;
        gc[k] = 0;
        A[k] = 1.0f;
        c0[k] = c1[k] = c2[k] = c3[k] = c4[k] = d0[k] = d1[k] = d2[k] = 0.0f;
        if (e < RN && i >= 0 && i < nx && j >= 0 && j < nz) {
            f |= 1u;
            const int c = i * nz + j;
            gc[k] = c;
            if (i >= 1 && i < nx - 1 && j >= 1 && j < nz - 1) f |= 2u;
            if (i >= i0 && i < i0 + TX && j >= j0 && j < j0 + TZ) f |= 4u;
            const bool lo = r >= 1 && z >= 1, hi = r <= RX - 2 && z <= RZ - 2;
            if (FWD ? lo : hi) f |= 8u;
            if (FWD ? hi : lo) f |= 16u;
            A[k] = sponge_at(i, j, nx, nz, m.fw, m.tab);
            if (PSV) {
                c0[k] = G(m.iru)[c];
                c1[k] = G(m.irw)[c];
                c2[k] = G(m.lam)[c];
                c3[k] = G(m.mu)[c];
                c4[k] = G(m.mxz)[c];
            }
            if (SH) {
                d0[k] = G(m.ir)[c];
                d1[k] = G(m.myx)[c];
                d2[k] = G(m.myz)[c];
            }
        }
        fl[k] = f;
    }

    // load the region's state (the small shapes hold it in registers until
    // the lookups and the schedule below are issued, so that those overlap
    // the loads)
    constexpr int NR = HOIST ? NC : 1;
    float st[NR][PSV && SH ? 8 : (PSV ? 5 : 3)];
    {
        const gptr<float> in_u = G(sh->in[0]), in_v = G(sh->in[1]), in_w = G(sh->in[2]);
        const gptr<float> in_xx = G(sh->in[3]), in_zz = G(sh->in[4]), in_xz = G(sh->in[5]);
        const gptr<float> in_yx = G(sh->in[6]), in_yz = G(sh->in[7]);
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const int e = t + NT * k;
            if (e >= RN) continue;
            const bool g = fl[k] & 1u;
            const int c = gc[k];
            float x[PSV && SH ? 8 : (PSV ? 5 : 3)];
            int q = 0;
            if (PSV) {
                x[q++] = g ? in_u[c] : 0.0f;
                x[q++] = g ? in_w[c] : 0.0f;
                x[q++] = g ? in_xx[c] : 0.0f;
                x[q++] = g ? in_zz[c] : 0.0f;
                x[q++] = g ? in_xz[c] : 0.0f;
            }
            if (SH) {
                x[q++] = g ? in_v[c] : 0.0f;
                x[q++] = g ? in_yx[c] : 0.0f;
                x[q++] = g ? in_yz[c] : 0.0f;
            }
            if constexpr (HOIST) {
#pragma unroll
                for (int f = 0; f < q; f++) st[k][f] = x[f];
            } else {
                q = 0;
                if (PSV) {
                    sU[e] = x[q++];
                    sW[e] = x[q++];
                    sXX[e] = x[q++];
                    sZZ[e] = x[q++];
                    sXZ[e] = x[q++];
                }
                if (SH) {
                    sV[e] = x[q++];
                    sYX[e] = x[q++];
                    sYZ[e] = x[q++];
                }
            }
        }
    }

    // image accumulators of the owned tile cells
    const bool imaging = !FWD && sh->num_snaps > 0;
    const gptr<float> img_u = G(sh->image_u), img_v = G(sh->image_v), img_w = G(sh->image_w);
    float aU[NC], aW[NC], aV[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) {
        aU[k] = aW[k] = aV[k] = 0.0f;
        if (imaging && (fl[k] & 4u)) {
            if (PSV) {
                aU[k] = img_u[gc[k]];
                aW[k] = img_w[gc[k]];
            }
            if (SH) aV[k] = img_v[gc[k]];
        }
    }

    const int save = sh->save, isnap = sh->isnap, num_snaps = sh->num_snaps;
    const int snap_base = sh->snap_base, cint = sh->check_interval;
    const gptr<float> snap_u = G(sh->snap_u), snap_v = G(sh->snap_v), snap_w = G(sh->snap_w);
    const size_t sstride = (size_t)sh->snap_stride;
    const bool ssponge = sh->snap_sponge != 0;

    // The small shapes with one wave system load the forward wavefield of the
    // image a whole step before the image update uses it (the others a phase
    // before): pf0 / pf1 hold the next step's u / w (or v) of every slot.
    constexpr bool PIPE = HOIST && (PSV != SH);
    constexpr int NF = PIPE ? NC : 1;
    float pf0[NF], pf1[NF];
    bool img_next = false;
    auto load_forward = [&](size_t off) {
#pragma unroll
        for (int k = 0; k < NF; k++) {
            if (PSV) {
                pf0[k] = snap_u[off + gc[k]];
                pf1[k] = snap_w[off + gc[k]];
            } else {
                pf0[k] = snap_v[off + gc[k]];
            }
        }
    };
    const int s_lo = it_lo > it_first ? it_lo - it_first : 0;
    const int s_hi = it_end - it_first < K ? it_end - it_first : K;
    // backward with a snapshot every step in the forward's own slots: the
    // first step's forward wavefield (the slot the schedule below finds) is
    // requested before the barrier
    bool pre = false;
    if constexpr (PIPE && !FWD) {
        if (imaging && isnap == 1 && sh->snaps_regular) {
            const int T = nt - 1 - (it_first + s_lo);
            const int sl = T - 1 - snap_base;
            if (T > 0 && sl >= 0 && sl < num_snaps) {
                pre = true;
                load_forward((size_t)sl * sstride);
            }
        }
    }

    // The tile's injection groups and receivers, looked up once per launch
    // while the region's loads are in flight (lists no longer than kInjTab
    // groups / the workgroup's receivers; longer lists keep the per-step
    // loops below): each region cell's thread slot learns
    // its group (a row of the table of injected values, or the group of
    // several entries) and adds the injection right after its own velocity
    // update, with no separate injection phase and barrier; thread t records
    // receiver rb0 + t.  Only for shapes with at most two cells per thread:
    // their few workgroups per CU leave per-step global loads' latency
    // exposed, and they have registers to spare; the larger shapes keep the
    // per-step lookups.
    const int n_groups = sh->n_groups, n_rec = sh->n_rec;
    int ig0 = 0, ig1 = 0, rb0 = 0, rb1 = 0;
    if (n_groups > 0) {
        const cptr<int> tstart = C(sh->inj_tstart);
        ig0 = tstart[tile];
        ig1 = tstart[tile + 1];
    }
    if (n_rec > 0) {
        const cptr<int> rstart = C(sh->rec_tstart);
        rb0 = rstart[tile];
        rb1 = rstart[tile + 1];
    }
    const bool inj_one_pass = HOIST && ig1 - ig0 <= kInjTab, rec_one_pass = HOIST && rb1 - rb0 <= NT;
    // per slot: the table row of its cell's single-entry group, -1 - g for a
    // group g of several entries (summed per step in entry order), or kNone
    constexpr int kNone = -2147483647 - 1;
    // Wave w sets up items w, w + NW, ... of the list; their cells and codes
    // with scalar loads, whose latency runs on the scalar counter beside the
    // region's vector loads (indexes clamped into the list; the extra items
    // fill table rows no slot reads).  An item's region cell and code go to
    // the region map after the first barrier; a single entry's K injected
    // values go to the item's row of the table, loaded by lanes 0 .. IPW K - 1
    // of the wave (time index clamped to [0, nt) at the ends of the record:
    // steps outside the shot's range never read theirs).
    constexpr int NWV = NT / 64, IPW = (kInjTab + NWV - 1) / NWV;
    static_assert(!HOIST || IPW * K <= 64, "one lane per injected value of a wave's items");
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6), ln = t & 63;
    int q_e[HOIST ? IPW : 1], q_code[HOIST ? IPW : 1];
    if constexpr (HOIST) {
        const int n_items = inj_one_pass ? ig1 - ig0 : 0;
        int k0s[IPW];
#pragma unroll
        for (int m2 = 0; m2 < IPW; m2++) {
            q_e[m2] = -1;
            q_code[m2] = kNone;
        }
        if (n_items > 0) {  // uniform
            const cptr<int> cells = C(sh->inj_tcells);
            int ce_e[IPW], ce_k[IPW];
#pragma unroll
            for (int m2 = 0; m2 < IPW; m2++) {
                const int w2 = 2 * (ig0 + min(wv + NWV * m2, n_items - 1));
                ce_e[m2] = cells[w2];
                ce_k[m2] = cells[w2 + 1];
            }
#pragma unroll
            for (int m2 = 0; m2 < IPW; m2++) {
                const int q = wv + NWV * m2;
                k0s[m2] = max(ce_k[m2], 0);
                q_e[m2] = q < n_items ? ce_e[m2] : -1;
                q_code[m2] = ce_k[m2] >= 0 ? q : ce_k[m2];
            }
            // lane l < IPW K loads value s = l % K of the wave's item l / K
            if (ln < IPW * K) {
                const int m2 = ln / K, s = ln - m2 * K;
                int k0 = k0s[0];
#pragma unroll
                for (int m3 = 1; m3 < IPW; m3++) k0 = m2 == m3 ? k0s[m3] : k0;
                const int col_ = min(max(FWD ? it_first + s : nt - 1 - (it_first + s), 0), nt - 1);
                const size_t o = (size_t)k0 * (size_t)nt + col_;
                const int r = (wv + NWV * m2) * K + s;
                if constexpr (PSV) {
                    tabU[r] = G(sh->su)[o];
                    tabW[r] = G(sh->sw)[o];
                }
                if constexpr (SH) tabV[r] = G(sh->sv)[o];
            }
        }
    }
    if constexpr (HOIST) {
        if (inj_one_pass && ig1 > ig0) {  // uniform
#pragma unroll
            for (int k = 0; k < NC; k++)
                if (t + NT * k < RN) gmap[t + NT * k] = kNone;
        }
    }
    int r_l = -1, r_e = 0;  // this thread's receiver (sponge factor after the barrier)
    float r_A = 0.0f;
    if (rec_one_pass && rb0 + t < rb1) {
        r_l = C(sh->rec_tcells)[2 * (rb0 + t)];
        r_e = C(sh->rec_tcells)[2 * (rb0 + t) + 1];
    }


    // The launch's step schedule, worked out once by wave 0 (while the loads
    // are in flight) and handed to the others through LDS: the launch runs
    // steps s_lo .. s_hi - 1 of its K (the shot's [it_lo, it_end)).  Bit s of
    // mult: the step's time index (it forward, tb = nt - 1 - it backward) is a
    // multiple of isnap, where the snapshot slot q - 1 - snap_base moves on
    // (up forward, down backward); omask: the forward writes that slot, or the
    // backward images it (slot in range and its isnaps entry equal to tb);
    // cmask: finite-check step (it a multiple of cint, index it / cint).  Per
    // step the slot offset and the check index then follow from bit tests and
    // running counters, with no division and no table read.
    __shared__ int sched[8];
    if (t < 64) {
        const int it_start = it_first + s_lo;                 // >= it_lo >= 0, < it_end <= nt
        const int T = FWD ? it_start : nt - 1 - it_start;  // >= 0
        const int qt = T / isnap, r = T - qt * isnap;
        const int qn = qt + ((FWD && r != 0) ? 1 : 0);  // quotient at the first multiple
        unsigned mult = 0, omask = 0, cmask = 0;
        int sl = qn - 1 - snap_base;
        for (int b = s_lo + (FWD ? (r != 0 ? isnap - r : 0) : r); b < s_hi;
             b += isnap, sl += FWD ? 1 : -1) {
            mult |= 1u << b;
            const bool in = sl >= 0 && sl < num_snaps;
            if (FWD && save && in && it_first + b != 0) omask |= 1u << b;
            if (!FWD && imaging && in) {
                const int tb_ = nt - 1 - (it_first + b);
                if (tb_ > 0 && (sh->snaps_regular || C(sh->isnaps)[sl] == tb_)) omask |= 1u << b;
            }
        }
        int cn = 0;
        if (cint > 0) {
            const int qc = it_start / cint, rc = it_start - qc * cint;
            cn = qc + (rc != 0 ? 1 : 0);
            for (int b = s_lo + (rc != 0 ? cint - rc : 0); b < s_hi; b += cint) cmask |= 1u << b;
        }
        const long long q_off = (long long)(qn - 1 - snap_base) * (long long)sstride;
        if (t == 0) {
            sched[0] = (int)mult;
            sched[1] = (int)omask;
            sched[2] = (int)cmask;
            sched[3] = cn;
            sched[4] = (int)(unsigned)(q_off & 0xffffffffll);
            sched[5] = (int)(q_off >> 32);
        }
    }
    if constexpr (HOIST) {
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const int e = t + NT * k;
            if (e >= RN) continue;
            int q = 0;
            if (PSV) {
                sU[e] = st[k][q++];
                sW[e] = st[k][q++];
                sXX[e] = st[k][q++];
                sZZ[e] = st[k][q++];
                sXZ[e] = st[k][q++];
            }
            if (SH) {
                sV[e] = st[k][q++];
                sYX[e] = st[k][q++];
                sYZ[e] = st[k][q++];
            }
        }
    }
    __syncthreads();
    // the medium's coefficients count as defined here: computations the
    // compiler hoists out of the step loop (2 mu, dt * c, ...) then cannot
    // wait for their loads ahead of the state loads
#pragma unroll
    for (int k = 0; k < NC; k++) {
        if constexpr (PSV) asm volatile("" : "+v"(c0[k]), "+v"(c1[k]), "+v"(c2[k]), "+v"(c3[k]), "+v"(c4[k]));
        if constexpr (SH) asm volatile("" : "+v"(d0[k]), "+v"(d1[k]), "+v"(d2[k]));
    }
    if (r_l >= 0) r_A = sponge_at(ri0 + r_e / RZ, rj0 + r_e % RZ, nx, nz, m.fw, m.tab);
    const unsigned mult = (unsigned)__builtin_amdgcn_readfirstlane(sched[0]);
    const unsigned omask = (unsigned)__builtin_amdgcn_readfirstlane(sched[1]);
    const unsigned cmask = (unsigned)__builtin_amdgcn_readfirstlane(sched[2]);
    int cn = __builtin_amdgcn_readfirstlane(sched[3]);  // index of the next finite check
    long long q_off = (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane(sched[5]) << 32) |
                                  (unsigned)__builtin_amdgcn_readfirstlane(sched[4]));

    const size_t n = (size_t)nx * nz;
    const float sdt = FWD ? m.dt : -m.dt;
    const float idx = m.idx, idz = m.idz, dt = m.dt;
    const bool has_sm = m.sm != nullptr;

    int gk[HOIST ? NC : 1];
#pragma unroll
    for (int k = 0; k < (HOIST ? NC : 1); k++) gk[k] = kNone;
    if constexpr (HOIST) {
        if (inj_one_pass && ig1 > ig0) {  // uniform
#pragma unroll
            for (int m2 = 0; m2 < IPW; m2++)
                if (q_e[m2] >= 0 && ln == 0) gmap[q_e[m2]] = q_code[m2];
            __syncthreads();
#pragma unroll
            for (int k = 0; k < NC; k++)
                if (t + NT * k < RN) gk[k] = gmap[t + NT * k];
        }
    }

    const long long sstep = FWD ? (long long)sstride : -(long long)sstride;

    if constexpr (PIPE) {
        img_next = (omask >> s_lo) & 1u;
        if (img_next && !pre) load_forward((size_t)q_off);
    }

    bool started = false;  // a step ran in this launch: LDS velocities await the sponge
    // the step loop, compiled once per surface-matrix choice and, for the
    // small shapes, once with and once without the per-step list walks (taken
    // when a tile's lists are too long for the per-launch tables), so that the
    // executed loop's code is contiguous
    auto step_loop = [&](auto sm_loop, auto gen_loop) {
        constexpr bool SML = decltype(sm_loop)::value;
        constexpr bool GEN = !HOIST || decltype(gen_loop)::value;
        for (int s = s_lo; s < s_hi; s++) {
            const int it = it_first + s;
            const int tb = nt - 1 - it;  // backward time index
            const unsigned bit = 1u << s;
            // this step's snapshot to write (forward) or correlate (backward) at
            // slot offset s_off, finite-check index (-1: none)
            const size_t s_off = (size_t)q_off;
            if (mult & bit) q_off += sstep;
            const bool snap = FWD && (omask & bit);
            bool img = false;
            if constexpr (PIPE)
                img = img_next;
            else
                img = !FWD && (omask & bit);
            int chk = -1;
            if (cmask & bit) chk = cn++;
            // this step's injected values of the slots' single-entry groups, from
            // the launch's table, added after the velocity update
            constexpr int NG = HOIST ? NC : 1;
            float gvu[NG], gvv[NG], gvw[NG];
            const int col = FWD ? it : tb;
    #pragma unroll
            for (int k = 0; k < NG; k++) {
                gvu[k] = gvv[k] = gvw[k] = 0.0f;
                if (HOIST && gk[k] >= 0) {
                    const int o = gk[k] * K + s;
                    if constexpr (HOIST && PSV) {
                        gvu[k] = tabU[o];
                        gvw[k] = tabW[o];
                    }
                    if constexpr (HOIST && SH) gvv[k] = tabV[o];
                }
            }
            // the injection of slot k into its velocity (after the update, in
            // entry order, the expressions of the injection phase below)
            auto inject_u = [&](int k, float& u, float& w) {
                if (!HOIST || gk[k] == kNone) return;
                if (gk[k] >= 0) {
                    if (FWD) {
                        u = u + (gvu[k] * m.factor) * c0[k];
                        w = w + (gvw[k] * m.factor) * c1[k];
                    } else {
                        u = u + gvu[k];
                        w = w + gvw[k];
                    }
                    return;
                }
                const int g = -1 - gk[k];
                const int q0 = C(sh->start)[g], q1 = C(sh->start)[g + 1];
                for (int q2 = q0; q2 < q1; q2++) {
                    const size_t o = (size_t)C(sh->order)[q2] * (size_t)nt + col;
                    if (FWD) {
                        u = u + (C(sh->su)[o] * m.factor) * c0[k];
                        w = w + (C(sh->sw)[o] * m.factor) * c1[k];
                    } else {
                        u = u + C(sh->su)[o];
                        w = w + C(sh->sw)[o];
                    }
                }
            };
            auto inject_v = [&](int k, float& v) {
                if (!HOIST || gk[k] == kNone) return;
                if (gk[k] >= 0) {
                    v = FWD ? v + (gvv[k] * m.factor) * d0[k] : v + gvv[k];
                    return;
                }
                const int g = -1 - gk[k];
                const int q0 = C(sh->start)[g], q1 = C(sh->start)[g + 1];
                for (int q2 = q0; q2 < q1; q2++) {
                    const size_t o = (size_t)C(sh->order)[q2] * (size_t)nt + col;
                    v = FWD ? v + (C(sh->sv)[o] * m.factor) * d0[k] : v + C(sh->sv)[o];
                }
            };
            // SH imaging: the forward wavefield of every slot is loaded here, a
            // whole phase before the image update uses it (cells outside the grid
            // read cell 0; only the tile's cells use the value)
            float fimg[SH && !PSV ? NC : 1], fimu[PSV && !SH ? NC : 1], fimw[PSV && !SH ? NC : 1];
            if constexpr (PIPE) {
                if (img) {
    #pragma unroll
                    for (int k = 0; k < NF; k++) {
                        if constexpr (PSV) {
                            fimu[k] = ssponge ? pf0[k] * A[k] : pf0[k];
                            fimw[k] = ssponge ? pf1[k] * A[k] : pf1[k];
                        } else {
                            fimg[k] = ssponge ? pf0[k] * A[k] : pf0[k];
                        }
                    }
                }
                // the next step's forward wavefield, in flight during this step
                // (its slot offset is q_off, moved past this step)
                img_next = s + 1 < K && ((omask >> (s + 1)) & 1u);
                if (img_next) load_forward((size_t)q_off);
            } else if constexpr (SH && !PSV) {
                if (img) {
    #pragma unroll
                    for (int k = 0; k < NC; k++) fimg[k] = snap_v[s_off + gc[k]];
                }
            }
            if constexpr (PSV && !SH && !PIPE) {
                if (img) {
    #pragma unroll
                    for (int k = 0; k < NC; k++) {
                        fimu[k] = snap_u[s_off + gc[k]];
                        fimw[k] = snap_w[s_off + gc[k]];
                    }
                }
            }
            // The image's forward wavefield of slot k, times the sponge where
            // the snapshots hold unsponged velocities.  Loads issued in this
            // step are sponged at the use, so that they are waited for a phase
            // after they were issued, not right away.
            auto fwd_img = [&](float f, int k) {
                if constexpr (PIPE) return f;
                else return ssponge ? f * A[k] : f;
            };

            // ---- velocity update (region)
            // SH-only kernels, without a surface matrix branch-free over every
            // slot (the result selected per cell); P-SV keeps per-cell branches
            // (its registers are the tighter limit)
            auto velocity_phase = [&](auto sm_tag) {
                constexpr bool SM = decltype(sm_tag)::value;  // stresses read through the surface matrix
                auto mk = [&](float x, int c) {
                    if constexpr (SM) return x * sm[c];
                    else return x;
                };
    #pragma unroll
                for (int k = 0; k < NC; k++) {
                    if (SM && !(fl[k] & 1u)) continue;
                    const int e = t + NT * k;
                    const int c = gc[k];
                    const bool upd = (fl[k] & 10u) == 10u;
                    float v = sV[e];
                    if (started) v = v * A[k];
                    if (!SM || upd) {
                        float syx_x, syz_z, vn;
                        if (FWD) {
                            syx_x = (sYX[e] - sYX[e - RZ]) * idx;
                            syz_z = (mk(sYZ[e], c) - mk(sYZ[e - 1], c - 1)) * idz;
                            vn = v + ((syx_x + syz_z) * dt) * d0[k];
                        } else {
                            syx_x = (sYX[e + RZ] - sYX[e]) * idx;
                            syz_z = (mk(sYZ[e + 1], c + 1) - mk(sYZ[e], c)) * idz;
                            vn = v + ((-(syx_x + syz_z)) * dt) * d0[k];
                        }
                        v = upd ? vn : v;
                    }
                    inject_v(k, v);
                    sV[e] = v;
                }
            };
            // P-SV: per-cell branches; the surface matrix is a template choice,
            // so the stress reads of the update issue together
            auto velocity_psv = [&](auto sm_tag) {
                constexpr bool SM = decltype(sm_tag)::value;
                auto msk = [&](float x, int c) {
                    if constexpr (SM) return x * sm[c];
                    else return x;
                };
    #pragma unroll
                for (int k = 0; k < NC; k++) {
                    if (!(fl[k] & 1u)) continue;
                    const int e = t + NT * k;
                    const int c = gc[k];
                    const bool upd = (fl[k] & 2u) && (fl[k] & 8u);
                    if (PSV) {
                        float u = sU[e], w = sW[e];
                        if (started) {
                            u = u * A[k];
                            w = w * A[k];
                        }
                        if (upd) {
                            float sxx_x, szz_z, sxz_x, sxz_z;
                            if (FWD) {
                                sxx_x = (sXX[e] - sXX[e - RZ]) * idx;
                                szz_z = (msk(sZZ[e], c) - msk(sZZ[e - 1], c - 1)) * idz;
                                sxz_x = (msk(sXZ[e], c) - msk(sXZ[e - RZ], c - nz)) * idx;
                                sxz_z = (msk(sXZ[e], c) - msk(sXZ[e - 1], c - 1)) * idz;
                                u = u + ((sxx_x + sxz_z) * dt) * c0[k];
                                w = w + ((sxz_x + szz_z) * dt) * c1[k];
                            } else {
                                sxx_x = (sXX[e + RZ] - sXX[e]) * idx;
                                szz_z = (msk(sZZ[e + 1], c + 1) - msk(sZZ[e], c)) * idz;
                                sxz_x = (msk(sXZ[e + RZ], c + nz) - msk(sXZ[e], c)) * idx;
                                sxz_z = (msk(sXZ[e + 1], c + 1) - msk(sXZ[e], c)) * idz;
                                u = u + ((-(sxx_x + sxz_z)) * dt) * c0[k];
                                w = w + ((-(sxz_x + szz_z)) * dt) * c1[k];
                            }
                        }
                        inject_u(k, u, w);
                        sU[e] = u;
                        sW[e] = w;
                    }
                    if (SH) {
                        float v = sV[e];
                        if (started) v = v * A[k];
                        if (upd) {
                            float syx_x, syz_z;
                            if (FWD) {
                                syx_x = (sYX[e] - sYX[e - RZ]) * idx;
                                syz_z = (msk(sYZ[e], c) - msk(sYZ[e - 1], c - 1)) * idz;
                                v = v + ((syx_x + syz_z) * dt) * d0[k];
                            } else {
                                syx_x = (sYX[e + RZ] - sYX[e]) * idx;
                                syz_z = (msk(sYZ[e + 1], c + 1) - msk(sYZ[e], c)) * idz;
                                v = v + ((-(syx_x + syz_z)) * dt) * d0[k];
                            }
                        }
                        inject_v(k, v);
                        sV[e] = v;
                    }
                }
            };
            // P-SV, up to four cells per thread, flat surface: branch-free over
            // every slot
            auto velocity_psv_bf = [&]() {
    #pragma unroll
                for (int k = 0; k < NC; k++) {
                    const int e = t + NT * k;
                    const bool upd = (fl[k] & 10u) == 10u;
                    float u = sU[e], w = sW[e];
                    if (started) {
                        u = u * A[k];
                        w = w * A[k];
                    }
                    float un, wn;
                    if (FWD) {
                        const float sxx_x = (sXX[e] - sXX[e - RZ]) * idx;
                        const float szz_z = (sZZ[e] - sZZ[e - 1]) * idz;
                        const float sxz_x = (sXZ[e] - sXZ[e - RZ]) * idx;
                        const float sxz_z = (sXZ[e] - sXZ[e - 1]) * idz;
                        un = u + ((sxx_x + sxz_z) * dt) * c0[k];
                        wn = w + ((sxz_x + szz_z) * dt) * c1[k];
                    } else {
                        const float sxx_x = (sXX[e + RZ] - sXX[e]) * idx;
                        const float szz_z = (sZZ[e + 1] - sZZ[e]) * idz;
                        const float sxz_x = (sXZ[e + RZ] - sXZ[e]) * idx;
                        const float sxz_z = (sXZ[e + 1] - sXZ[e]) * idz;
                        un = u + ((-(sxx_x + sxz_z)) * dt) * c0[k];
                        wn = w + ((-(sxz_x + szz_z)) * dt) * c1[k];
                    }
                    u = upd ? un : u;
                    w = upd ? wn : w;
                    inject_u(k, u, w);
                    sU[e] = u;
                    sW[e] = w;
                }
            };
            if constexpr (BFP) {
                if constexpr (SML)
                    velocity_psv(std::true_type{});
                else
                    velocity_psv_bf();
            } else if constexpr (PSV) {
                velocity_psv(std::integral_constant<bool, SML>{});
            } else {
                velocity_phase(std::integral_constant<bool, SML>{});
            }
            __syncthreads();

            // ---- injection into the region (one thread per cell, entries in order;
            // the small shapes inject in the velocity update above)
            if (GEN && ig1 > ig0 && !inj_one_pass) {
                const gptr<const float> su = G(sh->su), sv = G(sh->sv), sw = G(sh->sw);
                const size_t ntt = (size_t)nt;
                {
                    const gptr<const int> titems = G(sh->inj_titems), gstart = G(sh->start);
                    const gptr<const int> order = G(sh->order), loc = G(sh->loc);
                    for (int q = ig0 + t; q < ig1; q += NT) {
                        const int g = titems[q];
                        const int e0 = gstart[g], e1 = gstart[g + 1];
                        const int k0 = order[e0];
                        const int ci = loc[2 * k0], cj = loc[2 * k0 + 1];
                        const int e = (ci - ri0) * RZ + (cj - rj0);
                        const int c = ci * nz + cj;
                        if (PSV) {
                            float x = sU[e], zz = sW[e];
                            for (int q2 = e0; q2 < e1; q2++) {
                                const size_t o = (size_t)order[q2] * ntt + col;
                                if (FWD) {
                                    x = x + (su[o] * m.factor) * G(m.iru)[c];
                                    zz = zz + (sw[o] * m.factor) * G(m.irw)[c];
                                } else {
                                    x = x + su[o];
                                    zz = zz + sw[o];
                                }
                            }
                            sU[e] = x;
                            sW[e] = zz;
                        }
                        if (SH) {
                            float y = sV[e];
                            for (int q2 = e0; q2 < e1; q2++) {
                                const size_t o = (size_t)order[q2] * ntt + col;
                                y = FWD ? y + (sv[o] * m.factor) * G(m.ir)[c] : y + sv[o];
                            }
                            sV[e] = y;
                        }
                    }
                }
                __syncthreads();
            }

            // ---- stress update + sponge (region), outputs (tile)
            int code = 0;
            auto stress_phase = [&](auto sm_tag) {
                constexpr bool SM = decltype(sm_tag)::value;
                auto mk = [&](float x, int c) {
                    if constexpr (SM) return x * sm[c];
                    else return x;
                };
    #pragma unroll
                for (int k = 0; k < NC; k++) {
                    if (SM && !(fl[k] & 1u)) continue;
                    const int e = t + NT * k;
                    const int c = gc[k];
                    const bool upd = (fl[k] & 18u) == 18u;
                    float syx = sYX[e], syz = mk(sYZ[e], c);
                    if (!SM || upd) {
                        float v_x, v_z;
                        if (FWD) {
                            v_x = (sV[e + RZ] - sV[e]) * idx;
                            v_z = (sV[e + 1] - sV[e]) * idz;
                        } else {
                            v_x = (sV[e] - sV[e - RZ]) * idx;
                            v_z = (sV[e] - sV[e - 1]) * idz;
                        }
                        const float yxn = syx + (sdt * d1[k]) * v_x;
                        const float yzn = syz + (sdt * d2[k]) * v_z;
                        syx = upd ? yxn : syx;
                        syz = upd ? yzn : syz;
                    }
                    sYX[e] = syx * A[k];
                    sYZ[e] = syz * A[k];
                    if (fl[k] & 4u) {  // the tile's cell: outputs
                        const float v = sV[e] * A[k];
                        if (snap) snap_v[s_off + c] = v;
                        if (img) aV[k] = aV[k] + fwd_img(fimg[k], k) * v;
                        if (chk >= 0 && not_finite_or_big(v) && code < 2) code = 2;
                    }
                }
            };
            auto stress_psv = [&](auto sm_tag) {
                constexpr bool SM = decltype(sm_tag)::value;
                auto msk = [&](float x, int c) {
                    if constexpr (SM) return x * sm[c];
                    else return x;
                };
    #pragma unroll
                for (int k = 0; k < NC; k++) {
                    if (!(fl[k] & 1u)) continue;
                    const int e = t + NT * k;
                    const int c = gc[k];
                    const bool upd = (fl[k] & 2u) && (fl[k] & 16u);
                    const bool own = fl[k] & 4u;
                    if (PSV) {
                        float sxx = sXX[e], szz = msk(sZZ[e], c), sxz = msk(sXZ[e], c);
                        if (upd) {
                            float u_x, u_z, w_x, w_z;
                            if (FWD) {
                                u_x = (sU[e + RZ] - sU[e]) * idx;
                                u_z = (sU[e + 1] - sU[e]) * idz;
                                w_x = (sW[e + RZ] - sW[e]) * idx;
                                w_z = (sW[e + 1] - sW[e]) * idz;
                            } else {
                                u_x = (sU[e] - sU[e - RZ]) * idx;
                                u_z = (sU[e] - sU[e - 1]) * idz;
                                w_x = (sW[e] - sW[e - RZ]) * idx;
                                w_z = (sW[e] - sW[e - 1]) * idz;
                            }
                            const float div_uw = u_x + w_z;
                            sxx = sxx + sdt * (c2[k] * div_uw + (2.0f * c3[k]) * u_x);
                            szz = szz + sdt * (c2[k] * div_uw + (2.0f * c3[k]) * w_z);
                            sxz = sxz + (sdt * c4[k]) * (u_z + w_x);
                        }
                        sXX[e] = sxx * A[k];
                        sZZ[e] = szz * A[k];
                        sXZ[e] = sxz * A[k];
                        if (own) {
                            const float u = sU[e] * A[k], w = sW[e] * A[k];
                            if (snap) {
                                snap_u[s_off + c] = u;
                                snap_w[s_off + c] = w;
                            }
                            if (img) {
                                if constexpr (!SH) {
                                    aU[k] = aU[k] + fwd_img(fimu[k], k) * u;
                                    aW[k] = aW[k] + fwd_img(fimw[k], k) * w;
                                } else {
                                    float fu = snap_u[s_off + c];
                                    float fw = snap_w[s_off + c];
                                    if (ssponge) {
                                        fu = fu * A[k];
                                        fw = fw * A[k];
                                    }
                                    aU[k] = aU[k] + fu * u;
                                    aW[k] = aW[k] + fw * w;
                                }
                            }
                            if (chk >= 0) {
                                if (not_finite_or_big(u) && code < 1) code = 1;
                                if (not_finite_or_big(w)) code = 3;
                            }
                        }
                    }
                    if (SH) {
                        float syx = sYX[e], syz = msk(sYZ[e], c);
                        if (upd) {
                            float v_x, v_z;
                            if (FWD) {
                                v_x = (sV[e + RZ] - sV[e]) * idx;
                                v_z = (sV[e + 1] - sV[e]) * idz;
                            } else {
                                v_x = (sV[e] - sV[e - RZ]) * idx;
                                v_z = (sV[e] - sV[e - 1]) * idz;
                            }
                            syx = syx + (sdt * d1[k]) * v_x;
                            syz = syz + (sdt * d2[k]) * v_z;
                        }
                        sYX[e] = syx * A[k];
                        sYZ[e] = syz * A[k];
                        if (own) {
                            const float v = sV[e] * A[k];
                            if (snap) snap_v[s_off + c] = v;
                            if (img) {
                                float fv = snap_v[s_off + c];
                                if (ssponge) fv = fv * A[k];
                                aV[k] = aV[k] + fv * v;
                            }
                            if (chk >= 0 && not_finite_or_big(v) && code < 2) code = 2;
                        }
                    }
                }
            };
            // P-SV, up to four cells per thread, flat surface: branch-free over
            // every slot; the outputs (snapshots, image, finite check) only for the
            // tile's cells
            auto stress_psv_bf = [&]() {
    #pragma unroll
                for (int k = 0; k < NC; k++) {
                    const int e = t + NT * k;
                    const int c = gc[k];
                    const bool upd = (fl[k] & 18u) == 18u;
                    float sxx = sXX[e], szz = sZZ[e], sxz = sXZ[e];
                    float u_x, u_z, w_x, w_z;
                    if (FWD) {
                        u_x = (sU[e + RZ] - sU[e]) * idx;
                        u_z = (sU[e + 1] - sU[e]) * idz;
                        w_x = (sW[e + RZ] - sW[e]) * idx;
                        w_z = (sW[e + 1] - sW[e]) * idz;
                    } else {
                        u_x = (sU[e] - sU[e - RZ]) * idx;
                        u_z = (sU[e] - sU[e - 1]) * idz;
                        w_x = (sW[e] - sW[e - RZ]) * idx;
                        w_z = (sW[e] - sW[e - 1]) * idz;
                    }
                    const float div_uw = u_x + w_z;
                    const float xxn = sxx + sdt * (c2[k] * div_uw + (2.0f * c3[k]) * u_x);
                    const float zzn = szz + sdt * (c2[k] * div_uw + (2.0f * c3[k]) * w_z);
                    const float xzn = sxz + (sdt * c4[k]) * (u_z + w_x);
                    sxx = upd ? xxn : sxx;
                    szz = upd ? zzn : szz;
                    sxz = upd ? xzn : sxz;
                    sXX[e] = sxx * A[k];
                    sZZ[e] = szz * A[k];
                    sXZ[e] = sxz * A[k];
                    const float u = sU[e] * A[k], w = sW[e] * A[k];
                    if constexpr (PIPE) {  // the image of a slot outside the tile is never written
                        if (img) {
                            aU[k] = aU[k] + fwd_img(fimu[k], k) * u;
                            aW[k] = aW[k] + fwd_img(fimw[k], k) * w;
                        }
                    }
                    if (fl[k] & 4u) {
                        if constexpr (!PIPE) {
                            if (img) {
                                aU[k] = aU[k] + fwd_img(fimu[k], k) * u;
                                aW[k] = aW[k] + fwd_img(fimw[k], k) * w;
                            }
                        }
                        if (snap) {
                            snap_u[s_off + c] = u;
                            snap_w[s_off + c] = w;
                        }
                        if (chk >= 0) {
                            if (not_finite_or_big(u) && code < 1) code = 1;
                            if (not_finite_or_big(w)) code = 3;
                        }
                    }
                }
            };
            if constexpr (BFP) {
                if constexpr (SML)
                    stress_psv(std::true_type{});
                else
                    stress_psv_bf();
            } else if constexpr (PSV) {
                stress_psv(std::integral_constant<bool, SML>{});
            } else {
                stress_phase(std::integral_constant<bool, SML>{});
            }
            if (code) atomicMax(sh->check_flags + chk, code + (FWD ? 0 : 3));

            // ---- receivers of the tile, after the sponge
            if (rb1 > rb0) {
                const gptr<float> ru = G(sh->rec_u), rv = G(sh->rec_v), rw = G(sh->rec_w);
                const size_t row = (size_t)(FWD ? it : tb) * n_rec;
                if (!GEN || rec_one_pass) {
                    if (r_l >= 0) {
                        if (PSV) {
                            ru[row + r_l] = sU[r_e] * r_A;
                            rw[row + r_l] = sW[r_e] * r_A;
                        }
                        if (SH) rv[row + r_l] = sV[r_e] * r_A;
                    }
                } else {
                    const int b = rb0, end = rb1;
                    const gptr<const int> ritems = G(sh->rec_titems), rloc = G(sh->rec_loc);
                    for (int q = b + t; q < end; q += NT) {
                        const int l = ritems[q];
                        const int ci = rloc[2 * l], cj = rloc[2 * l + 1];
                        const int e = (ci - ri0) * RZ + (cj - rj0);
                        const float Ac = sponge_at(ci, cj, nx, nz, m.fw, m.tab);
                        if (PSV) {
                            ru[row + l] = sU[e] * Ac;
                            rw[row + l] = sW[e] * Ac;
                        }
                        if (SH) rv[row + l] = sV[e] * Ac;
                    }
                }
            }
            started = true;
            __syncthreads();
        }
    };
    auto run_loop = [&](auto sm_loop) {
        if constexpr (HOIST) {
            if (!inj_one_pass || !rec_one_pass)
                step_loop(sm_loop, std::true_type{});
            else
                step_loop(sm_loop, std::false_type{});
        } else {
            step_loop(sm_loop, std::true_type{});
        }
    };
    if (has_sm)
        run_loop(std::true_type{});
    else
        run_loop(std::false_type{});

    // ---- write the tile's state (post-sponge velocities) and image
    const gptr<float> o_u = G(sh->out[0]), o_v = G(sh->out[1]), o_w = G(sh->out[2]);
    const gptr<float> o_xx = G(sh->out[3]), o_zz = G(sh->out[4]), o_xz = G(sh->out[5]);
    const gptr<float> o_yx = G(sh->out[6]), o_yz = G(sh->out[7]);
#pragma unroll
    for (int k = 0; k < NC; k++) {
        if ((fl[k] & 5u) != 5u) continue;
        const int e = t + NT * k;
        const int c = gc[k];
        if (PSV) {
            o_u[c] = sU[e] * A[k];
            o_w[c] = sW[e] * A[k];
            o_xx[c] = sXX[e];
            o_zz[c] = sZZ[e];
            o_xz[c] = sXZ[e];
            if (imaging) {
                img_u[c] = aU[k];
                img_w[c] = aW[k];
            }
        }
        if (SH) {
            o_v[c] = sV[e] * A[k];
            o_yx[c] = sYX[e];
            o_yz[c] = sYZ[e];
            if (imaging) img_v[c] = aV[k];
        }
    }
}

// ---------------------------------------------------------------------------
// host side

// configurations compiled per wave system
static const BlockCfg kCfgSH[] = {{16, 64, 256, 4},   {16, 64, 256, 8},    {16, 128, 512, 4},
                                  {16, 128, 512, 8},  {32, 128, 1024, 4},  {32, 128, 1024, 8},
                                  {24, 128, 1024, 8}, {32, 96, 1024, 8},   {48, 64, 1024, 8},
                                  {32, 48, 512, 8},   {24, 64, 512, 8},    {16, 32, 1024, 4},
                                  {16, 32, 1024, 8},  {16, 16, 1024, 8}};
static const BlockCfg kCfgPSV[] = {{16, 64, 256, 2},  {16, 64, 256, 4},    {16, 128, 512, 4},
                                   {16, 128, 512, 8}, {32, 128, 1024, 4},  {48, 64, 1024, 4},
                                   {32, 64, 512, 4},  {24, 56, 512, 4},    {16, 32, 1024, 4},
                                   {16, 32, 1024, 8}, {16, 16, 1024, 8}};
static const BlockCfg kCfgBoth[] = {{16, 64, 256, 2}, {16, 64, 256, 4}, {16, 128, 512, 4},
                                    {16, 32, 1024, 4}};
// (the 16 x 32 and 16 x 16 tiles with 1024 threads give every thread one or
// two region cells: the shortest step on grids too small to fill the GPU; 16
// x 16 with 576 threads and K = 4, 8 x 16 with 384 threads and K = 4, and 16 x
// 16 with K = 8 and 512 or 256 threads (two or four cells per thread) were
// measured slower on the 200 x 200 grid; 24 x 56 with 512 threads holds four
// cells per thread, branch-free, two workgroups per CU: the batch and salt
// jobs' choice; 32 x 64 with 1024 threads and K = 8 (four cells, one
// workgroup per CU) was slower there, and the tuner kept 24 x 56 over 28 x 48
// and 32 x 40 with 512 threads, whose regions hold slightly fewer halo cells
// per tile cell but shorter columns)

static void cfg_list(int systems, const BlockCfg** list, int* n) {
    if (systems == TRTM_SH) {
        *list = kCfgSH;
        *n = sizeof(kCfgSH) / sizeof(BlockCfg);
    } else if (systems == TRTM_PSV) {
        *list = kCfgPSV;
        *n = sizeof(kCfgPSV) / sizeof(BlockCfg);
    } else {
        *list = kCfgBoth;
        *n = sizeof(kCfgBoth) / sizeof(BlockCfg);
    }
}

// defaults: the fastest shapes measured on MI355X for the 1200 x 600 SH and
// 200 x 200 / 2000 x 1000 P-SV workloads
static BlockCfg default_cfg(int systems) {
    if (systems == TRTM_SH) return {32, 48, 512, 8};
    if (systems == TRTM_PSV) return {32, 128, 1024, 4};
    return {16, 128, 512, 4};
}

static BlockCfg default_wave_cfg(int systems) {
    if (systems == TRTM_SH) return {16, 56, 64, 4};
    if (systems == TRTM_PSV) return {16, 56, 64, 4};
    return {8, 56, 64, 4};
}

// first compiled configuration of the requested kernel family matching every
// nonzero request field
int block_choose(const trtm_step_desc* d, BlockCfg* out) {
    const bool wave = d->kernel == TRTM_KERNEL_WAVE;
    if (d->tile_x <= 0 && d->tile_z <= 0 && d->block_steps <= 0) {
        *out = wave ? default_wave_cfg(d->systems) : default_cfg(d->systems);
        return TRTM_OK;
    }
    const BlockCfg* list;
    int n;
    if (wave)
        wave_configs(d->systems, &list, &n);
    else
        cfg_list(d->systems, &list, &n);
    for (int i = 0; i < n; i++) {
        const BlockCfg& c = list[i];
        if ((d->tile_x <= 0 || c.tx == d->tile_x) && (d->tile_z <= 0 || c.tz == d->tile_z) &&
            (d->block_steps <= 0 || c.k == d->block_steps)) {
            *out = c;
            return TRTM_OK;
        }
    }
    return fail(TRTM_ERR_UNSUPPORTED,
                "%s kernel has no tile_x=%d tile_z=%d block_steps=%d build for systems=%d",
                wave ? "wave" : "blocked", d->tile_x, d->tile_z, d->block_steps, d->systems);
}

static size_t align256b(size_t x) { return (x + 255) & ~size_t(255); }

// workspace of one shot: second state buffer, tile lists sized for every
// compiled configuration of both kernel families, device copy of isnaps, two
// shot-table entries
struct BlockLayout {
    size_t alt, inj_start, inj_cursor, inj_items, inj_cells, rec_start, rec_cursor, rec_items,
        rec_cells, maps, isnaps,
        shot, total;
    int tiles_x, tiles_z;
};

static BlockLayout block_layout(const trtm_step_desc* d, BlockCfg c) {
    BlockLayout L;
    const size_t n = (size_t)d->nx * d->nz;
    L.tiles_x = (d->nx + c.tx - 1) / c.tx;
    L.tiles_z = (d->nz + c.tz - 1) / c.tz;
    // LDS-tiled kernel: lists per tile, sized for the compiled shape with the
    // most tiles on this grid; a cell lies in at most 2 x 2 regions when
    // 2K <= min(TX, TZ)
    size_t tiles_max = 0;
    for (const BlockCfg* lst : {kCfgSH, kCfgPSV, kCfgBoth}) {
        const int cnt = lst == kCfgSH ? (int)(sizeof(kCfgSH) / sizeof(BlockCfg))
                        : lst == kCfgPSV ? (int)(sizeof(kCfgPSV) / sizeof(BlockCfg))
                                         : (int)(sizeof(kCfgBoth) / sizeof(BlockCfg));
        for (int k = 0; k < cnt; k++)
            tiles_max = std::max(tiles_max, (size_t)((d->nx + lst[k].tx - 1) / lst[k].tx) *
                                                ((d->nz + lst[k].tz - 1) / lst[k].tz));
    }
    size_t ib = tiles_max + 1, rb = tiles_max + 1;
    size_t ii = 4 * (size_t)std::max(d->n_inj_groups, 0) * sizeof(int);
    size_t ri = (size_t)std::max(d->n_rec, 0) * sizeof(int);
    size_t mb = 0;
    // wavefront kernel: lists per tile row, per-lane maps
    const BlockCfg* list;
    int nw;
    wave_configs(d->systems, &list, &nw);
    for (int k = 0; k < nw; k++) {
        const ListSizes w = wave_list_sizes(d, list[k]);
        ib = std::max(ib, w.inj_buckets);
        rb = std::max(rb, w.rec_buckets);
        ii = std::max(ii, w.inj_item_bytes);
        ri = std::max(ri, w.rec_item_bytes);
        mb = std::max(mb, w.map_bytes);
    }
    size_t off = 0;
    L.alt = off;        off = align256b(off + 8 * n * sizeof(float));
    L.inj_start = off;  off = align256b(off + ib * sizeof(int));
    L.inj_cursor = off; off = align256b(off + ib * sizeof(int));
    L.inj_items = off;  off = align256b(off + ii);
    L.inj_cells = off;  off = align256b(off + 2 * ii);
    L.rec_start = off;  off = align256b(off + rb * sizeof(int));
    L.rec_cursor = off; off = align256b(off + rb * sizeof(int));
    L.rec_items = off;  off = align256b(off + ri);
    L.rec_cells = off;  off = align256b(off + 2 * ri);
    L.maps = off;       off = align256b(off + mb);
    L.isnaps = off;     off = align256b(off + (size_t)std::max(d->num_snaps, 0) * sizeof(int));
    L.shot = off;       off = align256b(off + 2 * sizeof(ShotDev));
    L.total = off;
    return L;
}

int64_t block_workspace_bytes(const trtm_step_desc* d) {
    return (int64_t)block_layout(d, BlockCfg{16, 64, 256, 1}).total;
}

// mode 0 counts, mode 1 fills.  Groups go to every tile whose region (tile +
// halo K) holds their cell; receivers to the tile that holds their cell.
// Mode 1 also writes per placed item its region cell e (row-major in the
// tile's region, RZ = tz + 2K per row) with, for groups, the data row of a
// single entry or -1 - g (icells), for receivers the receiver index (rcells).
__global__ void k_block_lists(int mode, int tx, int tz, int K, int tiles_x, int tiles_z,
                              int n_groups, const int* loc, const int* order, const int* start,
                              int* icount, const int* istart, int* iitems, int* icells, int n_rec,
                              const int* rloc, int* rcount, const int* rstart, int* ritems,
                              int* rcells) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n_groups) {
        const int k0 = order[start[q]];
        const int ci = loc[2 * k0], cj = loc[2 * k0 + 1];
        for (int ti = (ci - K - tx + 1) / tx - 1; ti <= (ci + K) / tx + 1; ti++)
            for (int tk = (cj - K - tz + 1) / tz - 1; tk <= (cj + K) / tz + 1; tk++) {
                if (ti < 0 || ti >= tiles_x || tk < 0 || tk >= tiles_z) continue;
                const int ri0 = ti * tx - K, rj0 = tk * tz - K;
                if (ci < ri0 || ci >= ri0 + tx + 2 * K || cj < rj0 || cj >= rj0 + tz + 2 * K)
                    continue;
                const int tile = ti * tiles_z + tk;
                const int pos = atomicAdd(icount + tile, 1);
                if (mode == 1) {
                    const int w = istart[tile] + pos;
                    iitems[w] = q;
                    icells[2 * w] = (ci - ri0) * (tz + 2 * K) + (cj - rj0);
                    icells[2 * w + 1] = start[q + 1] - start[q] == 1 ? k0 : -1 - q;
                }
            }
    }
    if (q < n_rec) {
        const int ti = rloc[2 * q] / tx, tk = rloc[2 * q + 1] / tz;
        const int tile = ti * tiles_z + tk;
        const int pos = atomicAdd(rcount + tile, 1);
        if (mode == 1) {
            const int w = rstart[tile] + pos;
            ritems[w] = q;
            rcells[2 * w] = q;
            rcells[2 * w + 1] = (rloc[2 * q] - (ti * tx - K)) * (tz + 2 * K) + (rloc[2 * q + 1] - (tk * tz - K));
        }
    }
}

// The same lists in one workgroup per shot, with counts, cursors and scans in
// LDS (dynamic, 4 * tiles + 2 ints): one launch instead of eight.  Only the
// list outputs go to global memory.
constexpr int kListsOneMaxTiles = 3800;  // LDS stays under 64 KB

__global__ __launch_bounds__(1024) void k_block_lists_one(int tx, int tz, int K, int tiles_x,
                                                          int tiles_z, int n_groups,
                                                          const int* loc, const int* order,
                                                          const int* start, int* istart_g,
                                                          int* iitems, int* icells, int n_rec,
                                                          const int* rloc, int* rstart_g,
                                                          int* ritems, int* rcells) {
    extern __shared__ int lds[];
    __shared__ int part[1024];
    const int t = threadIdx.x, tiles = tiles_x * tiles_z;
    int* icnt = lds;
    int* rcnt = lds + tiles;
    int* ist = lds + 2 * tiles;          // tiles + 1 entries
    int* rst = lds + 3 * tiles + 1;      // tiles + 1 entries
    auto zero = [&]() {
        for (int i = t; i < tiles; i += 1024) icnt[i] = rcnt[i] = 0;
        __syncthreads();
    };
    // mode 0 counts, mode 1 places (same tile rule as k_block_lists)
    auto pass = [&](int mode) {
        const int m = n_groups > n_rec ? n_groups : n_rec;
        for (int q = t; q < m; q += 1024) {
            if (q < n_groups) {
                const int k0 = order[start[q]];
                const int ci = loc[2 * k0], cj = loc[2 * k0 + 1];
                for (int ti = (ci - K - tx + 1) / tx - 1; ti <= (ci + K) / tx + 1; ti++)
                    for (int tk = (cj - K - tz + 1) / tz - 1; tk <= (cj + K) / tz + 1; tk++) {
                        if (ti < 0 || ti >= tiles_x || tk < 0 || tk >= tiles_z) continue;
                        const int ri0 = ti * tx - K, rj0 = tk * tz - K;
                        if (ci < ri0 || ci >= ri0 + tx + 2 * K || cj < rj0 || cj >= rj0 + tz + 2 * K)
                            continue;
                        const int tile = ti * tiles_z + tk;
                        const int pos = atomicAdd(icnt + tile, 1);
                        if (mode == 1) {
                            const int w = ist[tile] + pos;
                            iitems[w] = q;
                            icells[2 * w] = (ci - ri0) * (tz + 2 * K) + (cj - rj0);
                            icells[2 * w + 1] = start[q + 1] - start[q] == 1 ? k0 : -1 - q;
                        }
                    }
            }
            if (q < n_rec) {
                const int ti = rloc[2 * q] / tx, tk = rloc[2 * q + 1] / tz;
                const int tile = ti * tiles_z + tk;
                const int pos = atomicAdd(rcnt + tile, 1);
                if (mode == 1) {
                    const int w = rst[tile] + pos;
                    ritems[w] = q;
                    rcells[2 * w] = q;
                    rcells[2 * w + 1] = (rloc[2 * q] - (ti * tx - K)) * (tz + 2 * K) +
                                        (rloc[2 * q + 1] - (tk * tz - K));
                }
            }
        }
        __syncthreads();
    };
    // exclusive scan of cnt into st[0..tiles] (LDS) and out[0..tiles] (global)
    auto scan = [&](const int* cnt, int* st, int* out) {
        const int chunk = (tiles + 1023) / 1024;
        const int b = t * chunk, e = min(tiles, b + chunk);
        int sum = 0;
        for (int k = b; k < e; k++) sum += cnt[k];
        part[t] = sum;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int v = t >= off ? part[t - off] : 0;
            __syncthreads();
            part[t] += v;
            __syncthreads();
        }
        int run = t > 0 ? part[t - 1] : 0;
        for (int k = b; k < e; k++) {
            st[k] = run;
            out[k] = run;
            run += cnt[k];
        }
        if (t == 1023) st[tiles] = out[tiles] = part[1023];
        __syncthreads();
    };
    zero();
    pass(0);
    scan(icnt, ist, istart_g);
    scan(rcnt, rst, rstart_g);
    zero();
    pass(1);
}

// Host tables to device memory without waiting: each launch carries up to
// 3.5 KB of the table in its kernel arguments (copied at launch), so the host
// buffer may be released as soon as the launches are issued.
struct UploadChunk {
    unsigned int w[896];
};

__global__ void k_upload(UploadChunk p, int nwords, unsigned int* dst) {
    for (int i = threadIdx.x; i < nwords; i += blockDim.x) dst[i] = p.w[i];
}

static hipError_t upload_async(void* dst, const void* src, size_t bytes, hipStream_t s) {
    const size_t words = bytes / 4;  // tables are whole 32-bit words
    const unsigned int* from = static_cast<const unsigned int*>(src);
    for (size_t off = 0; off < words; off += 896) {
        UploadChunk c;
        const size_t n = std::min<size_t>(896, words - off);
        std::memcpy(c.w, from + off, n * 4);
        hipLaunchKernelGGL(k_upload, dim3(1), dim3(256), 0, s, c, (int)n,
                           static_cast<unsigned int*>(dst) + off);
    }
    return hipGetLastError();
}

static int build_shot(const trtm_step_desc* d, bool fwd, bool save, BlockCfg cfg, int it_lo,
                      int it_hi, hipStream_t s, ShotDev* sd) {
    const int tx = cfg.tx, K = cfg.k;
    const BlockLayout L = block_layout(d, cfg);
    char* ws = static_cast<char*>(d->workspace);
    const int tiles = L.tiles_x * L.tiles_z;
    int* istart = reinterpret_cast<int*>(ws + L.inj_start);
    int* icur = reinterpret_cast<int*>(ws + L.inj_cursor);
    int* iitems = reinterpret_cast<int*>(ws + L.inj_items);
    int* rstart = reinterpret_cast<int*>(ws + L.rec_start);
    int* rcur = reinterpret_cast<int*>(ws + L.rec_cursor);
    int* ritems = reinterpret_cast<int*>(ws + L.rec_items);
    int* icells = reinterpret_cast<int*>(ws + L.inj_cells);
    int* rcells = reinterpret_cast<int*>(ws + L.rec_cells);
    const int ng = d->n_inj_groups;
    const int nr = (d->rec_u || d->rec_v) ? d->n_rec : 0;
    hipError_t e;
    if (is_wave(cfg)) {
        const int rc = wave_build_lists(d, cfg, istart, icur, ws + L.inj_items, rstart, rcur,
                                        ws + L.rec_items, ws + L.maps, s);
        if (rc) return rc;
    } else if (tiles <= kListsOneMaxTiles) {
        const size_t lds = (4 * (size_t)tiles + 2) * sizeof(int);
        hipLaunchKernelGGL(k_block_lists_one, dim3(1), dim3(1024), lds, s, tx, cfg.tz, K,
                           L.tiles_x, L.tiles_z, ng, d->inj_loc, d->inj_order,
                           d->inj_group_start, istart, iitems, icells, nr, d->rec_loc, rstart,
                           ritems, rcells);
        e = hipGetLastError();
        if (e != hipSuccess)
            return fail(TRTM_ERR_HIP, "blocked tile lists: %s", hipGetErrorString(e));
    } else {
        const dim3 tb(256), tg((tiles + 255) / 256), lg((std::max(ng, nr) + 255) / 256);
        hipLaunchKernelGGL(k_fill_int_b, tg, tb, 0, s, icur, tiles, 0);
        hipLaunchKernelGGL(k_fill_int_b, tg, tb, 0, s, rcur, tiles, 0);
        if (ng > 0 || nr > 0)
            hipLaunchKernelGGL(k_block_lists, lg, tb, 0, s, 0, tx, cfg.tz, K, L.tiles_x, L.tiles_z,
                               ng, d->inj_loc, d->inj_order, d->inj_group_start, icur,
                               (const int*)nullptr, (int*)nullptr, (int*)nullptr, nr, d->rec_loc,
                               rcur, (const int*)nullptr, (int*)nullptr, (int*)nullptr);
        hipLaunchKernelGGL(k_scan_b, dim3(1), dim3(1024), 0, s, icur, istart, tiles);
        hipLaunchKernelGGL(k_scan_b, dim3(1), dim3(1024), 0, s, rcur, rstart, tiles);
        hipLaunchKernelGGL(k_fill_int_b, tg, tb, 0, s, icur, tiles, 0);
        hipLaunchKernelGGL(k_fill_int_b, tg, tb, 0, s, rcur, tiles, 0);
        if (ng > 0 || nr > 0)
            hipLaunchKernelGGL(k_block_lists, lg, tb, 0, s, 1, tx, cfg.tz, K, L.tiles_x,
                               L.tiles_z, ng, d->inj_loc, d->inj_order, d->inj_group_start, icur,
                               (const int*)istart, iitems, icells, nr, d->rec_loc, rcur,
                               (const int*)rstart, ritems, rcells);
        e = hipGetLastError();
        if (e != hipSuccess)
            return fail(TRTM_ERR_HIP, "blocked tile lists: %s", hipGetErrorString(e));
    }

    ShotDev& x = *sd;
    float* state[8] = {d->u, d->v, d->w, d->sxx, d->szz, d->sxz, d->syx, d->syz};
    float* alt = reinterpret_cast<float*>(ws + L.alt);
    const size_t n = (size_t)d->nx * d->nz;
    for (int k = 0; k < 8; k++) {
        x.in[k] = state[k];
        x.out[k] = alt + k * n;
    }
    x.nt = d->nt;
    x.it_end = it_hi;
    x.it_lo = it_lo;
    x.it_off = d->step_offset;
    x.save = save ? 1 : 0;
    x.isnap = d->isnap > 0 ? d->isnap : 1;
    x.num_snaps = (save || !fwd) ? d->num_snaps : 0;
    if (!fwd && (!d->isnaps_host || d->num_snaps <= 0)) x.num_snaps = 0;
    x.snap_base = d->snap_base;
    x.check_interval = d->check_interval;
    x.n_groups = ng;
    x.n_rec = nr;
    x.snaps_regular = 0;
    x.loc = d->inj_loc;
    x.order = d->inj_order;
    x.start = d->inj_group_start;
    x.su = d->inj_u;
    x.sv = d->inj_v;
    x.sw = d->inj_w;
    x.rec_loc = d->rec_loc;
    x.rec_u = d->rec_u;
    x.rec_v = d->rec_v;
    x.rec_w = d->rec_w;
    x.inj_tstart = istart;
    x.inj_titems = iitems;
    x.rec_tstart = rstart;
    x.rec_titems = ritems;
    x.inj_tcells = icells;
    x.rec_tcells = rcells;
    x.imap_slot = x.imap = x.rmap_slot = x.rmap = nullptr;
    if (is_wave(cfg)) {
        const WaveMaps wm = wave_maps(d, cfg, ws + L.maps);
        x.imap_slot = wm.imap_slot;
        x.imap = wm.imap;
        x.rmap_slot = wm.rmap_slot;
        x.rmap = wm.rmap;
    }
    x.snap_stride = d->snap_stride > 0 ? d->snap_stride : (long long)d->nx * d->nz;
    x.snap_sponge = d->snap_sponge;
    x.snap_u = d->snap_u;
    x.snap_v = d->snap_v;
    x.snap_w = d->snap_w;
    x.isnaps = nullptr;
    if (!fwd && x.num_snaps > 0) {
        int* dev = reinterpret_cast<int*>(ws + L.isnaps);
        e = upload_async(dev, d->isnaps_host, sizeof(int) * d->num_snaps, s);
        if (e != hipSuccess) return fail(TRTM_ERR_HIP, "isnaps copy: %s", hipGetErrorString(e));
        x.isnaps = dev;
        // the slots of the forward's own schedule: the kernel then needs no
        // isnaps reads to find the imaging steps
        bool regular = true;
        for (int k = 0; k < x.num_snaps && regular; k++)
            regular = (long long)d->isnaps_host[k] == (long long)(k + 1 + x.snap_base) * x.isnap;
        x.snaps_regular = regular ? 1 : 0;
    }
    x.image_u = d->image_u;
    x.image_v = d->image_v;
    x.image_w = d->image_w;
    x.check_flags = d->check_flags;
    return TRTM_OK;
}

template <bool FWD, bool PSV, bool SH, int TX, int TZ, int NT, int K>
static void launch_cfg(const KArgs& m, const ShotDev* shots, int n_shots, int it0, int tiles_x,
                       int tiles_z, hipStream_t s) {
    const int tiles = tiles_x * tiles_z;
    const dim3 grid(8 * ((tiles + 7) / 8), n_shots);
    hipLaunchKernelGGL((k_block<FWD, PSV, SH, TX, TZ, NT, K>), grid, dim3(NT), 0, s, m, shots,
                       it0, tiles_x, tiles_z);
}

template <bool FWD, bool PSV, bool SH>
static bool launch_sys(BlockCfg c, const KArgs& m, const ShotDev* shots, int n_shots, int it0,
                       int tiles_x, int tiles_z, hipStream_t s) {
#define TRTM_CFG(TXV, TZV, NTV, KV)                                                          \
    if (c.tx == TXV && c.tz == TZV && c.nt == NTV && c.k == KV) {                            \
        launch_cfg<FWD, PSV, SH, TXV, TZV, NTV, KV>(m, shots, n_shots, it0, tiles_x, tiles_z, s); \
        return true;                                                                         \
    }
    if constexpr (PSV && SH) {
        TRTM_CFG(16, 64, 256, 2) TRTM_CFG(16, 64, 256, 4) TRTM_CFG(16, 128, 512, 4)
        TRTM_CFG(16, 32, 1024, 4)
    } else if constexpr (PSV) {
        TRTM_CFG(16, 64, 256, 2) TRTM_CFG(16, 64, 256, 4) TRTM_CFG(16, 128, 512, 4)
        TRTM_CFG(16, 128, 512, 8) TRTM_CFG(32, 128, 1024, 4) TRTM_CFG(48, 64, 1024, 4)
        TRTM_CFG(32, 64, 512, 4) TRTM_CFG(24, 56, 512, 4) TRTM_CFG(16, 32, 1024, 4)
        TRTM_CFG(16, 32, 1024, 8) TRTM_CFG(16, 16, 1024, 8)
    } else {
        TRTM_CFG(16, 64, 256, 4) TRTM_CFG(16, 64, 256, 8) TRTM_CFG(16, 128, 512, 4)
        TRTM_CFG(16, 128, 512, 8) TRTM_CFG(32, 128, 1024, 4) TRTM_CFG(32, 128, 1024, 8)
        TRTM_CFG(24, 128, 1024, 8) TRTM_CFG(32, 96, 1024, 8) TRTM_CFG(48, 64, 1024, 8)
        TRTM_CFG(32, 48, 512, 8) TRTM_CFG(24, 64, 512, 8) TRTM_CFG(16, 32, 1024, 4)
        TRTM_CFG(16, 32, 1024, 8) TRTM_CFG(16, 16, 1024, 8)
    }
#undef TRTM_CFG
    return false;
}

static bool launch_block(bool fwd, int systems, BlockCfg c, const KArgs& m, const ShotDev* shots,
                         int n_shots, int it0, int tiles_x, int tiles_z, hipStream_t s) {
    if (is_wave(c)) return launch_wave(fwd, systems, c, m, shots, n_shots, it0, tiles_x, tiles_z, s);
    const bool psv = systems & TRTM_PSV, sh = systems & TRTM_SH;
    if (fwd) {
        if (psv && sh) return launch_sys<true, true, true>(c, m, shots, n_shots, it0, tiles_x, tiles_z, s);
        if (psv) return launch_sys<true, true, false>(c, m, shots, n_shots, it0, tiles_x, tiles_z, s);
        return launch_sys<true, false, true>(c, m, shots, n_shots, it0, tiles_x, tiles_z, s);
    }
    if (psv && sh) return launch_sys<false, true, true>(c, m, shots, n_shots, it0, tiles_x, tiles_z, s);
    if (psv) return launch_sys<false, true, false>(c, m, shots, n_shots, it0, tiles_x, tiles_z, s);
    return launch_sys<false, false, true>(c, m, shots, n_shots, it0, tiles_x, tiles_z, s);
}

// The launches of one call, prepared: per-shot tile lists and the two
// shot-table halves are written to device memory by work queued on the stream
// before them, so issuing them touches no host memory (and can be captured
// into a graph).
struct BlockRun {
    BlockCfg cfg;
    int n = 0, systems = 0, it_begin = 0, tiles_x = 0, tiles_z = 0;
    bool fwd = true;
    KArgs m;
    ShotDev* shots_dev = nullptr;
    std::vector<char> active;                       // per launch: a shot takes part
    std::vector<std::pair<float*, float*>> copies;  // state copied back (to, from)
    size_t bytes = 0;
};

// Global steps [it_begin, it_end) for n shots; shot q runs its own steps
// [it_begin, it_end) + step_offset, clipped to [0, nt).  One launch covers K
// global steps.  State ping-pongs between the caller's buffers and the shot's
// workspace, with each shot's phase set by the first launch it takes part in.
// shots_dev: device room for 2n ShotDev entries.  Builds the tile lists and
// writes the shot tables on stream s (the tables travel in kernel arguments,
// so nothing waits for the device).
static int block_prepare(const trtm_step_desc* descs, int n, bool fwd, bool save,
                         int32_t it_begin, int32_t it_end, ShotDev* shots_dev, const KArgs& m,
                         hipStream_t s, BlockRun* r) {
    const trtm_step_desc* d0 = &descs[0];
    BlockCfg cfg;
    int rc = block_choose(d0, &cfg);
    if (rc) return rc;
    const int K = cfg.k;
    const int n_launch = (it_end - it_begin + K - 1) / K;
    std::vector<ShotDev> even(n), odd(n);
    std::vector<int> first(n, -1), count(n, 0);
    for (int q = 0; q < n; q++) {
        const trtm_step_desc* d = &descs[q];
        const int lo = std::max(0, it_begin + d->step_offset);
        const int hi = std::min(d->nt, it_end + d->step_offset);
        rc = build_shot(d, fwd, save, cfg, lo, hi, s, &even[q]);
        if (rc) return rc;
        for (int j = 0; j < n_launch; j++) {
            const int a = it_begin + j * K + d->step_offset;
            if (std::max(a, lo) < std::min(a + K, hi)) {
                if (first[q] < 0) first[q] = j;
                count[q]++;
            }
        }
        ShotDev sw = even[q];
        for (int k = 0; k < 8; k++) std::swap(sw.in[k], sw.out[k]);
        odd[q] = sw;
        if (first[q] > 0 && first[q] % 2 == 1) std::swap(even[q], odd[q]);
    }
    const BlockLayout L = block_layout(d0, cfg);
    static_assert(sizeof(ShotDev) % 4 == 0, "shot table of whole words");
    hipError_t e = upload_async(shots_dev, even.data(), sizeof(ShotDev) * n, s);
    if (e == hipSuccess) e = upload_async(shots_dev + n, odd.data(), sizeof(ShotDev) * n, s);
    if (e != hipSuccess) return fail(TRTM_ERR_HIP, "shot table copy: %s", hipGetErrorString(e));

    r->cfg = cfg;
    r->n = n;
    r->systems = d0->systems;
    r->it_begin = it_begin;
    r->tiles_x = L.tiles_x;
    r->tiles_z = L.tiles_z;
    r->fwd = fwd;
    r->m = m;
    r->shots_dev = shots_dev;
    r->active.assign(n_launch, 0);
    for (int j = 0; j < n_launch; j++)
        for (int q = 0; q < n && !r->active[j]; q++)
            r->active[j] = first[q] >= 0 && j >= first[q] && j < first[q] + count[q];
    // a shot that takes part in an odd number of launches ends with its latest
    // state in the workspace: copied back
    r->bytes = (size_t)d0->nx * d0->nz * sizeof(float);
    const bool psv = d0->systems & TRTM_PSV, sh = d0->systems & TRTM_SH;
    const bool act[8] = {psv, sh, psv, psv, psv, psv, sh, sh};
    r->copies.clear();
    for (int q = 0; q < n; q++) {
        if (count[q] % 2 == 0) continue;
        const trtm_step_desc* d = &descs[q];
        float* state[8] = {d->u, d->v, d->w, d->sxx, d->szz, d->sxz, d->syx, d->syz};
        float* alt_base = reinterpret_cast<float*>(static_cast<char*>(d->workspace) + L.alt);
        const size_t nn = (size_t)d->nx * d->nz;
        for (int k = 0; k < 8; k++)
            if (act[k]) r->copies.emplace_back(state[k], alt_base + k * nn);
    }
    return TRTM_OK;
}

// the prepared launches and state copies, device work only
static int block_launch(const BlockRun& r, hipStream_t s) {
    const int K = r.cfg.k;
    for (int j = 0; j < (int)r.active.size(); j++) {
        if (!r.active[j]) continue;
        const ShotDev* tab = r.shots_dev + (j % 2 ? r.n : 0);
        if (!launch_block(r.fwd, r.systems, r.cfg, r.m, tab, r.n, r.it_begin + j * K, r.tiles_x,
                          r.tiles_z, s))
            return fail(TRTM_ERR_UNSUPPORTED, "no blocked kernel for the chosen configuration");
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(TRTM_ERR_HIP, "blocked launch: %s", hipGetErrorString(e));
    for (const auto& c : r.copies) {
        e = hipMemcpyAsync(c.first, c.second, r.bytes, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return fail(TRTM_ERR_HIP, "state copy: %s", hipGetErrorString(e));
    }
    return TRTM_OK;
}

int run_block(const trtm_step_desc* descs, int n, bool fwd, bool save, int32_t it_begin,
              int32_t it_end, ShotDev* shots_dev, const KArgs& m, hipStream_t s) {
    BlockRun r;
    const int rc = block_prepare(descs, n, fwd, save, it_begin, it_end, shots_dev, m, s, &r);
    return rc ? rc : block_launch(r, s);
}

// Prepares the launches on stream s, then records them into a graph on a
// stream of the library's own (default priority; the caller's stream may be
// the null stream, which cannot capture) and instantiates it.  Nothing runs.
int block_capture(const trtm_step_desc* descs, int n, bool fwd, bool save, int32_t it_begin,
                  int32_t it_end, ShotDev* shots_dev, const KArgs& m, hipStream_t s,
                  hipGraph_t* graph, hipGraphExec_t* exec) {
    BlockRun r;
    int rc = block_prepare(descs, n, fwd, save, it_begin, it_end, shots_dev, m, s, &r);
    if (rc) return rc;
    // the lists and tables are in place before the plan can be launched on
    // any stream
    hipError_t w = hipStreamSynchronize(s);
    if (w != hipSuccess) return fail(TRTM_ERR_HIP, "plan preparation: %s", hipGetErrorString(w));
    hipStream_t cap;
    hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(TRTM_ERR_HIP, "capture stream: %s", hipGetErrorString(e));
    e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        hipStreamDestroy(cap);
        return fail(TRTM_ERR_HIP, "begin capture: %s", hipGetErrorString(e));
    }
    rc = block_launch(r, cap);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cap, &g);  // ends the capture on every path
    hipStreamDestroy(cap);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(TRTM_ERR_HIP, "end capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(g);
        return fail(TRTM_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    }
    *graph = g;
    return TRTM_OK;
}

int block_config_query(const trtm_step_desc* d, int32_t* out) {
    BlockCfg c;
    const int rc = block_choose(d, &c);
    if (rc) return rc;
    out[0] = c.tx;
    out[1] = c.tz;
    out[2] = c.nt;
    out[3] = c.k;
    return TRTM_OK;
}

int block_config_list(int systems, int kernel, int32_t* out, int max_n) {
    const BlockCfg* list;
    int n;
    if (kernel == TRTM_KERNEL_WAVE)
        wave_configs(systems, &list, &n);
    else
        cfg_list(systems, &list, &n);
    for (int i = 0; i < n && i < max_n; i++) {
        out[4 * i + 0] = list[i].tx;
        out[4 * i + 1] = list[i].tz;
        out[4 * i + 2] = list[i].nt;
        out[4 * i + 3] = list[i].k;
    }
    return n;
}

int64_t block_shot_table_bytes(int n) { return (int64_t)(2 * n * sizeof(ShotDev) + 256); }

ShotDev* block_shot_slot(const trtm_step_desc* d) {
    // a single shot keeps its two-entry table in its own workspace
    const BlockLayout L = block_layout(d, BlockCfg{16, 64, 256, 1});
    return reinterpret_cast<ShotDev*>(static_cast<char*>(d->workspace) + L.shot);
}

}  // namespace trtm
