// Exact discrete adjoint of the forward step, for FWI gradients.
//
// Forward step n (forward_modeling.py:539-551), as linear maps of the state x
// for a fixed medium m:  V (velocity update, interior), S_n (source term),
// S (stress update, interior), P (sponge), then the record y_n = R x_{n+1}.
// With J = sum_n j(y_n) and p = dJ/dx_{n+1}, one adjoint step is
//     p += R^T dj/dy_n;  p = P^T p;  p += S^T p  (gradients of lam, mu, mxz,
//     myx, myz);  source gradient (inv_rho_*);  p += V^T p  (gradients of
//     inv_rho, inv_rho_u, inv_rho_w)
// after which p = dJ/dx_n.  The coefficient gradients need the stresses the
// velocity update read and the velocities the stress update read; the forward
// pass records them as a per-step history (recomputed per checkpoint segment).
//
// The transposes are written in gather form: every output cell sums the
// contributions of its own and neighbouring cells, with the interior masks of
// the forward updates, so each kernel writes only its own cell.
#include <algorithm>
#include <stdarg.h>
#include <string>
#include <vector>

#include "trtm_fwi.hpp"
#include "trtm_real.hpp"

namespace trtm {

int fail(int code, const char* fmt, ...);

template <class R>
struct AA {
    int nx, nz, nt, fw, n_rec, n_rec_groups, n_src_groups;
    R dt, idx, idz, factor;
    const R* tab;
    const R *lam, *mu, *mxz, *myx, *myz, *ir, *iru, *irw;
    R *pu, *pv, *pw, *psxx, *pszz, *psxz, *psyx, *psyz;
    const int *rec_loc, *rec_order, *rec_start;
    const R *res_u, *res_v, *res_w;
    const int *src_loc, *src_order, *src_start;
    const R *wav_u, *wav_v, *wav_w;
    R *g_lam, *g_mu, *g_mxz, *g_myx, *g_myz, *g_ir, *g_iru, *g_irw;
    // history plane offsets (in cells) of sxx, szz, sxz, u, w, syx, syz, v
    long long h_sxx, h_szz, h_sxz, h_u, h_w, h_syx, h_syz, h_v;
    const int *src_map, *rec_map;  // cell -> source / receiver group, -1 none
    const R* hist;                 // history of the current segment
    long long hist_step;           // cells per history slot
};

// Per cell: the data adjoint of step n at the receivers (entries in order),
// the sponge transpose + stress-update transpose (new velocity adjoint), the
// gradients of lam, mu, mxz (P-SV) / myx, myz (SH), then the source-term
// gradient of inv_rho_u / inv_rho_w / inv_rho (u += (wavelet * factor) *
// inv_rho_u at a source cell).  The velocity adjoint of a cell is read and
// written only by its own thread.
template <class R, bool PSV, bool SH, class S>
__global__ __launch_bounds__(256) void k_adj_stress(S src, int n, int slot) {
    const AA<R> a = src.get();
    const int j = blockIdx.x * kBlockZ + threadIdx.x;
    const int i = blockIdx.y * kBlockX + threadIdx.y;
    const int nx = a.nx, nz = a.nz;
    if (i >= nx || j >= nz || n >= a.nt) return;
    const R* __restrict__ h = a.hist + (size_t)slot * a.hist_step;
    const size_t c = (size_t)i * nz + j;
    R pu0 = PSV ? a.pu[c] : R(0), pw0 = PSV ? a.pw[c] : R(0), pv0 = SH ? a.pv[c] : R(0);
    const int rg = a.rec_map ? a.rec_map[c] : -1;
    if (rg >= 0) {
        const size_t row = (size_t)n * a.n_rec;
        for (int e = a.rec_start[rg]; e < a.rec_start[rg + 1]; e++) {
            const int l = a.rec_order[e];
            if (PSV) {
                pu0 += a.res_u[row + l];
                pw0 += a.res_w[row + l];
            }
            if (SH) pv0 += a.res_v[row + l];
        }
    }
    const R A = sponge_at_t<R>(i, j, nx, nz, a.fw, a.tab);
    auto inner = [&](int ii, int jj) { return ii >= 1 && ii < nx - 1 && jj >= 1 && jj < nz - 1; };
    auto sp = [&](const R* f, int ii, int jj) {
        return sponge_at_t<R>(ii, jj, nx, nz, a.fw, a.tab) * f[(size_t)ii * nz + jj];
    };
    const R dt = a.dt;
    if (PSV) {
        auto qx = [&](int ii, int jj) -> R {
            if (!inner(ii, jj)) return R(0);
            const size_t k = (size_t)ii * nz + jj;
            const R axx = dt * sp(a.psxx, ii, jj), azz = dt * sp(a.pszz, ii, jj);
            return (a.lam[k] + R(2) * a.mu[k]) * axx + a.lam[k] * azz;
        };
        auto qz = [&](int ii, int jj) -> R {
            if (!inner(ii, jj)) return R(0);
            const size_t k = (size_t)ii * nz + jj;
            const R axx = dt * sp(a.psxx, ii, jj), azz = dt * sp(a.pszz, ii, jj);
            return a.lam[k] * axx + (a.lam[k] + R(2) * a.mu[k]) * azz;
        };
        auto ss = [&](int ii, int jj) -> R {
            if (!inner(ii, jj)) return R(0);
            return a.mxz[(size_t)ii * nz + jj] * dt * sp(a.psxz, ii, jj);
        };
        const R pu = A * pu0 + a.idx * (qx(i - 1, j) - qx(i, j)) + a.idz * (ss(i, j - 1) - ss(i, j));
        const R pw = A * pw0 + a.idx * (ss(i - 1, j) - ss(i, j)) + a.idz * (qz(i, j - 1) - qz(i, j));
        a.pu[c] = pu;
        a.pw[c] = pw;
        if (inner(i, j)) {
            const R axx = dt * A * a.psxx[c], azz = dt * A * a.pszz[c], axz = dt * A * a.psxz[c];
            const R* hu = h + a.h_u;
            const R* hw = h + a.h_w;
            const R ux = (hu[c + nz] - hu[c]) * a.idx, uz = (hu[c + 1] - hu[c]) * a.idz;
            const R wx = (hw[c + nz] - hw[c]) * a.idx, wz = (hw[c + 1] - hw[c]) * a.idz;
            a.g_lam[c] += (axx + azz) * (ux + wz);
            a.g_mu[c] += R(2) * (axx * ux + azz * wz);
            a.g_mxz[c] += axz * (uz + wx);
        }
    }
    if (SH) {
        auto qy = [&](int ii, int jj) -> R {
            if (!inner(ii, jj)) return R(0);
            return a.myx[(size_t)ii * nz + jj] * dt * sp(a.psyx, ii, jj);
        };
        auto qv = [&](int ii, int jj) -> R {
            if (!inner(ii, jj)) return R(0);
            return a.myz[(size_t)ii * nz + jj] * dt * sp(a.psyz, ii, jj);
        };
        a.pv[c] = A * pv0 + a.idx * (qy(i - 1, j) - qy(i, j)) + a.idz * (qv(i, j - 1) - qv(i, j));
        if (inner(i, j)) {
            const R* hv = h + a.h_v;
            const R vx = (hv[c + nz] - hv[c]) * a.idx, vz = (hv[c + 1] - hv[c]) * a.idz;
            a.g_myx[c] += dt * A * a.psyx[c] * vx;
            a.g_myz[c] += dt * A * a.psyz[c] * vz;
        }
    }
    const int sg = a.src_map ? a.src_map[c] : -1;
    if (sg >= 0) {
        for (int e = a.src_start[sg]; e < a.src_start[sg + 1]; e++) {
            const size_t o = (size_t)a.src_order[e] * a.nt + n;
            if (PSV) {
                a.g_iru[c] += a.pu[c] * a.wav_u[o] * a.factor;
                a.g_irw[c] += a.pw[c] * a.wav_w[o] * a.factor;
            }
            if (SH) a.g_ir[c] += a.pv[c] * a.wav_v[o] * a.factor;
        }
    }
}
// sponge transpose on the stresses + velocity-update transpose: new stress
// adjoint, and the gradients of inv_rho_u, inv_rho_w (P-SV) / inv_rho (SH)
template <class R, bool PSV, bool SH, class S>
__global__ __launch_bounds__(256) void k_adj_velocity(S src, int n, int slot) {
    const AA<R> a = src.get();
    const int j = blockIdx.x * kBlockZ + threadIdx.x;
    const int i = blockIdx.y * kBlockX + threadIdx.y;
    const int nx = a.nx, nz = a.nz;
    if (i >= nx || j >= nz || n >= a.nt) return;
    const R* __restrict__ h = a.hist + (size_t)slot * a.hist_step;
    const size_t c = (size_t)i * nz + j;
    const R A = sponge_at_t<R>(i, j, nx, nz, a.fw, a.tab);
    auto inner = [&](int ii, int jj) { return ii >= 1 && ii < nx - 1 && jj >= 1 && jj < nz - 1; };
    const R dt = a.dt;
    if (PSV) {
        auto bu = [&](int ii, int jj) -> R {
            if (!inner(ii, jj)) return R(0);
            const size_t k = (size_t)ii * nz + jj;
            return dt * a.iru[k] * a.pu[k];
        };
        auto bw = [&](int ii, int jj) -> R {
            if (!inner(ii, jj)) return R(0);
            const size_t k = (size_t)ii * nz + jj;
            return dt * a.irw[k] * a.pw[k];
        };
        const R bu0 = bu(i, j), bw0 = bw(i, j);
        a.psxx[c] = A * a.psxx[c] + a.idx * (bu0 - bu(i + 1, j));
        a.psxz[c] = A * a.psxz[c] + a.idz * (bu0 - bu(i, j + 1)) + a.idx * (bw0 - bw(i + 1, j));
        a.pszz[c] = A * a.pszz[c] + a.idz * (bw0 - bw(i, j + 1));
        if (inner(i, j)) {
            const R* hxx = h + a.h_sxx;
            const R* hzz = h + a.h_szz;
            const R* hxz = h + a.h_sxz;
            const R sxx_x = (hxx[c] - hxx[c - nz]) * a.idx, sxz_z = (hxz[c] - hxz[c - 1]) * a.idz;
            const R sxz_x = (hxz[c] - hxz[c - nz]) * a.idx, szz_z = (hzz[c] - hzz[c - 1]) * a.idz;
            a.g_iru[c] += a.pu[c] * dt * (sxx_x + sxz_z);
            a.g_irw[c] += a.pw[c] * dt * (sxz_x + szz_z);
        }
    }
    if (SH) {
        auto bv = [&](int ii, int jj) -> R {
            if (!inner(ii, jj)) return R(0);
            const size_t k = (size_t)ii * nz + jj;
            return dt * a.ir[k] * a.pv[k];
        };
        const R bv0 = bv(i, j);
        a.psyx[c] = A * a.psyx[c] + a.idx * (bv0 - bv(i + 1, j));
        a.psyz[c] = A * a.psyz[c] + a.idz * (bv0 - bv(i, j + 1));
        if (inner(i, j)) {
            const R* hyx = h + a.h_syx;
            const R* hyz = h + a.h_syz;
            const R syx_x = (hyx[c] - hyx[c - nz]) * a.idx, syz_z = (hyz[c] - hyz[c - 1]) * a.idz;
            a.g_ir[c] += a.pv[c] * dt * (syx_x + syz_z);
        }
    }
}

template <class R, bool PSV, bool SH, class S>
static hipError_t adjoint_step(S src, int nx, int nz, int shots, int n, int slot, hipStream_t s) {
    const dim3 block(kBlockZ, kBlockX);
    const dim3 grid((nz + kBlockZ - 1) / kBlockZ, (nx + kBlockX - 1) / kBlockX, shots);
    hipLaunchKernelGGL((k_adj_stress<R, PSV, SH, S>), grid, block, 0, s, src, n, slot);
    hipLaunchKernelGGL((k_adj_velocity<R, PSV, SH, S>), grid, block, 0, s, src, n, slot);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Forward step of the FWI pass (forward_modeling.py:539-551, flat free
// surface), three launches: velocity update + source injection (+ history of
// the pre-step stresses and the injected velocities), stress update + stress
// sponge, velocity sponge + recording.  Same expressions, in the same order, as
// the reference-shaped kernels.

template <class R>
struct FF {
    int nx, nz, nt, fw, n_rec;
    R dt, idx, idz, factor;
    const R* tab;
    const R *lam, *mu, *mxz, *myx, *myz, *ir, *iru, *irw;
    R *u, *v, *w, *sxx, *szz, *sxz, *syx, *syz;
    const int *src_map, *rec_map;
    const int *src_order, *src_start, *rec_order, *rec_start;
    const R *wav_u, *wav_v, *wav_w;
    R *rec_u, *rec_v, *rec_w;
    R* hist;  // history of the current segment or nullptr
    long long hist_step;
    long long h_sxx, h_szz, h_sxz, h_u, h_w, h_syx, h_syz, h_v;
};

template <class R, bool PSV, bool SH, class S>
__global__ __launch_bounds__(256) void k_ff_velocity(S src, int it, int slot) {
    const FF<R> a = src.get();
    const int j = blockIdx.x * kBlockZ + threadIdx.x;
    const int i = blockIdx.y * kBlockX + threadIdx.y;
    const int nx = a.nx, nz = a.nz;
    if (i >= nx || j >= nz || it >= a.nt) return;
    R* const hh = (a.hist && slot >= 0) ? a.hist + (size_t)slot * a.hist_step : nullptr;
    const size_t c = (size_t)i * nz + j;
    const bool inner = i >= 1 && i < nx - 1 && j >= 1 && j < nz - 1;
    const int sg = a.src_map ? a.src_map[c] : -1;
    if (PSV) {
        if (hh) {
            hh[a.h_sxx + c] = a.sxx[c];
            hh[a.h_szz + c] = a.szz[c];
            hh[a.h_sxz + c] = a.sxz[c];
        }
        R u = a.u[c], w = a.w[c];
        if (inner) {
            const R sxx_x = (a.sxx[c] - a.sxx[c - nz]) * a.idx;
            const R szz_z = (a.szz[c] - a.szz[c - 1]) * a.idz;
            const R sxz_x = (a.sxz[c] - a.sxz[c - nz]) * a.idx;
            const R sxz_z = (a.sxz[c] - a.sxz[c - 1]) * a.idz;
            u = u + ((sxx_x + sxz_z) * a.dt) * a.iru[c];
            w = w + ((sxz_x + szz_z) * a.dt) * a.irw[c];
        }
        if (sg >= 0)
            for (int e = a.src_start[sg]; e < a.src_start[sg + 1]; e++) {
                const size_t o = (size_t)a.src_order[e] * a.nt + it;
                u = u + (a.wav_u[o] * a.factor) * a.iru[c];
                w = w + (a.wav_w[o] * a.factor) * a.irw[c];
            }
        a.u[c] = u;
        a.w[c] = w;
        if (hh) {
            hh[a.h_u + c] = u;
            hh[a.h_w + c] = w;
        }
    }
    if (SH) {
        if (hh) {
            hh[a.h_syx + c] = a.syx[c];
            hh[a.h_syz + c] = a.syz[c];
        }
        R v = a.v[c];
        if (inner) {
            const R syx_x = (a.syx[c] - a.syx[c - nz]) * a.idx;
            const R syz_z = (a.syz[c] - a.syz[c - 1]) * a.idz;
            v = v + ((syx_x + syz_z) * a.dt) * a.ir[c];
        }
        if (sg >= 0)
            for (int e = a.src_start[sg]; e < a.src_start[sg + 1]; e++)
                v = v + (a.wav_v[(size_t)a.src_order[e] * a.nt + it] * a.factor) * a.ir[c];
        a.v[c] = v;
        if (hh) hh[a.h_v + c] = v;
    }
}

template <class R, bool PSV, bool SH, class S>
__global__ __launch_bounds__(256) void k_ff_stress(S src, int it) {
    const FF<R> a = src.get();
    const int j = blockIdx.x * kBlockZ + threadIdx.x;
    const int i = blockIdx.y * kBlockX + threadIdx.y;
    const int nx = a.nx, nz = a.nz;
    if (i >= nx || j >= nz || it >= a.nt) return;
    const size_t c = (size_t)i * nz + j;
    const R A = sponge_at_t<R>(i, j, nx, nz, a.fw, a.tab);
    const bool inner = i >= 1 && i < nx - 1 && j >= 1 && j < nz - 1;
    const R sdt = a.dt;
    if (PSV) {
        R sxx = a.sxx[c], szz = a.szz[c], sxz = a.sxz[c];
        if (inner) {
            const R u_x = (a.u[c + nz] - a.u[c]) * a.idx;
            const R u_z = (a.u[c + 1] - a.u[c]) * a.idz;
            const R w_x = (a.w[c + nz] - a.w[c]) * a.idx;
            const R w_z = (a.w[c + 1] - a.w[c]) * a.idz;
            const R div_uw = u_x + w_z;
            const R lam = a.lam[c], mu = a.mu[c];
            sxx = sxx + sdt * (lam * div_uw + (R(2) * mu) * u_x);
            szz = szz + sdt * (lam * div_uw + (R(2) * mu) * w_z);
            sxz = sxz + (sdt * a.mxz[c]) * (u_z + w_x);
        }
        a.sxx[c] = sxx * A;
        a.szz[c] = szz * A;
        a.sxz[c] = sxz * A;
    }
    if (SH) {
        R syx = a.syx[c], syz = a.syz[c];
        if (inner) {
            const R v_x = (a.v[c + nz] - a.v[c]) * a.idx;
            const R v_z = (a.v[c + 1] - a.v[c]) * a.idz;
            syx = syx + (sdt * a.myx[c]) * v_x;
            syz = syz + (sdt * a.myz[c]) * v_z;
        }
        a.syx[c] = syx * A;
        a.syz[c] = syz * A;
    }
}

template <class R, bool PSV, bool SH, class S>
__global__ __launch_bounds__(256) void k_ff_post(S src, int it) {
    const FF<R> a = src.get();
    const int j = blockIdx.x * kBlockZ + threadIdx.x;
    const int i = blockIdx.y * kBlockX + threadIdx.y;
    const int nx = a.nx, nz = a.nz;
    if (i >= nx || j >= nz || it >= a.nt) return;
    const size_t c = (size_t)i * nz + j;
    const R A = sponge_at_t<R>(i, j, nx, nz, a.fw, a.tab);
    R u = 0, w = 0, v = 0;
    if (PSV) {
        u = a.u[c] * A;
        w = a.w[c] * A;
        a.u[c] = u;
        a.w[c] = w;
    }
    if (SH) {
        v = a.v[c] * A;
        a.v[c] = v;
    }
    const int rg = a.rec_map ? a.rec_map[c] : -1;
    if (rg >= 0) {
        const size_t row = (size_t)it * a.n_rec;
        for (int e = a.rec_start[rg]; e < a.rec_start[rg + 1]; e++) {
            const int l = a.rec_order[e];
            if (PSV && a.rec_u) {
                a.rec_u[row + l] = u;
                a.rec_w[row + l] = w;
            }
            if (SH && a.rec_v) a.rec_v[row + l] = v;
        }
    }
}

template <class R, bool PSV, bool SH, class S>
static hipError_t forward_step(S src, int nx, int nz, int shots, int it, int slot, hipStream_t s) {
    const dim3 block(kBlockZ, kBlockX);
    const dim3 grid((nz + kBlockZ - 1) / kBlockZ, (nx + kBlockX - 1) / kBlockX, shots);
    hipLaunchKernelGGL((k_ff_velocity<R, PSV, SH, S>), grid, block, 0, s, src, it, slot);
    hipLaunchKernelGGL((k_ff_stress<R, PSV, SH, S>), grid, block, 0, s, src, it);
    hipLaunchKernelGGL((k_ff_post<R, PSV, SH, S>), grid, block, 0, s, src, it);
    return hipGetLastError();
}

// cell -> group of a plan (one thread per group)
__global__ void k_group_map(int n_groups, int nz, const int* loc, const int* order,
                            const int* start, int* map) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    const int k0 = order[start[g]];
    map[(size_t)loc[2 * k0] * nz + loc[2 * k0 + 1]] = g;
}

__global__ void k_fill_i32(int* p, size_t n, int value) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = value;
}

// coefficient space -> (lam, mu, rho) -> (vp, vs, rho), double precision
__global__ __launch_bounds__(256) void k_chain(int nx, int nz, const double* mu, const double* rho,
                                               const double* vp, const double* vs,
                                               const double* g_lam, const double* g_mu,
                                               const double* g_mxz, const double* g_myx,
                                               const double* g_myz, const double* g_ir,
                                               const double* g_iru, const double* g_irw,
                                               double* o_vp, double* o_vs, double* o_rho) {
    const int j = blockIdx.x * kBlockZ + threadIdx.x;
    const int i = blockIdx.y * kBlockX + threadIdx.y;
    if (i >= nx || j >= nz) return;
    auto at = [&](const double* f, int ii, int jj) { return f[(size_t)ii * nz + jj]; };
    auto inner = [&](int ii, int jj) { return ii >= 1 && ii < nx - 1 && jj >= 1 && jj < nz - 1; };
    const double m0 = at(mu, i, j), r0 = at(rho, i, j);
    double gmu = at(g_mu, i, j);
    // mxz(i', j') = 4 / sum 1/mu over the 2x2 block at (i', j'): d/dmu = mxz^2 / (4 mu^2)
    for (int di = -1; di <= 0; di++)
        for (int dj = -1; dj <= 0; dj++) {
            const int ii = i + di, jj = j + dj;
            if (!inner(ii, jj)) continue;
            const double s = 1.0 / at(mu, ii, jj) + 1.0 / at(mu, ii + 1, jj) +
                             1.0 / at(mu, ii, jj + 1) + 1.0 / at(mu, ii + 1, jj + 1);
            const double H = 4.0 / s;
            gmu += at(g_mxz, ii, jj) * H * H / (4.0 * m0 * m0);
        }
    // myx(i', j) = 2 / (1/mu(i') + 1/mu(i'+1)), myz(i, j') likewise in z: d/dmu = H^2 / (2 mu^2)
    for (int di = -1; di <= 0; di++) {
        const int ii = i + di;
        if (!inner(ii, j)) continue;
        const double H = 2.0 / (1.0 / at(mu, ii, j) + 1.0 / at(mu, ii + 1, j));
        gmu += at(g_myx, ii, j) * H * H / (2.0 * m0 * m0);
    }
    for (int dj = -1; dj <= 0; dj++) {
        const int jj = j + dj;
        if (!inner(i, jj)) continue;
        const double H = 2.0 / (1.0 / at(mu, i, jj) + 1.0 / at(mu, i, jj + 1));
        gmu += at(g_myz, i, jj) * H * H / (2.0 * m0 * m0);
    }
    // inv_rho = 1/rho; inv_rho_u(i', j) = 2 / (rho(i') + rho(i'+1)); inv_rho_w likewise in z
    double grho = -at(g_ir, i, j) / (r0 * r0);
    for (int di = -1; di <= 0; di++) {
        const int ii = i + di;
        if (!inner(ii, j)) continue;
        const double q = 2.0 / (at(rho, ii, j) + at(rho, ii + 1, j));
        grho += -0.5 * at(g_iru, ii, j) * q * q;
    }
    for (int dj = -1; dj <= 0; dj++) {
        const int jj = j + dj;
        if (!inner(i, jj)) continue;
        const double q = 2.0 / (at(rho, i, jj) + at(rho, i, jj + 1));
        grho += -0.5 * at(g_irw, i, jj) * q * q;
    }
    // mu = rho vs^2, lam = rho (vp^2 - 2 vs^2)
    const double p = at(vp, i, j), sv = at(vs, i, j), gl = at(g_lam, i, j);
    const size_t c = (size_t)i * nz + j;
    o_vp[c] = gl * 2.0 * r0 * p;
    o_vs[c] = gmu * 2.0 * r0 * sv - gl * 4.0 * r0 * sv;
    o_rho[c] = grho + gmu * sv * sv + gl * (p * p - 2.0 * sv * sv);
}

template <class R>
__global__ __launch_bounds__(256) void k_prepare_t(int nx, int nz, const R* __restrict__ mu,
                                                   const R* __restrict__ rho, R* myx, R* myz,
                                                   R* mxz, R* rho_u, R* rho_w, R* inv_rho,
                                                   R* inv_rho_u, R* inv_rho_w) {
    const int j = blockIdx.x * kBlockZ + threadIdx.x;
    const int i = blockIdx.y * kBlockX + threadIdx.y;
    if (i >= nx || j >= nz) return;
    const size_t c = (size_t)i * nz + j;
    R hx = 0, hz = 0, hxz = 0, ru = 0, rw = 0;
    if (i >= 1 && i < nx - 1 && j >= 1 && j < nz - 1) {
        const R a = mu[c], b = mu[c + nz], cc = mu[c + 1], d = mu[c + nz + 1];
        hx = R(2) * a * b / (a + b);
        hz = R(2) * a * cc / (a + cc);
        const R product = a * b * cc * d;
        const R sum_inv_prod = b * cc * d + a * cc * d + a * b * d + a * b * cc;
        hxz = R(4) * product / sum_inv_prod;
        const R r = rho[c];
        ru = R(0.5) * (r + rho[c + nz]);
        rw = R(0.5) * (r + rho[c + 1]);
    }
    myx[c] = hx;
    myz[c] = hz;
    mxz[c] = hxz;
    rho_u[c] = ru;
    rho_w[c] = rw;
    inv_rho[c] = R(1) / rho[c];
    inv_rho_u[c] = R(1) / ru;
    inv_rho_w[c] = R(1) / rw;
}

// ---------------------------------------------------------------------------

static int check_fwi(const trtm_fwi_desc* d, int it_begin, int it_end) {
    if (!d) return fail(TRTM_ERR_ARG, "descriptor is NULL");
    if (d->struct_size != (int32_t)sizeof(trtm_fwi_desc))
        return fail(TRTM_ERR_ARG, "struct_size %d != %d (ABI mismatch)", d->struct_size,
                    (int)sizeof(trtm_fwi_desc));
    if (d->real_bytes != 4 && d->real_bytes != 8) return fail(TRTM_ERR_ARG, "real_bytes must be 4 or 8");
    if (d->nx < 2 || d->nz < 2 || d->nt < 1 || d->fw < 0 || d->fw > d->nx || d->fw > d->nz)
        return fail(TRTM_ERR_ARG, "bad grid");
    if (it_begin < 0 || it_end > d->nt || it_begin > it_end)
        return fail(TRTM_ERR_ARG, "step range [%d, %d) outside [0, %d]", it_begin, it_end, d->nt);
    if (d->systems <= 0 || (d->systems & ~(TRTM_PSV | TRTM_SH)))
        return fail(TRTM_ERR_ARG, "bad systems %d", d->systems);
    if (!d->sponge_table || !d->lam || !d->mu || !d->mxz || !d->inv_rho_u || !d->inv_rho_w ||
        !d->myx || !d->myz || !d->inv_rho)
        return fail(TRTM_ERR_ARG, "medium pointer is NULL");
    if ((d->n_src_groups > 0 || d->n_rec_groups > 0) && !d->cell_map)
        return fail(TRTM_ERR_ARG, "cell_map is NULL (call trtm_fwi_prepare)");
    if (d->n_src_groups > 0 && (!d->src_order || !d->src_group_start))
        return fail(TRTM_ERR_ARG, "source plan is NULL");
    if ((d->rec_u || d->rec_v) && d->n_rec > 0 && (d->n_rec_groups <= 0 || !d->rec_order || !d->rec_group_start))
        return fail(TRTM_ERR_ARG, "recording needs the receiver plan");
    if (d->kernel != TRTM_KERNEL_AUTO && d->kernel != TRTM_KERNEL_SIMPLE &&
        d->kernel != TRTM_KERNEL_BLOCKED)
        return fail(TRTM_ERR_ARG, "kernel must be TRTM_KERNEL_AUTO, _SIMPLE or _BLOCKED");
    if (d->kernel == TRTM_KERNEL_BLOCKED && (d->real_bytes != 4 || !d->workspace || d->nx < 3 ||
                                             d->nz < 3))
        return fail(TRTM_ERR_ARG, "the blocked FWI kernel needs float, the workspace and a grid "
                                  "of at least 3 x 3");
    return TRTM_OK;
}

// blocked kernels: float with a workspace, unless one step per launch is asked
// for, or (automatic choice) the adjoint is given the full history layout,
// which the one-step-per-launch adjoint reads
static bool use_block(const trtm_fwi_desc* d, bool adjoint) {
    if (d->kernel == TRTM_KERNEL_SIMPLE) return false;
    if (adjoint && d->kernel == TRTM_KERNEL_AUTO && d->hist_every == 0) return false;
    return d->real_bytes == 4 && d->workspace != nullptr && d->nx >= 3 && d->nz >= 3;
}

// the compact history layout is written and read by the blocked kernels only
static int check_layout(const trtm_fwi_desc* d, bool adjoint) {
    if (d->hist_every < 0) return fail(TRTM_ERR_ARG, "hist_every must be >= 0");
    if (d->hist && d->hist_every > 0 && !use_block(d, adjoint))
        return fail(TRTM_ERR_ARG, "the compact history (hist_every > 0) needs the blocked "
                                  "kernels (float, the workspace, a grid of at least 3 x 3)");
    return TRTM_OK;
}

static int hist_fields(int systems) {
    return ((systems & TRTM_PSV) ? 5 : 0) + ((systems & TRTM_SH) ? 3 : 0);
}

static void hist_offsets(int systems, long long n, long long* o) {
    // sxx, szz, sxz, u, w, syx, syz, v
    int f = 0;
    for (int k = 0; k < 8; k++) o[k] = 0;
    if (systems & TRTM_PSV)
        for (int k = 0; k < 5; k++) o[k] = (f++) * n;
    if (systems & TRTM_SH)
        for (int k = 5; k < 8; k++) o[k] = (f++) * n;
}

template <class R>
static FF<R> make_ff(const trtm_fwi_desc* d) {
    FF<R> a;
    a.nx = d->nx; a.nz = d->nz; a.nt = d->nt; a.fw = d->fw; a.n_rec = d->n_rec;
    a.dt = (R)d->dt; a.idx = (R)d->inv_dx; a.idz = (R)d->inv_dz; a.factor = (R)d->src_factor;
    a.tab = (const R*)d->sponge_table;
    a.lam = (const R*)d->lam; a.mu = (const R*)d->mu; a.mxz = (const R*)d->mxz;
    a.myx = (const R*)d->myx; a.myz = (const R*)d->myz; a.ir = (const R*)d->inv_rho;
    a.iru = (const R*)d->inv_rho_u; a.irw = (const R*)d->inv_rho_w;
    a.u = (R*)d->u; a.v = (R*)d->v; a.w = (R*)d->w;
    a.sxx = (R*)d->sxx; a.szz = (R*)d->szz; a.sxz = (R*)d->sxz;
    a.syx = (R*)d->syx; a.syz = (R*)d->syz;
    const long long n = (long long)d->nx * d->nz;
    a.src_map = d->n_src_groups > 0 ? d->cell_map : nullptr;
    a.rec_map = (d->n_rec_groups > 0 && (d->rec_u || d->rec_v)) ? d->cell_map + n : nullptr;
    a.src_order = d->src_order; a.src_start = d->src_group_start;
    a.rec_order = d->rec_order; a.rec_start = d->rec_group_start;
    a.wav_u = (const R*)d->wav_u; a.wav_v = (const R*)d->wav_v; a.wav_w = (const R*)d->wav_w;
    a.rec_u = (R*)d->rec_u; a.rec_v = (R*)d->rec_v; a.rec_w = (R*)d->rec_w;
    long long o[8];
    hist_offsets(d->systems, n, o);
    a.h_sxx = o[0]; a.h_szz = o[1]; a.h_sxz = o[2]; a.h_u = o[3]; a.h_w = o[4];
    a.h_syx = o[5]; a.h_syz = o[6]; a.h_v = o[7];
    a.hist = (R*)d->hist;
    a.hist_step = (long long)hist_fields(d->systems) * n;
    return a;
}

template <class R>
static AA<R> make_aa(const trtm_fwi_desc* d) {
    AA<R> a;
    a.nx = d->nx; a.nz = d->nz; a.nt = d->nt; a.fw = d->fw;
    a.n_rec = d->n_rec; a.n_rec_groups = d->n_rec_groups; a.n_src_groups = d->n_src_groups;
    a.dt = (R)d->dt; a.idx = (R)d->inv_dx; a.idz = (R)d->inv_dz; a.factor = (R)d->src_factor;
    a.tab = (const R*)d->sponge_table;
    a.lam = (const R*)d->lam; a.mu = (const R*)d->mu; a.mxz = (const R*)d->mxz;
    a.myx = (const R*)d->myx; a.myz = (const R*)d->myz; a.ir = (const R*)d->inv_rho;
    a.iru = (const R*)d->inv_rho_u; a.irw = (const R*)d->inv_rho_w;
    a.pu = (R*)d->pu; a.pv = (R*)d->pv; a.pw = (R*)d->pw;
    a.psxx = (R*)d->psxx; a.pszz = (R*)d->pszz; a.psxz = (R*)d->psxz;
    a.psyx = (R*)d->psyx; a.psyz = (R*)d->psyz;
    a.rec_loc = d->rec_loc; a.rec_order = d->rec_order; a.rec_start = d->rec_group_start;
    a.res_u = (const R*)d->res_u; a.res_v = (const R*)d->res_v; a.res_w = (const R*)d->res_w;
    a.src_loc = d->src_loc; a.src_order = d->src_order; a.src_start = d->src_group_start;
    a.wav_u = (const R*)d->wav_u; a.wav_v = (const R*)d->wav_v; a.wav_w = (const R*)d->wav_w;
    a.g_lam = (R*)d->g_lam; a.g_mu = (R*)d->g_mu; a.g_mxz = (R*)d->g_mxz;
    a.g_myx = (R*)d->g_myx; a.g_myz = (R*)d->g_myz; a.g_ir = (R*)d->g_ir;
    a.g_iru = (R*)d->g_iru; a.g_irw = (R*)d->g_irw;
    const long long n = (long long)d->nx * d->nz;
    long long o[8];
    hist_offsets(d->systems, n, o);
    a.h_sxx = o[0]; a.h_szz = o[1]; a.h_sxz = o[2]; a.h_u = o[3]; a.h_w = o[4];
    a.h_syx = o[5]; a.h_syz = o[6]; a.h_v = o[7];
    a.src_map = d->n_src_groups > 0 ? d->cell_map : nullptr;
    a.rec_map = d->n_rec_groups > 0 ? d->cell_map + n : nullptr;
    a.hist = (const R*)d->hist;
    a.hist_step = (long long)hist_fields(d->systems) * n;
    return a;
}

template <class R, class S>
static int forward_loop(S src, const trtm_fwi_desc* d0, int shots, int it_begin, int it_end,
                        hipStream_t s) {
    const bool psv = d0->systems & TRTM_PSV, sh = d0->systems & TRTM_SH;
    for (int it = it_begin; it < it_end; it++) {
        const int slot = it - d0->hist_base;
        const int sl = (d0->hist && slot >= 0 && slot < d0->hist_len) ? slot : -1;
        hipError_t e;
        if (psv && sh) e = forward_step<R, true, true, S>(src, d0->nx, d0->nz, shots, it, sl, s);
        else if (psv) e = forward_step<R, true, false, S>(src, d0->nx, d0->nz, shots, it, sl, s);
        else e = forward_step<R, false, true, S>(src, d0->nx, d0->nz, shots, it, sl, s);
        if (e != hipSuccess) return fail(TRTM_ERR_HIP, "fwi forward step %d: %s", it, hipGetErrorString(e));
    }
    return TRTM_OK;
}

template <class R, class S>
static int adjoint_loop(S src, const trtm_fwi_desc* d0, int shots, int it_begin, int it_end,
                        hipStream_t s) {
    const bool psv = d0->systems & TRTM_PSV, sh = d0->systems & TRTM_SH;
    for (int it = it_end - 1; it >= it_begin; it--) {
        const int slot = it - d0->hist_base;
        if (slot < 0 || slot >= d0->hist_len)
            return fail(TRTM_ERR_ARG, "adjoint step %d has no history (base %d, length %d)", it,
                        d0->hist_base, d0->hist_len);
        hipError_t e;
        if (psv && sh) e = adjoint_step<R, true, true, S>(src, d0->nx, d0->nz, shots, it, slot, s);
        else if (psv) e = adjoint_step<R, true, false, S>(src, d0->nx, d0->nz, shots, it, slot, s);
        else e = adjoint_step<R, false, true, S>(src, d0->nx, d0->nz, shots, it, slot, s);
        if (e != hipSuccess) return fail(TRTM_ERR_HIP, "adjoint step %d: %s", it, hipGetErrorString(e));
    }
    return TRTM_OK;
}

static_assert(sizeof(ByArray<AA<double>>) <= 4096 && sizeof(ByArray<FF<double>>) <= 4096,
              "per-shot kernel arguments exceed 4 KB");
static_assert(sizeof(ByArray<AA<float>>) <= 4096 && sizeof(ByArray<FF<float>>) <= 4096,
              "per-shot kernel arguments exceed 4 KB");

// One shot: parameters by value.  Several: groups of up to kShotsPerLaunch
// shots, their parameters an array kernel argument indexed by blockIdx.z.
// The groups of a step are launched back to back.
template <class R>
static int fwi_forward_t(const trtm_fwi_desc* d, int n, int it_begin, int it_end, hipStream_t s) {
    if (n == 1) return forward_loop<R>(ByValue<FF<R>>{make_ff<R>(d)}, d, 1, it_begin, it_end, s);
    std::vector<ByArray<FF<R>>> groups((n + kShotsPerLaunch - 1) / kShotsPerLaunch);
    for (int q = 0; q < n; q++) groups[q / kShotsPerLaunch].v[q % kShotsPerLaunch] = make_ff<R>(&d[q]);
    for (int it = it_begin; it < it_end; it++)
        for (size_t g = 0; g < groups.size(); g++) {
            const int m = std::min(kShotsPerLaunch, n - (int)g * kShotsPerLaunch);
            const int rc = forward_loop<R>(groups[g], d, m, it, it + 1, s);
            if (rc) return rc;
        }
    return TRTM_OK;
}

template <class R>
static int fwi_adjoint_t(const trtm_fwi_desc* d, int n, int it_begin, int it_end, hipStream_t s) {
    if (n == 1) return adjoint_loop<R>(ByValue<AA<R>>{make_aa<R>(d)}, d, 1, it_begin, it_end, s);
    std::vector<ByArray<AA<R>>> groups((n + kShotsPerLaunch - 1) / kShotsPerLaunch);
    for (int q = 0; q < n; q++) groups[q / kShotsPerLaunch].v[q % kShotsPerLaunch] = make_aa<R>(&d[q]);
    for (int it = it_end - 1; it >= it_begin; it--)
        for (size_t g = 0; g < groups.size(); g++) {
            const int m = std::min(kShotsPerLaunch, n - (int)g * kShotsPerLaunch);
            const int rc = adjoint_loop<R>(groups[g], d, m, it, it + 1, s);
            if (rc) return rc;
        }
    return TRTM_OK;
}

}  // namespace trtm

using namespace trtm;

extern "C" {

int trtm_prepare_medium_f64(int32_t nx, int32_t nz, const double* mu, const double* rho,
                            double* myx, double* myz, double* mxz, double* rho_u, double* rho_w,
                            double* inv_rho, double* inv_rho_u, double* inv_rho_w, void* stream) {
    if (nx < 1 || nz < 1 || !mu || !rho || !myx || !myz || !mxz || !rho_u || !rho_w || !inv_rho ||
        !inv_rho_u || !inv_rho_w)
        return fail(TRTM_ERR_ARG, "bad arguments");
    hipLaunchKernelGGL(k_prepare_t<double>,
                       dim3((nz + kBlockZ - 1) / kBlockZ, (nx + kBlockX - 1) / kBlockX),
                       dim3(kBlockZ, kBlockX), 0, (hipStream_t)stream, nx, nz, mu, rho, myx, myz,
                       mxz, rho_u, rho_w, inv_rho, inv_rho_u, inv_rho_w);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TRTM_OK : fail(TRTM_ERR_HIP, "prepare f64: %s", hipGetErrorString(e));
}

int trtm_fwi_prepare(const trtm_fwi_desc* d, void* stream) {
    if (!d || d->struct_size != (int32_t)sizeof(trtm_fwi_desc))
        return fail(TRTM_ERR_ARG, "descriptor is NULL or of another ABI");
    if (d->nx < 2 || d->nz < 2) return fail(TRTM_ERR_ARG, "bad grid");
    if (!d->cell_map) return fail(TRTM_ERR_ARG, "cell_map is NULL");
    if ((d->n_src_groups > 0 && (!d->src_loc || !d->src_order || !d->src_group_start)) ||
        (d->n_rec_groups > 0 && (!d->rec_loc || !d->rec_order || !d->rec_group_start)))
        return fail(TRTM_ERR_ARG, "plan pointer is NULL");
    const size_t n = (size_t)d->nx * d->nz;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((2 * n + 255) / 256)), dim3(256), 0, s,
                       d->cell_map, 2 * n, -1);
    if (d->n_src_groups > 0)
        hipLaunchKernelGGL(k_group_map, dim3((d->n_src_groups + 255) / 256), dim3(256), 0, s,
                           d->n_src_groups, d->nz, d->src_loc, d->src_order, d->src_group_start,
                           d->cell_map);
    if (d->n_rec_groups > 0)
        hipLaunchKernelGGL(k_group_map, dim3((d->n_rec_groups + 255) / 256), dim3(256), 0, s,
                           d->n_rec_groups, d->nz, d->rec_loc, d->rec_order, d->rec_group_start,
                           d->cell_map + n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TRTM_OK : fail(TRTM_ERR_HIP, "fwi prepare: %s", hipGetErrorString(e));
}

}  // extern "C"

static int check_fwd(const trtm_fwi_desc* d, int it_begin, int it_end) {
    int rc = check_fwi(d, it_begin, it_end);
    if (rc) return rc;
    if ((d->systems & TRTM_PSV) && (!d->u || !d->w || !d->sxx || !d->szz || !d->sxz))
        return fail(TRTM_ERR_ARG, "P-SV state pointer is NULL");
    if ((d->systems & TRTM_SH) && (!d->v || !d->syx || !d->syz))
        return fail(TRTM_ERR_ARG, "SH state pointer is NULL");
    return TRTM_OK;
}

static int check_adj(const trtm_fwi_desc* d, int it_begin, int it_end) {
    int rc = check_fwi(d, it_begin, it_end);
    if (rc) return rc;
    if (!d->hist) return fail(TRTM_ERR_ARG, "adjoint needs the forward history");
    if ((d->systems & TRTM_PSV) && (!d->pu || !d->pw || !d->psxx || !d->pszz || !d->psxz ||
                                    !d->g_lam || !d->g_mu || !d->g_mxz || !d->g_iru || !d->g_irw))
        return fail(TRTM_ERR_ARG, "P-SV adjoint pointer is NULL");
    if ((d->systems & TRTM_SH) && (!d->pv || !d->psyx || !d->psyz || !d->g_myx || !d->g_myz ||
                                   !d->g_ir))
        return fail(TRTM_ERR_ARG, "SH adjoint pointer is NULL");
    if (d->n_rec_groups > 0 && (!d->rec_order || !d->rec_group_start ||
                                ((d->systems & TRTM_PSV) && (!d->res_u || !d->res_w)) ||
                                ((d->systems & TRTM_SH) && !d->res_v)))
        return fail(TRTM_ERR_ARG, "receiver plan or data adjoint is NULL");
    return TRTM_OK;
}

// shots of a batch: same grid, precision, wave systems and history window;
// each shot's own record length bounds its steps
static int check_batch(const trtm_fwi_desc* d, int n, int it_begin, int it_end, bool adjoint) {
    if (!d || n < 1) return fail(TRTM_ERR_ARG, "no shots");
    int nt_max = 0;
    for (int q = 0; q < n; q++) {
        const trtm_fwi_desc* e = &d[q];
        const int hi = std::min(it_end, e->nt > 0 ? e->nt : it_end);
        const int rc = adjoint ? check_adj(e, std::min(it_begin, hi), hi) : check_fwd(e, std::min(it_begin, hi), hi);
        if (rc) {
            const std::string why = trtm_last_error();
            return fail(rc, "shot %d: %s", q, why.c_str());
        }
        if (e->nx != d->nx || e->nz != d->nz || e->real_bytes != d->real_bytes ||
            e->systems != d->systems || e->hist_base != d->hist_base ||
            e->hist_len != d->hist_len || (e->hist == nullptr) != (d->hist == nullptr))
            return fail(TRTM_ERR_ARG, "shot %d does not share the grid, precision, systems and "
                                      "history window of shot 0", q);
        if (use_block(e, adjoint) != use_block(d, adjoint) || e->tile_x != d->tile_x ||
            e->tile_z != d->tile_z || e->block_steps != d->block_steps ||
            e->hist_every != d->hist_every)
            return fail(TRTM_ERR_ARG, "shot %d does not share the step kernel of shot 0", q);
        nt_max = std::max(nt_max, e->nt);
    }
    if (it_begin < 0 || it_end > nt_max || it_begin > it_end)
        return fail(TRTM_ERR_ARG, "step range [%d, %d) outside [0, %d]", it_begin, it_end, nt_max);
    return TRTM_OK;
}

extern "C" {

// the adjoint reads the history of every step it takes
static int check_window(const trtm_fwi_desc* d, int it_begin, int it_end) {
    if (it_begin < it_end && (it_begin < d->hist_base || it_end > d->hist_base + d->hist_len))
        return fail(TRTM_ERR_ARG, "adjoint steps [%d, %d) outside the history [%d, %d)", it_begin,
                    it_end, d->hist_base, d->hist_base + d->hist_len);
    return TRTM_OK;
}

int trtm_fwi_forward(const trtm_fwi_desc* d, int32_t it_begin, int32_t it_end, void* stream) {
    int rc = check_fwd(d, it_begin, it_end);
    if (rc || (rc = check_layout(d, false))) return rc;
    if (use_block(d, false)) return fwi_block_forward(d, 1, it_begin, it_end, (hipStream_t)stream);
    if (d->real_bytes == 4) return fwi_forward_t<float>(d, 1, it_begin, it_end, (hipStream_t)stream);
    return fwi_forward_t<double>(d, 1, it_begin, it_end, (hipStream_t)stream);
}

int trtm_fwi_adjoint(const trtm_fwi_desc* d, int32_t it_begin, int32_t it_end, void* stream) {
    int rc = check_adj(d, it_begin, it_end);
    if (rc || (rc = check_layout(d, true))) return rc;
    if (use_block(d, true)) {
        if ((rc = check_window(d, it_begin, it_end))) return rc;
        return fwi_block_adjoint(d, 1, it_begin, it_end, (hipStream_t)stream);
    }
    if (d->real_bytes == 4) return fwi_adjoint_t<float>(d, 1, it_begin, it_end, (hipStream_t)stream);
    return fwi_adjoint_t<double>(d, 1, it_begin, it_end, (hipStream_t)stream);
}

int trtm_fwi_forward_batch(const trtm_fwi_desc* descs, int32_t n, int32_t it_begin,
                           int32_t it_end, void* stream) {
    int rc = check_batch(descs, n, it_begin, it_end, false);
    if (rc || (rc = check_layout(descs, false))) return rc;
    if (use_block(descs, false)) return fwi_block_forward(descs, n, it_begin, it_end, (hipStream_t)stream);
    if (descs->real_bytes == 4)
        return fwi_forward_t<float>(descs, n, it_begin, it_end, (hipStream_t)stream);
    return fwi_forward_t<double>(descs, n, it_begin, it_end, (hipStream_t)stream);
}

int trtm_fwi_forward_scattered_batch(const trtm_fwi_desc* descs, int32_t n, int32_t it_begin,
                                     int32_t it_end, void* stream) {
    int rc = check_batch(descs, n, it_begin, it_end, false);
    if (rc) return rc;
    for (int q = 0; q < n; q++) {
        const trtm_fwi_desc* d = &descs[q];
        if (d->real_bytes != 4 || !d->workspace || !d->scat_workspace)
            return fail(TRTM_ERR_ARG, "shot %d: the scattered forward is float32 with both "
                                      "workspaces", q);
        if (d->hist) return fail(TRTM_ERR_ARG, "shot %d: the scattered forward keeps no history", q);
        if ((d->systems & TRTM_PSV) &&
            (!d->du || !d->dw || !d->dsxx || !d->dszz || !d->dsxz || !d->dif_lam || !d->dif_mu ||
             !d->dif_mxz || !d->dif_inv_rho_u || !d->dif_inv_rho_w))
            return fail(TRTM_ERR_ARG, "shot %d: P-SV scattered state or medium difference is NULL",
                        q);
        if ((d->systems & TRTM_SH) &&
            (!d->dv || !d->dsyx || !d->dsyz || !d->dif_myx || !d->dif_myz || !d->dif_inv_rho))
            return fail(TRTM_ERR_ARG, "shot %d: SH scattered state or medium difference is NULL", q);
    }
    return fwi_scat_forward(descs, n, it_begin, it_end, (hipStream_t)stream);
}

int trtm_fwi_adjoint_batch(const trtm_fwi_desc* descs, int32_t n, int32_t it_begin,
                           int32_t it_end, void* stream) {
    int rc = check_batch(descs, n, it_begin, it_end, true);
    if (rc || (rc = check_layout(descs, true))) return rc;
    if (use_block(descs, true)) {
        if ((rc = check_window(descs, it_begin, it_end))) return rc;
        return fwi_block_adjoint(descs, n, it_begin, it_end, (hipStream_t)stream);
    }
    if (descs->real_bytes == 4)
        return fwi_adjoint_t<float>(descs, n, it_begin, it_end, (hipStream_t)stream);
    return fwi_adjoint_t<double>(descs, n, it_begin, it_end, (hipStream_t)stream);
}

int64_t trtm_fwi_workspace_bytes(const trtm_fwi_desc* d) {
    if (!d || d->nx < 1 || d->nz < 1 || (d->real_bytes != 4 && d->real_bytes != 8))
        return fail(TRTM_ERR_ARG, "bad descriptor");
    return (int64_t)8 * d->nx * d->nz * d->real_bytes;
}

int trtm_fwi_block_configs(int32_t systems, int32_t pass, int32_t* out, int32_t max_n) {
    if (!out || max_n < 0 || pass < 0 || pass > 2) return fail(TRTM_ERR_ARG, "bad arguments");
    return fwi_block_configs(systems, pass, out, max_n);
}

int trtm_fwi_chain_gradient(int32_t nx, int32_t nz, const double* mu, const double* rho,
                            const double* vp, const double* vs, const double* g_lam,
                            const double* g_mu, const double* g_mxz, const double* g_myx,
                            const double* g_myz, const double* g_ir, const double* g_iru,
                            const double* g_irw, double* g_vp, double* g_vs, double* g_rho,
                            void* stream) {
    if (nx < 2 || nz < 2 || !mu || !rho || !vp || !vs || !g_lam || !g_mu || !g_mxz || !g_myx ||
        !g_myz || !g_ir || !g_iru || !g_irw || !g_vp || !g_vs || !g_rho)
        return fail(TRTM_ERR_ARG, "bad arguments");
    hipLaunchKernelGGL(k_chain, dim3((nz + kBlockZ - 1) / kBlockZ, (nx + kBlockX - 1) / kBlockX),
                       dim3(kBlockZ, kBlockX), 0, (hipStream_t)stream, nx, nz, mu, rho, vp, vs,
                       g_lam, g_mu, g_mxz, g_myx, g_myz, g_ir, g_iru, g_irw, g_vp, g_vs, g_rho);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TRTM_OK : fail(TRTM_ERR_HIP, "chain: %s", hipGetErrorString(e));
}

}  // extern "C"
