// Reference-shaped step: one launch per group of reference kernels.
//
//   k_surface   _set_surface_boundary            (forward_modeling.py:412-419)
//   k_velocity  _update_velocity[_backward]      (forward_modeling.py:309-347,
//                                                 backward_modeling.py:312-345)
//   k_inject    _add_source_kernel / _add_observed_data_kernel
//                                                (forward_modeling.py:421-443,
//                                                 backward_modeling.py:452-460)
//   k_stress    _update_stress[_backward] + sponge on the stresses
//                                                (forward_modeling.py:349-402,
//                                                 backward_modeling.py:347-398)
//   k_post      sponge on the velocities, snapshot, image, finite check
//                                                (forward_modeling.py:386-402,455-502;
//                                                 backward_modeling.py:382-398,417-450,481-488)
//   k_record    _record_seismogram_kernel / _record_synthetic_source_kernel
//
// The velocity sponge cannot join k_stress: the stress update reads the
// neighbours' velocities, which must still be the unsponged values.
// This path is the plain baseline the fused kernel is checked against.
#include "trtm_real.hpp"

namespace trtm {


template <class R, bool PSV, bool SH>
__global__ __launch_bounds__(256) void k_surface(KA<R> a) {
    const int j = blockIdx.x * kBlockZ + threadIdx.x;
    const int i = blockIdx.y * kBlockX + threadIdx.y;
    if (i >= a.nx || j >= a.nz) return;
    const size_t c = (size_t)i * a.nz + j;
    const R s = a.sm[c];
    if (SH) a.syz[c] = a.syz[c] * s;
    if (PSV) {
        a.sxz[c] = a.sxz[c] * s;
        a.szz[c] = a.szz[c] * s;
    }
}

template <class R, bool FWD, bool PSV, bool SH>
__global__ __launch_bounds__(256) void k_velocity(KA<R> a) {
    const int j = blockIdx.x * kBlockZ + threadIdx.x;
    const int i = blockIdx.y * kBlockX + threadIdx.y;
    if (i < 1 || i >= a.nx - 1 || j < 1 || j >= a.nz - 1) return;
    const size_t nz = (size_t)a.nz;
    const size_t c = (size_t)i * nz + j;
    if (PSV) {
        R sxx_x, szz_z, sxz_x, sxz_z;
        if (FWD) {
            sxx_x = (a.sxx[c] - a.sxx[c - nz]) * a.idx;
            szz_z = (a.szz[c] - a.szz[c - 1]) * a.idz;
            sxz_x = (a.sxz[c] - a.sxz[c - nz]) * a.idx;
            sxz_z = (a.sxz[c] - a.sxz[c - 1]) * a.idz;
            a.u[c] = a.u[c] + ((sxx_x + sxz_z) * a.dt) * a.iru[c];
            a.w[c] = a.w[c] + ((sxz_x + szz_z) * a.dt) * a.irw[c];
        } else {
            sxx_x = (a.sxx[c + nz] - a.sxx[c]) * a.idx;
            szz_z = (a.szz[c + 1] - a.szz[c]) * a.idz;
            sxz_x = (a.sxz[c + nz] - a.sxz[c]) * a.idx;
            sxz_z = (a.sxz[c + 1] - a.sxz[c]) * a.idz;
            a.u[c] = a.u[c] + ((-(sxx_x + sxz_z)) * a.dt) * a.iru[c];
            a.w[c] = a.w[c] + ((-(sxz_x + szz_z)) * a.dt) * a.irw[c];
        }
    }
    if (SH) {
        R syx_x, syz_z;
        if (FWD) {
            syx_x = (a.syx[c] - a.syx[c - nz]) * a.idx;
            syz_z = (a.syz[c] - a.syz[c - 1]) * a.idz;
            a.v[c] = a.v[c] + ((syx_x + syz_z) * a.dt) * a.ir[c];
        } else {
            syx_x = (a.syx[c + nz] - a.syx[c]) * a.idx;
            syz_z = (a.syz[c + 1] - a.syz[c]) * a.idz;
            a.v[c] = a.v[c] + ((-(syx_x + syz_z)) * a.dt) * a.ir[c];
        }
    }
}

// One thread per distinct cell; entries of a cell are added in entry order.
template <class R, bool FWD, bool PSV, bool SH>
__global__ __launch_bounds__(256) void k_inject(KA<R> a, IA<R> q) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= q.n_groups) return;
    const int e0 = q.start[g], e1 = q.start[g + 1];
    const int k0 = q.order[e0];
    const size_t c = (size_t)q.loc[2 * k0] * a.nz + q.loc[2 * k0 + 1];
    const size_t nt = (size_t)a.nt;
    if (PSV) {
        R x = a.u[c], z = a.w[c];
        for (int e = e0; e < e1; e++) {
            const size_t o = (size_t)q.order[e] * nt + q.col;
            if (FWD) {
                x = x + (q.su[o] * a.factor) * a.iru[c];
                z = z + (q.sw[o] * a.factor) * a.irw[c];
            } else {
                x = x + q.su[o];
                z = z + q.sw[o];
            }
        }
        a.u[c] = x;
        a.w[c] = z;
    }
    if (SH) {
        R y = a.v[c];
        for (int e = e0; e < e1; e++) {
            const size_t o = (size_t)q.order[e] * nt + q.col;
            y = FWD ? y + (q.sv[o] * a.factor) * a.ir[c] : y + q.sv[o];
        }
        a.v[c] = y;
    }
}

template <class R, bool FWD, bool PSV, bool SH>
__global__ __launch_bounds__(256) void k_stress(KA<R> a) {
    const int j = blockIdx.x * kBlockZ + threadIdx.x;
    const int i = blockIdx.y * kBlockX + threadIdx.y;
    if (i >= a.nx || j >= a.nz) return;
    const size_t nz = (size_t)a.nz;
    const size_t c = (size_t)i * nz + j;
    const R A = sponge_at_t<R>(i, j, a.nx, a.nz, a.fw, a.tab);
    const bool inner = i >= 1 && i < a.nx - 1 && j >= 1 && j < a.nz - 1;
    const R sdt = FWD ? a.dt : -a.dt;
    if (PSV) {
        R sxx = a.sxx[c], szz = a.szz[c], sxz = a.sxz[c];
        if (inner) {
            R u_x, u_z, w_x, w_z;
            if (FWD) {
                u_x = (a.u[c + nz] - a.u[c]) * a.idx;
                u_z = (a.u[c + 1] - a.u[c]) * a.idz;
                w_x = (a.w[c + nz] - a.w[c]) * a.idx;
                w_z = (a.w[c + 1] - a.w[c]) * a.idz;
            } else {
                u_x = (a.u[c] - a.u[c - nz]) * a.idx;
                u_z = (a.u[c] - a.u[c - 1]) * a.idz;
                w_x = (a.w[c] - a.w[c - nz]) * a.idx;
                w_z = (a.w[c] - a.w[c - 1]) * a.idz;
            }
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
            R v_x, v_z;
            if (FWD) {
                v_x = (a.v[c + nz] - a.v[c]) * a.idx;
                v_z = (a.v[c + 1] - a.v[c]) * a.idz;
            } else {
                v_x = (a.v[c] - a.v[c - nz]) * a.idx;
                v_z = (a.v[c] - a.v[c - 1]) * a.idz;
            }
            syx = syx + (sdt * a.myx[c]) * v_x;
            syz = syz + (sdt * a.myz[c]) * v_z;
        }
        a.syx[c] = syx * A;
        a.syz[c] = syz * A;
    }
}

template <class R, bool PSV, bool SH>
__global__ __launch_bounds__(256) void k_post(KA<R> a, SIO<R> p) {
    const int j = blockIdx.x * kBlockZ + threadIdx.x;
    const int i = blockIdx.y * kBlockX + threadIdx.y;
    if (i >= a.nx || j >= a.nz) return;
    const size_t n = (size_t)a.nx * a.nz;
    const size_t c = (size_t)i * a.nz + j;
    const R A = sponge_at_t<R>(i, j, a.nx, a.nz, a.fw, a.tab);
    int code = 0;
    if (PSV) {
        const R u = a.u[c] * A, w = a.w[c] * A;
        a.u[c] = u;
        a.w[c] = w;
        if (p.snap_slot >= 0) {
            p.snap_u[(size_t)p.snap_slot * n + c] = u;
            p.snap_w[(size_t)p.snap_slot * n + c] = w;
        }
        if (p.image_slot >= 0) {
            p.image_u[c] = p.image_u[c] + p.snap_u[(size_t)p.image_slot * n + c] * u;
            p.image_w[c] = p.image_w[c] + p.snap_w[(size_t)p.image_slot * n + c] * w;
        }
        if (p.check_slot >= 0) {
            if (bad_value<R>(u)) code = p.check_base + 1;
            if (bad_value<R>(w)) code = p.check_base + 3;
        }
    }
    if (SH) {
        const R v = a.v[c] * A;
        a.v[c] = v;
        if (p.snap_slot >= 0) p.snap_v[(size_t)p.snap_slot * n + c] = v;
        if (p.image_slot >= 0)
            p.image_v[c] = p.image_v[c] + p.snap_v[(size_t)p.image_slot * n + c] * v;
        if (p.check_slot >= 0 && bad_value<R>(v) && code < p.check_base + 2)
            code = p.check_base + 2;
    }
    if (code) atomicMax(p.check_flags + p.check_slot, code);
}

template <class R, bool PSV, bool SH>
__global__ __launch_bounds__(256) void k_record(KA<R> a, SIO<R> p) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= p.n_rec) return;
    const size_t c = (size_t)p.rec_loc[2 * l] * a.nz + p.rec_loc[2 * l + 1];
    const size_t o = (size_t)p.rec_row * p.n_rec + l;
    if (PSV) {
        p.rec_u[o] = a.u[c];
        p.rec_w[o] = a.w[c];
    }
    if (SH) p.rec_v[o] = a.v[c];
}

template <class R, bool FWD, bool PSV, bool SH>
static hipError_t simple_step(const KA<R>& a, const IA<R>& q, const SIO<R>& p, hipStream_t s) {
    const dim3 block(kBlockZ, kBlockX);
    const dim3 grid((a.nz + kBlockZ - 1) / kBlockZ, (a.nx + kBlockX - 1) / kBlockX);
    if (a.sm) hipLaunchKernelGGL((k_surface<R, PSV, SH>), grid, block, 0, s, a);
    hipLaunchKernelGGL((k_velocity<R, FWD, PSV, SH>), grid, block, 0, s, a);
    if (q.n_groups > 0)
        hipLaunchKernelGGL((k_inject<R, FWD, PSV, SH>), dim3((q.n_groups + 255) / 256), dim3(256),
                           0, s, a, q);
    hipLaunchKernelGGL((k_stress<R, FWD, PSV, SH>), grid, block, 0, s, a);
    hipLaunchKernelGGL((k_post<R, PSV, SH>), grid, block, 0, s, a, p);
    if (p.n_rec > 0 && p.rec_row >= 0)
        hipLaunchKernelGGL((k_record<R, PSV, SH>), dim3((p.n_rec + 255) / 256), dim3(256), 0, s,
                           a, p);
    return hipGetLastError();
}

template <class R>
hipError_t launch_simple_step_t(bool fwd, int systems, const KA<R>& a, const IA<R>& q,
                                const SIO<R>& p, hipStream_t s) {
    const bool psv = systems & TRTM_PSV, sh = systems & TRTM_SH;
    if (fwd) {
        if (psv && sh) return simple_step<R, true, true, true>(a, q, p, s);
        if (psv) return simple_step<R, true, true, false>(a, q, p, s);
        if (sh) return simple_step<R, true, false, true>(a, q, p, s);
    } else {
        if (psv && sh) return simple_step<R, false, true, true>(a, q, p, s);
        if (psv) return simple_step<R, false, true, false>(a, q, p, s);
        if (sh) return simple_step<R, false, false, true>(a, q, p, s);
    }
    return hipSuccess;
}


// float entry used by the product path: the KArgs / InjArgs / StepIO layouts
// are the float instances of KA / IA / SIO
hipError_t launch_simple_step(bool fwd, int systems, const KArgs& a, const InjArgs& q,
                              const StepIO& p, hipStream_t s) {
    KA<float> ka;
    ka.nx = a.nx; ka.nz = a.nz; ka.nt = a.nt; ka.fw = a.fw;
    ka.dt = a.dt; ka.idx = a.idx; ka.idz = a.idz; ka.factor = a.factor;
    ka.tab = a.tab; ka.lam = a.lam; ka.mu = a.mu; ka.mxz = a.mxz; ka.myx = a.myx; ka.myz = a.myz;
    ka.ir = a.ir; ka.iru = a.iru; ka.irw = a.irw; ka.sm = a.sm;
    ka.u = a.u; ka.v = a.v; ka.w = a.w; ka.sxx = a.sxx; ka.szz = a.szz; ka.sxz = a.sxz;
    ka.syx = a.syx; ka.syz = a.syz;
    IA<float> qa;
    qa.n_groups = q.n_groups; qa.col = q.col; qa.loc = q.loc; qa.order = q.order;
    qa.start = q.start; qa.su = q.su; qa.sv = q.sv; qa.sw = q.sw;
    SIO<float> pa;
    pa.snap_slot = p.snap_slot; pa.image_slot = p.image_slot; pa.check_slot = p.check_slot;
    pa.check_base = p.check_base; pa.rec_row = p.rec_row;
    pa.snap_u = p.snap_u; pa.snap_v = p.snap_v; pa.snap_w = p.snap_w;
    pa.image_u = p.image_u; pa.image_v = p.image_v; pa.image_w = p.image_w;
    pa.check_flags = p.check_flags; pa.n_rec = p.n_rec; pa.rec_loc = p.rec_loc;
    pa.rec_u = p.rec_u; pa.rec_v = p.rec_v; pa.rec_w = p.rec_w;
    return launch_simple_step_t<float>(fwd, systems, ka, qa, pa, s);
}

}  // namespace trtm
