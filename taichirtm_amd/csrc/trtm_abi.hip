// C ABI of the gfx950 elastic FD / RTM path: argument checks, set-up kernels
// and per-step dispatch.  Declarations and reference citations are in
// include/taichirtm_hip.h.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "trtm_common.hpp"

namespace trtm {

hipError_t launch_simple_step(bool fwd, int systems, const KArgs& a, const InjArgs& q,
                              const StepIO& p, hipStream_t s);
hipError_t launch_fused_step(bool fwd, int systems, const KArgs& a, const KArgs& out,
                             const InjArgs& q, const StepIO& p, const TilePlan& plan,
                             hipStream_t s);
int64_t fused_workspace_bytes(const trtm_step_desc* d);
int fused_prepare(const trtm_step_desc* d, bool fwd, hipStream_t s, TilePlan* plan);
struct ShotDev;
int64_t block_workspace_bytes(const trtm_step_desc* d);
int64_t block_shot_table_bytes(int n);
int block_config_query(const trtm_step_desc* d, int32_t* out);
int block_config_list(int systems, int kernel, int32_t* out, int max_n);
ShotDev* block_shot_slot(const trtm_step_desc* d);
int run_block(const trtm_step_desc* descs, int n, bool fwd, bool save, int32_t it_begin,
              int32_t it_end, ShotDev* shots_dev, const KArgs& m, hipStream_t s);
int block_capture(const trtm_step_desc* descs, int n, bool fwd, bool save, int32_t it_begin,
                  int32_t it_end, ShotDev* shots_dev, const KArgs& m, hipStream_t s,
                  hipGraph_t* graph, hipGraphExec_t* exec);

static thread_local std::string g_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

#define TRTM_HIP(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(TRTM_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));    \
    } while (0)

// --- medium preparation (forward_modeling.py:243-307) -------------------------

__global__ __launch_bounds__(256) void k_prepare(int nx, int nz, const float* __restrict__ mu,
                                                 const float* __restrict__ rho, float* myx,
                                                 float* myz, float* mxz, float* rho_u,
                                                 float* rho_w, float* inv_rho, float* inv_rho_u,
                                                 float* inv_rho_w) {
    const int j = blockIdx.x * kBlockZ + threadIdx.x;
    const int i = blockIdx.y * kBlockX + threadIdx.y;
    if (i >= nx || j >= nz) return;
    const size_t c = (size_t)i * nz + j;
    float hx = 0.0f, hz = 0.0f, hxz = 0.0f, ru = 0.0f, rw = 0.0f;
    if (i >= 1 && i < nx - 1 && j >= 1 && j < nz - 1) {
        const float a = mu[c], b = mu[c + nz], cc = mu[c + 1], d = mu[c + nz + 1];
        hx = 2.0f * a * b / (a + b);
        hz = 2.0f * a * cc / (a + cc);
        const float product = a * b * cc * d;
        const float sum_inv_prod = b * cc * d + a * cc * d + a * b * d + a * b * cc;
        hxz = 4.0f * product / sum_inv_prod;
        const float r = rho[c];
        ru = 0.5f * (r + rho[c + nz]);
        rw = 0.5f * (r + rho[c + 1]);
    }
    myx[c] = hx;
    myz[c] = hz;
    mxz[c] = hxz;
    rho_u[c] = ru;
    rho_w[c] = rw;
    inv_rho[c] = 1.0f / rho[c];
    inv_rho_u[c] = 1.0f / ru;
    inv_rho_w[c] = 1.0f / rw;
}

__global__ __launch_bounds__(256) void k_sponge_fill(int nx, int nz, int fw,
                                                     const float* __restrict__ tab, float* out) {
    const int j = blockIdx.x * kBlockZ + threadIdx.x;
    const int i = blockIdx.y * kBlockX + threadIdx.y;
    if (i >= nx || j >= nz) return;
    out[(size_t)i * nz + j] = sponge_at(i, j, nx, nz, fw, tab);
}

__global__ __launch_bounds__(256) void k_check_finite(int nx, int nz, const float* u,
                                                      const float* v, const float* w, int base,
                                                      int* flag) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)nx * nz) return;
    int code = 0;
    if (not_finite_or_big(u[c])) code = base + 1;
    if (not_finite_or_big(v[c])) code = base + 2;
    if (not_finite_or_big(w[c])) code = base + 3;
    if (code) atomicMax(flag, code);
}

// 64x64 tiles through LDS; +1 column avoids bank conflicts on the transposed read.
__global__ __launch_bounds__(256) void k_transpose(int rows, int cols, const float* __restrict__ src,
                                                   float* __restrict__ dst) {
    __shared__ float tile[64][65];
    const int c0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
    for (int k = threadIdx.y; k < 64; k += 4) {
        const int r = r0 + k, c = c0 + threadIdx.x;
        if (r < rows && c < cols) tile[k][threadIdx.x] = src[(size_t)r * cols + c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 64; k += 4) {
        const int c = c0 + k, r = r0 + threadIdx.x;
        if (r < rows && c < cols) dst[(size_t)c * rows + r] = tile[threadIdx.x][k];
    }
}

static dim3 grid2d(int nx, int nz) {
    return dim3((nz + kBlockZ - 1) / kBlockZ, (nx + kBlockX - 1) / kBlockX);
}

// --- descriptor checks --------------------------------------------------------

static int check_desc(const trtm_step_desc* d, int32_t it_begin, int32_t it_end) {
    if (!d) return fail(TRTM_ERR_ARG, "descriptor is NULL");
    if (d->struct_size != (int32_t)sizeof(trtm_step_desc))
        return fail(TRTM_ERR_ARG, "struct_size %d != %d (ABI mismatch)", d->struct_size,
                    (int)sizeof(trtm_step_desc));
    if (d->nx < 2 || d->nz < 2 || d->nt < 1)
        return fail(TRTM_ERR_ARG, "bad grid nx=%d nz=%d nt=%d", d->nx, d->nz, d->nt);
    if (d->fw < 0 || d->fw > d->nx || d->fw > d->nz)
        return fail(TRTM_ERR_ARG, "absorbing_frame %d outside [0, min(nx, nz)]", d->fw);
    if (it_begin < 0 || it_end > d->nt || it_begin > it_end)
        return fail(TRTM_ERR_ARG, "step range [%d, %d) outside [0, %d]", it_begin, it_end, d->nt);
    if (d->systems & ~(TRTM_PSV | TRTM_SH)) return fail(TRTM_ERR_ARG, "bad systems %d", d->systems);
    if (!d->sponge_table) return fail(TRTM_ERR_ARG, "sponge_table is NULL");
    if (d->systems & TRTM_PSV) {
        if (!d->u || !d->w || !d->sxx || !d->szz || !d->sxz || !d->lam || !d->mu || !d->mxz ||
            !d->inv_rho_u || !d->inv_rho_w)
            return fail(TRTM_ERR_ARG, "P-SV system active but a P-SV pointer is NULL");
    }
    if (d->systems & TRTM_SH) {
        if (!d->v || !d->syx || !d->syz || !d->myx || !d->myz || !d->inv_rho)
            return fail(TRTM_ERR_ARG, "SH system active but an SH pointer is NULL");
    }
    if (d->surface_matrix && (((d->systems & TRTM_PSV) && (!d->sxz || !d->szz)) ||
                              ((d->systems & TRTM_SH) && !d->syz)))
        return fail(TRTM_ERR_ARG, "surface_matrix needs the active systems' sxz, szz, syz");
    if (d->n_inj < 0 || d->n_inj_groups < 0 || d->n_inj_groups > d->n_inj)
        return fail(TRTM_ERR_ARG, "bad injection counts");
    if (d->n_inj_groups > 0) {
        if (!d->inj_loc || !d->inj_order || !d->inj_group_start)
            return fail(TRTM_ERR_ARG, "injection plan pointers are NULL");
        if ((d->systems & TRTM_PSV) && (!d->inj_u || !d->inj_w))
            return fail(TRTM_ERR_ARG, "P-SV injection data is NULL");
        if ((d->systems & TRTM_SH) && !d->inj_v) return fail(TRTM_ERR_ARG, "SH injection data is NULL");
    }
    if (d->n_rec < 0) return fail(TRTM_ERR_ARG, "n_rec < 0");
    if (d->n_rec > 0 && d->rec_u) {
        if (!d->rec_loc) return fail(TRTM_ERR_ARG, "rec_loc is NULL");
        if ((d->systems & TRTM_PSV) && !d->rec_w) return fail(TRTM_ERR_ARG, "rec_w is NULL");
        if ((d->systems & TRTM_SH) && !d->rec_v) return fail(TRTM_ERR_ARG, "rec_v is NULL");
    }
    if (d->check_interval < 0) return fail(TRTM_ERR_ARG, "check_interval < 0");
    if (d->check_interval > 0 && !d->check_flags) return fail(TRTM_ERR_ARG, "check_flags is NULL");
    if (d->kernel < TRTM_KERNEL_AUTO || d->kernel > TRTM_KERNEL_WAVE)
        return fail(TRTM_ERR_ARG, "bad kernel selector %d", d->kernel);
    if (d->snap_stride < 0 || (d->snap_stride > 0 && d->snap_stride < (int64_t)d->nx * d->nz) ||
        (d->snap_sponge != 0 && d->snap_sponge != 1))
        return fail(TRTM_ERR_ARG, "snap_stride must be 0 or >= nx*nz, snap_sponge 0 or 1");
    if ((d->snap_stride > 0 || d->snap_sponge) && d->kernel != TRTM_KERNEL_BLOCKED)
        return fail(TRTM_ERR_UNSUPPORTED,
                    "a snapshot stride or sponge needs the blocked kernel (TRTM_KERNEL_BLOCKED)");
    return TRTM_OK;
}

static KArgs make_kargs(const trtm_step_desc* d) {
    KArgs a;
    a.nx = d->nx; a.nz = d->nz; a.nt = d->nt; a.fw = d->fw;
    a.dt = d->dt; a.idx = d->inv_dx; a.idz = d->inv_dz; a.factor = d->src_factor;
    a.tab = d->sponge_table;
    a.lam = d->lam; a.mu = d->mu; a.mxz = d->mxz; a.myx = d->myx; a.myz = d->myz;
    a.ir = d->inv_rho; a.iru = d->inv_rho_u; a.irw = d->inv_rho_w;
    a.sm = d->surface_matrix;
    a.u = d->u; a.v = d->v; a.w = d->w;
    a.sxx = d->sxx; a.szz = d->szz; a.sxz = d->sxz; a.syx = d->syx; a.syz = d->syz;
    return a;
}

// kernel the library runs for d: AUTO takes the blocked kernel when the
// workspace allows it, else the fused one, else the simple one
static int pick_kernel(const trtm_step_desc* d) {
    if (d->kernel != TRTM_KERNEL_AUTO) return d->kernel;
    if (d->workspace && d->workspace_bytes >= block_workspace_bytes(d)) return TRTM_KERNEL_BLOCKED;
    if (d->workspace && d->workspace_bytes >= fused_workspace_bytes(d)) return TRTM_KERNEL_FUSED;
    return TRTM_KERNEL_SIMPLE;
}

static int check_io(const trtm_step_desc* d, bool fwd, bool save) {
    if (save || !fwd) {
        if (d->isnap < 1 && (save || d->num_snaps > 0))
            return fail(TRTM_ERR_ARG, "isnap must be >= 1");
        if (d->num_snaps > 0) {
            if ((d->systems & TRTM_PSV) && (!d->snap_u || !d->snap_w))
                return fail(TRTM_ERR_ARG, "snapshot buffers are NULL");
            if ((d->systems & TRTM_SH) && !d->snap_v) return fail(TRTM_ERR_ARG, "snap_v is NULL");
            if (!fwd) {
                if (!d->isnaps_host) return fail(TRTM_ERR_ARG, "isnaps_host is NULL");
                if ((d->systems & TRTM_PSV) && (!d->image_u || !d->image_w))
                    return fail(TRTM_ERR_ARG, "image buffers are NULL");
                if ((d->systems & TRTM_SH) && !d->image_v) return fail(TRTM_ERR_ARG, "image_v is NULL");
            }
        }
    }
    return TRTM_OK;
}

static int run_steps(const trtm_step_desc* d, int32_t it_begin, int32_t it_end, bool fwd,
                     bool save, hipStream_t s) {
    int rc = check_desc(d, it_begin, it_end);
    if (rc) return rc;
    rc = check_io(d, fwd, save);
    if (rc) return rc;
    if (d->systems == 0 || it_begin == it_end) return TRTM_OK;  // nothing can change

    const int kernel = pick_kernel(d);
    if (d->step_offset != 0)
        return fail(TRTM_ERR_ARG, "step_offset is for batch calls; use 0 here");
    if (kernel == TRTM_KERNEL_BLOCKED || kernel == TRTM_KERNEL_WAVE) {
        if (!d->workspace || d->workspace_bytes < block_workspace_bytes(d))
            return fail(TRTM_ERR_ARG, "blocked kernel needs %lld workspace bytes",
                        (long long)block_workspace_bytes(d));
        return run_block(d, 1, fwd, save, it_begin, it_end, block_shot_slot(d), make_kargs(d), s);
    }
    const bool fused = kernel == TRTM_KERNEL_FUSED;
    if (fused && (!d->workspace || d->workspace_bytes < fused_workspace_bytes(d)))
        return fail(TRTM_ERR_ARG, "fused kernel needs %lld workspace bytes",
                    (long long)fused_workspace_bytes(d));
    TilePlan plan{};
    if (fused) {
        rc = fused_prepare(d, fwd, s, &plan);
        if (rc) return rc;
    }

    KArgs a = make_kargs(d);
    InjArgs q;
    q.n_groups = d->n_inj_groups;
    q.loc = d->inj_loc; q.order = d->inj_order; q.start = d->inj_group_start;
    q.su = d->inj_u; q.sv = d->inj_v; q.sw = d->inj_w;
    StepIO p;
    p.snap_u = d->snap_u; p.snap_v = d->snap_v; p.snap_w = d->snap_w;
    p.image_u = d->image_u; p.image_v = d->image_v; p.image_w = d->image_w;
    p.check_flags = d->check_flags;
    p.check_base = fwd ? 0 : 3;
    p.n_rec = d->rec_u || d->rec_v ? d->n_rec : 0;
    p.rec_loc = d->rec_loc;
    p.rec_u = d->rec_u; p.rec_v = d->rec_v; p.rec_w = d->rec_w;

    // fused path ping-pongs between the caller's state and a workspace copy
    KArgs alt = a;
    if (fused) {
        float* base = static_cast<float*>(d->workspace);
        const size_t n = (size_t)d->nx * d->nz;
        float** dst[8] = {&alt.u, &alt.v, &alt.w, &alt.sxx, &alt.szz, &alt.sxz, &alt.syx, &alt.syz};
        for (int k = 0; k < 8; k++) *dst[k] = base + k * n;
    }
    KArgs cur = a, nxt = alt;
    for (int it = it_begin; it < it_end; it++) {
        const int t = d->nt - 1 - it;
        q.col = fwd ? it : t;
        p.rec_row = fwd ? it : t;
        p.snap_slot = -1;
        p.image_slot = -1;
        if (fwd && save && it % d->isnap == 0 && it != 0) {
            const int slot = it / d->isnap - 1;
            if (slot < d->num_snaps) p.snap_slot = slot;
        }
        if (!fwd && d->num_snaps > 0 && t > 0 && t % d->isnap == 0) {
            const int slot = t / d->isnap - 1;
            if (slot >= 0 && slot < d->num_snaps && d->isnaps_host[slot] == t) p.image_slot = slot;
        }
        p.check_slot = (d->check_interval > 0 && it % d->check_interval == 0)
                           ? it / d->check_interval : -1;
        hipError_t e;
        if (fused) {
            e = launch_fused_step(fwd, d->systems, cur, nxt, q, p, plan, s);
            std::swap(cur, nxt);
        } else {
            e = launch_simple_step(fwd, d->systems, a, q, p, s);
        }
        if (e != hipSuccess) return fail(TRTM_ERR_HIP, "step %d launch: %s", it, hipGetErrorString(e));
    }
    if (fused && cur.u != a.u) {  // odd number of steps: latest state is in the workspace
        const size_t bytes = (size_t)d->nx * d->nz * sizeof(float);
        float* from[8] = {cur.u, cur.v, cur.w, cur.sxx, cur.szz, cur.sxz, cur.syx, cur.syz};
        float* to[8] = {a.u, a.v, a.w, a.sxx, a.szz, a.sxz, a.syx, a.syz};
        const bool active[8] = {bool(d->systems & TRTM_PSV), bool(d->systems & TRTM_SH),
                                bool(d->systems & TRTM_PSV), bool(d->systems & TRTM_PSV),
                                bool(d->systems & TRTM_PSV), bool(d->systems & TRTM_PSV),
                                bool(d->systems & TRTM_SH),  bool(d->systems & TRTM_SH)};
        for (int k = 0; k < 8; k++)
            if (active[k]) TRTM_HIP(hipMemcpyAsync(to[k], from[k], bytes, hipMemcpyDeviceToDevice, s));
    }
    return TRTM_OK;
}

}  // namespace trtm

using namespace trtm;

extern "C" {

int trtm_abi_version(void) { return TRTM_ABI_VERSION; }

const char* trtm_last_error(void) { return g_error.c_str(); }

int trtm_prepare_medium(int32_t nx, int32_t nz, const float* mu, const float* rho, float* myx,
                        float* myz, float* mxz, float* rho_u, float* rho_w, float* inv_rho,
                        float* inv_rho_u, float* inv_rho_w, void* stream) {
    if (nx < 1 || nz < 1) return fail(TRTM_ERR_ARG, "bad grid %d x %d", nx, nz);
    if (!mu || !rho || !myx || !myz || !mxz || !rho_u || !rho_w || !inv_rho || !inv_rho_u ||
        !inv_rho_w)
        return fail(TRTM_ERR_ARG, "NULL pointer");
    hipLaunchKernelGGL(k_prepare, grid2d(nx, nz), dim3(kBlockZ, kBlockX), 0,
                       (hipStream_t)stream, nx, nz, mu, rho, myx, myz, mxz, rho_u, rho_w, inv_rho,
                       inv_rho_u, inv_rho_w);
    TRTM_HIP(hipGetLastError());
    return TRTM_OK;
}

int trtm_sponge_table(int32_t fw, float a, float* table_host) {
    if (fw < 0 || !table_host) return fail(TRTM_ERR_ARG, "bad sponge table arguments");
    const float a_sq = a * a;  // forward_modeling.py:209, float32
    for (int k = 0; k <= fw; k++) table_host[k] = expf((-a_sq * (float)k) * (float)k);
    return TRTM_OK;
}

int trtm_sponge_fill(int32_t nx, int32_t nz, int32_t fw, const float* table, float* coeff,
                     void* stream) {
    if (nx < 1 || nz < 1 || fw < 0 || fw > nx || fw > nz || !table || !coeff)
        return fail(TRTM_ERR_ARG, "bad sponge fill arguments (fw must be <= min(nx, nz))");
    hipLaunchKernelGGL(k_sponge_fill, grid2d(nx, nz), dim3(kBlockZ, kBlockX), 0,
                       (hipStream_t)stream, nx, nz, fw, table, coeff);
    TRTM_HIP(hipGetLastError());
    return TRTM_OK;
}

int trtm_plan_injection(int32_t nx, int32_t nz, int32_t n, const int32_t* loc_host,
                        int32_t* order_host, int32_t* group_start_host, int32_t* n_groups) {
    if (n < 0 || (n > 0 && (!loc_host || !order_host)) || !group_start_host || !n_groups)
        return fail(TRTM_ERR_ARG, "bad injection plan arguments");
    for (int k = 0; k < n; k++) {
        const int i = loc_host[2 * k], j = loc_host[2 * k + 1];
        if (i < 0 || i >= nx || j < 0 || j >= nz)
            return fail(TRTM_ERR_ARG, "location %d = (%d, %d) outside the %d x %d grid", k, i, j,
                        nx, nz);
    }
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    auto key = [&](int k) { return (int64_t)loc_host[2 * k] * nz + loc_host[2 * k + 1]; };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key(x) < key(y); });
    int g = 0;
    for (int e = 0; e < n; e++) {
        if (e == 0 || key(order[e]) != key(order[e - 1])) group_start_host[g++] = e;
        order_host[e] = order[e];
    }
    group_start_host[g] = n;
    *n_groups = g;
    return TRTM_OK;
}

int64_t trtm_workspace_bytes(const trtm_step_desc* d) {
    if (!d || d->struct_size != (int32_t)sizeof(trtm_step_desc)) return -1;
    if (d->kernel == TRTM_KERNEL_SIMPLE) return 0;
    if (d->kernel == TRTM_KERNEL_FUSED) return fused_workspace_bytes(d);
    return std::max(block_workspace_bytes(d), fused_workspace_bytes(d));
}

int trtm_block_configs(int32_t systems, int32_t kernel, int32_t* out, int32_t max_n) {
    if (systems <= 0 || (systems & ~(TRTM_PSV | TRTM_SH)) || !out || max_n < 0 ||
        (kernel != TRTM_KERNEL_BLOCKED && kernel != TRTM_KERNEL_WAVE))
        return fail(TRTM_ERR_ARG, "bad arguments");
    return block_config_list(systems, kernel, out, max_n);
}

int trtm_block_config(const trtm_step_desc* d, int32_t* out) {
    if (!d || !out || d->struct_size != (int32_t)sizeof(trtm_step_desc))
        return fail(TRTM_ERR_ARG, "bad arguments");
    return block_config_query(d, out);
}

int64_t trtm_batch_workspace_bytes(int32_t n) {
    return n < 1 ? -1 : block_shot_table_bytes(n);
}

// arguments of a batch call (every shot on shot 0's medium and grid)
static int check_batch_call(const trtm_step_desc* descs, int32_t n, void* bws, int64_t bws_bytes,
                            int32_t it_begin, int32_t it_end, bool fwd, bool save) {
    if (!descs || n < 1) return fail(TRTM_ERR_ARG, "empty batch");
    if (!bws || bws_bytes < block_shot_table_bytes(n))
        return fail(TRTM_ERR_ARG, "batch workspace needs %lld bytes",
                    (long long)block_shot_table_bytes(n));
    const trtm_step_desc* d0 = &descs[0];
    for (int q = 0; q < n; q++) {
        const trtm_step_desc* d = &descs[q];
        const int lo = std::min(std::max(0, it_begin + d->step_offset), d->nt);
        const int hi = std::max(lo, std::min(it_end + d->step_offset, d->nt));
        int rc = check_desc(d, lo, hi);
        if (rc) return fail(rc, "shot %d: %s", q, g_error.c_str());
        rc = check_io(d, fwd, save);
        if (rc) return fail(rc, "shot %d: %s", q, g_error.c_str());
        if (d->nx != d0->nx || d->nz != d0->nz || d->fw != d0->fw || d->systems != d0->systems ||
            d->dt != d0->dt || d->inv_dx != d0->inv_dx || d->inv_dz != d0->inv_dz ||
            d->src_factor != d0->src_factor || d->sponge_table != d0->sponge_table ||
            d->lam != d0->lam || d->mu != d0->mu || d->mxz != d0->mxz || d->myx != d0->myx ||
            d->myz != d0->myz || d->inv_rho != d0->inv_rho || d->inv_rho_u != d0->inv_rho_u ||
            d->inv_rho_w != d0->inv_rho_w || d->surface_matrix != d0->surface_matrix ||
            d->tile_x != d0->tile_x || d->block_steps != d0->block_steps ||
            d->tile_z != d0->tile_z)
            return fail(TRTM_ERR_ARG, "shot %d does not share the medium and grid of shot 0", q);
        if (!d->workspace || d->workspace_bytes < block_workspace_bytes(d))
            return fail(TRTM_ERR_ARG, "shot %d workspace needs %lld bytes", q,
                        (long long)block_workspace_bytes(d));
    }
    return TRTM_OK;
}

static int run_batch(const trtm_step_desc* descs, int32_t n, void* bws, int64_t bws_bytes,
                     int32_t it_begin, int32_t it_end, bool fwd, bool save, hipStream_t s) {
    const int rc = check_batch_call(descs, n, bws, bws_bytes, it_begin, it_end, fwd, save);
    if (rc) return rc;
    const trtm_step_desc* d0 = &descs[0];
    if (d0->systems == 0 || it_begin >= it_end) return TRTM_OK;
    return run_block(descs, n, fwd, save, it_begin, it_end, static_cast<ShotDev*>(bws),
                     make_kargs(d0), s);
}

int trtm_forward_steps_batch(const trtm_step_desc* descs, int32_t n, void* batch_workspace,
                             int64_t batch_workspace_bytes, int32_t it_begin, int32_t it_end,
                             int32_t save, void* stream) {
    return run_batch(descs, n, batch_workspace, batch_workspace_bytes, it_begin, it_end, true,
                     save != 0, (hipStream_t)stream);
}

int trtm_backward_steps_batch(const trtm_step_desc* descs, int32_t n, void* batch_workspace,
                              int64_t batch_workspace_bytes, int32_t it_begin, int32_t it_end,
                              void* stream) {
    return run_batch(descs, n, batch_workspace, batch_workspace_bytes, it_begin, it_end, false,
                     false, (hipStream_t)stream);
}

struct trtm_plan {
    hipGraph_t graph;
    hipGraphExec_t exec;  // null: nothing to run
};

int trtm_plan_create(const trtm_step_desc* descs, int32_t n, void* batch_workspace,
                     int64_t batch_workspace_bytes, int32_t it_begin, int32_t it_end,
                     int32_t forward, int32_t save, void* stream, trtm_plan** plan) {
    if (!plan) return fail(TRTM_ERR_ARG, "plan is NULL");
    *plan = nullptr;
    const bool fwd = forward != 0, sv = fwd && save != 0;
    if (descs && n >= 1 && descs[0].kernel != TRTM_KERNEL_BLOCKED &&
        descs[0].kernel != TRTM_KERNEL_WAVE && descs[0].kernel != TRTM_KERNEL_AUTO)
        return fail(TRTM_ERR_UNSUPPORTED, "plans capture the blocked and wave kernels only");
    int rc = check_batch_call(descs, n, batch_workspace, batch_workspace_bytes, it_begin, it_end,
                              fwd, sv);
    if (rc) return rc;
    trtm_plan* p = new trtm_plan{nullptr, nullptr};
    if (descs[0].systems != 0 && it_begin < it_end) {
        rc = block_capture(descs, n, fwd, sv, it_begin, it_end,
                           static_cast<ShotDev*>(batch_workspace), make_kargs(&descs[0]),
                           (hipStream_t)stream, &p->graph, &p->exec);
        if (rc) {
            delete p;
            return rc;
        }
    }
    *plan = p;
    return TRTM_OK;
}

int trtm_plan_launch(trtm_plan* plan, void* stream) {
    if (!plan) return fail(TRTM_ERR_ARG, "plan is NULL");
    if (plan->exec) TRTM_HIP(hipGraphLaunch(plan->exec, (hipStream_t)stream));
    return TRTM_OK;
}

int trtm_plan_destroy(trtm_plan* plan) {
    if (!plan) return TRTM_OK;
    if (plan->exec) hipGraphExecDestroy(plan->exec);
    if (plan->graph) hipGraphDestroy(plan->graph);
    delete plan;
    return TRTM_OK;
}

int trtm_forward_steps(const trtm_step_desc* d, int32_t it_begin, int32_t it_end, int32_t save,
                       void* stream) {
    return run_steps(d, it_begin, it_end, true, save != 0, (hipStream_t)stream);
}

int trtm_backward_steps(const trtm_step_desc* d, int32_t it_begin, int32_t it_end, void* stream) {
    return run_steps(d, it_begin, it_end, false, false, (hipStream_t)stream);
}

}  // extern "C"

// --- allocation-free forward steps ----------------------------------------------
//
// The workspace of trtm_forward_steps (second state buffer, per-tile lists,
// shot table) comes from this file-scope device array instead of the caller.
// Each device holds its own copy until the process ends.  Host code never uses
// the array's name as a device address: the address is looked up once per
// device with hipGetSymbolAddress and kept in g_scratch_dev.
static __device__ __align__(256) unsigned char scratch[TRTM_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_dev[kMaxDevices] = {};

// device address of `scratch` on the current device (looked up on the first
// call per device)
int scratch_address(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(TRTM_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(TRTM_ERR_UNSUPPORTED, "device %d beyond the %d this library tracks", dev,
                    kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(TRTM_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                        hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return TRTM_OK;
}
}  // namespace

extern "C" {

int trtm_forward_steps_noalloc(const trtm_step_desc* d, int32_t it_begin, int32_t it_end,
                               int32_t save, void* stream) {
    if (!d || d->struct_size != (int32_t)sizeof(trtm_step_desc))
        return fail(TRTM_ERR_ARG, "bad descriptor");
    if (d->workspace || d->workspace_bytes)
        return fail(TRTM_ERR_ARG, "the workspace comes from the library's static scratch; "
                                  "pass workspace = NULL, workspace_bytes = 0");
    const int64_t scratch_bytes = TRTM_NOALLOC_SCRATCH_BYTES;
    trtm_step_desc w = *d;
    w.workspace = nullptr;
    w.workspace_bytes = 0;
    if (w.kernel != TRTM_KERNEL_SIMPLE) {
        // an explicitly chosen kernel whose workspace exceeds the scratch is
        // rejected before anything is queued; TRTM_KERNEL_AUTO falls back to
        // the fused kernel, then the simple one, exactly as with a caller's
        // workspace of TRTM_NOALLOC_SCRATCH_BYTES
        const int64_t need = (w.kernel == TRTM_KERNEL_BLOCKED || w.kernel == TRTM_KERNEL_WAVE)
                                 ? block_workspace_bytes(&w)
                             : w.kernel == TRTM_KERNEL_FUSED ? fused_workspace_bytes(&w)
                                                             : 0;
        if (need > scratch_bytes)
            return fail(TRTM_ERR_ARG,
                        "this grid needs %lld workspace bytes; the static scratch holds %lld",
                        (long long)need, (long long)scratch_bytes);
        void* p = nullptr;
        const int rc = scratch_address(&p);
        if (rc) return rc;
        w.workspace = p;
        w.workspace_bytes = scratch_bytes;
    }
    return run_steps(&w, it_begin, it_end, true, save != 0, (hipStream_t)stream);
}

int trtm_check_finite(int32_t nx, int32_t nz, const float* u, const float* v, const float* w,
                      int32_t base, int32_t* flag, void* stream) {
    if (nx < 1 || nz < 1 || !u || !v || !w || !flag) return fail(TRTM_ERR_ARG, "bad arguments");
    const size_t n = (size_t)nx * nz;
    hipLaunchKernelGGL(k_check_finite, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, nx, nz, u, v, w, base, flag);
    TRTM_HIP(hipGetLastError());
    return TRTM_OK;
}

int trtm_transpose(int32_t rows, int32_t cols, const float* src, float* dst, void* stream) {
    if (rows < 0 || cols < 0 || (rows * (int64_t)cols > 0 && (!src || !dst)))
        return fail(TRTM_ERR_ARG, "bad transpose arguments");
    if (rows == 0 || cols == 0) return TRTM_OK;
    hipLaunchKernelGGL(k_transpose, dim3((cols + 63) / 64, (rows + 63) / 64), dim3(64, 4), 0,
                       (hipStream_t)stream, rows, cols, src, dst);
    TRTM_HIP(hipGetLastError());
    return TRTM_OK;
}

}  // extern "C"
